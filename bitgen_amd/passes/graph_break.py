"""Fusion policy (``bitgen/passes/graph_break.py:368-407``).

* -1 ("fuse all", the reference's default for its best results): the whole
  group program is one fused subgraph.  The reference then guards its final
  store with ``if (idx < n_unit_basic)`` (:398-407); the hip kernels bound
  their stores themselves, so the instructions are left unchanged.
* 0, 1, 2 (the reference's ablation baselines): graph breaks split the
  top-level program into subgraphs that hand streams to each other through
  global memory.  Placement as in the reference (:373-396):
    2 -- around every while loop, ScanThru (and MatchStar, the same
         carry-propagating form) and Advance: only bitwise operations fuse;
    1 -- before every instruction but stream loads and stores: nothing fuses;
    0 -- around while loops and ScanThru / MatchStar: shifts fuse with the
         bitwise operations around them.
  The hip backend runs such a program at stream level
  (``codegen/gen_stream.py``): one tile-parallel kernel per subgraph over
  streams in a global workspace.
"""
from ..inst import BsGraphBreak, BsInstType

_CARRY = (BsInstType.SCANTHRU, BsInstType.MATCHSTAR)
_AROUND = {
    2: (BsInstType.WHILE, BsInstType.ADVANCE) + _CARRY,
    0: (BsInstType.WHILE,) + _CARRY,
}
_UNSPLIT = (BsInstType.STREAMSTORE, BsInstType.STREAMLOAD, BsInstType.SKIPZERO, BsInstType.LABEL,
            BsInstType.GRAPHBREAK)


class graph_break_pass:
    def __init__(self, graph_break_type=-1):
        if graph_break_type not in (-1, 0, 1, 2):
            raise ValueError(f"unknown graph break type {graph_break_type}")
        self.pass_name = "graph_break"
        self.graph_break_type = graph_break_type

    def run(self, insts, var_name_map):
        mode = self.graph_break_type
        if mode == -1:
            return insts, var_name_map
        out = []
        for inst in insts:
            if mode == 1:
                if inst.type not in _UNSPLIT:
                    out.append(BsGraphBreak())
                out.append(inst)
            elif inst.type in _AROUND[mode]:
                out += [BsGraphBreak(), inst, BsGraphBreak()]
            else:
                out.append(inst)
        return out, var_name_map
