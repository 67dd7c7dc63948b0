"""Fusion policy (``bitgen/passes/graph_break.py:368-580``).

Type -1 ("fuse all", the reference's default for its best results) keeps the
whole group program as one fused subgraph; the reference then guards its final
store with ``if (idx < n_unit_basic)`` (:398-407).  The HIP backend always
executes one fused program per (group, tile) and bounds its stores in the
kernel, so type -1 leaves the instructions unchanged.  Types 0/1/2 split the
program into subgraphs that spill streams to global ``tmp_streams``; they exist
only for the reference's ablation and are not implemented by the HIP backend.
"""


class graph_break_pass:
    def __init__(self, graph_break_type=-1):
        self.pass_name = "graph_break"
        self.graph_break_type = graph_break_type

    def run(self, insts, var_name_map):
        if self.graph_break_type == -1:
            return insts, var_name_map
        raise NotImplementedError(
            f"--pass-graph-break={self.graph_break_type}: only -1 (fuse all) is supported by the hip backend"
        )
