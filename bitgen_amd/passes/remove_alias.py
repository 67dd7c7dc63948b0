"""Copy propagation (``bitgen/passes/remove_alias.py:10-52``).

``x = y`` is deleted and every read of x becomes a read of y when both x and
y are written exactly once in the whole program (y may also be a basis stream
or a constant), y is written before the copy and x is not read before it, so
x and y hold the same value wherever x can be read.
``output.matches`` and ``basis*`` are kept, as in the reference; loop-carried
``test*``/``accum*`` variables are written more than once and so are kept by
the same rule.
"""
from ..inst import BsInstType, walk
from .pass_utils import def_counts, is_const

_KEEP_PREFIX = ("basis",)
_KEEP = ("output.matches", "bs_result_stream")


class remove_alias_pass:
    def __init__(self):
        self.pass_name = "remove_alias"
        self.removed = 0

    def run(self, insts, var_name_map):
        changed = True
        while changed:
            changed = False
            defs = def_counts(insts)
            def_at, first_use = {}, {}
            for idx, inst in enumerate(walk(insts)):
                for u in inst.uses():
                    first_use.setdefault(u, idx)
                for d in inst.defs():
                    def_at.setdefault(d, idx)
            mapping = {}
            dead = set()
            for idx, inst in enumerate(walk(insts)):
                if inst.type != BsInstType.ASSIGN:
                    continue
                x, y = inst.ret, inst.operands[0]
                name = var_name_map.get(x, "")
                if name in _KEEP or name.startswith(_KEEP_PREFIX):
                    continue
                if defs.get(x, 0) != 1:
                    continue
                if not (is_const(var_name_map, y) or defs.get(y, 0) <= 1):
                    continue
                if x == y or x in mapping.values():
                    continue
                # program order: y written before the copy, x never read
                # before it (no loop-carried reads of either)
                if not is_const(var_name_map, y) and def_at.get(y, -1) >= idx:
                    continue
                if first_use.get(x, idx + 1) < idx:
                    continue
                mapping[x] = y
                dead.add(id(inst))
            if not mapping:
                break

            def resolve(v):
                seen = set()
                while v in mapping and v not in seen:
                    seen.add(v)
                    v = mapping[v]
                return v

            full = {k: resolve(k) for k in mapping}
            insts = self._rewrite(insts, full, dead)
            self.removed += len(dead)
            changed = True
        return insts, var_name_map

    def _rewrite(self, insts, mapping, dead):
        out = []
        for inst in insts:
            if id(inst) in dead:
                continue
            inst.replace_uses(mapping)
            if inst.type == BsInstType.WHILE:
                inst.body = self._rewrite(inst.body, mapping, dead)
            elif inst.type == BsInstType.IF:
                inst.body_1 = self._rewrite(inst.body_1, mapping, dead)
                if inst.body_2 is not None:
                    inst.body_2 = self._rewrite(inst.body_2, mapping, dead)
            out.append(inst)
        return out
