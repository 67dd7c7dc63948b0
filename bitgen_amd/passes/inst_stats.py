"""Static instruction statistics (``bitgen/passes/inst_stats.py:10-42``).

Counts per-kind instructions of the final program into
``global_timer.inst_result`` (keys ``<kind>``: counts summed over groups).
"""
from ..inst import BsInstType, walk
from ..tool import global_timer


def inst_counts(insts):
    c = {}
    for inst in walk(insts):
        key = inst.type.name.lower()
        if inst.type == BsInstType.ADVANCE:
            key = "advance_left" if inst.is_left else "advance_right"
        c[key] = c.get(key, 0) + 1
    return c


class inst_stats_pass:
    def __init__(self):
        self.pass_name = "inst_stats"

    def run(self, insts, var_name_map):
        res = global_timer.inst_result
        for k, v in inst_counts(insts).items():
            res[k] = res.get(k, 0) + v
        return insts, var_name_map
