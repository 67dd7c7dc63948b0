"""Shared helpers for the IR passes (def/use bookkeeping, fresh variables).

Counterpart of ``bitgen/passes/pass_utils.py`` (build_graph 52-102,
update_insts, break_insts 105-118) restated over this package's IR.
"""
from ..inst import BsInstType, walk


def get_var_by_name(var_name, var_name_map):
    for k, v in var_name_map.items():
        if v == var_name:
            return k
    return None


def add_new_var(var_name_map, name):
    """Add a variable whose name is unique in the map (reference naming style
    ``<base>_advance<k>`` with ``_<n>`` suffixes, cc_advance.py:106-119)."""
    existing = set(var_name_map.values())
    cand = name
    n = 0
    while cand in existing:
        n += 1
        cand = f"{name}_{n}"
    vid = max(var_name_map) + 1 if var_name_map else 0
    var_name_map[vid] = cand
    return vid


def const_id(var_name_map, value):
    """Id of the constant stream "0" or "1", created if missing."""
    vid = get_var_by_name(value, var_name_map)
    if vid is None:
        vid = add_new_var(var_name_map, value)
        var_name_map[vid] = value
    return vid


def is_const(var_name_map, v):
    return var_name_map.get(v) in ("0", "1")


def def_counts(insts):
    counts = {}
    for inst in walk(insts):
        for d in inst.defs():
            counts[d] = counts.get(d, 0) + 1
    return counts


def block_defs(insts):
    out = set()
    for inst in walk(insts):
        out.update(inst.defs())
    return out


def block_uses(insts):
    out = set()
    for inst in walk(insts):
        out.update(inst.uses())
    return out


def top_level_last_use(insts):
    """var -> index of the last top-level instruction that reads it (uses
    nested in a loop/branch count at the enclosing top-level index)."""
    last = {}
    for i, inst in enumerate(insts):
        uses = set(inst.uses())
        if inst.type == BsInstType.WHILE:
            uses |= block_uses(inst.body)
        elif inst.type == BsInstType.IF:
            uses |= block_uses(inst.body_1)
            uses |= block_uses(inst.body_2 or [])
        for u in uses:
            last[u] = i
    return last


def break_insts(insts):
    groups, cur = [], []
    for inst in insts:
        if inst.type == BsInstType.GRAPHBREAK:
            if cur:
                groups.append(cur)
                cur = []
        else:
            cur.append(inst)
    if cur:
        groups.append(cur)
    return groups
