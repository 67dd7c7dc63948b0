"""Rewrite fixed-point star loops into ``MatchStar``.

icgrep's ``-disable-matchstar`` output (``bitgen/tool/pcre2pablo.py:41``)
writes ``C*`` as a while loop; on the GPU a whole-stream fixed point cannot be
evaluated tile by tile, but the loop computes exactly MatchStar(m, C)
(Cameron et al., PACT 2014), which the HIP backend evaluates with carry
look-back.  The recognised shape is::

    test = m
    accum = m
    while test:
        t = (test & C)            # C defined outside the loop
        adv = pablo.Advance(t, 1)
        nacc = (~accum)
        new = (adv & nacc)
        accum = (accum | adv)
        test = new

(operands of & and | in either order, statements in any dependency-respecting
order).  The fixed point leaves accum = every position reachable from m through
a run of C, i.e. (((m & C) + C) ^ C) | m, and test = 0.  The loop is replaced
by ``accum = pablo.MatchStar(m, C)`` (plus ``test = 0`` when test is read
afterwards).  Loops of any other shape are left alone.
"""
from ..inst import BsInstType, BsAssign, BsMatchStar
from .pass_utils import block_defs, const_id


class star_rewrite_pass:
    def __init__(self):
        self.pass_name = "star_rewrite"
        self.rewritten = 0

    @staticmethod
    def _match_loop(loop):
        body = loop.body
        if len(body) != 6:
            return None
        test = loop.condition
        by_type = {}
        for inst in body:
            by_type.setdefault(inst.type, []).append(inst)
        ands = by_type.get(BsInstType.AND, [])
        advs = by_type.get(BsInstType.ADVANCE, [])
        nots = by_type.get(BsInstType.NOT, [])
        ors = by_type.get(BsInstType.OR, [])
        asg = by_type.get(BsInstType.ASSIGN, [])
        if not (len(ands) == 2 and len(advs) == 1 and len(nots) == 1 and len(ors) == 1 and len(asg) == 1):
            return None
        adv, nacc_i, or_i, test_i = advs[0], nots[0], ors[0], asg[0]
        if adv.is_left or adv.amount != 1 or test_i.ret != test:
            return None
        accum = nacc_i.operands[0]
        if or_i.ret != accum or set(or_i.operands) != {accum, adv.ret}:
            return None
        t_i = next((a for a in ands if a.ret == adv.operands[0]), None)
        new_i = next((a for a in ands if a is not t_i), None)
        if t_i is None or new_i is None or test not in t_i.operands:
            return None
        cls = t_i.operands[1] if t_i.operands[0] == test else t_i.operands[0]
        if set(new_i.operands) != {adv.ret, nacc_i.ret} or test_i.operands[0] != new_i.ret:
            return None
        if cls in (test, accum) or cls in block_defs(body):
            return None
        # dependency order inside the body
        pos = {id(i): k for k, i in enumerate(body)}
        if not (pos[id(t_i)] < pos[id(adv)] < pos[id(new_i)] < pos[id(test_i)]
                and pos[id(adv)] < pos[id(or_i)] and pos[id(nacc_i)] < pos[id(new_i)]
                and pos[id(nacc_i)] < pos[id(or_i)]):
            return None
        temps = {t_i.ret, adv.ret, nacc_i.ret, new_i.ret}
        if len(temps) != 4 or temps & {test, accum, cls}:
            return None
        return test, accum, cls, temps

    def _rewrite_list(self, insts, var_name_map, uses_outside):
        out = []
        i = 0
        while i < len(insts):
            inst = insts[i]
            if inst.type == BsInstType.WHILE:
                inst.body = self._rewrite_list(inst.body, var_name_map, uses_outside)
                m = self._match_loop(inst)
                if m is not None and len(out) >= 2:
                    test, accum, cls, temps = m
                    a, b = out[-2], out[-1]
                    pre = {x.ret: x for x in (a, b) if x.type == BsInstType.ASSIGN}
                    outside = uses_outside(inst)
                    if (set(pre) == {test, accum} and pre[test].operands[0] == pre[accum].operands[0]
                            and pre[test].operands[0] not in (test, accum)
                            and not (temps & outside)):
                        src = pre[test].operands[0]
                        out = out[:-2]
                        out.append(BsMatchStar([src, cls], accum))
                        if test in outside:
                            out.append(BsAssign([const_id(var_name_map, "0")], test))
                        self.rewritten += 1
                        i += 1
                        continue
            elif inst.type == BsInstType.IF:
                inst.body_1 = self._rewrite_list(inst.body_1, var_name_map, uses_outside)
                if inst.body_2:
                    inst.body_2 = self._rewrite_list(inst.body_2, var_name_map, uses_outside)
            out.append(inst)
            i += 1
        return out

    def run(self, insts, var_name_map):
        from collections import Counter
        from ..inst import walk

        total = Counter()
        for inst in walk(insts):
            total.update(inst.uses())

        def uses_outside(loop):
            inner = Counter()
            for inst in walk([loop]):
                inner.update(inst.uses())
            return {v for v, n in total.items() if n > inner.get(v, 0)}

        return self._rewrite_list(list(insts), var_name_map, uses_outside), var_name_map
