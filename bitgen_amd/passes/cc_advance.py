"""Shift rebalancing ("cc advance").

Follows the idea of ``bitgen/passes/cc_advance.py:59-378``: marker
``Advance`` instructions are deleted and every variable carries an *advance
count* k meaning "stored value = true value shifted back by k positions"
(true = Advance(stored, k)).  Where the operands of an operation have
different counts they are aligned, and the final result is re-aligned before
the store (the reference never does the last step, so its result stream is
offset by the residual count; here the stored result is exact).

Alignment rules -- each is exact on positions 0..n of an n-byte input:

* ``a & b``: the operand with the smaller count is shifted toward lower
  positions (look-ahead, ``BSAdvanceLeftFunction``) by the difference; the
  result takes the larger count.  The bits lost at the stream start are
  below the larger count, where the true AND is zero anyway.
* ``a | b``, ``a ^ b``: the operand with the larger count is advanced (right
  shift, ``BSAdvanceRightFunctionSync``) down to the smaller count, like the
  reference's "result OR" (cc_advance.py:259-283); the result takes the
  smaller count.
* ``~a`` with a nonzero count: a is re-aligned to count 0 first (the true
  complement has ones below k that a shifted value cannot hold).
* ``MatchStar(M, C)`` / ``ScanThru(M, C)``: C is look-ahead shifted up to M's
  count when its count is smaller (M is zero below its count, so no carry
  enters from there), otherwise advanced down to it.
* values read inside a ``while``/``if`` block are re-aligned to count 0 in
  place before the block and the block is left untouched (the reference
  disables the pass inside loop bodies as well, cc_advance.py:73-176).

Constants "0"/"1" are shift invariant on the positions that reach the output.
The ``--pass-cc-advanced-max`` knob of the reference bounds its shared-memory
batches (cc_advance.py:33-57); the HIP backend has no such batches, so the
knob is accepted and recorded only.
"""
from ..inst import (
    BsInstType, BsAssign, BsAdvance, BsNot, ADV_LEFT, ADV_RIGHT, walk,
)
from .pass_utils import is_const


class cc_advance_pass:
    def __init__(self):
        self.pass_name = "cc_advance"
        self.advance_count = {}
        self.shift_cache = {}   # (var, amount, left) -> shifted var id
        self.cache_keys = {}    # var -> cache keys naming it (as source or result)
        self.names = None
        self.inserted_left = 0
        self.inserted_right = 0

    def get_advance_count(self, var_id):
        return self.advance_count.get(var_id, 0)

    # -- helpers -------------------------------------------------------------
    def _invalidate(self, var):
        for key in self.cache_keys.pop(var, ()):
            self.shift_cache.pop(key, None)

    def _new_var(self, vnm, name):
        if self.names is None:
            self.names = set(vnm.values())
            self.next_id = max(vnm) + 1 if vnm else 0
        cand, n = name, 0
        while cand in self.names:
            n += 1
            cand = f"{name}_{n}"
        vid = self.next_id
        self.next_id += 1
        vnm[vid] = cand
        self.names.add(cand)
        return vid

    def _shift(self, out, vnm, var, amount, left):
        key = (var, amount, left)
        if key in self.shift_cache:
            return self.shift_cache[key]
        name = f"{vnm[var]}_advance{amount}" if left else f"{vnm[var]}_realign{amount}"
        new = self._new_var(vnm, name)
        out.append(BsAdvance([var, amount], new, ADV_LEFT if left else ADV_RIGHT))
        if left:
            self.inserted_left += 1
        else:
            self.inserted_right += 1
        self.shift_cache[key] = new
        self.cache_keys.setdefault(var, []).append(key)
        self.cache_keys.setdefault(new, []).append(key)
        return new

    def _define(self, ret, count):
        self._invalidate(ret)
        self.advance_count[ret] = count

    # -- the pass -------------------------------------------------------------
    def advance(self, insts, vnm):
        out = []
        for inst in insts:
            t = inst.type
            cnt = lambda v: 0 if is_const(vnm, v) else self.get_advance_count(v)  # noqa: E731
            if t == BsInstType.STREAMLOAD:
                out.append(inst)
                self._define(inst.ret, 0)
            elif t == BsInstType.ADVANCE:
                if inst.is_left:
                    raise NotImplementedError("cc_advance: input already contains look-ahead shifts")
                src = inst.operands[0]
                c = cnt(src) + inst.amount
                out.append(BsAssign([src], inst.ret))
                self._define(inst.ret, c)
            elif t == BsInstType.ASSIGN:
                c = cnt(inst.operands[0])
                out.append(inst)
                self._define(inst.ret, c)
            elif t == BsInstType.AND:
                a, b = inst.operands
                ca, cb = cnt(a), cnt(b)
                if ca != cb:
                    lo_i = 0 if ca < cb else 1
                    lo = inst.operands[lo_i]
                    if not is_const(vnm, lo):
                        inst.operands[lo_i] = self._shift(out, vnm, lo, abs(ca - cb), left=True)
                out.append(inst)
                self._define(inst.ret, max(ca, cb))
            elif t in (BsInstType.OR, BsInstType.XOR):
                a, b = inst.operands
                ca, cb = cnt(a), cnt(b)
                if ca != cb:
                    hi_i = 0 if ca > cb else 1
                    inst.operands[hi_i] = self._shift(out, vnm, inst.operands[hi_i], abs(ca - cb), left=False)
                out.append(inst)
                self._define(inst.ret, min(ca, cb))
            elif t == BsInstType.NOT:
                a = inst.operands[0]
                if cnt(a) > 0:
                    inst.operands[0] = self._shift(out, vnm, a, cnt(a), left=False)
                out.append(inst)
                self._define(inst.ret, 0)
            elif t in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU):
                m, c = inst.operands
                cm, cc = cnt(m), cnt(c)
                if cc < cm and not is_const(vnm, c):
                    inst.operands[1] = self._shift(out, vnm, c, cm - cc, left=True)
                elif cc > cm:
                    inst.operands[1] = self._shift(out, vnm, c, cc - cm, left=False)
                out.append(inst)
                self._define(inst.ret, cm)
            elif t == BsInstType.SEL:
                for i in range(2):
                    v = inst.operands[i]
                    if cnt(v) > 0:
                        inst.operands[i] = self._shift(out, vnm, v, cnt(v), left=False)
                if cnt(inst.condition) > 0:
                    inst.condition = self._shift(out, vnm, inst.condition, cnt(inst.condition), left=False)
                out.append(inst)
                self._define(inst.ret, 0)
            elif t in (BsInstType.WHILE, BsInstType.IF):
                used = set()
                defined = set()
                for sub in walk([inst]):
                    used.update(sub.uses())
                    defined.update(sub.defs())
                for v in sorted(used):
                    k = cnt(v)
                    if k > 0:
                        # in place: the block may reassign v (loop-carried)
                        out.append(BsAdvance([v, k], v, ADV_RIGHT))
                        self.inserted_right += 1
                        self._define(v, 0)
                out.append(inst)
                for v in defined:
                    self._define(v, 0)
            elif t == BsInstType.STREAMSTORE:
                a = inst.operands[0]
                if cnt(a) > 0:
                    inst.operands[0] = self._shift(out, vnm, a, cnt(a), left=False)
                out.append(inst)
            elif t in (BsInstType.GRAPHBREAK, BsInstType.STR, BsInstType.LABEL, BsInstType.SKIPZERO):
                out.append(inst)
            else:
                raise NotImplementedError(f"cc_advance: {inst}")
        return out

    def run(self, insts, var_name_map):
        if any(i.type == BsInstType.GRAPHBREAK for i in insts):
            raise NotImplementedError("cc_advance runs before graph breaks are inserted (single subgraph only)")
        return self.advance(list(insts), var_name_map), var_name_map
