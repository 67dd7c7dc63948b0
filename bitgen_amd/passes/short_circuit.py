"""Zero-block skipping ("short circuit").

Same goal as ``bitgen/passes/short_circuit.py:91-388``: along chains of ANDs
on marker streams, test whether the running value is all zero in the current
unit of execution and, if so, jump over the rest of the chain.  The reference
finds chains with ``nx.all_simple_paths`` from the basis loads (exponential,
:23-64), zero-initialises every variable defined in the skipped range
(:332-350) and only guards names starting with ``at`` against escaping
(:172-206).  This pass instead proves each jump exact:

* a value is a *marker* when its computation involves a shift or carry
  (Advance in either direction, MatchStar, ScanThru); an AND with a marker
  operand is a chain step and its depth is one more than the deepest chain
  step among its operands;
* checks go after chain steps of depth ``start+1, start+1+interval, ...``
  (``--pass-short-circuit-start``, ``--pass-short-circuit-interval``);
* from a check on v, the skipped range grows instruction by instruction while
  tracking Z, the values that are zero whenever v is zero (AND with a Z
  operand, Advance of Z, OR/XOR/copy of Z-only operands).  NOT is never in
  Z, and a range stops before any MatchStar/ScanThru: a carry can enter from
  outside the unit, and on the GPU every tile must publish each carry
  operation's look-back status.  The range may end at instruction e only if
  every value defined in it and read after e is in Z;
* those values are set to zero before the check, so skipping and executing
  give the same result.

Only straight-line top-level code is scanned (loops and branches end a range).
"""
from ..inst import BsInstType, BsAssign, BsSkipZero, BsLabel
from .pass_utils import def_counts, top_level_last_use, const_id, is_const
from .. import config as cfg

_SHIFTY = (BsInstType.ADVANCE, BsInstType.MATCHSTAR, BsInstType.SCANTHRU)
# A skipped range never contains a carry operation: every tile must reach each
# MatchStar/ScanThru so its successor's carry look-back can complete.
_STOP = (BsInstType.WHILE, BsInstType.IF, BsInstType.SKIPZERO, BsInstType.LABEL,
         BsInstType.STREAMSTORE, BsInstType.STREAMLOAD, BsInstType.GRAPHBREAK,
         BsInstType.STR, BsInstType.SEL, BsInstType.MATCHSTAR, BsInstType.SCANTHRU,
         BsInstType.ADD)


class short_circuit_pass:
    def __init__(self, profile=False, start=None, interval=None, min_region=3):
        self.pass_name = "short_circuit"
        self.profile = profile
        self.start = start
        self.interval = interval
        self.min_region = min_region
        self.checks = 0
        self.skippable_insts = 0

    def _params(self):
        start = self.start if self.start is not None else cfg.get_config("pass_short_circuit_start", 0)
        interval = self.interval if self.interval is not None else cfg.get_config("pass_short_circuit_interval", 1)
        interval = max(1, int(interval))
        return max(0, int(start)), interval

    def run(self, insts, vnm):
        if cfg.get_config("pass_short_circuit_interval_disable", 0):
            return insts, vnm
        start, interval = self._params()
        insts = list(insts)
        n = len(insts)
        defs = def_counts(insts)
        last_use = top_level_last_use(insts)

        # marker analysis and chain depth (top level, program order)
        marker = set()
        depth_of = {}
        for inst in insts:
            t = inst.type
            if t in (BsInstType.WHILE, BsInstType.IF):
                from ..inst import walk
                for sub in walk([inst]):
                    marker.update(sub.defs())
                continue
            uses = inst.uses()
            if t in _SHIFTY or any(u in marker for u in uses):
                marker.update(inst.defs())
            if t == BsInstType.AND and any(u in marker for u in uses):
                depth_of[inst.ret] = 1 + max((depth_of.get(u, 0) for u in uses), default=0)

        regions = []
        for i, inst in enumerate(insts):
            if inst.type != BsInstType.AND or defs.get(inst.ret, 0) != 1:
                continue
            d = depth_of.get(inst.ret, 0)
            if d < start + 1 or (d - start - 1) % interval != 0:
                continue
            e = self._region(insts, i, defs, last_use, vnm)
            if e is None:
                continue
            regions.append((i, e))

        if not regions:
            return insts, vnm

        zero = const_id(vnm, "0")
        before = {}   # index -> list of instructions inserted before it
        after = {}    # index -> labels placed after it
        for label, (i, e) in enumerate(regions):
            v = insts[i].ret
            live = []
            for j in range(i + 1, e + 1):
                for x in insts[j].defs():
                    if last_use.get(x, -1) > e and x not in live:
                        live.append(x)
            pre = [BsAssign([zero], x) for x in live]
            pre.append(BsSkipZero(v, label))
            before.setdefault(i + 1, []).extend(pre)
            after.setdefault(e, []).append(BsLabel(label))
            self.checks += 1
            self.skippable_insts += e - i
        out = []
        for j, inst in enumerate(insts):
            out.extend(before.get(j, []))
            out.append(inst)
            # inner (later-opened) ranges close first
            out.extend(reversed(after.get(j, [])))
        return out, vnm

    def _region(self, insts, i, defs, last_use, vnm):
        v = insts[i].ret
        Z = {v}
        defined = []
        best = None
        useful = 0
        horizon = last_use.get(v, i)
        for j in range(i + 1, len(insts)):
            if j > horizon:
                break
            inst = insts[j]
            t = inst.type
            if t in _STOP:
                break
            ds = inst.defs()
            if any(defs.get(x, 0) != 1 for x in ds):
                break
            ops = inst.uses()
            zi = False
            if t == BsInstType.AND:
                zi = any(o in Z for o in ops)
            elif t in (BsInstType.ADVANCE, BsInstType.ASSIGN):
                zi = ops[0] in Z
            elif t in (BsInstType.OR, BsInstType.XOR):
                zi = all(o in Z or (is_const(vnm, o) and vnm[o] == "0") for o in ops)
            if zi:
                Z.update(ds)
                for x in ds:
                    horizon = max(horizon, last_use.get(x, j))
            defined.extend(ds)
            useful += 1
            if all(x in Z or last_use.get(x, -1) <= j for x in defined):
                best = j if useful >= self.min_region else best
        return best
