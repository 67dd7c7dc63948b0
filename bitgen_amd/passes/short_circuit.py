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
from ..inst import BsInstType, BsAssign, BsOr, BsSkipZero, BsLabel
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
        live_of = []
        for (i, e) in regions:
            live = []
            for j in range(i + 1, e + 1):
                for x in insts[j].defs():
                    if last_use.get(x, -1) > e and x not in live:
                        live.append(x)
            live_of.append(live)
        folded = self._fold_accumulations(insts, regions, live_of, defs, last_use)

        before = {}   # index -> list of instructions inserted before it
        after = {}    # index -> instructions / labels placed after it
        zeroed = {}   # value -> check index of the range that zero-initialises it
        for label, (i, e) in enumerate(regions):
            v = insts[i].ret
            live = live_of[label]
            # x is written once, inside this range; a range opened earlier
            # that already zeroes x has not reached x's definition yet, so x is
            # still zero here
            pre = [BsAssign([zero], x) for x in live if not (x in zeroed and zeroed[x] < i)]
            for x in live:
                zeroed.setdefault(x, i)
            pre.append(BsSkipZero(v, label))
            before.setdefault(i + 1, []).extend(pre)
            after.setdefault(e, []).append(BsLabel(label))
            self.checks += 1
            self.skippable_insts += e - i
        for e, (first_check, init, acc) in folded.items():
            before.setdefault(first_check + 1, []).insert(0, init)
            after[e].insert(0, acc)
        out = []
        for j, inst in enumerate(insts):
            if j - 1 in folded:
                continue  # accumulation moved into the range ending at j - 1
            out.extend(before.get(j, []))
            out.append(inst)
            # the moved accumulation first, then labels (inner ranges close first)
            tail = after.get(j, [])
            if j in folded:
                out.append(tail[0])
                tail = tail[1:]
            out.extend(reversed(tail))
        return out, vnm

    @staticmethod
    def _fold_accumulations(insts, regions, live_of, defs, last_use):
        """``y = a | z`` right after ranges whose only result z feeds it becomes
        ``y = a`` before the first check and ``y = y | z`` at the end of the
        ranges: skipping then leaves y = a = a | 0 in place, so the
        accumulator needs no zero-initialised z and no copy at the join."""
        ends = {}
        for k, (i, e) in enumerate(regions):
            ends.setdefault(e, []).append(k)
        folded = {}
        for e, ks in ends.items():
            if e + 1 >= len(insts):
                continue
            acc = insts[e + 1]
            if acc.type != BsInstType.OR or defs.get(acc.ret, 0) != 1:
                continue
            first = min(regions[k][0] for k in ks)
            # no other range may span the accumulation
            if any(i < e + 1 <= ee and ee != e for (i, ee) in regions):
                continue
            a, b = acc.operands
            for z, other in ((a, b), (b, a)):
                if (all(z in live_of[k] for k in ks) and last_use.get(z) == e + 1
                        and other not in {d for j in range(first + 1, e + 1) for d in insts[j].defs()}
                        and other != acc.ret and z != other):
                    for k in ks:
                        live_of[k].remove(z)
                    y = acc.ret
                    folded[e] = (first, BsAssign([other], y), BsOr([y, z], y))
                    break
        return folded

    def _region(self, insts, i, defs, last_use, vnm):
        import heapq

        v = insts[i].ret
        Z = {v}
        pending = set()   # defined in the range, not in Z, read after the current end
        heap = []         # (last use, value) of pending values
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
            else:
                for x in ds:
                    lu = last_use.get(x, -1)
                    if lu > j:
                        pending.add(x)
                        heapq.heappush(heap, (lu, x))
            while heap and heap[0][0] <= j:
                pending.discard(heapq.heappop(heap)[1])
            useful += 1
            if not pending and useful >= self.min_region:
                best = j
        return best
