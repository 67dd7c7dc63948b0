"""CPU model of the deferred-settle work loop (carry_settle="defer").

The generated loop (GroupProgram.work_loop) is restated here step for step:
tile queues (bg_queue_open / issue / finish / close), the choice of the
iteration (a tile to settle at once, a new tile, the pending tile), the
phase-0 pass (publishes every op's status, or on an abort the ops before
the abort point and deferred marks after it), the phase-1 settle (resolves
the carry-ins from the predecessor's statuses and replays when needed), and
the status rows the device code addresses.  Many waves run it under random
interleavings; every status access is bounds-checked against the status
buffer the host allocates (n_tiles x n_carry words), and the model checks
that every tile is passed once, settled once, that waits only ever target
smaller tiles, and that the loop terminates (no wave waits forever) with
every wave making exactly one failing take.
"""
import random

import pytest

from tests.test_codegen_abi import _gprog

NONE = 0xFFFFFFFF
UNPUB, AGG, INC, DEF = 0, 1, 2, 3


class Device:
    def __init__(self, n_tiles, n_carry, n_queues, n_blocks, waves_per_block, rng, abort_p, agg_p):
        self.n, self.nc = n_tiles, n_carry
        self.status = [UNPUB] * (n_tiles * n_carry)   # kinds only (values do not matter here)
        self.counter = [0] * 64
        self.nq = min(n_queues, n_blocks) or 1
        self.rng = rng
        self.abort_p, self.agg_p = abort_p, agg_p
        self.passed = [0] * n_tiles
        self.settled = [0] * n_tiles
        self.replayed = [0] * n_tiles
        self.waves = [Wave(self, b % self.nq) for b in range(n_blocks) for _ in range(waves_per_block)]

    def st(self, t, k):
        i = t * self.nc + k
        assert 0 <= t < self.n and 0 <= k < self.nc and 0 <= i < len(self.status), (t, k)
        return i

    def take(self, q):
        """bg_queue_issue + bg_queue_finish with take = 1: tile k*nq + q."""
        k = self.counter[q]
        self.counter[q] += 1
        t = k * self.nq + q
        return t, t < self.n


class Wave:
    """One wave running the generated loop as a generator: yields while it
    waits (the scheduler runs another wave), returns when the loop exits."""

    def __init__(self, dev, q):
        self.dev, self.q = dev, q
        self.failing_takes = 0
        self.waits_on = None

    def run(self):
        d = self.dev
        cur, more = d.take(self.q)
        if not more:
            self.failing_takes += 1
        pend = now = NONE
        while True:
            if now != NONE:
                t, ph, now = now, 1, NONE
            elif more:
                t, ph = cur, 0
                nxt = d.take(self.q)            # issued before the pass
            elif pend != NONE:
                t, ph, pend = pend, 1, NONE
            else:
                break
            assert 0 <= t < d.n
            if ph == 0:
                ok = yield from self.pass_(t)
                cur, more = nxt
                if not more:
                    self.failing_takes += 1
                if not ok:
                    now = t
                else:
                    now, pend = pend, t
            else:
                yield from self.settle(t)

    def pass_(self, t):
        """Phase 0: one speculative pass.  Verify may wait on the predecessor
        (an op's status not published yet); in the sweep phase a deferred
        status aborts instead of waiting."""
        d = self.dev
        d.passed[t] += 1
        abort_at = d.nc
        if t > 0 and d.rng.random() < d.abort_p:
            k = d.rng.randrange(d.nc)
            while True:  # bg_verify -> bg_lookback on op k of tile t-1
                kind = d.status[d.st(t - 1, k)]
                if kind == DEF:
                    break
                if kind != UNPUB:
                    break
                self.waits_on = t - 1
                yield
            self.waits_on = None
            abort_at = k
        for k in range(abort_at):
            # tile 0 is exact: inclusive statuses only
            d.status[d.st(t, k)] = AGG if t > 0 and d.rng.random() < d.agg_p else INC
        if abort_at < d.nc:
            for k in range(abort_at, d.nc):   # bg_sweep_end: deferred marks
                d.status[d.st(t, k)] = DEF
            return False
        return True

    def settle(self, t):
        """Phase 1: bg_settle_begin -- the carry-in of every op the tile
        published comes from the predecessor (waiting while its status is not
        published or deferred); a replay publishes every op (waiting op by
        op for the predecessor where the tile had not published)."""
        d = self.dev
        d.settled[t] += 1
        if t == 0:
            return
        for k in range(d.nc):
            # own status, then the predecessor's (look-back through
            # aggregates ends at an inclusive status or tile 0)
            d.status[d.st(t, k)]
            j = t - 1
            while True:
                kind = d.status[d.st(j, k)]
                if kind == INC:
                    break
                if kind == AGG:
                    j -= 1
                    continue
                self.waits_on = j
                yield
            self.waits_on = None
            assert j < t
        if d.rng.random() < 0.1 or any(d.status[d.st(t, k)] == DEF for k in range(d.nc)):
            d.replayed[t] += 1
        for k in range(d.nc):
            d.status[d.st(t, k)] = INC


def simulate(seed, n_tiles, n_carry, n_queues, n_blocks, wpb, abort_p, agg_p):
    rng = random.Random(seed)
    dev = Device(n_tiles, n_carry, n_queues, n_blocks, wpb, rng, abort_p, agg_p)
    live = {i: w.run() for i, w in enumerate(dev.waves)}
    idle_rounds = 0
    while live:
        order = list(live)
        rng.shuffle(order)
        progressed = False
        for i in order:
            w = dev.waves[i]
            before = (sum(dev.passed), sum(dev.settled), w.waits_on)
            try:
                next(live[i])
            except StopIteration:
                del live[i]
                progressed = True
                continue
            if (sum(dev.passed), sum(dev.settled), w.waits_on) != before or w.waits_on is None:
                progressed = True
        idle_rounds = 0 if progressed else idle_rounds + 1
        assert idle_rounds < 1000, "no wave can make progress: the loop would wait forever"
    return dev


@pytest.mark.parametrize("seed", range(40))
def test_deferred_loop_model(seed):
    rng = random.Random(seed)
    n_tiles = rng.choice([1, 2, 3, 7, 64, 200, 1000])
    n_carry = rng.choice([1, 3, 33, 64])
    n_blocks = rng.choice([1, 2, 5, 40])
    wpb = rng.choice([1, 2, 4])
    dev = simulate(seed, n_tiles, n_carry, rng.choice([1, 8, 64]), n_blocks, wpb,
                   abort_p=rng.choice([0.0, 0.05, 0.5]), agg_p=rng.choice([0.0, 0.3, 1.0]))
    # every tile passed once (phase 0) and settled once (phase 1)
    assert dev.passed == [1] * n_tiles
    assert dev.settled == [1] * n_tiles
    # every status word ends inclusive (published for the next launch's epoch)
    assert all(k == INC for k in dev.status)
    # one failing take per wave: the queue protocol's close condition
    assert all(w.failing_takes == 1 for w in dev.waves)


def test_model_follows_the_generated_loop():
    """The statements the model restates are the generated loop's."""
    p = _gprog(["ab[^\\n]*cd", "x.*y"], carry_settle="defer")
    loop = p.source()[p.source().index("__global__"):]
    for stmt in ("if (now != 0xFFFFFFFFu) { t = now; ph = 1; now = 0xFFFFFFFFu; }",
                 "else if (more) { t = cur; ph = 0; next = bg_queue_issue(Q); }",
                 "else if (pend != 0xFFFFFFFFu) { t = pend; ph = 1; pend = 0xFFFFFFFFu; }",
                 "more = bg_queue_finish(Q, next, cur);",
                 "if (!bg_ok) now = t;",
                 "else { now = pend; pend = t; }",
                 "bg_queue_open(a.ctl, a.n_queues, a.n_tiles, 1)"):
        assert stmt in loop, stmt
