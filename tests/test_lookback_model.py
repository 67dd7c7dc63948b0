"""CPU model of the decoupled look-back walk (bg_lookback) and its
write-backs, on the status words of one carry op.

Each tile either passes its carry-in on (its status is an aggregate) or
decides its carry-out (inclusive status with the value).  A walk from tile t
reads its predecessor, behind an aggregate the 64 statuses below it at once,
and so on down to an inclusive status (tile 0 is always inclusive).  Once
found, the value is written back as an inclusive status into the
aggregates of the walk's last step (next to the inclusive status).  Every
aggregate between t and the inclusive status passes the same carry on, so
the written value is each one's true carry-out.

The model runs walks in random order over random aggregate / inclusive
patterns and checks that each walk returns the true carry-in of its tile
and that every written status holds the tile's true carry-out."""
import random

import pytest

AGG, INC = 1, 2


def walk(st, t, stats):
    """bg_lookback from tile t (statuses published for every tile): returns
    the carry into t; st[i] = (kind, value)."""
    j = t - 1
    kind, val = st[j]
    while True:
        if kind == INC:
            return val
        # tile j passes its carry-in on: read the 64 tiles before it
        w = [st[j - 1 - lane] if j - 1 - lane >= 0 else (INC, 0) for lane in range(64)]
        stats["steps"] += 1
        aggs = sum(1 << lane for lane in range(64) if w[lane][0] == AGG)
        if aggs == (1 << 64) - 1:
            j -= 64
            kind, val = st[j]
            continue
        f = next(lane for lane in range(64) if not (aggs >> lane) & 1)
        kind, val = w[f]
        jlast = j
        j -= 1 + f
        if kind == INC:
            for lane in range(f):
                st[jlast - 1 - lane] = (INC, val)


def truth(pattern):
    """True carry-out of every tile: an inclusive tile decides its own, an
    aggregate passes on its predecessor's (tile 0: decided)."""
    out = []
    for i, (kind, val) in enumerate(pattern):
        out.append(val if kind == INC or i == 0 else out[-1])
    return out


@pytest.mark.parametrize("seed", range(8))
def test_walks_return_true_carries_and_write_back_true_values(seed):
    rng = random.Random(seed)
    n = rng.choice([100, 700, 3000])
    p_inc = rng.choice([0.5, 0.05, 0.002])
    pattern = [(INC, rng.randrange(2))] + [(INC, rng.randrange(2)) if rng.random() < p_inc else (AGG, 0)
                                           for _ in range(n - 1)]
    out = truth(pattern)
    st = list(pattern)
    stats = {"steps": 0}
    order = list(range(1, n))
    rng.shuffle(order)
    for t in order:
        assert walk(st, t, stats) == out[t - 1], t
        for i, (kind, val) in enumerate(st):
            assert kind == AGG or val == out[i], (i, kind, val)

