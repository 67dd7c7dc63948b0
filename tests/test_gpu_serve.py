"""Continuous-batching server (GPU): every job's result equals the oracle's,
and equals the same job served alone."""
import random
import threading
import time

import numpy as np
import pytest

from bitgen_amd.batch import BatchPatternSet
from bitgen_amd.serve import MatchServer, job_footprint
from tests.gpu_helpers import run_oracle
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SETS = {
    0: [["ab[0-9]{1,3}c", "x[^\\n]*yz"], ["hello", "w.*d"]],
    1: [["q(a|b)z*[xy]{2}"], ["a.{0,40}b", "dd[^a]*c"], ["abcd", "^a[0-9]*$"]],
    2: [["zz[0-9]+q", "hel+o"]],
}


def _data(rng, g):
    alpha = np.frombuffer(b"abcdxyzqhelowr0123 \n", dtype=np.uint8)
    n = rng.choice([0, 5, 100, 3000, 9000, 50000, 120000])
    return g.choice(alpha, size=n)


@pytest.fixture(scope="module")
def sets(gpu):
    return {k: BatchPatternSet(v).load(0) for k, v in SETS.items()}


def _check(set_id, data, res):
    ref = run_oracle(SETS[set_id], data)
    got = O.mask_to_positions(res, data.size + 1)
    assert got.shape == ref.shape and np.array_equal(got, ref), (set_id, data.size)


def test_served_results_match_oracle_and_isolation(sets):
    rng = random.Random(5)
    g = np.random.default_rng(5)
    jobs = [(rng.choice(list(SETS)), _data(rng, g)) for _ in range(120)]
    srv = MatchServer(sets, n_streams=3, max_batch_jobs=16).start()
    futs = []

    def producer():
        for sid, d in jobs:
            futs.append(srv.submit(sid, d))
            time.sleep(rng.random() * 2e-4)

    t = threading.Thread(target=producer)
    t.start()
    t.join()
    results = [f.result(timeout=60) for f in futs]
    srv.stop()
    assert srv.batches > 1  # jobs arrived over time and were batched continuously
    for (sid, d), r in zip(jobs, results):
        _check(sid, d, r)
    # the same jobs, each served alone, give identical streams
    alone = MatchServer(sets, n_streams=1, max_batch_jobs=1).start()
    for (sid, d), r in list(zip(jobs, results))[:25]:
        r1 = alone.submit(sid, d).result(timeout=60)
        assert np.array_equal(O.mask_to_positions(r1, d.size + 1), O.mask_to_positions(r, d.size + 1))
    alone.stop()


def test_admission_cap_queues_instead_of_failing(sets):
    g = np.random.default_rng(9)
    datas = [g.choice(np.frombuffer(b"abcd\n", dtype=np.uint8), size=20000) for _ in range(12)]
    one = job_footprint(20000, len(SETS[1]))
    srv = MatchServer(sets, hbm_cap_bytes=int(one * 2.5), n_streams=3).start()
    futs = [srv.submit(1, d) for d in datas]
    res = [f.result(timeout=60) for f in futs]
    srv.stop()
    assert srv.batches >= 6  # at most two jobs fit under the cap at a time
    for d, r in zip(datas, res):
        _check(1, d, r)
    # a job larger than the cap is rejected, not hung
    small = MatchServer(sets, hbm_cap_bytes=1000).start()
    with pytest.raises(ValueError):
        small.submit(0, datas[0]).result(timeout=10)
    small.stop()


def test_warm_then_serve(sets):
    """warm() runs one batch of every set on every slot before serving; the
    warm batches are not recorded and later jobs are exact."""
    rng = random.Random(8)
    g = np.random.default_rng(8)
    srv = MatchServer(sets, n_streams=3, max_batch_jobs=8).warm(_data(rng, g))
    assert srv.batches == 0 and not srv.completed and srv.in_flight_bytes == 0
    srv.start()
    jobs = [(rng.choice(list(SETS)), _data(rng, g)) for _ in range(30)]
    futs = [srv.submit(sid, d) for sid, d in jobs]
    results = [f.result(timeout=60) for f in futs]
    srv.stop()
    assert len(srv.completed) == len(jobs)
    for (sid, d), r in zip(jobs, results):
        _check(sid, d, r)
