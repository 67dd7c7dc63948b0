"""Batched form (one persistent launch per pattern set over many uneven
buffers, optionally graph-captured): bit-exact against the oracle (GPU)."""
import random

import numpy as np
import pytest

from bitgen_amd.batch import BatchPatternSet, BatchPlan
from tests.gpu_helpers import run_oracle
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SETS = [
    [["ab[0-9]{1,3}c", "x[^\\n]*yz"], ["hello", "w.*d"]],
    [["q(a|b)z*[xy]{2}"], ["a.{0,40}b", "dd[^a]*c"], ["abcd", "^a[0-9]*$"]],
]


def _buffers(seed, n):
    rng = random.Random(seed)
    alpha = np.frombuffer(b"abcdxyzqhelowr0123 \n", dtype=np.uint8)
    g = np.random.default_rng(seed)
    sizes = [rng.choice([0, 1, 63, 64, 500, 4000, 9000, 40000, 200000]) for _ in range(n)]
    return [g.choice(alpha, size=s) for s in sizes]


def _check(plan, sets, buffers, jobs):
    import torch

    torch.cuda.synchronize()
    plan.check_errors()
    for (si, b) in jobs:
        got = plan.results(si, b).cpu().numpy()
        ref = run_oracle(SETS[si], buffers[b])
        n = buffers[b].size
        for g in range(len(SETS[si])):
            assert np.array_equal(O.mask_to_positions(got[g], n + 1), ref[g]), (si, b, g, n)


def test_batched_matches_oracle_eager_and_graph(gpu):
    import torch

    sets = [BatchPatternSet(s).load(0) for s in SETS]
    buffers = _buffers(1, 14)
    jobs = [(si, b) for si in range(len(sets)) for b in range(len(buffers))]
    print(f"buffers={[b.size for b in buffers]}", flush=True)
    plan = BatchPlan(sets, buffers, jobs)
    plan.enqueue()
    _check(plan, sets, buffers, jobs)
    # graph replays with eager work between them; every replay must match
    plan.result.fill_(0x5A)
    plan.capture()
    for rep in range(3):
        plan.result.fill_(0)
        plan.run_graph()
        torch.cuda.synchronize()
        _check(plan, sets, buffers, jobs)
        _ = (plan.bytes[:128].to(torch.int32) * 3).sum().item()  # unrelated eager work
    plan.destroy()
    for s in sets:
        s.free()


def test_batched_partial_job_list(gpu):
    """A subset of the (set, buffer) pairs, few persistent blocks (many tiles
    per wave), carries settled inline (one launch per kernel)."""
    sets = [BatchPatternSet(s, {"carry_settle": "inline"}).load(0) for s in SETS]
    buffers = _buffers(2, 9)
    jobs = [(0, 0), (0, 3), (1, 3), (1, 8), (0, 8)]
    plan = BatchPlan(sets, buffers, jobs, grid_blocks=7)  # few persistent blocks: many tiles per wave
    plan.enqueue()
    _check(plan, sets, buffers, jobs)
    plan.destroy()


def test_batched_merged_kernels(gpu):
    """Batch kernels that compute two or three groups each: per-group
    streams of every job stay bit-exact."""
    sets = [BatchPatternSet(s, {"groups_per_kernel": k}).load(0) for s, k in zip(SETS, (2, 3))]
    assert [s.n_kernels for s in sets] == [1, 1]
    buffers = _buffers(3, 10)
    jobs = [(si, b) for si in range(len(sets)) for b in range(len(buffers))]
    print(f"buffers={[b.size for b in buffers]}", flush=True)
    plan = BatchPlan(sets, buffers, jobs)
    plan.enqueue()
    _check(plan, sets, buffers, jobs)
    plan.destroy()
    for s in sets:
        s.free()


def test_batched_bigram_filter(gpu):
    """Batch kernels with byte-pair filters: the filter table is built per
    tile from that tile's job, so jobs of other buffers never leak into it."""
    sets = [BatchPatternSet(s, {"bigram_filter": 1, "groups_per_kernel": k}).load(0) for s, k in zip(SETS, (1, 2))]
    buffers = _buffers(5, 12)
    jobs = [(si, b) for si in range(len(sets)) for b in range(len(buffers))]
    print(f"buffers={[b.size for b in buffers]}", flush=True)
    plan = BatchPlan(sets, buffers, jobs)
    plan.enqueue()
    _check(plan, sets, buffers, jobs)
    plan.destroy()
    for s in sets:
        s.free()



SHARED_SETS = [
    [["hello", "ab[0-9]+c"], ["x[yz]q", "w.*d"]],
    [["abcd", "zz[^\\n]*yy"], ["q0q0", "dcba(xy|zq)"], ["[ab]cd", "ca.*bd"]],
]


def test_batched_shared_filter_eager_and_graph(gpu):
    """Sets whose kernels share their byte-pair answers (two or more kernels
    with filters): one batched filter pre-pass per set over the batch's
    global tiles (each mapped to its job's buffer and tile), then the set's
    kernels read their rows; eager and graph replays (the pre-pass is a
    kernel node), every job bit-exact."""
    import torch

    sets = [BatchPatternSet(s, {"bigram_filter": 1, "shared_filter": 1, "groups_per_kernel": k}).load(0)
            for s, k in zip(SHARED_SETS, (1, 2))]
    assert all(s.filter_row > 0 and s.filter_kernel is not None for s in sets)
    buffers = _buffers(9, 14)
    # a partial job list: not every buffer in every set
    jobs = [(si, b) for si in range(len(sets)) for b in range(len(buffers)) if (si + b) % 3]
    print(f"buffers={[b.size for b in buffers]}", flush=True)
    plan = BatchPlan(sets, buffers, jobs)

    def check():
        torch.cuda.synchronize()
        plan.check_errors()
        for (si, b) in jobs:
            got = plan.results(si, b).cpu().numpy()
            ref = run_oracle(SHARED_SETS[si], buffers[b])
            n = buffers[b].size
            for g in range(len(SHARED_SETS[si])):
                assert np.array_equal(O.mask_to_positions(got[g], n + 1), ref[g]), (si, b, g, n)

    plan.enqueue()
    check()
    plan.capture()
    for rep in range(3):
        plan.result.fill_(0)
        plan.run_graph()
        check()
        _ = (plan.bytes[:128].to(torch.int32) * 3).sum().item()  # unrelated eager work
    plan.destroy()
    for s in sets:
        s.free()


@pytest.mark.parametrize("settle", ["launch", "defer"])
def test_batched_phased_graph(gpu, settle):
    """Batch kernels in the sweep + settle form (both launches of every
    kernel captured into the graph) and in the deferred-settle form (one
    launch); eager and every replay bit-exact."""
    import torch

    sets = [BatchPatternSet(s, {"carry_settle": settle, "groups_per_kernel": k}).load(0)
            for s, k in zip(SETS, (1, 2))]
    buffers = _buffers(7, 12)
    jobs = [(si, b) for si in range(len(sets)) for b in range(len(buffers))]
    print(f"buffers={[b.size for b in buffers]}", flush=True)
    plan = BatchPlan(sets, buffers, jobs)
    plan.enqueue()
    _check(plan, sets, buffers, jobs)
    plan.capture()
    for rep in range(3):
        plan.result.fill_(0x5A)
        plan.run_graph()
        torch.cuda.synchronize()
        _check(plan, sets, buffers, jobs)
    plan.destroy()
    for s in sets:
        s.free()


LOOP_SETS = [
    [["a(bc)*d", "xy"], ["hello", "w.*d"]],
    [["(ab)+c"], ["q0q0"], ["(a|bb)+c", "zz[^\\n]*yy"]],
]


def test_batched_stream_level_groups_eager(gpu):
    """Sets with groups whose loops are decided on the whole stream: the
    tile-parallel groups run as batch kernels, the stream-level groups run
    their plan job by job through the plain form (eagerly); every job
    bit-exact; such a plan is refused for graph capture."""
    import torch

    sets = [BatchPatternSet(s).load(0) for s in LOOP_SETS]
    assert all(s.has_stream_groups for s in sets)
    buffers = _buffers(13, 10)
    for b in buffers[:4]:
        if b.size > 40:
            b[10:17] = np.frombuffer(b"abcbcbd", dtype=np.uint8)
            b[25:32] = np.frombuffer(b"abababc", dtype=np.uint8)
    jobs = [(si, b) for si in range(len(sets)) for b in range(len(buffers)) if (si + b) % 4]
    print(f"buffers={[b.size for b in buffers]}", flush=True)
    plan = BatchPlan(sets, buffers, jobs)
    assert not plan.graph_capturable
    for rep in range(2):
        plan.result.fill_(0x5A)
        plan.enqueue()
        torch.cuda.synchronize()
        plan.check_errors()
        for (si, b) in jobs:
            got = plan.results(si, b).cpu().numpy()
            ref = run_oracle(LOOP_SETS[si], buffers[b])
            n = buffers[b].size
            for g in range(len(LOOP_SETS[si])):
                assert np.array_equal(O.mask_to_positions(got[g], n + 1), ref[g]), (rep, si, b, g, n)
    with pytest.raises(NotImplementedError, match="enqueue"):
        plan.capture()
    plan.destroy()
    for s in sets:
        s.free()
