"""The benchmark's workloads at their stated sizes and in the launch shapes the
benchmark runs, bit-exact against the oracle (GPU).

* IDS: 1,000 regexes in 16 groups over the 100 MiB buffer, in the shape the
  autotuner picks for it (4 groups per kernel, 5 waves per SIMD, byte-pair
  filters); every group's whole stream is compared.
* Every autotuner candidate on a 16 MiB prefix of the IDS buffer, and on
  2 MiB of the genomic text, so no shape the tuner can choose is unchecked.
* Batched uneven: the 2 GiB plan (3 sets x 4,096 log-normal buffers) as one
  graph replay; a stratified sample of jobs (every set, the largest and
  smallest buffers, random ones) against the oracle.
* Served: the 8 resident sets (32 to 2,048 regexes) behind the
  continuous-batching server on a Poisson schedule; sampled jobs against the
  oracle, and a few served again alone.
"""
import concurrent.futures
import threading
import time

import numpy as np
import pytest

from bitgen_amd import autotune, compiler, workloads
from bitgen_amd.bitstream import transpose_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

USUAL = autotune.options_of(autotune.USUAL)


def oracle_streams(groups, data, threads=16):
    """Whole-stream oracle of each group's pre-pass program, [G, W] words
    masked to the positions 0..n."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    basis = O.transpose(data, O.n_words_for(data.size))
    progs = [O.flatten(*compiler.group_programs(g)[:2]) for g in groups]
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        outs = list(ex.map(lambda p: O.run_records(p[0], p[1], basis), progs))
    return np.stack([O.mask_to_positions(o, data.size + 1) for o in outs])


def hip_streams(gpu, groups, data, options):
    """The plain form in the benchmark's launch shape: [G, W] words."""
    import torch

    ks = compiler.compile_pattern_set(groups, options).load(gpu.index or 0)
    x = torch.from_numpy(np.ascontiguousarray(data)).to(gpu).reshape(1, 1, -1)
    basis, _ = transpose_bytes(x, data.size)
    del x
    W = basis.size(-1)
    result = torch.empty((1, 1, ks.n_groups, W), dtype=torch.uint64, device=gpu)
    result.fill_(0x5A)  # stale garbage must be overwritten
    status = ks.alloc_status(1, W)
    ks.run(basis, result, status)
    torch.cuda.synchronize()
    ks.check_errors()
    out = result[0, 0].cpu().numpy()
    ks.free()
    return np.stack([O.mask_to_positions(r, data.size + 1) for r in out])


def assert_same(got, ref, label, groups):
    assert got.shape == ref.shape, (label, got.shape, ref.shape)
    for g in range(ref.shape[0]):
        if not np.array_equal(got[g], ref[g]):
            diff = np.nonzero(got[g] != ref[g])[0]
            raise AssertionError(f"{label}: group {g} differs in {diff.size} words (first {diff[0]}); "
                                 f"popcount hip {O.popcount(got[g])} oracle {O.popcount(ref[g])}; "
                                 f"{len(groups[g])} regexes")


@pytest.fixture(scope="module")
def ids_workload():
    regexes, n_groups, buf = workloads.ids()
    return compiler.group_regex_in_list(regexes, n_groups), buf


def test_ids_100MiB_benchmark_shape(gpu, ids_workload):
    groups, buf = ids_workload
    print(f"IDS 100 MiB: {len(groups)} groups, options {USUAL}", flush=True)
    got = hip_streams(gpu, groups, buf, USUAL)
    ref = oracle_streams(groups, buf)
    assert O.popcount(ref) > 1000
    assert_same(got, ref, "IDS 100 MiB", groups)


@pytest.fixture(scope="module")
def ids_sample(ids_workload):
    groups, buf = ids_workload
    sample = np.ascontiguousarray(buf[:16 * 2**20])
    return groups, sample, oracle_streams(groups, sample)


@pytest.mark.parametrize("cand", autotune.candidates(16), ids=lambda c: "gpk%d_mw%d_q%d_bg%d_bt%d" % c)
def test_ids_every_tuner_candidate(gpu, ids_sample, cand):
    groups, sample, ref = ids_sample
    print(f"IDS 16 MiB candidate {cand}", flush=True)
    got = hip_streams(gpu, groups, sample, autotune.options_of(cand))
    assert_same(got, ref, f"IDS candidate {cand}", groups)


@pytest.fixture(scope="module")
def genomic_sample():
    regexes, n_groups, buf = workloads.genomic(n_bytes=2 * 2**20)
    groups = compiler.group_regex_in_list(regexes, n_groups)
    return groups, buf, oracle_streams(groups, buf)


@pytest.mark.parametrize("cand", autotune.candidates(32), ids=lambda c: "gpk%d_mw%d_q%d_bg%d_bt%d" % c)
def test_genomic_every_tuner_candidate(gpu, genomic_sample, cand):
    groups, buf, ref = genomic_sample
    print(f"genomic 2 MiB candidate {cand}", flush=True)
    got = hip_streams(gpu, groups, buf, autotune.options_of(cand))
    assert O.popcount(ref) > 0
    assert_same(got, ref, f"genomic candidate {cand}", groups)


def test_batched_uneven_2GiB(gpu):
    import torch

    from bitgen_amd.batch import BatchPatternSet, BatchPlan

    sets, buffers, jobs = workloads.batched()
    set_groups = [compiler.group_regex_in_list(s, workloads.groups_for(len(s))) for s in sets]
    sizes = np.array([b.size for b in buffers])
    print(f"batched: {len(buffers)} buffers, {sizes.sum()} bytes, {len(jobs)} jobs, options {USUAL}", flush=True)
    bsets = [BatchPatternSet(g, USUAL).load(0) for g in set_groups]
    plan = BatchPlan(bsets, buffers, jobs)
    plan.capture()
    for _ in range(2):  # the second replay must give the same streams
        plan.result.fill_(0x5A)
        plan.run_graph()
        torch.cuda.synchronize()
        plan.check_errors()
    order = np.argsort(sizes)
    rng = np.random.default_rng(44)
    picks = {int(order[-1]), int(order[-2]), int(order[0])} | {int(b) for b in rng.choice(len(buffers), 6, replace=False)}
    checked = 0
    for si in range(len(sets)):
        for b in sorted(picks):
            got = plan.results(si, b).cpu().numpy()
            got = np.stack([O.mask_to_positions(r, sizes[b] + 1) for r in got])
            ref = oracle_streams(set_groups[si], buffers[b])
            assert_same(got, ref, f"batched set {si} buffer {b} ({sizes[b]} bytes)", set_groups[si])
            checked += O.popcount(ref)
    print(f"{len(picks) * len(sets)} jobs compared, {checked} match positions", flush=True)
    assert checked > 0
    plan.destroy()
    for s in bsets:
        s.free()


def test_served_real_sets(gpu):
    from bitgen_amd.batch import BatchPatternSet
    from bitgen_amd.serve import MatchServer

    set_groups = [compiler.group_regex_in_list(r, workloads.groups_for(len(r))) for r in workloads.served_sets()]
    # the server's kernels build their byte-pair tables themselves
    bsets = {i: BatchPatternSet(g, dict(USUAL, shared_filter=0)).load(0) for i, g in enumerate(set_groups)}
    arrivals, set_idx, bufs = workloads.serving(400, 5000.0, seed=23)
    print(f"served: 8 sets {[len(g) for g in set_groups]} groups, {len(bufs)} jobs, "
          f"{sum(b.size for b in bufs)} bytes", flush=True)
    srv = MatchServer(bsets).warm(bufs[0]).start()
    futs = [None] * len(bufs)
    t0 = time.perf_counter()

    def producer():
        for i, (a, s, b) in enumerate(zip(arrivals, set_idx, bufs)):
            delay = t0 + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futs[i] = srv.submit(int(s), b)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    results = [f.result(timeout=120) for f in futs]
    srv.stop()
    assert srv.batches > 1
    rng = np.random.default_rng(23)
    sizes = np.array([b.size for b in bufs])
    picks = sorted({int(np.argmax(sizes))} | {int(i) for i in rng.choice(len(bufs), 40, replace=False)})
    for i in picks:
        s = int(set_idx[i])
        ref = oracle_streams(set_groups[s], bufs[i])
        got = np.stack([O.mask_to_positions(r, sizes[i] + 1) for r in results[i]])
        assert_same(got, ref, f"served job {i} (set {s}, {sizes[i]} bytes)", set_groups[s])
    # the same jobs served alone give the same streams
    alone = MatchServer(bsets, n_streams=1, max_batch_jobs=1).start()
    for i in picks[:8]:
        r1 = alone.submit(int(set_idx[i]), bufs[i]).result(timeout=120)
        assert np.array_equal(O.mask_to_positions(r1, sizes[i] + 1), O.mask_to_positions(results[i], sizes[i] + 1))
    alone.stop()
    for s in bsets.values():
        s.free()
