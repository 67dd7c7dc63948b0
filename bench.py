"""Benchmark of the hip bitstream path on one MI355X.

Workload (the multi-pattern IDS configuration): 1,000 seeded regexes (600
literals of 4-32 bytes, 300 literal + class + bounded repeat, 100 with .* or
[^\\r\\n]*) balanced into 16 groups, over one 100 MiB buffer (70% printable
ASCII, 30% uniform bytes, planted hits at 1e-4 per byte).  One step = the
whole hot path over the resident buffer: byte->basis transpose, then the 16
fused group kernels (one launch per group), i.e. one (pattern set, buffer) job.

Output: one JSON line (see the keys below).  ``value`` = scanned GB/s =
buffer bytes per step / step time (inputs already in HBM).  ``roofline`` is
for the dominant kernels (the group kernels): algorithmic bytes per launch =
basis read (N bytes) + match stream write (ceil((N+1)/64)*8 bytes), divided
by the average group-kernel duration measured with HIP events on the launch
stream.  ``cpu_baseline`` times the oracle (the C executor of the same
pre-pass IR) on a bounded sample of the same workload on this host.

Two more measurements ride on the same line (the metric names them):

* ``batched``: the batched-uneven workload -- pattern sets of 64, 256 and
  1,024 regexes (groups of about 64) x 4,096 buffers of log-normal length
  (median 16 KiB, clipped to 1 KiB-8 MiB), every (set, buffer) pair a job --
  run as one graph-captured batch (one transpose launch plus one persistent
  launch per group, all buffers) and as plain per-job launches (one
  transpose plus one launch per group for every job); scanned GB/s of each
  and their ratio.
* ``served``: Poisson arrivals of jobs naming one of 8 resident pattern sets
  (32 to 2,048 regexes) with log-normal buffers (median 4 KiB, max 64 MiB),
  through the continuous-batching server and through synchronous per-job
  plain launches on the same arrival schedule; p50/p99 latency (arrival to
  result on the host) and GB/s of each.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mib", type=int, default=100, help="buffer size in MiB")
    ap.add_argument("--regexes", type=int, default=1000)
    ap.add_argument("--groups", type=int, default=16)
    ap.add_argument("--queues", type=int, default=0, help="tile queues per launch (0 = runtime default)")
    ap.add_argument("--grid-blocks", type=int, default=0, help="persistent blocks per launch (0 = all that fit)")
    ap.add_argument("--filter-grid-blocks", type=int, default=0,
                    help="blocks of the byte-pair filter pre-pass (0 = all that fit at once)")
    ap.add_argument("--take", type=int, default=0, help="tiles per queue take (0 = runtime default)")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="compiler option override (e.g. pass_short_circuit_interval=2); skips the autotuner")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape instead of the autotuner's")
    ap.add_argument("--tune-sample-mib", type=int, default=16, help="sample the autotuner times candidates on")
    ap.add_argument("--cpu-sample-mib", type=int, default=4,
                    help="prefix of the IDS buffer the CPU baseline scans (per run; 2 warm-ups + 3 timed runs, "
                         "at 1 thread and at all cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="check the GPU result against the oracle on a slice")
    ap.add_argument("--batched-buffers", type=int, default=4096)
    ap.add_argument("--batched-mib", type=int, default=2048, help="total bytes of the batched workload's buffers")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=10000, help="jobs per arrival rate")
    ap.add_argument("--served-rates", default="2000,5000,10000,20000,50000",
                    help="Poisson arrival rates (jobs/s), comma separated")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--genomic-mib", type=int, default=1024, help="DNA text of the genomic line (MiB)")
    ap.add_argument("--no-genomic", action="store_true")
    ap.add_argument("--prepare", action="store_true",
                    help="compile every code object a default run loads (no GPU needed) and exit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (three times) and exit: the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain device-to-device copy of known size: reads and writes exactly
    mib MiB per launch, so counter readings of it give the counter's scale."""
    import torch

    n = mib * 2**20
    src = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n}), flush=True)


def prepare(args):
    """Compile (hipcc only, no GPU) every code object a default run loads:
    each autotuner candidate of the IDS and genomic sets, and the batched and
    served sets in batch and plain form under the autotuner's usual choice
    (other choices compile when first loaded)."""
    import concurrent.futures

    from bitgen_amd import autotune, compiler, workloads
    from bitgen_amd.batch import BatchPatternSet

    t0 = time.perf_counter()
    ids_groups = compiler.group_regex_in_list(workloads.ids_regexes(args.regexes, 3), args.groups)
    gen_groups = compiler.group_regex_in_list(workloads.genomic_regexes(512, 5), 32)
    tasks = [(compiler.compile_pattern_set, g, autotune.options_of(c))
             for g in (ids_groups, gen_groups) for c in autotune.candidates(len(g))]
    # the batched and served sets take the IDS set's tuned shape: prepare the
    # usual one (another choice compiles when first loaded)
    opts = autotune.options_of(autotune.USUAL)
    for regs in workloads.batched_sets():
        g = compiler.group_regex_in_list(regs, workloads.groups_for(len(regs)))
        tasks += [(compiler.compile_pattern_set, g, opts), (BatchPatternSet, g, opts)]
    for regs in workloads.served_sets():
        # served sets build their byte-pair tables in the kernels
        g = compiler.group_regex_in_list(regs, workloads.groups_for(len(regs)))
        tasks += [(compiler.compile_pattern_set, g, opts), (BatchPatternSet, g, dict(opts, shared_filter=0))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=3) as ex:
        for f in [ex.submit(fn, g, o) for fn, g, o in tasks]:
            f.result()
    _log(f"prepared {len(tasks)} pattern-set forms ({time.perf_counter() - t0:.0f}s)")


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def _host_threads():
    try:
        return int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        return os.cpu_count() or 1


def cpu_baseline(groups, buf, sample_bytes):
    """The oracle (C executor of the same pre-pass IR, whole streams) over a
    prefix of the buffer, timed like the reference times its CPU path
    (scripts/run_icgrep.py:20-21: 2 warm-ups, then the median of 3 runs), at
    one thread and with one thread per group over all host threads (ctypes
    releases the GIL, so the groups run in parallel)."""
    import concurrent.futures

    import numpy as np

    from bitgen_amd import compiler
    from oracle import oracle as O

    sample = np.ascontiguousarray(buf[:sample_bytes])
    W = O.n_words_for(sample.size)
    basis = O.transpose(sample, W)
    progs = [O.flatten(*compiler.group_programs(g)[:2]) for g in groups]
    nproc = _host_threads()

    def timed(threads):
        def one(p):
            return O.run_records(p[0], p[1], basis)

        runs = []
        with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
            for r in range(5):
                t0 = time.perf_counter()
                list(ex.map(one, progs))
                if r >= 2:
                    runs.append(time.perf_counter() - t0)
        return float(np.median(runs))

    all_threads = min(nproc, len(progs))
    t1 = timed(1)
    tn = timed(all_threads)
    return {
        "value": round(sample.size / tn / 1e9, 6),
        "unit": "GB/s",
        "cores": all_threads,
        "kind": "port",
        "value_1_thread": round(sample.size / t1 / 1e9, 6),
        "host_threads": nproc,
        "cpu_model": _cpu_model(),
        "sample": f"{sample.size / 2**20:.0f} MiB prefix of the IDS buffer, all {len(groups)} groups; oracle C "
                  f"executor of the pre-pass IR; 2 warm-ups + median of 3 runs; {all_threads} threads "
                  f"(one per group) {tn:.3f} s per run, 1 thread {t1:.3f} s per run",
    }


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def genomic_measurement(args, steps=3):
    """Genomic motif scan: 512 IUPAC motifs (half gapped) in 32 groups over
    synthetic DNA (1 GiB by default), plain form, launch shape from the
    autotuner (on a sample of the text); whole-pass GB/s over the resident
    text."""
    import numpy as np
    import torch

    from bitgen_amd import compiler, workloads
    from bitgen_amd.bitstream import transpose_bytes

    t0 = time.perf_counter()
    n_bytes = args.genomic_mib * 2**20
    regexes, n_groups, buf = workloads.genomic(n_bytes=n_bytes)
    groups = compiler.group_regex_in_list(regexes, n_groups)
    options, tuning = {}, None
    if not args.no_tune:
        from bitgen_amd import autotune

        tuning = autotune.tune(groups, buf[:args.tune_sample_mib * 2**20])
        options = dict(tuning["options"])
    ks = compiler.compile_pattern_set(groups, options).load(0)
    if tuning:
        ks.n_queues = tuning["queues"]
    dev = torch.device("cuda", 0)
    staged = torch.from_numpy(buf).to(dev).reshape(1, 1, -1)
    del buf
    basis, staged = transpose_bytes(staged, n_bytes)
    W = basis.size(-1)
    result = torch.empty((1, 1, ks.n_groups, W), dtype=torch.uint64, device=dev)
    status = ks.alloc_status(1, W)
    _log(f"genomic workload ready: {len(regexes)} regexes, {ks.n_groups} groups, {n_bytes} bytes "
         f"({time.perf_counter() - t0:.1f}s)")
    stream = torch.cuda.current_stream()

    def step():
        transpose_bytes(staged, n_bytes, out=basis)
        ks.run(basis, result, status)

    step()
    torch.cuda.synchronize()
    ks.check_errors()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(steps)]
    for e in ev:
        e[0].record(stream)
        step()
        e[1].record(stream)
    torch.cuda.synchronize()
    ks.check_errors()
    ms = float(np.median([e[0].elapsed_time(e[1]) for e in ev]))
    out = {"workload": "genomic", "regexes": len(regexes), "groups": ks.n_groups, "kernels": ks.n_kernels,
           "text_bytes": n_bytes, "ms_per_pass": round(ms, 3), "gbs": round(n_bytes / (ms / 1e3) / 1e9, 3),
           "options": options or "default (one kernel per group)", "steps": steps,
           "autotune": [(c["groups_per_kernel"], c["min_waves"], c.get("bigram_filter", 0), c.get("block_threads", 256),
                         c["ms"]) for c in tuning["candidates"]] if tuning else None}
    _log(f"genomic: {out}")
    ks.free()
    del basis, staged, result, status
    torch.cuda.empty_cache()
    return out


def batched_measurement(args):
    """Batched-uneven workload: one graph-captured batch vs plain per-job
    launches over the same resident buffers (see the module docstring)."""
    import numpy as np
    import torch

    from bitgen_amd import compiler, workloads
    from bitgen_amd.backend.hip import runtime as rt
    from bitgen_amd.batch import BatchPatternSet, BatchPlan, words_for

    t0 = time.perf_counter()
    sets, buffers, jobs = workloads.batched(n_buffers=args.batched_buffers, total_bytes=args.batched_mib * 2**20)
    set_groups = [compiler.group_regex_in_list(s, workloads.groups_for(len(s))) for s in sets]
    _log(f"batched workload ready: {len(buffers)} buffers, {sum(b.size for b in buffers)} bytes, sets "
         f"{[len(s) for s in sets]} regexes in {[len(g) for g in set_groups]} groups ({time.perf_counter() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    opts = dict(getattr(args, "tuned_options", {}) or {})
    bsets = []
    for g in set_groups:
        bsets.append(BatchPatternSet(g, opts).load(0))
        _log(f"batched: set of {len(g)} groups loaded ({time.perf_counter() - t0:.1f}s)")
    plan = BatchPlan(bsets, buffers, jobs)
    scanned = plan.scanned_bytes()
    stream = torch.cuda.current_stream()
    plan.capture()
    for _ in range(2):
        plan.run_graph()
    torch.cuda.synchronize()
    plan.check_errors()
    reps = 5
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.run_graph()
    e1.record(stream)
    torch.cuda.synchronize()
    plan.check_errors()
    batched_ms = e0.elapsed_time(e1) / reps
    _log(f"batched: {batched_ms:.3f} ms per batch")
    # per-launch durations (eager, HIP events on the launch stream between
    # the launches the graph holds) for the roofline of the batch kernels:
    # algorithmic bytes of a (set, kernel) launch = its jobs' basis reads
    # (N bytes each) + their match streams (n_out * 8 W bytes each)
    sptr_b = rt.current_stream_ptr(stream)
    n_launch = len(plan.queues)
    # a set with shared byte-pair filters runs its filter pre-pass before its
    # first kernel, timed on its own (events 2 + n_launch + set)
    n_sets = len(plan.sets)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2 + n_launch + 2 * n_sets)] for _ in range(3)]
    for ev in evs:
        ev[0].record(stream)
        plan.enqueue_transpose(sptr_b)
        ev[1].record(stream)
        seen = set()
        for i, q in enumerate(plan.queues):
            ps = plan.sets[q["set"]]
            f = plan.filters.get(q["set"])
            if f is not None and q["set"] not in seen:
                seen.add(q["set"])
                fq = f["queue"]
                ev[2 + n_launch + 2 * q["set"]].record(stream)
                rt.filter_batch_launch(ps.filter_kernel, fq["jobs"].data_ptr(), fq["tile_job"].data_ptr(),
                                       fq["total_tiles"], f["buf"].data_ptr(), sptr_b)
                ev[3 + n_launch + 2 * q["set"]].record(stream)
            # the kernel's launch is timed from here (after any pre-pass)
            start = torch.cuda.Event(enable_timing=True)
            start.record(stream)
            if q["total_tiles"]:
                rt.batch_launch(ps.kernels[q["kernel"]], q["args"], plan._grid(ps, q["total_tiles"]),
                                ps.block_threads, sptr_b)
            ev[2 + i].record(stream)
            q.setdefault("_starts", []).append(start)
    torch.cuda.synchronize()
    plan.check_errors()
    launch_ms = [float(np.median([s.elapsed_time(ev[2 + i]) for s, ev in zip(q.pop("_starts"), evs)]))
                 for i, q in enumerate(plan.queues)]
    filter_ms = sum(float(np.median([ev[2 + n_launch + 2 * si].elapsed_time(ev[3 + n_launch + 2 * si])
                                     for ev in evs])) for si in plan.filters)
    alg = [sum(plan.lengths[e["buffer"]] + e["n_out"] * 8 * e["W"] for e in q["entries"]) for q in plan.queues]
    dom = int(np.argmax(launch_ms))
    batch_roofline = {
        "bound": "hbm",
        "achieved": round(sum(alg) / (sum(launch_ms) / 1e3) / 1e9, 3),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(sum(alg) / (sum(launch_ms) / 1e3) / 1e9 / HBM_PEAK_GBS, 6),
        "traffic": None,
        "kernels": "all batch launches of the replay (sum of bytes / sum of durations)",
        "dominant": {"set": plan.queues[dom]["set"], "kernel": plan.queues[dom]["kernel"],
                     "ms": round(launch_ms[dom], 4), "alg_bytes": int(alg[dom]),
                     "frac": round(alg[dom] / (launch_ms[dom] / 1e3) / 1e9 / HBM_PEAK_GBS, 6)},
        "transpose_ms": round(float(np.median([ev[0].elapsed_time(ev[1]) for ev in evs])), 4),
        # the byte-pair filter pre-passes of the sets that share them (not
        # part of the kernels' bytes or durations above)
        "filter_ms": round(filter_ms, 4),
    }

    # plain form: per (set, buffer) job one transpose and one launch per group,
    # inputs resident on the device (16-byte row stride, as the transpose needs)
    psets = []
    for g in set_groups:
        psets.append(compiler.compile_pattern_set(g, opts).load(0))
        psets[-1].n_queues = args.queues
    _log("batched: plain form loaded")
    sptr = rt.current_stream_ptr(stream)
    staged, basis, results, status = [], [], {}, {}
    for b in buffers:
        n = int(b.size)
        stride = (n + 15) // 16 * 16
        x = torch.zeros(max(stride, 16), dtype=torch.uint8, device=dev)
        x[:n] = torch.from_numpy(b).to(dev)
        staged.append(x)
        basis.append(torch.empty((8, words_for(n)), dtype=torch.uint64, device=dev))
    for (si, bi) in jobs:
        W = words_for(int(buffers[bi].size))
        results[(si, bi)] = torch.empty((psets[si].n_groups, W), dtype=torch.uint64, device=dev)
        status[(si, bi)] = psets[si].alloc_status(1, W)

    def plain_pass():
        for (si, bi) in jobs:
            ks = psets[si]
            n = int(buffers[bi].size)
            W = words_for(n)
            rt.transpose(staged[bi].data_ptr(), n, staged[bi].numel(), basis[bi].data_ptr(), W, 1, sptr)
            r, st = results[(si, bi)], status[(si, bi)]
            ks.launch_filter(basis[bi].data_ptr(), W, sptr)
            for k in range(ks.n_kernels):
                ks.launch(k, basis[bi].data_ptr(), r.data_ptr() + ks.first_group[k] * W * 8,
                          st.data_ptr() + k * st.size(2) * 4, W, sptr)

    plain_pass()
    torch.cuda.synchronize()
    e0.record(stream)
    plain_pass()
    e1.record(stream)
    torch.cuda.synchronize()
    for ks in psets:
        ks.check_errors()
    plain_ms = e0.elapsed_time(e1)
    _log(f"plain per-job: {plain_ms:.3f} ms for {len(jobs)} jobs")
    # the two forms agree bit for bit on a sample of jobs
    rng = np.random.default_rng(0)
    for k in rng.choice(len(jobs), size=min(64, len(jobs)), replace=False):
        si, bi = jobs[int(k)]
        if not torch.equal(plan.results(si, bi), results[(si, bi)]):
            raise AssertionError(f"batched and plain results differ for set {si} buffer {bi}")
    plan.destroy()
    return {
        "workload": "batched_uneven",
        "sets_regexes": [len(s) for s in sets],
        "sets_groups": [len(g) for g in set_groups],
        "buffers": len(buffers),
        "jobs": len(jobs),
        "scanned_bytes": int(scanned),
        "batched_ms": round(batched_ms, 4),
        "batched_gbs": round(scanned / (batched_ms / 1e3) / 1e9, 3),
        "batched_pct_hbm_peak": round(100 * scanned / (batched_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "plain_ms": round(plain_ms, 4),
        "plain_gbs": round(scanned / (plain_ms / 1e3) / 1e9, 3),
        "plain_launches": len(jobs) + sum(psets[si].n_kernels + (psets[si].filter_kernel is not None)
                                          for si, _ in jobs),
        "launch_shape": {k: opts.get(k, compiler.DEFAULT_OPTIONS[k])
                         for k in ("groups_per_kernel", "min_waves", "bigram_filter", "block_threads")},
        "speedup": round(plain_ms / batched_ms, 3),
        "roofline": batch_roofline,
        "form": "batched = one graph replay (1 transpose + 1 persistent launch per group); "
                "plain = 1 transpose + 1 launch per group per job",
    }


def served_measurement(args):
    """Continuous-batching server vs synchronous per-job plain launches on
    the same Poisson arrival schedules (see the module docstring)."""
    import threading

    import numpy as np
    import torch

    from bitgen_amd import compiler, workloads
    from bitgen_amd.batch import BatchPatternSet
    from bitgen_amd.serve import MatchServer, run_plain_job

    sizes = (32, 64, 128, 256, 512, 1024, 1536, 2048)
    t0 = time.perf_counter()
    set_groups = [compiler.group_regex_in_list(workloads.ids_regexes(n, 100 + i), workloads.groups_for(n))
                  for i, n in enumerate(sizes)]
    opts = dict(getattr(args, "tuned_options", {}) or {})
    bsets, psets = {}, {}
    for i, g in enumerate(set_groups):
        # the server's native staging launches kernels that build their own
        # byte-pair tables
        bsets[i] = BatchPatternSet(g, dict(opts, shared_filter=0)).load(0)
        psets[i] = compiler.compile_pattern_set(g, opts).load(0)
        _log(f"served: set {i} ({len(g)} groups) loaded ({time.perf_counter() - t0:.1f}s)")
    _log(f"served pattern sets loaded: {[len(g) for g in set_groups]} groups ({time.perf_counter() - t0:.1f}s)")
    dev = torch.device("cuda", 0)
    out = {"workload": "served", "sets_regexes": list(sizes), "sets_groups": [len(g) for g in set_groups],
           "hbm_cap_bytes": 200 * 10**9, "rates": []}
    for rate in [float(r) for r in args.served_rates.split(",") if r]:
        arrivals, set_idx, bufs = workloads.serving(args.served_jobs, rate, seed=11 + int(rate))
        nbytes = int(sum(b.size for b in bufs))
        # first launches of every set's kernels on every slot's stream are
        # paid before the clock starts
        srv = MatchServer(bsets, hbm_cap_bytes=200 * 10**9).warm(bufs[0]).start()
        futs = [None] * len(bufs)

        def producer():
            t_start = time.perf_counter()
            for i, (a, s, b) in enumerate(zip(arrivals, set_idx, bufs)):
                delay = t_start + a - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                futs[i] = srv.submit(int(s), b)

        th = threading.Thread(target=producer)
        t_first = time.perf_counter()
        th.start()
        th.join()
        for f in futs:
            f.result(timeout=600)
        srv.stop()
        lat = srv.latencies()
        t_last = max(j.done for j in srv.completed)
        served_gbs = nbytes / (t_last - t_first) / 1e9

        # per-job plain launches, FIFO on one host thread, same arrivals
        lat_plain = np.empty(len(bufs))
        for i in psets:
            run_plain_job(psets[i], bufs[0], dev)
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for i, (a, s, b) in enumerate(zip(arrivals, set_idx, bufs)):
            now = time.perf_counter()
            if now < t_start + a:
                time.sleep(t_start + a - now)
            run_plain_job(psets[int(s)], b, dev)
            lat_plain[i] = time.perf_counter() - (t_start + a)
        plain_gbs = nbytes / (time.perf_counter() - t_start) / 1e9
        blog = np.array(srv.batch_log) if srv.batch_log else np.zeros((1, 3))
        row = {
            "rate_per_s": rate,
            "jobs": len(bufs),
            "bytes": nbytes,
            "served_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "served_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "served_max_ms": round(float(lat.max()) * 1e3, 3),
            "served_gbs": round(served_gbs, 4),
            "served_batches": srv.batches,
            "served_jobs_per_batch_mean": round(float(blog[:, 0].mean()), 2),
            "served_host_stage_ms_p50_p99": [round(float(np.percentile(blog[:, 1], q)) * 1e3, 3) for q in (50, 99)],
            "served_batch_device_ms_p50_p99": [round(float(np.percentile(blog[:, 2], q)) * 1e3, 3) for q in (50, 99)],
            "plain_p50_ms": round(float(np.percentile(lat_plain, 50)) * 1e3, 3),
            "plain_p99_ms": round(float(np.percentile(lat_plain, 99)) * 1e3, 3),
            "plain_gbs": round(plain_gbs, 4),
        }
        _log(f"served at {rate:.0f}/s: {row}")
        out["rates"].append(row)
    for ks in psets.values():
        ks.free()
    for s in bsets.values():
        s.free()
    return out


# MI355X VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD
# every 4 cycles at 2.4 GHz (giga wave-instructions per second)
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4


def profiled_counters(args, options):
    """Counters of the group kernels measured by rocprofv3 for the default
    workload, from the committed counter summary: {"traffic": HBM bytes per
    launch (FETCH_SIZE and WRITE_SIZE scaled by a plain copy kernel),
    "valu": [SQ_INSTS_VALU per launch, per kernel]}; None for other workloads
    or another launch shape than the one profiled."""
    if (args.mib, args.regexes, args.groups) != (100, 1000, 16):
        return None
    path = os.path.join(ROOT, "profiles", "bg_group_ids_100MiB_counters.md")
    if not os.path.exists(path):
        return None
    found, shape = {}, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            found["traffic"] = int(float(line.split("=")[1]))
        elif line.startswith("valu_insts_per_launch"):
            found["valu"] = json.loads(line.split("=", 1)[1])
        elif line.startswith("salu_insts_per_launch"):
            found["salu"] = json.loads(line.split("=", 1)[1])
        elif line.startswith("shape ="):
            shape = dict(kv.split("=") for kv in line.split("=", 1)[1].split())
    keys = ("groups_per_kernel", "min_waves", "bigram_filter", "block_threads", "carry_settle", "shared_filter")
    want = {k: str(options.get(k)) for k in keys}
    return found if shape == want else None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; sharding over several GPUs is designed "
                                   "(by group set or input range) but not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.prepare:
        prepare(args)
        return
    import numpy as np
    import torch

    from bitgen_amd import compiler, workloads
    from bitgen_amd.bitstream import transpose_bytes

    t_setup = time.perf_counter()
    n_bytes = args.mib * 2**20
    regexes, n_groups, buf = workloads.ids(n_bytes=n_bytes, n_regex=args.regexes, n_groups=args.groups)
    groups = compiler.group_regex_in_list(regexes, n_groups)
    print(f"[bench] workload ready: {len(regexes)} regexes, {len(groups)} groups, {n_bytes} bytes "
          f"({time.perf_counter() - t_setup:.1f}s)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    options = {}
    for kv in args.option:
        k, v = kv.split("=", 1)
        options[k] = int(v) if v.lstrip("-").isdigit() else v
    tuning = None
    if not options and not args.no_tune:
        from bitgen_amd import autotune

        tuning = autotune.tune(groups, buf[:args.tune_sample_mib * 2**20])
        options = dict(tuning["options"])
        if not args.queues:
            args.queues = tuning["queues"]
        _log(f"autotuner: {tuning['options']} queues {tuning['queues']} "
             f"({[(c['groups_per_kernel'], c['min_waves'], c.get('bigram_filter', 0), c.get('block_threads', 256), c['ms']) for c in tuning['candidates']]})")
    args.tuned_options = options
    cs = compiler.compile_pattern_set(groups, options)
    print(f"[bench] compiled {len(groups)} groups into {len(cs.metas)} kernels ({time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    dev = torch.device("cuda", 0)
    ks = cs.load(0)
    ks.n_queues = args.queues
    ks.grid_blocks = args.grid_blocks
    ks.filter_grid_blocks = args.filter_grid_blocks
    ks.take = args.take
    from bitgen_amd.backend.hip import runtime as rt_

    occ = [rt_.kernel_occupancy(k, 256)[0] for k in ks.kernels]
    print(f"[bench] blocks per CU (256 threads) per group kernel: {occ}; info {ks.info()[0]}", file=sys.stderr,
          flush=True)
    x = torch.from_numpy(buf).to(dev).reshape(1, 1, -1)
    basis, staged = transpose_bytes(x, n_bytes)
    W = basis.size(-1)
    result = torch.zeros((1, 1, ks.n_groups, W), dtype=torch.uint64, device=dev)
    status = ks.alloc_status(1, W)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    K = ks.n_kernels

    from bitgen_amd.backend.hip import runtime as rt

    sptr = rt.current_stream_ptr(stream)

    def step(events=None):
        if events is not None:
            events[0].record(stream)
        transpose_bytes(staged, n_bytes, out=basis)
        if events is not None:
            events[1].record(stream)
        ks.launch_filter(basis.data_ptr(), W, sptr)  # byte-pair answers of a set with shared filters
        if events is not None:
            events[2].record(stream)
        for k in range(K):
            ks.launch(k, basis.data_ptr(), result.data_ptr() + ks.first_group[k] * W * 8,
                      status.data_ptr() + k * status.size(2) * 4, W, sptr)
            if events is not None:
                events[3 + k].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ks.check_errors()
    print("[bench] warm-up done", file=sys.stderr, flush=True)

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3 + K)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for s in range(args.steps):
        step(ev[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ks.check_errors()

    trans_ms = [e[0].elapsed_time(e[1]) for e in ev]
    filter_ms = [e[1].elapsed_time(e[2]) for e in ev]
    kernel_ms = [[ev[s][2 + k].elapsed_time(ev[s][3 + k]) for k in range(K)] for s in range(args.steps)]
    step_ms = [e[0].elapsed_time(e[2 + K]) for e in ev]
    ms_per_step = float(np.median(step_ms))
    value = n_bytes / (ms_per_step / 1e3) / 1e9
    # algorithmic bytes of a kernel launch: basis read (N bytes) + its groups'
    # match streams written (8 W bytes each)
    alg_bytes = [n_bytes + m.get("n_out", 1) * W * 8 for m in ks.metas]
    mean_kernels_s = float(np.mean([sum(r) for r in kernel_ms])) / 1e3
    achieved = sum(alg_bytes) / mean_kernels_s / 1e9
    counters = profiled_counters(args, cs.options)
    per_kernel = [float(np.median([kernel_ms[s][k] for s in range(args.steps)])) for k in range(K)]

    count = None
    if args.check:
        from bitgen_amd.bitstream import Bitstream

        count = Bitstream(result.cpu().numpy(), n_positions=n_bytes + 1).get_count()

    out = {
        "metric": "scanned GB/s (% HBM peak) batched vs plain; served p50/p99 latency",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "ids_plain",
            "regexes": len(regexes),
            "groups": ks.n_groups,
            "kernels": K,
            "groups_per_kernel": cs.options["groups_per_kernel"],
            "buffer_bytes": n_bytes,
            "form": "plain (one transpose + one launch per kernel for the buffer)",
            "parallelism": "single GPU",
        },
        "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 4),
        "transpose_ms": round(float(np.median(trans_ms)), 4),
        # the set's byte-pair filter pre-pass (kernels with shared filters;
        # 0 otherwise): part of the step, not of the group kernels' roofline
        "filter_ms": round(float(np.median(filter_ms)), 4) if ks.filter_kernel is not None else 0.0,
        "kernel_ms": [round(v, 4) for v in per_kernel],
        "wall_s": round(wall, 3),
        "autotune": ({"chosen": {"groups_per_kernel": cs.options["groups_per_kernel"],
                                 "min_waves": cs.options["min_waves"], "queues": args.queues,
                                 "shared_filter": cs.options.get("shared_filter", 0)},
                      "sample_bytes": tuning["sample_bytes"], "candidates": tuning["candidates"]}
                     if tuning else None),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": counters.get("traffic") if counters else None,
        },
        # the same launches against the VALU issue peak: SQ_INSTS_VALU per
        # launch (committed counter profile of this launch shape) / the
        # launches' measured time
        "roofline_valu": ({
            "bound": "valu",
            "achieved": round(sum(counters["valu"]) / 1e9 / mean_kernels_s, 3),
            "peak": VALU_PEAK_GINST,
            "unit": "G wave-instructions/s",
            "frac": round(sum(counters["valu"]) / 1e9 / mean_kernels_s / VALU_PEAK_GINST, 6),
        } if counters and "valu" in counters and len(counters["valu"]) == K else None),
        # scalar instructions: one scalar unit per CU issues at most one per
        # cycle (the same 614 G/s); with the byte-pair tables shared, the
        # group kernels issue more SALU than VALU instructions
        "roofline_salu": ({
            "bound": "salu",
            "achieved": round(sum(counters["salu"]) / 1e9 / mean_kernels_s, 3),
            "peak": VALU_PEAK_GINST,
            "unit": "G scalar instructions/s",
            "frac": round(sum(counters["salu"]) / 1e9 / mean_kernels_s / VALU_PEAK_GINST, 6),
        } if counters and "salu" in counters and len(counters["salu"]) == K else None),
    }
    if count is not None:
        out["match_count"] = count
    ks.free()
    if not args.no_genomic:
        out["genomic"] = genomic_measurement(args)
    if not args.no_batched:
        out["batched"] = batched_measurement(args)
    if not args.no_served:
        out["served"] = served_measurement(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(groups, buf, args.cpu_sample_mib * 2**20)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
