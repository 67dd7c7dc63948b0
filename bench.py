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
"""
import argparse
import json
import os
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
    ap.add_argument("--take", type=int, default=0, help="tiles per queue take (0 = runtime default)")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="compiler option override (e.g. pass_short_circuit_interval=2)")
    ap.add_argument("--cpu-sample-mib", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="check the GPU result against the oracle on a slice")
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


def cpu_baseline(groups, buf, sample_bytes):
    """Oracle (pre-pass IR, whole-stream C executor) over a prefix sample,
    one thread per group (ctypes releases the GIL)."""
    import concurrent.futures

    from bitgen_amd import compiler
    from oracle import oracle as O

    sample = buf[:sample_bytes]
    W = O.n_words_for(sample.size)
    basis = O.transpose(sample, W)
    progs = [O.flatten(*compiler.group_programs(g)[:2]) for g in groups]
    threads = min(len(progs), 16, os.cpu_count() or 1)

    def one(p):
        return O.run_records(p[0], p[1], basis)

    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, progs[:1]))  # warm-up
        t0 = time.perf_counter()
        list(ex.map(one, progs))
        dt = time.perf_counter() - t0
    return {
        "value": round(sample.size / dt / 1e9, 6),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample.size / 2**20:.0f} MiB prefix of the same buffer, all {len(groups)} groups, "
                  f"oracle C executor, one thread per group ({dt:.2f} s wall)",
    }


def profiled_traffic(args):
    """HBM bytes per group-kernel launch measured by rocprofv3 (FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy kernel) for the default workload, from
    the committed counter summary; None for other workloads."""
    if (args.mib, args.regexes, args.groups) != (100, 1000, 16):
        return None
    path = os.path.join(ROOT, "profiles", "bg_group_ids_100MiB_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return int(float(line.split("=")[1]))
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; multi-GPU sharding is described in DESIGN.md "
                                   "but not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
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
    cs = compiler.compile_pattern_set(groups, options)
    print(f"[bench] compiled {len(groups)} group kernels ({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    dev = torch.device("cuda", 0)
    ks = cs.load(0)
    ks.n_queues = args.queues
    ks.grid_blocks = args.grid_blocks
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
    G = ks.n_groups

    from bitgen_amd.backend.hip import runtime as rt

    sptr = rt.current_stream_ptr(stream)

    def step(events=None):
        if events is not None:
            events[0].record(stream)
        transpose_bytes(staged, n_bytes, out=basis)
        if events is not None:
            events[1].record(stream)
        for g in range(G):
            ks.launch(g, basis.data_ptr(), result.data_ptr() + g * W * 8,
                      status.data_ptr() + g * status.size(2) * 4, W, sptr)
            if events is not None:
                events[2 + g].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ks.check_errors()
    print("[bench] warm-up done", file=sys.stderr, flush=True)

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2 + G)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for s in range(args.steps):
        step(ev[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ks.check_errors()

    trans_ms = [e[0].elapsed_time(e[1]) for e in ev]
    group_ms = [[ev[s][1 + g].elapsed_time(ev[s][2 + g]) for g in range(G)] for s in range(args.steps)]
    step_ms = [e[0].elapsed_time(e[1 + G]) for e in ev]
    ms_per_step = float(np.median(step_ms))
    value = n_bytes / (ms_per_step / 1e3) / 1e9
    mean_group_ms = float(np.mean(group_ms))
    alg_bytes_launch = n_bytes + W * 8  # basis read + match stream write
    achieved = alg_bytes_launch / (mean_group_ms / 1e3) / 1e9
    per_group = [float(np.median([group_ms[s][g] for s in range(args.steps)])) for g in range(G)]

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
            "groups": G,
            "buffer_bytes": n_bytes,
            "form": "plain (one launch per group per buffer)",
            "parallelism": "single GPU",
        },
        "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 4),
        "transpose_ms": round(float(np.median(trans_ms)), 4),
        "group_kernel_ms": [round(v, 4) for v in per_group],
        "wall_s": round(wall, 3),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": profiled_traffic(args),
        },
    }
    if count is not None:
        out["match_count"] = count
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(groups, buf, args.cpu_sample_mib * 2**20)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
