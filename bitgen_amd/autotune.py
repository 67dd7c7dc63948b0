"""Autotuner of the plain and batched group kernels, cached per pattern set.

For one pattern set (its regex groups) the tuner compiles each candidate
launch shape, times it on the GPU over a sample buffer, and keeps the
fastest.  The candidates cover the knobs that decide how the hot kernel maps
onto gfx950:

* ``groups_per_kernel`` -- groups merged into one program: one launch then
  shares the basis loads, the character-class header, the tile queue and the
  carry look-back of several groups (fewer, longer launches; more registers);
* ``min_waves`` -- waves per SIMD the register allocator must leave room for
  (more waves hide more latency; past the program's natural register count
  values spill to scratch);
* ``queues`` -- tile-queue counters the launch spreads its tile takes over;
* ``bigram_filter`` -- per-tile byte-pair filters in front of the literal
  first stages of the regexes (a tile whose bytes hold none of a stage's
  pairs skips that regex without evaluating its character classes); 1:
  each kernel builds the tile's pair table, 2: one filter pre-pass per set
  builds it once and answers every kernel's filters (``shared_filter``);
* ``block_threads`` -- waves per block (x 64 threads): the bytes a block
  scans per round of takes, and the granularity the CUs are filled in.

The result is cached under ``cache_dir/tune/<set digest>.json`` together
with every candidate's time, keyed by the pattern set, the sample size, the
device name and the code digest of the compiler, so a later call with the
same set returns at once.  Timing needs a GPU; compiling every candidate
beforehand (``prepare``) needs only hipcc.
"""
import hashlib
import json
import os
import time

from . import config as cfg
from . import compiler

# candidate launch shapes: (groups_per_kernel, min_waves, queues,
# bigram_filter (2 = shared by the set's kernels), block_threads).  A tile is one wave whatever the block size,
# so block_threads sets the waves per block (and the bytes a block scans per
# round: 4 KiB per wave); words per lane stay 1 (a 64-bit word per lane is the
# wave64 shift and carry unit).
# Shared filters are listed before per-kernel tables: on the 16 MiB sample
# the pre-pass's fixed cost narrows their lead to about 2 % (a tie), while on
# a 100 MiB buffer they are 8 % faster (IDS on MI355X), so a tie goes to them.
CANDIDATES = [
    (1, 0, 64, 0, 256),
    (2, 5, 64, 0, 256),
    (4, 5, 64, 0, 256),
    (4, 0, 64, 0, 256),
    (8, 5, 64, 0, 256),
    (4, 5, 64, 2, 256),
    (8, 5, 64, 2, 256),
    (1, 0, 64, 1, 256),
    (4, 5, 64, 1, 256),
    (4, 5, 64, 1, 128),
    (4, 5, 64, 1, 512),
]


# the shape chosen for the IDS set on MI355X so far; compiled ahead of time
# for the sets that reuse the IDS choice
USUAL = (4, 5, 64, 2, 256)

# candidates timed within this fraction of the fastest are ties (the spread
# of repeated timings on MI355X is 1-3%); a tie goes to the candidate listed
# first, so the choice does not flip between runs on noise
TIE = 0.02


def candidates(n_groups):
    """Distinct candidates for a set of n_groups groups (a merge wider than
    the set is the same kernel as merging all of it)."""
    seen, out = set(), []
    for gpk, mw, q, bg, bt in CANDIDATES:
        cand = (max(1, min(gpk, n_groups)), mw, q, bg, bt)
        if cand not in seen:
            seen.add(cand)
            out.append(cand)
    return out


def options_of(candidate, base=None):
    gpk, mw, _q, bg, bt = candidate
    o = dict(base or {})
    o.update(groups_per_kernel=gpk, min_waves=mw, bigram_filter=int(bool(bg)), block_threads=bt,
             shared_filter=int(bg == 2))
    return o


def _key(groups, base_options, sample_bytes, device_name):
    h = hashlib.sha256(json.dumps([groups, sorted((base_options or {}).items()), sample_bytes, device_name,
                                   compiler.code_digest()]).encode())
    return h.hexdigest()[:24]


def _path(key):
    root = cfg.get_config("cache_dir", ".cache_bitgen")
    d = os.path.join(root, "tune")
    os.makedirs(d, exist_ok=True)
    return os.path.join(d, f"{key}.json")


def pick(rows):
    """The row (candidate timings in CANDIDATES order) the tuner chooses:
    the first within TIE of the fastest."""
    fastest = min(r["ms"] for r in rows)
    return next(r for r in rows if r["ms"] <= fastest * (1 + TIE))


def prepare(groups, base_options=None):
    """Compile every candidate of a set (CPU only: hipcc)."""
    for cand in candidates(len(groups)):
        compiler.compile_pattern_set(groups, options_of(cand, base_options))


def tune(groups, sample, base_options=None, device=0, reps=5, use_cache=True):
    """Best candidate for ``groups`` on ``sample`` (uint8 array) on the GPU:
    the first candidate (in CANDIDATES order) timed within TIE of the
    fastest.

    Returns a dict: "options" (compile options of the winner), "queues",
    "ms" (its time per pass over the sample) and "candidates" (every
    candidate's shape and time)."""
    import numpy as np
    import torch

    from .backend.hip import runtime as rt
    from .bitstream import transpose_bytes

    groups = [list(g) for g in groups]
    dev = torch.device("cuda", device)
    name = torch.cuda.get_device_name(dev)
    key = _key(groups, base_options, int(np.asarray(sample).size), name)
    path = _path(key)
    if use_cache and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    data = np.ascontiguousarray(sample, dtype=np.uint8)
    x = torch.from_numpy(data).to(dev).reshape(1, 1, -1)
    basis, _ = transpose_bytes(x, data.size)
    W = basis.size(-1)
    stream = torch.cuda.current_stream(dev)
    sptr = rt.current_stream_ptr(stream)
    rows = []
    for cand in candidates(len(groups)):
        t0 = time.perf_counter()
        ks = compiler.compile_pattern_set(groups, options_of(cand, base_options)).load(device)
        ks.n_queues = cand[2]
        result = torch.empty((1, 1, ks.n_groups, W), dtype=torch.uint64, device=dev)
        status = ks.alloc_status(1, W)

        def one_pass():
            ks.launch_filter(basis.data_ptr(), W, sptr)
            for k in range(ks.n_kernels):
                ks.launch(k, basis.data_ptr(), result.data_ptr() + ks.first_group[k] * W * 8,
                          status.data_ptr() + k * status.size(2) * 4, W, sptr)

        one_pass()
        torch.cuda.synchronize(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            one_pass()
        e1.record(stream)
        torch.cuda.synchronize(dev)
        ks.check_errors()
        ms = e0.elapsed_time(e1) / reps
        info = ks.info()
        rows.append({"groups_per_kernel": cand[0], "min_waves": cand[1], "queues": cand[2],
                     "bigram_filter": cand[3], "block_threads": cand[4], "ms": round(ms, 4),
                     "kernels": ks.n_kernels, "vgprs": max(i["vgprs"] for i in info),
                     "scratch_bytes": max(i["scratch_bytes"] for i in info),
                     "setup_s": round(time.perf_counter() - t0, 2)})
        ks.free()
    best = pick(rows)
    out = {
        "options": options_of((best["groups_per_kernel"], best["min_waves"], best["queues"],
                                best["bigram_filter"], best["block_threads"]), base_options),
        "queues": best["queues"],
        "ms": best["ms"],
        "sample_bytes": int(data.size),
        "device": name,
        "candidates": rows,
    }
    with open(path + ".tmp", "w") as f:
        json.dump(out, f)
    os.replace(path + ".tmp", path)
    return out
