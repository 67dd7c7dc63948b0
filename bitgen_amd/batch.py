"""Batched execution: many regex groups x many buffers of uneven length.

The reference launches one grid of (groups x inputs) blocks per pattern set
over copies of one input (``gen_cuda.py:441-473``, ``run_bitgen.py:239-245``);
uneven buffers need one launch each.  Here all buffers of a batch are
transposed by one launch (a descriptor per buffer) and every (kernel, buffer)
job of a pattern set runs in one persistent launch of the set's batch kernel,
whose waves take 4,096-position tiles from a device work queue
(``tile_job``): one persistent launch per kernel of each set covers all
buffers (a kernel computes one group, or several groups merged into one
program).  The sequence (one transpose launch + one batch launch per kernel)
can be captured into a graph and replayed.

Layout (all device-resident, allocated once per plan):
  bytes   u8  [sum of 64-byte-aligned buffer lengths]
  basis   u64 [sum over buffers of 8 * W_b]
  result  u64 [sum over jobs of n_out * W_b]  (per job: its groups' streams)
  status  u32 [sum over jobs of T_j * C_k]
  jobs    bg_job_t[n_jobs], tile_job u32[total tiles]    (per set and kernel)
"""
import ctypes
import os

import numpy as np

from .backend.hip import runtime as rt
from .codegen.gen_hip import FILTER_BATCH_KERNEL, batch_kernel_name, batch_source, compile_sources, filter_source
from .compiler import (DEFAULT_OPTIONS, compile_pattern_set, kernel_meta, kernel_programs, load_cached_set,
                       store_cached_set, _set_dir)

WAVE_WORDS = 64


def words_for(n):
    return (n + 1 + 63) // 64


def n_tiles(meta, W):
    s0 = 64 - meta["HR"]
    if W <= s0:
        return 1
    return 1 + (W - s0 + meta["S"] - 1) // meta["S"]


class BatchPatternSet:
    """The groups of one pattern set compiled into persistent batch kernels
    (one kernel per ``groups_per_kernel`` consecutive groups, merged into one
    program; separate code objects, compiled in parallel).

    A group whose program needs whole streams between its steps (general
    while loops, graph breaks: meta kind "stream") has no batch kernel: a
    BatchPlan runs its plan job by job through the plain form of the set
    (``plain_set``), eagerly -- each while loop's condition is decided on
    the host -- so a plan holding such a set is not captured in a graph."""

    def __init__(self, groups, options=None, block_threads=256):
        o = dict(DEFAULT_OPTIONS)
        o.update(options or {})
        self.groups = [list(g) for g in groups]
        self.options = o
        # the set's block size (compile option) unless the caller names one
        block_threads = int(o.get("block_threads") or block_threads)
        self.block_threads = block_threads
        self.modules = None
        self.kernels = None
        self.filter_module = None
        self.filter_kernel = None
        self.plain = None       # the plain form (sets with stream-level groups)
        self.plain_set = None
        d = _set_dir(self.groups, dict(o, form="batch", block=block_threads))
        cached = load_cached_set(d, "batch.json")
        if cached is not None:
            self.metas = cached
            self.hsacos = [None if m.get("kind") == "stream" else m["hsaco"] for m in cached]
            if self.has_stream_groups:
                self.plain = compile_pattern_set(self.groups, o)
            return
        programs = kernel_programs(self.groups, o)
        self.metas = [kernel_meta(p) for p in programs]
        if self.has_stream_groups:
            self.plain = compile_pattern_set(self.groups, o)
        stream_kernels = {k for k, m in enumerate(self.metas) if m.get("kind") == "stream"}
        programs = [p for k, p in enumerate(programs) if k not in stream_kernels]
        sources = {}
        for p in programs:
            path = os.path.join(d, f"batch_{p.gid}.hip")
            src = batch_source(p, block_threads)
            with open(path, "w") as f:
                f.write(src)
            sources[path] = src
        # shared byte-pair answers: the set's filter pre-pass (its batched kernel)
        row = max((p.filter_row for p in programs if getattr(p, "shared_filter", False)), default=0)
        fpath = os.path.join(d, "filter.hip")
        if row:
            src = filter_source(programs)
            with open(fpath, "w") as f:
                f.write(src)
            sources[fpath] = src
        hs = compile_sources(sources)
        self.hsacos = []
        for k, m in enumerate(self.metas):
            if k in stream_kernels:
                # the group's plan runs through the plain form; its meta keeps
                # the plain code object (so the cache entry stays checkable)
                m["hsaco"] = self.plain.metas[k]["hsaco"]
                self.hsacos.append(None)
                continue
            h = os.path.abspath(hs[os.path.join(d, f"batch_{m['gid']}.hip")])
            m["hsaco"] = h
            self.hsacos.append(h)
            if row:
                m["filter"] = {"hsaco": os.path.abspath(hs[fpath]), "kernel": FILTER_BATCH_KERNEL, "row": row}
        store_cached_set(d, "batch.json", self.metas)

    @property
    def n_groups(self):
        return len(self.groups)

    @property
    def n_kernels(self):
        return len(self.metas)

    def load(self, device=0):
        rt.init(device)
        if self.modules is None:
            self.modules = [rt.Module(h) if h else None for h in self.hsacos]
            self.kernels = [m.kernel(batch_kernel_name(k)) if m is not None else None
                            for k, m in enumerate(self.modules)]
            if self.plain is not None:
                self.plain_set = self.plain.load(device)
            f = next((m["filter"] for m in self.metas if m.get("filter")), None)
            if f is not None:
                self.filter_module = rt.Module(f["hsaco"])
                self.filter_kernel = self.filter_module.kernel(f["kernel"])
        return self

    @property
    def has_stream_groups(self):
        """True when a group runs its plan at stream level (no batch kernel)."""
        return any(m.get("kind") == "stream" for m in self.metas)

    @property
    def filter_row(self):
        """Answer words per tile of the set's shared byte-pair filters (0: none)."""
        return max((m.get("filter_row", 0) for m in self.metas), default=0)

    def info(self):
        return [m.info(batch_kernel_name(k)) if m is not None else None for k, m in enumerate(self.modules)]

    def free(self):
        if self.modules is not None:
            for m in self.modules:
                if m is not None:
                    m.free()
            self.modules = None
            self.kernels = None
        if self.plain_set is not None:
            self.plain_set.free()
            self.plain_set = None
        if self.filter_module is not None:
            self.filter_module.free()
            self.filter_module = None
            self.filter_kernel = None


class BatchPlan:
    """Device buffers and work queues for a fixed batch of jobs.

    ``sets``: list of loaded BatchPatternSet; ``buffers``: list of uint8
    arrays; ``jobs``: list of (set index, buffer index) pairs.  Each pair
    expands to one job per kernel of the set."""

    def __init__(self, sets, buffers, jobs, device=0, grid_blocks=None, n_queues=0, take=0):
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", device)
        self.sets = sets
        self.n_buffers = len(buffers)
        self.lengths = [int(np.asarray(b).size) for b in buffers]
        self.words = [words_for(n) for n in self.lengths]
        # bytes: 64-byte aligned offsets
        offs, o = [], 0
        for n in self.lengths:
            offs.append(o)
            o += (n + 63) // 64 * 64
        self.byte_off = offs
        host = np.zeros(max(o, 64), dtype=np.uint8)
        for b, off in zip(buffers, offs):
            a = np.asarray(b, dtype=np.uint8)
            host[off:off + a.size] = a
        self.bytes = torch.from_numpy(host).to(self.dev)
        boffs, o = [], 0
        for W in self.words:
            boffs.append(o)
            o += 8 * W
        self.basis_off = boffs
        self.basis = torch.zeros(max(o, 1), dtype=torch.uint64, device=self.dev)
        # one work queue per (set, kernel): the kernel's jobs over the buffers;
        # a job's result is n_out consecutive streams (its kernel's groups)
        self.jobs = list(jobs)
        res_total = 0
        stat_total = 0
        self.queues = []   # (set, kernel) -> queue state
        for si, ps in enumerate(sets):
            bufs = [b for (s, b) in self.jobs if s == si]
            for kid, meta in enumerate(ps.metas):
                C = meta["status_slots"]
                n_out = meta.get("n_out", 1)
                first = meta.get("first_group", kid)
                entries, tile = [], 0
                for b in bufs:
                    W = self.words[b]
                    T = n_tiles(meta, W)
                    entries.append(dict(set=si, kernel=kid, first_group=first, n_out=n_out, buffer=b, W=W, T=T,
                                        C=C, tile_start=tile, res_off=res_total, stat_off=stat_total))
                    res_total += W * n_out
                    stat_total += T * C
                    tile += T
                self.queues.append(dict(set=si, kernel=kid, entries=entries, total_tiles=tile))
        self.result = torch.zeros(max(res_total, 1), dtype=torch.uint64, device=self.dev)
        self.status = torch.zeros(max(stat_total, 1), dtype=torch.int32, device=self.dev)
        # sets with shared byte-pair filters: one answer buffer per set, rows
        # by the global tiles of a sharing kernel's queue (every sharing
        # kernel of the set has the same tiles per job)
        self.filters = {}
        for si, ps in enumerate(sets):
            row = ps.filter_row
            if not row:
                continue
            if ps.filter_kernel is None:
                raise ValueError(f"set {si} shares its byte-pair filters: load() it first")
            kid = next(k for k, m in enumerate(ps.metas) if m.get("filter_row"))
            q = next(q for q in self.queues if q["set"] == si and q["kernel"] == kid)
            buf = torch.empty(max(1, q["total_tiles"] * row), dtype=torch.int32, device=self.dev)
            self.filters[si] = dict(buf=buf, queue=q)
        for q in self.queues:
            entries = q["entries"]
            arr = (rt.Job * max(1, len(entries)))()
            for k, e in enumerate(entries):
                arr[k].basis = self.basis.data_ptr() + 8 * self.basis_off[e["buffer"]]
                arr[k].result = self.result.data_ptr() + 8 * e["res_off"]
                arr[k].status = self.status.data_ptr() + 4 * e["stat_off"]
                arr[k].n_words = e["W"]
                arr[k].group = e["first_group"]
                arr[k].n_carry = e["C"]
                arr[k].tile_start = e["tile_start"]
                arr[k].n_tiles = e["T"]
            q["jobs"] = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(self.dev)
            tj = (np.repeat(np.arange(len(entries), dtype=np.int64), [e["T"] for e in entries])
                  if entries else np.zeros(1, dtype=np.int64))
            q["tile_job"] = torch.from_numpy(tj.astype(np.int32)).to(self.dev)
            q["ctl"] = rt.new_ctl(self.dev)
            args = rt.BatchArgs()
            args.jobs = q["jobs"].data_ptr()
            args.tile_job = q["tile_job"].data_ptr()
            args.ctl = q["ctl"].data_ptr()
            args.n_jobs = len(entries)
            args.total_tiles = q["total_tiles"]
            args.n_queues = n_queues
            args.take = take
            q["args"] = args
        for si, f in self.filters.items():
            for q in self.queues:
                if q["set"] == si and sets[si].metas[q["kernel"]].get("filter_row"):
                    q["args"].filter = f["buf"].data_ptr()
        # one transpose launch for all buffers (a descriptor per buffer)
        td = (rt.TDesc * max(1, self.n_buffers))()
        ws = 0
        for b in range(self.n_buffers):
            td[b].byte_off = self.byte_off[b]
            td[b].n_bytes = self.lengths[b]
            td[b].basis_off = self.basis_off[b]
            td[b].n_words = self.words[b]
            td[b].word_start = ws
            ws += self.words[b]
        self.total_words = ws
        self.tdesc = torch.from_numpy(np.frombuffer(bytes(td), dtype=np.uint8).copy()).to(self.dev)
        self.grid_blocks = grid_blocks
        self.graph = None
        self._stream_status = {}  # status words of stream-level groups, per (set, W)

    # -- execution -------------------------------------------------------------
    def _grid(self, ps, total_tiles):
        """Persistent blocks (0 = the runtime's default: as many as fit on the
        device at once, never more than one wave per tile)."""
        if self.grid_blocks:
            wpb = ps.block_threads // 64
            return max(1, min(self.grid_blocks, (total_tiles + wpb - 1) // wpb))
        return 0

    def enqueue_transpose(self, stream):
        rt.transpose_batch(self.bytes.data_ptr(), self.basis.data_ptr(), self.tdesc.data_ptr(), self.n_buffers,
                           self.total_words, stream)

    def enqueue_match(self, stream):
        done = set()
        for q in self.queues:
            ps = self.sets[q["set"]]
            f = self.filters.get(q["set"])
            if f is not None and q["set"] not in done:
                # the set's byte-pair answers, before its kernels
                done.add(q["set"])
                fq = f["queue"]
                if fq["total_tiles"]:
                    rt.filter_batch_launch(ps.filter_kernel, fq["jobs"].data_ptr(), fq["tile_job"].data_ptr(),
                                           fq["total_tiles"], f["buf"].data_ptr(), stream)
            if ps.metas[q["kernel"]].get("kind") == "stream":
                self._run_stream_jobs(ps, q)
            elif q["total_tiles"]:
                rt.batch_launch(ps.kernels[q["kernel"]], q["args"], self._grid(ps, q["total_tiles"]),
                                ps.block_threads, stream)

    def _run_stream_jobs(self, ps, q):
        """A stream-level group over each of its jobs: the group's plan
        (subgraph kernels, copies, while loops decided on the host) through
        the set's plain form, on the current torch stream, over the job's
        basis rows into its result row."""
        torch = self.torch
        k = q["kernel"]
        for e in q["entries"]:
            W = e["W"]
            st = self._stream_status.get((id(ps), W))
            if st is None:
                st = ps.plain_set.alloc_status(1, W)
                self._stream_status[(id(ps), W)] = st
            boff = self.basis_off[e["buffer"]]
            basis = self.basis[boff:boff + 8 * W].view(8, W)
            row = self.result[e["res_off"]:e["res_off"] + W]
            ps.plain_set._run_stream(k, basis, row, st.data_ptr() + k * st.size(2) * 4, W,
                                     torch.cuda.current_stream(self.dev))

    @property
    def graph_capturable(self):
        """False when a set has stream-level groups (their loops are decided
        on the host between launches)."""
        return not any(ps.has_stream_groups for ps in self.sets)

    def enqueue(self, stream=None):
        """Transpose + every set's kernels on ``stream`` (a torch stream or
        None: the current one; sets with stream-level groups run them on the
        current torch stream, so pass None or make ``stream`` current)."""
        s = rt.current_stream_ptr(stream)
        self.enqueue_transpose(s)
        self.enqueue_match(s)

    def capture(self, stream=None):
        """Capture transpose + batch launches into a graph (replay with run_graph)."""
        if not self.graph_capturable:
            raise NotImplementedError("a set with stream-level groups (while loops, graph breaks) decides its loops "
                                      "on the host: run the plan with enqueue(), not as a graph")
        torch = self.torch
        cap = torch.cuda.Stream(device=self.dev)
        cap.wait_stream(torch.cuda.current_stream(self.dev))
        s = rt.current_stream_ptr(cap)
        rt.graph_begin(s)
        try:
            self.enqueue_transpose(s)
            self.enqueue_match(s)
        finally:
            self.graph = rt.graph_end(s)
        torch.cuda.current_stream(self.dev).wait_stream(cap)
        return self.graph

    def run_graph(self, stream=None):
        if self.graph is None:
            self.capture()
        self.graph.launch(rt.current_stream_ptr(stream))

    def check_errors(self):
        """Raise if a carry look-back of the last run gave up (call after a
        sync); the flags are cleared first, so later replays report only
        their own errors."""
        bad = [i for i, q in enumerate(self.queues) if int(q["ctl"][rt.CTL_ERROR].item()) != 0]
        if bad:
            for i in bad:
                self.queues[i]["ctl"][rt.CTL_ERROR].zero_()
            raise rt.HipRuntimeError(f"carry look-back timed out in batched launches {bad}; results are invalid")

    def results(self, set_index, buffer_index):
        """Match streams [G, W] of one (set, buffer) job (device tensor)."""
        rows = [e for q in self.queues if q["set"] == set_index for e in q["entries"] if e["buffer"] == buffer_index]
        if not rows:
            raise KeyError((set_index, buffer_index))
        W = rows[0]["W"]
        rows.sort(key=lambda e: e["first_group"])
        return self.torch.cat([self.result[e["res_off"]:e["res_off"] + W * e["n_out"]].view(e["n_out"], W)
                               for e in rows])

    def scanned_bytes(self):
        return sum(self.lengths[b] for (_, b) in self.jobs)

    def destroy(self):
        if self.graph is not None:
            self.graph.destroy()
            self.graph = None
