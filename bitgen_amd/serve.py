"""Match serving with continuous batching.

Jobs (pattern-set id + byte buffer) arrive at any time through ``submit``;
a scheduler thread keeps the pattern sets' batch kernels loaded on the GPU,
admits waiting jobs while their HBM footprint fits under a cap, packs the
admitted jobs into one batch (a batched transpose plus one persistent launch
per kernel of every set involved) on one of several streams, and completes
jobs when their batch's event fires.  Jobs that arrive while a batch is
running are admitted into the next batch on another stream, so their work
overlaps on the device.  Each job's result is its own buffer's match streams,
so it does not depend on which other jobs shared the batch.

The reference has no serving component; this is the north star's addition.
"""
import ctypes
import gc
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from .backend.hip import runtime as rt
from .batch import words_for

@dataclass
class MatchJob:
    set_id: int
    data: np.ndarray
    arrival: float
    future: Future = field(default_factory=Future)
    footprint: int = 0
    done: float = 0.0


def job_footprint(n_bytes, n_groups):
    """Device bytes a job holds while in flight: input, basis, match streams,
    look-back status (upper bound)."""
    W = words_for(n_bytes)
    return (n_bytes + 64) + 64 * W + 8 * W * n_groups + 4 * (W // 60 + 2) * 8 * n_groups


def admit(waiting, in_flight_bytes, cap, max_jobs, max_bytes):
    """Admission in arrival order: a job joins the batch while its footprint
    fits under the cap next to everything in flight, up to max_jobs jobs and
    max_bytes input bytes (a single job larger than max_bytes still goes
    alone).  Returns (batch, still waiting, new in-flight bytes)."""
    batch, rest, nbytes = [], [], 0
    for job in waiting:
        fits = (in_flight_bytes + job.footprint <= cap and len(batch) < max_jobs
                and (not batch or nbytes + job.data.size <= max_bytes))
        if fits:
            batch.append(job)
            nbytes += job.data.size
            in_flight_bytes += job.footprint
        else:
            rest.append(job)
    return batch, rest, in_flight_bytes


def grid_caps(occupancy, n_streams):
    """Persistent-grid cap of every (set, kernel) launch and the sets whose
    batches must run alone.

    ``occupancy``: {(set id, kernel id): (blocks of the kernel one CU holds,
    CUs)}.  Batches on different slots run at the same time, and a
    persistent launch needs a resident block for every one of its tile
    queues (a tile waits for its predecessor, which only a resident block of
    that queue can take): launches that each hold part of the device can
    starve each other.  A launch therefore takes at most its 1/n_streams
    share of the blocks a CU holds of its kernel, so every slot's launch fits
    beside the others.  A kernel that fits fewer blocks per CU than there are
    streams has no such share: its set is exclusive -- its batches run while
    no other batch is on the device, with the whole device as the cap.
    Returns ({(set, kernel): max blocks}, set of exclusive set ids)."""
    n = max(1, int(n_streams))
    caps, exclusive = {}, set()
    for (sid, kid), (per_cu, cus) in occupancy.items():
        if n > 1 and per_cu < n:
            exclusive.add(sid)
    for (sid, kid), (per_cu, cus) in occupancy.items():
        per_cu, cus = max(1, int(per_cu)), max(1, int(cus))
        caps[(sid, kid)] = per_cu * cus if (sid in exclusive or n == 1) else (per_cu // n) * cus
    return caps, exclusive


class _Slot:
    """One stream with its in-flight batch and its own buffers, reused from
    batch to batch and grown (never shrunk) when a batch needs more: a
    pinned staging buffer the native runtime packs the batch into and its
    device copy, device buffers for basis, match streams and look-back
    status, a pinned buffer the match streams (and the batch's error flag)
    come back to, and one control block that every launch of the slot uses
    (so the epochs of its status words never repeat: the status buffer is
    zeroed only when it is allocated).  A slot's buffers are touched only
    while it holds no batch, so reuse never races the GPU."""

    # initial capacities (elements): enough for a batch of a few MiB, so
    # steady serving never allocates (pinned allocations take milliseconds)
    INITIAL = {"stage": 16 << 20, "dev_stage": 16 << 20, "basis": 1 << 20, "result": 4 << 20,
               "status": 1 << 20, "host_out": 4 << 20}
    DTYPES = {"stage": "uint8", "dev_stage": "uint8", "basis": "uint64", "result": "uint64", "status": "int32",
              "host_out": "uint64"}

    def __init__(self, torch, dev, max_queues, max_jobs, max_kernels):
        self.torch = torch
        self.dev = dev
        self.stream = torch.cuda.Stream(device=dev)
        self.stream_ptr = rt.current_stream_ptr(self.stream)
        self.ctl = torch.zeros(rt.CTL_WORDS, dtype=torch.int32, device=dev)
        self.flag = torch.zeros(2, dtype=torch.int32, device=dev)
        self.queues = (rt.ServeQueue * max_queues)()
        self.job_result = np.zeros(max_jobs * max_kernels, dtype=np.uint64)
        self.layout = rt.ServeLayout()
        self.bufs = {}
        self.event = None
        self.jobs = None
        self.out_rows = None
        self.exclusive = False
        self.t_launch = 0.0
        for name, n in self.INITIAL.items():
            self.buf(name, n)

    def errors(self):
        """Error flag of the last batch (its copy has landed once the batch's
        event fired)."""
        return int(self.bufs["host_out"][self.layout.result_words].item() & 0xFFFFFFFF)

    def clear_errors(self):
        """Nothing to clear: the batch's launches moved the control block's
        error word into the batch's own flag."""

    def buf(self, name, n):
        """A buffer of at least n elements (contents undefined, but a new
        status buffer is zeroed)."""
        t = self.bufs.get(name)
        if t is None or t.numel() < n:
            size = max(int(n), int(1.5 * t.numel()) if t is not None else 0, 1024)
            dtype = getattr(self.torch, self.DTYPES[name])
            if name in ("stage", "host_out"):
                t = self.torch.empty(size, dtype=dtype, pin_memory=True)
            elif name == "status":
                t = self.torch.zeros(size, dtype=dtype, device=self.dev)
            else:
                t = self.torch.empty(size, dtype=dtype, device=self.dev)
            self.bufs[name] = t
        return t


class MatchServer:
    def __init__(self, pattern_sets, device=0, hbm_cap_bytes=200 * 10**9, max_batch_jobs=512,
                 max_batch_bytes=256 * 2**20, n_streams=3, grid_waves_per_cu=16, freeze_gc=True):
        """pattern_sets: {set_id: loaded BatchPatternSet}.  freeze_gc: move the
        objects alive at start() out of the garbage collector's reach while
        serving (gc.freeze; undone by stop())."""
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", device)
        rt.init(device)
        self.sets = dict(pattern_sets)
        self.freeze_gc = bool(freeze_gc)
        self.cap = int(hbm_cap_bytes)
        self.max_batch_jobs = max_batch_jobs
        self.max_batch_bytes = max_batch_bytes
        self.inbox = queue.Queue()
        self.waiting = []
        self.in_flight_bytes = 0
        self.completed = []
        self.batches = 0
        # per completed batch: (jobs, host staging seconds, launch-to-done seconds)
        self.batch_log = []
        self._stop = threading.Event()
        self._thread = None
        props = torch.cuda.get_device_properties(self.dev)
        self.grid_waves = props.multi_processor_count * grid_waves_per_cu
        occ = {}
        for sid, ps in self.sets.items():
            if ps.kernels is None:
                raise ValueError(f"pattern set {sid} is not loaded (call BatchPatternSet.load first)")
            if ps.has_stream_groups:
                raise ValueError(f"pattern set {sid} has groups that run at stream level (while loops, graph "
                                 "breaks); the server's native staging runs batch kernels only")
            if ps.filter_row:
                # the native staging launches the set's kernels only; their
                # byte-pair tables must be built in the kernels
                raise ValueError(f"pattern set {sid} shares its byte-pair filters across kernels (shared_filter); "
                                 "served sets are compiled with shared_filter=0")
            for kid, kern in enumerate(ps.kernels):
                occ[(sid, kid)] = rt.kernel_occupancy(kern, ps.block_threads)
        self.grid_blocks_max, self.exclusive_sets = grid_caps(occ, n_streams)
        # native tables of the resident kernels: one entry per (set, kernel)
        self.set_index = {sid: i for i, sid in enumerate(self.sets)}
        n_k = sum(len(ps.metas) for ps in self.sets.values())
        self.c_kernels = (rt.ServeKernel * max(1, n_k))()
        self.c_sets = (rt.ServeSet * max(1, len(self.sets)))()
        self.kernel_n_out = []
        k = 0
        for sid, ps in self.sets.items():
            self.c_sets[self.set_index[sid]] = rt.ServeSet(k, len(ps.metas))
            wpb = ps.block_threads // 64
            for kid, meta in enumerate(ps.metas):
                cap = max(1, min(self.grid_waves // wpb, self.grid_blocks_max[(sid, kid)]))
                self.c_kernels[k] = rt.ServeKernel(ps.kernels[kid], meta["HR"], meta["S"], meta["status_slots"],
                                                   meta.get("n_out", 1), ps.block_threads, cap)
                self.kernel_n_out.append(meta.get("n_out", 1))
                k += 1
        self.max_kernels = max(len(ps.metas) for ps in self.sets.values())
        self.slots = [_Slot(torch, self.dev, max(1, n_k), max_batch_jobs + len(self.sets), self.max_kernels)
                      for _ in range(n_streams)]
        self.failed = 0
        self.error = None  # exception that ended the scheduler thread, if any

    # -- client side -------------------------------------------------------------
    def submit(self, set_id, data):
        if set_id not in self.sets:
            raise KeyError(f"unknown pattern set {set_id}")
        if self.error is not None:
            raise RuntimeError("the match server's scheduler has stopped") from self.error
        arr = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray))
                                   else np.asarray(data, dtype=np.uint8))
        job = MatchJob(set_id, arr, time.perf_counter())
        job.footprint = job_footprint(arr.size, self.sets[set_id].n_groups)
        if job.footprint > self.cap:
            job.future.set_exception(ValueError(f"job needs {job.footprint} bytes, above the {self.cap}-byte cap"))
            return job.future
        self.inbox.put(job)
        return job.future

    def warm(self, data):
        """Before start(): run, on every slot, one batch holding one job per
        pattern set over ``data`` and wait for it.  The first launch of every
        kernel on every slot's stream is then paid before serving, and the
        warm batches are not recorded."""
        if self._thread is not None:
            raise RuntimeError("warm() runs before start()")
        self.torch.cuda.set_device(self.dev)
        arr = np.ascontiguousarray(np.asarray(data, dtype=np.uint8))
        for slot in self.slots:
            jobs = []
            for sid in self.sets:
                job = MatchJob(sid, arr, time.perf_counter())
                job.footprint = job_footprint(arr.size, self.sets[sid].n_groups)
                self.in_flight_bytes += job.footprint
                jobs.append(job)
            self._launch(slot, jobs)
            while slot.jobs is not None:
                if not self._complete():
                    time.sleep(50e-6)
        self.completed.clear()
        self.batch_log.clear()
        self.batches = 0
        return self

    def start(self):
        # a full collection of the interpreter's heap (torch alone holds
        # hundreds of thousands of objects) stalls the scheduler thread for
        # tens of milliseconds; objects alive at start move to the permanent
        # generation, so later collections only scan what serving creates
        if self.freeze_gc:
            gc.collect()
            gc.freeze()
        self._thread = threading.Thread(target=self._loop, name="bitgen-serve", daemon=True)
        self._thread.start()
        return self

    def _pending(self):
        return not self.inbox.empty() or self.waiting or any(s.jobs is not None for s in self.slots)

    def stop(self, drain=True):
        """Stop serving; with drain, after every submitted job completed.
        Returns also when the scheduler thread has ended on an error (its
        jobs' futures then hold that error)."""
        if drain:
            while self._thread is not None and self._thread.is_alive() and self._pending():
                time.sleep(0.001)
        self._stop.set()
        if self._thread:
            self._thread.join()
        if self.freeze_gc:
            gc.unfreeze()

    # -- scheduler -----------------------------------------------------------------
    def _loop(self):
        try:
            self._serve()
        except BaseException as e:  # noqa: BLE001 -- every pending future gets the error
            self.error = e
            self._fail_all(e)

    def _fail_all(self, exc):
        """Fail every job not yet completed (in flight, waiting, in the inbox)."""
        in_flight = [j for s in self.slots if s.jobs is not None for j in s.jobs]
        jobs = in_flight + self.waiting
        while True:
            try:
                jobs.append(self.inbox.get_nowait())
            except queue.Empty:
                break
        for s in self.slots:
            s.jobs = s.out_rows = s.event = None
        self.waiting = []
        self.in_flight_bytes -= sum(j.footprint for j in in_flight)
        for job in jobs:
            if not job.future.done():
                job.future.set_exception(exc)

    def _serve(self):
        torch = self.torch
        torch.cuda.set_device(self.dev)
        while not self._stop.is_set():
            progressed = self._complete()
            while True:
                try:
                    self.waiting.append(self.inbox.get_nowait())
                except queue.Empty:
                    break
            free = [s for s in self.slots if s.jobs is None]
            exclusive_busy = any(s.jobs is not None and s.exclusive for s in self.slots)
            if self.waiting and free and not exclusive_busy:
                batch = self._admit()
                excl = any(j.set_id in self.exclusive_sets for j in batch)
                if batch and excl and len(free) < len(self.slots):
                    # an exclusive batch waits until the device is drained;
                    # nothing else is launched meanwhile (arrival order holds)
                    self.in_flight_bytes -= sum(j.footprint for j in batch)
                    self.waiting = sorted(batch + self.waiting, key=lambda j: j.arrival)
                elif batch:
                    self._launch(free[0], batch, exclusive=excl)
                    progressed = True
            if not progressed:
                if self.waiting or any(s.jobs is not None for s in self.slots):
                    time.sleep(20e-6)
                else:
                    try:
                        self.waiting.append(self.inbox.get(timeout=0.002))
                    except queue.Empty:
                        pass

    def _admit(self):
        batch, self.waiting, self.in_flight_bytes = admit(
            self.waiting, self.in_flight_bytes, self.cap, self.max_batch_jobs, self.max_batch_bytes)
        return batch

    def _launch(self, slot, jobs, exclusive=False):
        """Stage a batch on a free slot in the native runtime: bg_serve_pack
        lays out input bytes and launch tables in the slot's pinned buffer,
        bg_serve_enqueue copies them over, transposes, launches every (set,
        kernel) of the batch and copies the match streams back, all on the
        slot's stream."""
        slot.t_start = time.perf_counter()
        n = len(jobs)
        L = rt.load_library()
        jt = np.empty(n, dtype=rt.SERVE_JOB_DTYPE)
        jt["data"] = [j.data.ctypes.data for j in jobs]
        jt["n_bytes"] = [j.data.size for j in jobs]
        jt["set"] = [self.set_index[j.set_id] for j in jobs]
        jt["reserved"] = 0
        if slot.job_result.size < n * self.max_kernels:
            slot.job_result = np.zeros(n * self.max_kernels, dtype=np.uint64)
        lay = slot.layout
        while True:
            b = slot.bufs
            rc = L.bg_serve_pack(self.c_sets, len(self.sets), self.c_kernels, len(self.kernel_n_out),
                                 jt.ctypes.data, n, b["stage"].data_ptr(), b["stage"].numel(),
                                 b["basis"].data_ptr(), b["basis"].numel(), b["result"].data_ptr(),
                                 b["result"].numel(), b["status"].data_ptr(), b["status"].numel(),
                                 ctypes.byref(lay), slot.queues, len(slot.queues), slot.job_result.ctypes.data,
                                 self.max_kernels)
            if rc != rt.ENOSPACE:
                rt._check(rc, "bg_serve_pack")
                break
            slot.buf("stage", lay.stage_bytes)
            slot.buf("dev_stage", lay.stage_bytes)
            slot.buf("basis", lay.basis_words)
            slot.buf("result", lay.result_words)
            slot.buf("status", lay.status_words)
        slot.buf("dev_stage", lay.stage_bytes)
        host_out = slot.buf("host_out", lay.result_words + 1)
        rt._check(L.bg_serve_enqueue(ctypes.byref(lay), slot.queues, self.c_kernels, slot.bufs["stage"].data_ptr(),
                                     slot.bufs["dev_stage"].data_ptr(), slot.bufs["basis"].data_ptr(),
                                     slot.bufs["result"].data_ptr(), slot.ctl.data_ptr(), slot.flag.data_ptr(),
                                     host_out.data_ptr(), slot.stream_ptr), "bg_serve_enqueue")
        ev = self.torch.cuda.Event()
        ev.record(slot.stream)
        # per job: words and the result offset of each of its set's kernels
        words = (jt["n_bytes"] + 1 + 63) // 64
        slot.event, slot.jobs = ev, jobs
        slot.out_rows = (slot.job_result[:n * self.max_kernels].copy(), words)
        slot.exclusive = exclusive
        slot.t_launch = time.perf_counter()
        self.batches += 1

    def _complete(self):
        progressed = False
        for slot in self.slots:
            if slot.jobs is None or not slot.event.query():
                continue
            if slot.errors():
                # a carry look-back gave up: this batch's results are
                # invalid, the next batches are not
                now = time.perf_counter()
                err = rt.HipRuntimeError(f"carry look-back timed out in a served batch of {len(slot.jobs)} jobs; "
                                         "its results are invalid")
                for job in slot.jobs:
                    job.done = now
                    self.in_flight_bytes -= job.footprint
                    job.future.set_exception(err)
                self.failed += len(slot.jobs)
                slot.clear_errors()
                slot.jobs = slot.out_rows = slot.event = None
                progressed = True
                continue
            out = slot.bufs["host_out"].numpy()
            offs, words = slot.out_rows
            now = time.perf_counter()
            self.batch_log.append((len(slot.jobs), slot.t_launch - slot.t_start, now - slot.t_launch))
            mk = self.max_kernels
            for i, job in enumerate(slot.jobs):
                W = int(words[i])
                cs = self.c_sets[self.set_index[job.set_id]]
                # the set's kernels in order: their groups are consecutive
                parts = []
                for kk in range(cs.n_kernels):
                    r, nk = int(offs[i * mk + kk]), self.kernel_n_out[cs.first_kernel + kk]
                    parts.append(out[r:r + nk * W].reshape(nk, W))
                res = parts[0].copy() if len(parts) == 1 else np.concatenate(parts)
                job.done = now
                self.in_flight_bytes -= job.footprint
                self.completed.append(job)
                job.future.set_result(res)
            slot.jobs = slot.out_rows = slot.event = None
            progressed = True
        return progressed

    def latencies(self):
        return np.array([j.done - j.arrival for j in self.completed])


def run_plain_job(group_set, data, device, stream=None):
    """Per-job plain form: transpose + one launch per group + copy back,
    synchronous (the baseline the served form is compared with)."""
    import torch

    from .bitstream import transpose_bytes

    x = torch.from_numpy(np.ascontiguousarray(data, dtype=np.uint8)).to(device).reshape(1, 1, -1)
    basis, _ = transpose_bytes(x, data.size)
    W = basis.size(-1)
    result = torch.empty((1, 1, group_set.n_groups, W), dtype=torch.uint64, device=device)
    status = group_set.alloc_status(1, W)
    group_set.run(basis, result, status)
    out = result.cpu().numpy()[0, 0]
    return out
