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
import gc
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from .backend.hip import runtime as rt
from .batch import words_for

_JOB_DTYPE = np.dtype([("basis", "<u8"), ("result", "<u8"), ("status", "<u8"), ("n_words", "<u8"),
                       ("group", "<u4"), ("n_carry", "<u4"), ("tile_start", "<u4"), ("n_tiles", "<u4")])
_TDESC_DTYPE = np.dtype([("byte_off", "<u8"), ("n_bytes", "<u8"), ("basis_off", "<u8"), ("n_words", "<u8"),
                         ("word_start", "<u8")])


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


_ZEROS = np.zeros(64, dtype=np.uint8)


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
    batch to batch and grown (never shrunk) when a batch needs more: pinned
    host staging for the input bytes and the launch tables, device buffers
    for input, basis, match streams, look-back status and tables, a pinned
    host buffer for the match streams, and one control block per (set,
    kernel) with a pinned copy of their error flags.  A slot's buffers are
    touched only while it holds no batch, so reuse never races the GPU."""

    # initial capacities (elements): enough for a batch of a few MiB, so
    # steady serving never allocates (pinned allocations take milliseconds)
    INITIAL = {"dev_in": 8 << 20, "basis": 1 << 20, "result": 4 << 20, "status": 1 << 20,
               "host_in": 8 << 20, "host_meta": 1 << 20, "dev_meta": 1 << 20, "host_out": 4 << 20}

    def __init__(self, torch, dev, n_ctl):
        self.torch = torch
        self.dev = dev
        self.stream = torch.cuda.Stream(device=dev)
        self.ctl = torch.zeros((max(1, n_ctl), rt.CTL_WORDS), dtype=torch.int32, device=dev)
        self.err_host = torch.zeros(max(1, n_ctl), dtype=torch.int32, pin_memory=True)
        self.bufs = {}
        self.event = None
        self.jobs = None
        self.out_rows = None
        self.exclusive = False
        self.t_launch = 0.0
        dtypes = {"dev_in": torch.uint8, "basis": torch.uint64, "result": torch.uint64, "status": torch.int32,
                  "host_in": torch.uint8, "host_meta": torch.uint8, "dev_meta": torch.uint8,
                  "host_out": torch.uint64}
        for name, n in self.INITIAL.items():
            self.buf(name, n, dtypes[name], pinned=name.startswith("host"))

    def errors(self):
        """Error flags of the last batch (its copy has landed once the
        batch's event fired)."""
        return int(self.err_host.abs().sum())

    def clear_errors(self):
        """Clear the sticky error flags of every control block of the slot
        (ordered on the slot's stream, before its next batch)."""
        with self.torch.cuda.stream(self.stream):
            self.ctl[:, rt.CTL_ERROR].zero_()
        self.err_host.zero_()

    def buf(self, name, n, dtype, pinned=False):
        """A buffer of at least n elements (contents undefined)."""
        t = self.bufs.get(name)
        if t is None or t.numel() < n:
            size = max(int(n), int(1.5 * t.numel()) if t is not None else 0, 1024)
            if pinned:
                t = self.torch.empty(size, dtype=dtype, pin_memory=True)
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
        # control-block row of every (set, kernel) pair, the same in each slot
        self.ctl_row = {}
        for sid, ps in self.sets.items():
            for kid in range(len(ps.metas)):
                self.ctl_row[(sid, kid)] = len(self.ctl_row)
        self.slots = [_Slot(torch, self.dev, len(self.ctl_row)) for _ in range(n_streams)]
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
            for kid, kern in enumerate(ps.kernels):
                occ[(sid, kid)] = rt.kernel_occupancy(kern, ps.block_threads)
        self.grid_blocks_max, self.exclusive_sets = grid_caps(occ, n_streams)
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
        """Stage a batch on a free slot: one copy of all input bytes and one
        of all launch tables, one batched transpose, one persistent launch
        per (set, kernel) involved, one copy of all match streams back."""
        torch = self.torch
        slot.t_start = time.perf_counter()
        n = len(jobs)
        lengths = np.array([j.data.size for j in jobs], dtype=np.int64)
        words = (lengths + 1 + 63) // 64
        padded = (lengths + 63) // 64 * 64
        byte_off = np.concatenate([[0], np.cumsum(padded)])[:-1]
        total_bytes = int(padded.sum()) or 64
        basis_off = np.concatenate([[0], np.cumsum(8 * words)])[:-1]
        # per (set, kernel) queues over this batch's jobs of that set; a job's
        # result holds its kernel's n_out group streams in order.  Everything
        # per job is vectorised: a batch can hold hundreds of jobs.
        queues, res_total, stat_total = [], 0, 0
        set_ids = np.array([j.set_id for j in jobs])
        for sid in dict.fromkeys(set_ids.tolist()):
            ps = self.sets[sid]
            idx = np.nonzero(set_ids == sid)[0]
            Wj = words[idx]
            for kid, meta in enumerate(ps.metas):
                C = meta["status_slots"]
                n_out = meta.get("n_out", 1)
                first = meta.get("first_group", kid)
                s0 = 64 - meta["HR"]
                T = np.where(Wj <= s0, 1, 1 + (Wj - s0 + meta["S"] - 1) // meta["S"]).astype(np.int64)
                ent = np.zeros(len(idx), dtype=_JOB_DTYPE)
                ent["n_words"] = Wj
                ent["group"] = first
                ent["n_carry"] = C
                ent["tile_start"] = np.concatenate([[0], np.cumsum(T)])[:-1]
                ent["n_tiles"] = T
                res_off = res_total + np.concatenate([[0], np.cumsum(Wj * n_out)])[:-1]
                stat_off = stat_total + np.concatenate([[0], np.cumsum(T * C)])[:-1]
                res_total += int(Wj.sum()) * n_out
                stat_total += int((T * C).sum())
                queues.append((sid, kid, idx, ent, res_off, stat_off, int(T.sum()), T, first, n_out))
        res_total, stat_total = max(res_total, 1), max(stat_total, 1)
        # device buffers of this slot (grown if needed), then the tables that
        # point into them, packed into one byte image with 256-byte sections
        dbytes = slot.buf("dev_in", total_bytes, torch.uint8)
        basis = slot.buf("basis", int(8 * words.sum()) or 1, torch.uint64)
        result = slot.buf("result", res_total, torch.uint64)
        status = slot.buf("status", stat_total, torch.int32)
        td = np.zeros(n, dtype=_TDESC_DTYPE)
        td["byte_off"] = byte_off
        td["n_bytes"] = lengths
        td["basis_off"] = basis_off
        td["n_words"] = words
        td["word_start"] = np.concatenate([[0], np.cumsum(words)])[:-1]
        sections = [td.view(np.uint8)]
        for (sid, kid, idx, ent, res_off, stat_off, total, T, first, n_out) in queues:
            ent["basis"] = basis.data_ptr() + 8 * basis_off[idx]
            ent["result"] = result.data_ptr() + 8 * res_off
            ent["status"] = status.data_ptr() + 4 * stat_off
            sections.append(ent.view(np.uint8))
            sections.append(np.repeat(np.arange(len(idx), dtype=np.int32), T).view(np.uint8))
        offs, o = [], 0
        for sec in sections:
            offs.append(o)
            o += (sec.size + 255) // 256 * 256
        meta_bytes = max(o, 256)
        host_in = slot.buf("host_in", total_bytes, torch.uint8, pinned=True)
        hv = host_in.numpy()
        # one concatenation into the pinned buffer: each job's bytes, then
        # zeros up to its 64-byte boundary
        pad = padded - lengths
        pieces = []
        for j, p in zip(jobs, pad.tolist()):
            pieces.append(j.data)
            if p:
                pieces.append(_ZEROS[:p])
        if pieces:
            np.concatenate(pieces, out=hv[:int(padded.sum())])
        host_meta = slot.buf("host_meta", meta_bytes, torch.uint8, pinned=True)
        mv = host_meta.numpy()
        for sec, off in zip(sections, offs):
            mv[off:off + sec.size] = sec
        dmeta = slot.buf("dev_meta", meta_bytes, torch.uint8)
        host_out = slot.buf("host_out", res_total, torch.uint64, pinned=True)
        with torch.cuda.stream(slot.stream):
            dbytes[:total_bytes].copy_(host_in[:total_bytes], non_blocking=True)
            dmeta[:meta_bytes].copy_(host_meta[:meta_bytes], non_blocking=True)
            # status words must start clear: regions move between (set,
            # kernel) pairs from batch to batch, and their epochs differ
            status[:stat_total].zero_()
            s = rt.current_stream_ptr(slot.stream)
            rt.transpose_batch(dbytes.data_ptr(), basis.data_ptr(), dmeta.data_ptr() + offs[0], n, int(words.sum()), s)
            for q, (sid, kid, idx, ent, res_off, stat_off, total, T, first, n_out) in enumerate(queues):
                ps = self.sets[sid]
                args = rt.BatchArgs()
                args.jobs = dmeta.data_ptr() + offs[1 + 2 * q]
                args.tile_job = dmeta.data_ptr() + offs[2 + 2 * q]
                args.ctl = slot.ctl[self.ctl_row[(sid, kid)]].data_ptr()
                args.n_jobs = len(idx)
                args.total_tiles = total
                wpb = ps.block_threads // 64
                blocks = max(1, min(self.grid_waves // wpb, self.grid_blocks_max[(sid, kid)],
                                    (total + wpb - 1) // wpb))
                rt.batch_launch(ps.kernels[kid], args, blocks, ps.block_threads, s)
            host_out[:res_total].copy_(result[:res_total], non_blocking=True)
            slot.err_host.copy_(slot.ctl[:, rt.CTL_ERROR], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        # per job: (result offset, first group, n_out) of each of its kernels
        parts = [[] for _ in jobs]
        for (sid, kid, idx, ent, res_off, stat_off, total, T, first, n_out) in queues:
            for i, r in zip(idx.tolist(), res_off.tolist()):
                parts[i].append((r, n_out))
        slot.event, slot.jobs, slot.out_rows = ev, jobs, (parts, words)
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
            parts, words = slot.out_rows
            now = time.perf_counter()
            self.batch_log.append((len(slot.jobs), slot.t_launch - slot.t_start, now - slot.t_launch))
            for i, job in enumerate(slot.jobs):
                W = int(words[i])
                # kernels were queued in group order, so the blocks are too
                res = np.concatenate([out[r:r + n * W].reshape(n, W) for r, n in parts[i]])
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
