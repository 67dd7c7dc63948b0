"""Match serving with continuous batching.

Jobs (pattern-set id + byte buffer) arrive at any time through ``submit``;
a scheduler thread keeps the pattern sets' batch kernels loaded on the GPU,
admits waiting jobs while their HBM footprint fits under a cap, packs the
admitted jobs into one batch (a batched transpose plus one persistent launch
per group of every set involved) on one of several streams, and completes
jobs when their batch's event fires.  Jobs that arrive while a batch is
running are admitted into the next batch on another stream, so their work
overlaps on the device.  Each job's result is its own buffer's match streams,
so it does not depend on which other jobs shared the batch.

The reference has no serving component; this is the north star's addition.
"""
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from .backend.hip import runtime as rt
from .batch import n_tiles, words_for

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


class _Slot:
    """One stream with its in-flight batch."""

    def __init__(self, torch, dev):
        self.stream = torch.cuda.Stream(device=dev)
        self.event = None
        self.jobs = None
        self.keep = None
        self.host_out = None
        self.out_rows = None


class MatchServer:
    def __init__(self, pattern_sets, device=0, hbm_cap_bytes=200 * 10**9, max_batch_jobs=512,
                 max_batch_bytes=256 * 2**20, n_streams=3, grid_waves_per_cu=16):
        """pattern_sets: {set_id: loaded BatchPatternSet}."""
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", device)
        rt.init(device)
        self.sets = dict(pattern_sets)
        self.cap = int(hbm_cap_bytes)
        self.max_batch_jobs = max_batch_jobs
        self.max_batch_bytes = max_batch_bytes
        self.inbox = queue.Queue()
        self.waiting = []
        self.in_flight_bytes = 0
        self.slots = [_Slot(torch, self.dev) for _ in range(n_streams)]
        self.completed = []
        self.batches = 0
        self._stop = threading.Event()
        self._thread = None
        props = torch.cuda.get_device_properties(self.dev)
        self.grid_waves = props.multi_processor_count * grid_waves_per_cu
        self.ctl = {}  # (set, group, slot) -> control block (tile queues, error flag, epoch)

    # -- client side -------------------------------------------------------------
    def submit(self, set_id, data):
        if set_id not in self.sets:
            raise KeyError(f"unknown pattern set {set_id}")
        arr = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray))
                                   else np.asarray(data, dtype=np.uint8))
        job = MatchJob(set_id, arr, time.perf_counter())
        job.footprint = job_footprint(arr.size, self.sets[set_id].n_groups)
        if job.footprint > self.cap:
            job.future.set_exception(ValueError(f"job needs {job.footprint} bytes, above the {self.cap}-byte cap"))
            return job.future
        self.inbox.put(job)
        return job.future

    def start(self):
        self._thread = threading.Thread(target=self._loop, name="bitgen-serve", daemon=True)
        self._thread.start()
        return self

    def stop(self, drain=True):
        if drain:
            while not self.inbox.empty() or self.waiting or any(s.jobs for s in self.slots):
                time.sleep(0.001)
        self._stop.set()
        if self._thread:
            self._thread.join()

    # -- scheduler -----------------------------------------------------------------
    def _loop(self):
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
            if self.waiting and free:
                batch = self._admit()
                if batch:
                    self._launch(free[0], batch)
                    progressed = True
            if not progressed:
                if self.waiting or any(s.jobs for s in self.slots):
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

    def _launch(self, slot, jobs):
        torch = self.torch
        n = len(jobs)
        lengths = np.array([j.data.size for j in jobs], dtype=np.int64)
        words = (lengths + 1 + 63) // 64
        byte_off = np.concatenate([[0], np.cumsum((lengths + 63) // 64 * 64)])[:-1]
        total_bytes = int(((lengths + 63) // 64 * 64).sum()) or 64
        host = torch.empty(total_bytes, dtype=torch.uint8, pin_memory=True)
        hv = host.numpy()
        for j, off in zip(jobs, byte_off):
            hv[off:off + j.data.size] = j.data
        basis_off = np.concatenate([[0], np.cumsum(8 * words)])[:-1]
        with torch.cuda.stream(slot.stream):
            dbytes = torch.empty(total_bytes, dtype=torch.uint8, device=self.dev)
            dbytes.copy_(host, non_blocking=True)
            basis = torch.empty(int(8 * words.sum()), dtype=torch.uint64, device=self.dev)
            # per (set, group) queues over this batch's jobs of that set
            queues, res_total, stat_total = [], 0, 0
            rows = []  # (job index, group, res_off, W)
            by_set = {}
            for i, j in enumerate(jobs):
                by_set.setdefault(j.set_id, []).append(i)
            for sid, idx in by_set.items():
                ps = self.sets[sid]
                for g, meta in enumerate(ps.metas):
                    C = max(1, meta["n_carry"])
                    T = np.array([n_tiles(meta, int(words[i])) for i in idx], dtype=np.int64)
                    starts = np.concatenate([[0], np.cumsum(T)])[:-1]
                    ent = np.zeros(len(idx), dtype=_JOB_DTYPE)
                    ent["n_words"] = words[idx]
                    ent["group"] = g
                    ent["n_carry"] = C
                    ent["tile_start"] = starts
                    ent["n_tiles"] = T
                    res_off = res_total + np.concatenate([[0], np.cumsum(words[idx])])[:-1]
                    stat_off = stat_total + np.concatenate([[0], np.cumsum(T * C)])[:-1]
                    res_total += int(words[idx].sum())
                    stat_total += int((T * C).sum())
                    queues.append((sid, g, idx, ent, res_off, stat_off, int(T.sum()), T))
                    rows.extend((i, g, int(r), int(words[i])) for i, r in zip(idx, res_off))
            result = torch.empty(max(res_total, 1), dtype=torch.uint64, device=self.dev)
            status = torch.zeros(max(stat_total, 1), dtype=torch.int32, device=self.dev)
            td = np.zeros(n, dtype=_TDESC_DTYPE)
            td["byte_off"] = byte_off
            td["n_bytes"] = lengths
            td["basis_off"] = basis_off
            td["n_words"] = words
            td["word_start"] = np.concatenate([[0], np.cumsum(words)])[:-1]
            tdesc = torch.from_numpy(td.view(np.uint8)).pin_memory().to(self.dev, non_blocking=True)
            s = rt.current_stream_ptr(slot.stream)
            rt.transpose_batch(dbytes.data_ptr(), basis.data_ptr(), tdesc.data_ptr(), n, int(words.sum()), s)
            keep = [host, dbytes, basis, result, status, tdesc]
            for (sid, g, idx, ent, res_off, stat_off, total, T) in queues:
                ps = self.sets[sid]
                ent["basis"] = basis.data_ptr() + 8 * basis_off[idx]
                ent["result"] = result.data_ptr() + 8 * res_off
                ent["status"] = status.data_ptr() + 4 * stat_off
                jt = torch.from_numpy(ent.view(np.uint8)).pin_memory().to(self.dev, non_blocking=True)
                tj_host = np.repeat(np.arange(len(idx), dtype=np.int32), T)
                tj = torch.from_numpy(tj_host).pin_memory().to(self.dev, non_blocking=True)
                key = (sid, g, id(slot))
                if key not in self.ctl:
                    self.ctl[key] = rt.new_ctl(self.dev)
                ctl = self.ctl[key]
                args = rt.BatchArgs()
                args.jobs = jt.data_ptr()
                args.tile_job = tj.data_ptr()
                args.ctl = ctl.data_ptr()
                args.n_jobs = len(idx)
                args.total_tiles = total
                wpb = ps.block_threads // 64
                blocks = max(1, min(self.grid_waves // wpb, (total + wpb - 1) // wpb))
                rt.batch_launch(ps.kernels[g], args, blocks, ps.block_threads, s)
                keep += [jt, tj]
            host_out = torch.empty(max(res_total, 1), dtype=torch.uint64, pin_memory=True)
            host_out.copy_(result, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(slot.stream)
        slot.event, slot.jobs, slot.keep, slot.host_out, slot.out_rows = ev, jobs, keep, host_out, rows
        self.batches += 1

    def _complete(self):
        progressed = False
        for slot in self.slots:
            if slot.jobs is None or not slot.event.query():
                continue
            out = slot.host_out.numpy()
            per_job = {}
            for (i, g, off, W) in slot.out_rows:
                per_job.setdefault(i, {})[g] = out[off:off + W]
            now = time.perf_counter()
            for i, job in enumerate(slot.jobs):
                gs = per_job.get(i, {})
                res = np.stack([gs[g] for g in sorted(gs)]).copy() if gs else np.zeros((0, 0), np.uint64)
                job.done = now
                self.in_flight_bytes -= job.footprint
                self.completed.append(job)
                job.future.set_result(res)
            for (sid, g, sl), ctl in self.ctl.items():
                if sl == id(slot) and int(ctl[rt.CTL_ERROR].item()) != 0:
                    raise rt.HipRuntimeError("carry look-back timed out in a served batch")
            slot.jobs = slot.keep = slot.host_out = slot.out_rows = slot.event = None
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
