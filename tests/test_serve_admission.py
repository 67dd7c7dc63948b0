"""Serving admission policy, grid caps and failure handling (CPU, no GPU
needed: the failure paths run against host-only fakes of a slot)."""
import queue
import threading

import numpy as np
import pytest

from bitgen_amd.backend.hip.runtime import HipRuntimeError
from bitgen_amd.serve import MatchJob, MatchServer, admit, grid_caps, job_footprint


def _job(n, fp=None):
    j = MatchJob(0, np.zeros(n, dtype=np.uint8), 0.0)
    j.footprint = fp if fp is not None else job_footprint(n, 2)
    return j


def test_admission_respects_cap_and_order():
    jobs = [_job(10, 100) for _ in range(5)]
    batch, rest, inflight = admit(jobs, 0, 250, 100, 10**9)
    assert batch == jobs[:2] and rest == jobs[2:] and inflight == 200


def test_admission_counts_jobs_already_in_flight():
    jobs = [_job(10, 100) for _ in range(3)]
    batch, rest, inflight = admit(jobs, 900, 1000, 100, 10**9)
    assert batch == jobs[:1] and inflight == 1000


def test_admission_batch_limits():
    jobs = [_job(1000, 1) for _ in range(10)]
    batch, rest, _ = admit(jobs, 0, 10**9, 4, 10**9)
    assert len(batch) == 4
    batch, rest, _ = admit(jobs, 0, 10**9, 100, 2500)
    assert len(batch) == 2
    big = [_job(5000, 1)]
    batch, rest, _ = admit(big, 0, 10**9, 100, 2500)
    assert batch == big  # one oversize job still runs alone


def test_admission_skips_a_blocked_job_without_reordering_the_rest():
    jobs = [_job(10, 100), _job(10, 900), _job(10, 100)]
    batch, rest, inflight = admit(jobs, 0, 300, 100, 10**9)
    assert batch == [jobs[0], jobs[2]] and rest == [jobs[1]] and inflight == 200


def test_footprint_grows_with_input_and_groups():
    assert job_footprint(10**6, 4) > job_footprint(10**6, 1) > job_footprint(10**5, 1)


# ---- grid caps and the failure path (host-only fakes, no GPU) -----------------

def test_grid_caps_share_the_device_between_streams():
    caps, excl = grid_caps({(0, 0): (6, 256), (0, 1): (5, 256), (1, 0): (9, 256)}, 3)
    assert excl == set()
    assert caps == {(0, 0): 2 * 256, (0, 1): 1 * 256, (1, 0): 3 * 256}
    # every slot's launch fits beside the others: n_streams shares <= a CU
    for (sid, kid), per in {(0, 0): 6, (0, 1): 5, (1, 0): 9}.items():
        assert 3 * caps[(sid, kid)] <= per * 256


def test_grid_caps_low_occupancy_set_runs_alone():
    caps, excl = grid_caps({(0, 0): (2, 256), (0, 1): (5, 256), (1, 0): (4, 256)}, 3)
    assert excl == {0}
    assert caps[(0, 0)] == 2 * 256 and caps[(0, 1)] == 5 * 256  # alone: the whole device
    assert caps[(1, 0)] == 1 * 256
    caps1, excl1 = grid_caps({(0, 0): (1, 256)}, 1)
    assert excl1 == set() and caps1[(0, 0)] == 256


class _FakeEvent:
    def query(self):
        return True


class _FakeSlot:
    def __init__(self, jobs, err):
        self.jobs = jobs
        self.event = _FakeEvent()
        self.out_rows = None
        self.exclusive = False
        self.err = err
        self.cleared = 0

    def errors(self):
        return self.err

    def clear_errors(self):
        self.err = 0
        self.cleared += 1


def _bare_server(slots):
    srv = MatchServer.__new__(MatchServer)
    srv.slots = slots
    srv.completed, srv.batch_log, srv.waiting = [], [], []
    srv.inbox = queue.Queue()
    srv.failed, srv.batches, srv.error = 0, 0, None
    srv.in_flight_bytes = sum(j.footprint for s in slots for j in (s.jobs or []))
    srv._stop = threading.Event()
    srv._thread = None
    srv.freeze_gc = False
    return srv


def test_error_flag_fails_only_that_batch_and_frees_its_slot():
    bad = [_job(10, 100), _job(20, 50)]
    slot = _FakeSlot(bad, err=1)
    srv = _bare_server([slot])
    assert srv._complete() is True
    for j in bad:
        with pytest.raises(HipRuntimeError):
            j.future.result(timeout=1)
    assert slot.jobs is None and slot.cleared == 1 and srv.failed == 2 and srv.in_flight_bytes == 0


def test_scheduler_death_fails_every_pending_job_and_stop_returns():
    in_flight = [_job(10, 100)]
    srv = _bare_server([_FakeSlot(in_flight, err=0)])
    waiting, queued = _job(10, 7), _job(10, 8)
    srv.waiting = [waiting]
    srv.inbox.put(queued)

    def boom():
        raise RuntimeError("scheduler fault")

    srv._serve = boom
    srv._thread = threading.Thread(target=srv._loop)
    srv._thread.start()
    srv.stop(drain=True)  # must return although work is "pending"
    for j in (in_flight[0], waiting, queued):
        with pytest.raises(RuntimeError, match="scheduler fault"):
            j.future.result(timeout=1)
    assert srv.in_flight_bytes == 0 and isinstance(srv.error, RuntimeError)
    with pytest.raises(RuntimeError):
        srv.sets = {0: None}
        srv.submit(0, b"abc")
