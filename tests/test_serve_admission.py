"""Serving admission policy (CPU, no GPU needed)."""
import numpy as np

from bitgen_amd.serve import MatchJob, admit, job_footprint


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
