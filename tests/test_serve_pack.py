"""bg_serve_pack (the native staging of a served batch) on the CPU: it makes
no device call, so its layout is checked here against a direct restatement
of what each part must hold -- input bytes, transpose descriptors, job
tables with device addresses, tile-to-job maps, result offsets, launch
sizes -- and against undersized buffers (BG_ENOSPACE with the sizes)."""
import ctypes
import random

import numpy as np
import pytest

from bitgen_amd.backend.hip import runtime as rt

JOB = np.dtype([("basis", "<u8"), ("result", "<u8"), ("status", "<u8"), ("n_words", "<u8"), ("group", "<u4"),
                ("n_carry", "<u4"), ("tile_start", "<u4"), ("n_tiles", "<u4")])
TDESC = np.dtype([("byte_off", "<u8"), ("n_bytes", "<u8"), ("basis_off", "<u8"), ("n_words", "<u8"),
                  ("word_start", "<u8")])
DEV_BASIS, DEV_RESULT, DEV_STATUS = 0x7F0000000000, 0x7E0000000000, 0x7D0000000000


def tiles(W, hr, s):
    s0 = 64 - hr
    return 1 if W <= s0 else 1 + (W - s0 + s - 1) // s


def _pack(kern, sets, datas, set_of, caps=(1 << 24, 1 << 22, 1 << 22, 1 << 22), qcap=64):
    L = rt.load_library()
    ks = (rt.ServeKernel * len(kern))()
    for i, (hr, s, slots, n_out, bt, cap) in enumerate(kern):
        ks[i] = rt.ServeKernel(None, hr, s, slots, n_out, bt, cap)
    ss = (rt.ServeSet * len(sets))(*[rt.ServeSet(a, b) for a, b in sets])
    jobs = np.zeros(len(datas), dtype=rt.SERVE_JOB_DTYPE)
    jobs["data"] = [d.ctypes.data for d in datas]
    jobs["n_bytes"] = [d.size for d in datas]
    jobs["set"] = set_of
    stage = np.zeros(caps[0], dtype=np.uint8)
    lay = rt.ServeLayout()
    qs = (rt.ServeQueue * qcap)()
    maxk = max(b for _, b in sets)
    job_result = np.zeros(max(1, len(datas) * maxk), dtype=np.uint64)
    rc = L.bg_serve_pack(ss, len(sets), ks, len(kern), jobs.ctypes.data, len(datas), stage.ctypes.data, caps[0],
                         DEV_BASIS, caps[1], DEV_RESULT, caps[2], DEV_STATUS, caps[3], ctypes.byref(lay), qs, qcap,
                         job_result.ctypes.data, maxk)
    return rc, lay, qs, stage, job_result, maxk


def test_pack_layout_matches_its_definition():
    rng = random.Random(5)
    g = np.random.default_rng(5)
    # (hr, s, status slots, n_out, block threads, grid cap)
    kern = [(1, 62, 3, 2, 256, 40), (2, 60, 1, 1, 256, 40), (1, 61, 7, 4, 128, 8)]
    sets = [(0, 2), (2, 1)]
    sizes = [0, 1, 63, 64, 65, 4000, 300_001, 17, 9000]
    datas = [g.integers(0, 256, size=n, dtype=np.uint8) for n in sizes]
    set_of = [rng.randrange(2) for _ in sizes]
    rc, lay, qs, stage, job_result, maxk = _pack(kern, sets, datas, set_of)
    assert rc == 0, rt.load_library().bg_last_error()
    words = [(n + 1 + 63) // 64 for n in sizes]
    # input bytes, 64-byte aligned per job, zero padded
    off = 0
    td = np.frombuffer(stage[lay.tdesc_off:lay.tdesc_off + TDESC.itemsize * len(sizes)].tobytes(), dtype=TDESC)
    ws = woff = 0
    for j, d in enumerate(datas):
        assert np.array_equal(stage[off:off + d.size], d)
        pad = (-d.size) % 64
        assert not stage[off + d.size:off + d.size + pad].any()
        assert (td[j]["byte_off"], td[j]["n_bytes"], td[j]["n_words"], td[j]["word_start"], td[j]["basis_off"]) == \
            (off, d.size, words[j], ws, woff)
        off += d.size + pad
        ws += words[j]
        woff += 8 * words[j]
    assert lay.basis_words == 8 * ws
    # queues: sets in order of their first job, kernels in set order
    order = list(dict.fromkeys(set_of))
    expect_q = [(s, k) for s in order for k in range(sets[s][1])]
    assert lay.n_queues == len(expect_q)
    res = stat = 0
    for q, (s, kk) in zip(qs, expect_q):
        kid = sets[s][0] + kk
        hr, S, slots, n_out, bt, cap = kern[kid]
        assert q.kernel == kid
        members = [j for j in range(len(sizes)) if set_of[j] == s]
        assert q.n_jobs == len(members)
        jt = np.frombuffer(stage[q.jobs_off:q.jobs_off + JOB.itemsize * q.n_jobs].tobytes(), dtype=JOB)
        tile = 0
        tj_expect = []
        for e, j in enumerate(members):
            T = tiles(words[j], hr, S)
            assert jt[e]["basis"] == DEV_BASIS + 8 * td[j]["basis_off"]
            assert jt[e]["result"] == DEV_RESULT + 8 * res
            assert jt[e]["status"] == DEV_STATUS + 4 * stat
            assert (jt[e]["n_words"], jt[e]["n_carry"], jt[e]["tile_start"], jt[e]["n_tiles"]) == \
                (words[j], slots, tile, T)
            assert job_result[j * maxk + kk] == res
            tj_expect += [e] * T
            res += words[j] * n_out
            stat += T * slots
            tile += T
        assert q.total_tiles == tile
        tj = np.frombuffer(stage[q.tile_job_off:q.tile_job_off + 4 * tile].tobytes(), dtype=np.uint32)
        assert tj.tolist() == tj_expect
        assert q.blocks == max(1, min(cap, -(-tile // (bt // 64))))
    assert lay.result_words == res and lay.status_words == stat
    assert lay.stage_bytes % 256 == 0 and lay.stage_bytes <= stage.size


def test_pack_reports_the_sizes_it_needs():
    kern = [(1, 62, 2, 1, 256, 10)]
    datas = [np.ones(100_000, dtype=np.uint8)]
    rc, lay, *_ = _pack(kern, [(0, 1)], datas, [0], caps=(1024, 1 << 20, 1 << 20, 1 << 20))
    assert rc == rt.ENOSPACE
    assert lay.stage_bytes > 100_000 and lay.result_words == (100_001 + 63) // 64
    rc2, lay2, *_ = _pack(kern, [(0, 1)], datas, [0], caps=(lay.stage_bytes, lay.basis_words, lay.result_words,
                                                            lay.status_words))
    assert rc2 == 0 and lay2.stage_bytes == lay.stage_bytes


def test_pack_rejects_bad_input():
    kern = [(1, 62, 2, 1, 256, 10)]
    rc, *_ = _pack(kern, [(0, 1)], [np.ones(10, dtype=np.uint8)], [1])
    assert rc == 1001  # BG_EINVAL: unknown set
    rc, *_ = _pack([(1, 62, 2, 1, 100, 10)], [(0, 1)], [np.ones(10, dtype=np.uint8)], [0])
    assert rc == 1001  # block size not a multiple of 64
