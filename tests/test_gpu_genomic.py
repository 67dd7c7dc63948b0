"""Genomic motif scan (GPU): 512 seeded IUPAC motifs, half of them gapped
(m1.{0,20}m2), in 32 groups over synthetic DNA (uniform bases, runs of N, a
line feed every 80 bytes).  Bit-exact against the oracle: in full on a
2 MiB text, and on random windows of the full 1 GiB text.

The window check is exact because every match stream of these programs is
local: a match ends at most 44 bytes after it starts, and the line-end
projection only carries a match to the end of its own line (lines are 80
bytes).  A position's value therefore depends only on the 256 bytes before
it, and the oracle run on a window that starts 256 bytes earlier gives the
same words as the whole-text run."""
import concurrent.futures

import numpy as np
import pytest

from bitgen_amd import compiler, workloads
from bitgen_amd.bitstream import transpose_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CONTEXT = 256  # bytes of text before a window that its words can depend on


def _oracle_groups(groups, data):
    W = O.n_words_for(data.size)
    basis = O.transpose(data, W)
    progs = [O.flatten(*compiler.group_programs(g)[:2]) for g in groups]
    with concurrent.futures.ThreadPoolExecutor(max_workers=16) as ex:
        outs = list(ex.map(lambda p: O.run_records(p[0], p[1], basis), progs))
    return np.stack([O.mask_to_positions(o, data.size + 1) for o in outs])


def _run_gpu(gpu, groups, buf):
    import torch

    ks = compiler.compile_pattern_set(groups).load(gpu.index or 0)
    x = torch.from_numpy(buf).to(gpu).reshape(1, 1, -1)
    basis, _ = transpose_bytes(x, buf.size)
    del x
    W = basis.size(-1)
    result = torch.empty((1, 1, ks.n_groups, W), dtype=torch.uint64, device=gpu)
    status = ks.alloc_status(1, W)
    ks.run(basis, result, status)
    torch.cuda.synchronize()
    ks.check_errors()
    ks.free()
    return result[0, 0]


def test_genomic_2MiB_all_groups(gpu):
    regexes, n_groups, buf = workloads.genomic(n_bytes=2 * 2**20)
    groups = compiler.group_regex_in_list(regexes, n_groups)
    print(f"genomic 2 MiB: {len(regexes)} regexes, {len(groups)} groups", flush=True)
    got = _run_gpu(gpu, groups, buf).cpu().numpy()
    ref = _oracle_groups(groups, buf)
    got = np.stack([O.mask_to_positions(r, buf.size + 1) for r in got])
    assert O.popcount(ref) > 0
    for g in range(len(groups)):
        assert np.array_equal(got[g], ref[g]), f"group {g}"


def test_genomic_1GiB_windows(gpu):
    regexes, n_groups, buf = workloads.genomic()
    groups = compiler.group_regex_in_list(regexes, n_groups)
    print(f"genomic 1 GiB: {len(regexes)} regexes, {len(groups)} groups, {buf.size} bytes", flush=True)
    res = _run_gpu(gpu, groups, buf)
    rng = np.random.default_rng(6)
    words_per_window = 1024  # 64 KiB of positions
    W = res.size(-1)
    starts = sorted(set(int(w) for w in rng.integers(CONTEXT // 64 + 1, W - words_per_window, size=12)))
    starts.append(W - words_per_window)  # the end of the text (and the position past it)
    total = 0
    for w0 in starts:
        lo = w0 * 64 - CONTEXT
        hi = min(buf.size, (w0 + words_per_window) * 64)
        ref = _oracle_groups(groups, buf[lo:hi])
        # window word j corresponds to text word w0 + j; its positions start at
        # byte w0*64, which is position CONTEXT of the window text
        shift = CONTEXT // 64
        ref_words = ref[:, shift:shift + words_per_window]
        got = res[:, w0:w0 + words_per_window].cpu().numpy()
        if hi == buf.size:
            got = np.stack([O.mask_to_positions(r, buf.size + 1 - w0 * 64) for r in got])
            ref_words = np.stack([O.mask_to_positions(r, buf.size + 1 - w0 * 64) for r in ref_words])
        assert got.shape == ref_words.shape
        for g in range(len(groups)):
            assert np.array_equal(got[g], ref_words[g]), f"window at word {w0}, group {g}"
        total += O.popcount(ref_words)
    print(f"{len(starts)} windows, {total} match positions compared", flush=True)
    assert total > 0
