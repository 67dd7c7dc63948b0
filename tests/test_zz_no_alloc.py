"""bg_group_launch_static (the launch without a workspace) against
bg_group_launch on the same kernels and buffers, through the C ABI (GPU).

Two calls in a row, the second over a smaller buffer: the static scratch
keeps the first call's status words, which the second must not read (they
carry an older epoch).  Both calls' streams equal bg_group_launch's bit for
bit, and equal the oracle's; the error flag the call hands back is 0."""
import numpy as np
import pytest

from bitgen_amd import compiler
from bitgen_amd.backend.hip import runtime as rt
from bitgen_amd.bitstream import transpose_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GROUPS = [["ab[0-9]{1,3}c", "x[^\\n]*yz", "hello"], ["w.*d", "q(a|b)z*[xy]{2}"]]


def _streams_both_ways(gpu, ks, data):
    import torch

    x = torch.from_numpy(np.ascontiguousarray(data)).to(gpu).reshape(1, 1, -1)
    basis, _ = transpose_bytes(x, data.size)
    W = basis.size(-1)
    ref = torch.full((ks.n_groups, W), 0x5A, dtype=torch.uint64, device=gpu)
    status = ks.alloc_status(1, W)
    stream = rt.current_stream_ptr()
    for k in range(ks.n_kernels):
        ks.launch(k, basis.data_ptr(), ref.data_ptr() + ks.first_group[k] * W * 8,
                  status.data_ptr() + k * status.size(2) * 4, W, stream)
    got = torch.full((ks.n_groups, W), 0x3C, dtype=torch.uint64, device=gpu)
    flags = torch.full((ks.n_kernels,), 7, dtype=torch.int32, device=gpu)
    for k, m in enumerate(ks.metas):
        rt.group_launch_static(ks.kernels[k], basis.data_ptr(), got.data_ptr() + ks.first_group[k] * W * 8, W,
                               ks.n_tiles(m, W), m["status_slots"], flags.data_ptr() + 4 * k, stream)
    torch.cuda.synchronize()
    ks.check_errors()
    return ref.cpu().numpy(), got.cpu().numpy(), flags.cpu().numpy()


def test_static_launch_matches_workspace_launch_twice(gpu):
    ks = compiler.compile_pattern_set(GROUPS, {"groups_per_kernel": 1}).load(0)
    rng = np.random.default_rng(31)
    alpha = np.frombuffer(b"abcdxyzqhelowr0123 \n", dtype=np.uint8)
    for n in (300_001, 20_000):
        data = rng.choice(alpha, size=n)
        print(f"n={n}", flush=True)
        ref, got, flags = _streams_both_ways(gpu, ks, data)
        assert np.array_equal(got, ref), n
        assert (flags == 0).all(), flags
        basis = O.transpose(data, O.n_words_for(n))
        for g, regs in enumerate(GROUPS):
            pre, vnm, _, _ = compiler.group_programs(regs)
            want = O.mask_to_positions(O.run(pre, vnm, basis), n + 1)
            assert np.array_equal(O.mask_to_positions(got[g], n + 1), want), (n, g)
    ks.free()


def test_static_launch_rejects_oversize(gpu):
    import torch

    ks = compiler.compile_pattern_set(GROUPS, {"groups_per_kernel": 1}).load(0)
    buf = torch.zeros(8, dtype=torch.uint64, device=gpu)
    flag = torch.zeros(1, dtype=torch.int32, device=gpu)
    with pytest.raises(rt.HipRuntimeError, match="static scratch holds"):
        rt.group_launch_static(ks.kernels[0], buf.data_ptr(), buf.data_ptr(), 1, rt.STATIC_STATUS_WORDS, 2,
                               flag.data_ptr(), rt.current_stream_ptr())
    ks.free()
