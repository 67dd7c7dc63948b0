"""Shared GPU-parity helpers: run pattern groups through the hip path and
through the oracle (pre-pass IR, whole-stream C executor) on the same bytes."""
import numpy as np

from bitgen_amd import compiler
from bitgen_amd.bitstream import transpose_bytes
from oracle import oracle as O


def run_hip(groups, data, device, options=None, ks=None, repeat=1):
    import torch

    cs = None
    if ks is None:
        cs = compiler.compile_pattern_set(groups, options)
        ks = cs.load(device.index or 0)
    data = np.ascontiguousarray(data, dtype=np.uint8)
    x = torch.from_numpy(data).to(device).reshape(1, 1, -1)
    basis, _ = transpose_bytes(x, data.size)
    W = basis.size(-1)
    result = torch.zeros((1, 1, ks.n_groups, W), dtype=torch.uint64, device=device)
    status = ks.alloc_status(1, W)
    outs = []
    for _ in range(repeat):
        result.fill_(0x5A)  # stale garbage must be overwritten
        ks.run(basis, result, status)
        torch.cuda.synchronize()
        ks.check_errors()
        outs.append(result.cpu().numpy()[0, 0].copy())
    return outs, ks


def run_oracle(groups, data, options=None):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    W = O.n_words_for(data.size)
    basis = O.transpose(data, W)
    res = []
    for regs in groups:
        pre, pre_vnm, _, _ = compiler.group_programs(regs, options)
        res.append(O.mask_to_positions(O.run(pre, pre_vnm, basis), data.size + 1))
    return np.stack(res)


def assert_parity(groups, data, device, options=None, repeat=1, label=""):
    outs, ks = run_hip(groups, data, device, options, repeat=repeat)
    ref = run_oracle(groups, data, options)
    n = np.asarray(data).size
    for r, got in enumerate(outs):
        got = O.mask_to_positions(got, n + 1)
        for g in range(len(groups)):
            if not np.array_equal(got[g], ref[g]):
                diff = np.nonzero(got[g] != ref[g])[0]
                raise AssertionError(
                    f"{label} group {g} rep {r}: {len(diff)} words differ (first word {diff[0]}), "
                    f"popcount hip {O.popcount(got[g])} oracle {O.popcount(ref[g])}; regexes {groups[g]}")
    ks.free()
    return ref
