"""Per-tile dynamic statistics (inst_stats_dyn, the reference's
--pass-inst-stats-dyn) on the GPU: the count of zero / byte-pair tests that
fell through in each tile's exact pass equals the CPU tile model's count for
the same program and bytes, in every carry settle form; every tile ran at
least one pass; results stay bit-exact."""
import random

import numpy as np
import pytest

from bitgen_amd import compiler
from oracle import oracle as O
from tests.gpu_helpers import run_oracle
from tests.tile_model import run_tiles

pytestmark = pytest.mark.gpu

POOL = ["ab[^\\n]*cd", "x.*y", "hello", "q[0-9]+z", "zz[0-9]{2,4}y", "a.{0,3}b", "dcba(xy|zq)", "c?abc",
        "bb[^a]*c", "xyz$", "^abc"]


@pytest.mark.parametrize("settle", ["inline", "launch", "defer"])
@pytest.mark.parametrize("seed", range(3))
def test_dyn_stats_match_tile_model(gpu, settle, seed):
    import torch

    from bitgen_amd.bitstream import transpose_bytes

    rng = random.Random(500 + seed)
    g = np.random.default_rng(500 + seed)
    n = rng.choice([5000, 70000, 300001])
    alpha = np.frombuffer(b"abcdxyzqhelo0123 \n", dtype=np.uint8)
    data = g.choice(alpha, size=n)
    for _ in range(rng.randint(1, 10)):
        lit = np.frombuffer(rng.choice([b"hello", b"abxcd", b"q12z", b"zz123y", b"dcbaxy"]), dtype=np.uint8)
        off = rng.randrange(0, n - lit.size)
        data[off:off + lit.size] = lit
    groups = [rng.sample(POOL, rng.randint(2, 4)) for _ in range(rng.randint(1, 4))]
    opts = dict(compiler.DEFAULT_OPTIONS, inst_stats_dyn=1, carry_settle=settle, bigram_filter=seed % 2,
                groups_per_kernel=1 + seed % 2)
    print(f"seed={seed} settle={settle} n={n} groups={groups}", flush=True)
    cs = compiler.compile_pattern_set(groups, opts)
    ks = cs.load(0)
    assert all(m["dyn_stats"] for m in ks.metas)
    dev = torch.device("cuda", 0)
    x = torch.from_numpy(data).to(dev).reshape(1, 1, -1)
    basis, _ = transpose_bytes(x, n)
    W = basis.size(-1)
    result = torch.zeros((1, 1, ks.n_groups, W), dtype=torch.uint64, device=dev)
    status = ks.alloc_status(1, W)
    for _ in range(2):  # the second run overwrites the first run's words
        ks.run(basis, result, status)
    torch.cuda.synchronize()
    ks.check_errors()
    dyn = ks.dyn_stats(status, W).cpu().numpy()[0]
    ref = run_oracle(groups, data)
    got = O.mask_to_positions(result.cpu().numpy()[0, 0], n + 1)
    for gi in range(len(groups)):
        assert np.array_equal(got[gi], ref[gi]), (settle, gi)
    host_basis = O.transpose(data, W)
    progs = compiler.kernel_programs(groups, opts)
    for k, prog in enumerate(progs):
        stats = []
        run_tiles(prog, host_basis, stats)
        T = len(stats)
        assert T == prog.n_tiles(W)
        assert list(dyn[k, 1:1 + T]) == stats, (settle, k)
        assert T <= dyn[k, 0] <= 3 * T  # one pass per tile, a few replays at most
    ks.free()
