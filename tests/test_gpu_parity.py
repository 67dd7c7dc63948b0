"""Bit-exact parity of the hip path against the oracle (GPU)."""
import random

import numpy as np
import pytest

from bitgen_amd import workloads
from oracle import oracle as O
from tests.gpu_helpers import assert_parity

pytestmark = pytest.mark.gpu

# unbounded repeats over single classes become MatchStar (fused kernels);
# repeats of longer sub-expressions (a(bc)*d, (ab)+c, (a|bb)+c) are while
# loops, which run at stream level
FUSED_POOL = ["ab", "a[bc]+d", "x.*y", "^q", "z$", "(ab|cd){1,3}e", "b[^\\n]*c", "a(bc){0,3}d", "c?d",
              "[0-9]{2,4}", "abcdx", "q(a|b)z*[xy]{2}", "a.{0,3}b", "(a|bb){1,4}c", "dd[^a]*c",
              "^a[0-9]*$", "x[yz]+q*0"]
STREAM_POOL = ["a(bc)*d", "(ab)+c", "(a|bb)+c"]
POOL = FUSED_POOL + STREAM_POOL


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 127, 128, 4095, 4096, 4097, 100003, 1 << 20])
def test_transpose_matches_oracle(gpu, n):
    import torch
    from bitgen_amd.bitstream import transpose_bytes

    rng = np.random.default_rng(n)
    data = rng.integers(0, 256, size=n, dtype=np.uint8)
    x = torch.from_numpy(data).to(gpu).reshape(1, 1, -1) if n else torch.zeros((1, 1, 0), dtype=torch.uint8, device=gpu)
    basis, _ = transpose_bytes(x, n)
    torch.cuda.synchronize()
    got = basis.cpu().numpy()[0, 0]
    ref = O.transpose(data, got.shape[-1])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("seed", range(12))
def test_random_groups(gpu, seed):
    rng = random.Random(seed)
    alpha = np.frombuffer(b"abcdxyzq0123\n", dtype=np.uint8)
    n = rng.choice([50, 1000, 4000, 9000, 70000, 300001])
    data = np.random.default_rng(seed).choice(alpha, size=n)
    groups = [rng.sample(POOL, rng.randint(1, 4)) for _ in range(rng.randint(1, 3))]
    print(f"seed={seed} n={n} groups={groups}", flush=True)
    assert_parity(groups, data, gpu, repeat=2, label=f"seed {seed}")


@pytest.mark.parametrize("seed,per_kernel,min_waves",
                         [(s, k, 0) for s in range(6) for k in (2, 3)] + [(s, 4, 5) for s in range(6)])
def test_merged_group_kernels(gpu, seed, per_kernel, min_waves):
    """Several groups merged into one kernel: every group's stream is still
    its own, bit for bit, at every tile of long and short inputs.  The
    4-group cases also force 5 waves per SIMD (the register budget the
    benchmark's shape spills under)."""
    rng = random.Random(50 + seed)
    alpha = np.frombuffer(b"abcdxyzq0123\n", dtype=np.uint8)
    n = rng.choice([50, 4000, 70000, 300001])
    data = np.random.default_rng(50 + seed).choice(alpha, size=n)
    groups = [rng.sample(POOL, rng.randint(1, 4)) for _ in range(rng.randint(2, 5) + (3 if per_kernel == 4 else 0))]
    print(f"seed={seed} per_kernel={per_kernel} min_waves={min_waves} n={n} groups={groups}", flush=True)
    assert_parity(groups, data, gpu, options={"groups_per_kernel": per_kernel, "min_waves": min_waves}, repeat=2,
                  label=f"merged seed {seed}")


# regexes that open with a two-byte literal stage (bigram filters apply), and
# some that do not
BIGRAM_POOL = ["hello", "ab[0-9]+c", "x[yz]q", "q0q0", "[ab]cd", "zz[^\\n]*yy", "ca.*bd", "a.{0,3}b",
               "dcba(xy|zq)", "c?abc", "^abc", "xyz$", "bb[0-9]{2,4}c", "\\x81\\x02", "[0-9]{2,4}"]


@pytest.mark.parametrize("seed,per_kernel,min_waves,shared",
                         [(s, k, 0, 0) for s in range(8) for k in (1, 2)] + [(s, 4, 5, 0) for s in range(8)]
                         + [(s, k, 5, 1) for s in range(8) for k in (1, 4)])
def test_bigram_filter_groups(gpu, seed, per_kernel, min_waves, shared):
    """Per-tile byte-pair filters in front of literal first stages: tiles
    that hold a stage's pair and tiles that do not (printable text with
    planted literals, bytes above 127 whose pairs share a hash with ASCII
    pairs), single and merged kernels, the pair tables built in each kernel
    or once per set by the filter pre-pass (shared), every group bit-exact."""
    rng = random.Random(200 + seed)
    g = np.random.default_rng(200 + seed)
    n = rng.choice([60, 5000, 70000, 300001])
    alpha = np.frombuffer(bytes(range(32, 127)) + b"\n\x81\x82\x02", dtype=np.uint8)
    data = g.choice(alpha, size=n)
    for _ in range(rng.randint(0, 12)):
        lit = np.frombuffer(rng.choice([b"hello", b"ab12c", b"xyq", b"q0q0", b"acd", b"zzqyy", b"\x81\x02"]),
                            dtype=np.uint8)
        off = rng.randrange(0, max(1, n - lit.size))
        data[off:off + lit.size] = lit[:n - off]
    groups = [rng.sample(BIGRAM_POOL, rng.randint(1, 5)) for _ in range(rng.randint(1, 4) + (4 if per_kernel == 4 else 0))]
    print(f"seed={seed} per_kernel={per_kernel} min_waves={min_waves} shared={shared} n={n} groups={groups}",
          flush=True)
    assert_parity(groups, data, gpu, options={"bigram_filter": 1, "groups_per_kernel": per_kernel,
                                              "min_waves": min_waves, "shared_filter": shared},
                  repeat=2, label=f"bigram seed {seed}")


@pytest.mark.parametrize("per_kernel", [1, 2])
def test_bigram_filter_after_star(gpu, per_kernel):
    """Filters behind a star (the stage after `.*` tests a pair the text
    rarely holds): the star's carry op still runs in every tile, so every
    tile publishes its carry status and the match stretches across tiles."""
    rng = np.random.default_rng(9)
    n = 400_000
    alpha = np.frombuffer(b"abcexyzq0123 ", dtype=np.uint8)  # no 'd', no LF
    data = rng.choice(alpha, size=n)
    for off, lit in ((1000, b"ca"), (150_000, b"bd"), (200_000, b"cab"), (390_000, b"xbd")):
        data[off:off + len(lit)] = np.frombuffer(lit, dtype=np.uint8)
    data[-1] = ord("\n")
    groups = [["ca.*bd", "x[yz]q"], ["zz[^\\n]*yy", "q0q0"], ["a[bc]+e.*bd"]]
    ref = assert_parity(groups, data, gpu, options={"bigram_filter": 1, "groups_per_kernel": per_kernel},
                        repeat=2, label=f"bigram after star gpk {per_kernel}")
    assert O.popcount(ref[0]) == 1


def test_plain_small_workload(gpu):
    regexes, n_groups, buf = workloads.plain_small()
    from bitgen_amd.compiler import group_regex_in_list

    groups = group_regex_in_list(regexes, n_groups)
    ref = assert_parity(groups, buf, gpu, label="plain_small")
    assert O.popcount(ref) > 0


@pytest.mark.parametrize("settle", ["inline", "launch", "defer"])
def test_carry_crosses_many_tiles(gpu, settle):
    # one 300 KB line: MatchStar/ScanThru carries and the line-end projection
    # propagate through ~75 tiles
    body = np.full(300_000, ord("x"), dtype=np.uint8)
    body[0] = ord("a")
    body[-2] = ord("b")
    body[-1] = ord("\n")
    body[150_000] = ord("a")
    groups = [["a.*b", "a[^\\n]*b"], ["ax*b"], ["q"]]
    ref = assert_parity(groups, body, gpu, options={"carry_settle": settle}, repeat=2, label=f"long line {settle}")
    assert O.popcount(ref[0]) == 1 and O.popcount(ref[1]) == 1 and O.popcount(ref[2]) == 0


CARRY_FORMS = {"speculate": {"lookback": "speculate", "carry_settle": "inline"},
               "wait": {"lookback": "wait"},
               "launch": {"lookback": "speculate", "carry_settle": "launch"},
               "defer": {"lookback": "speculate", "carry_settle": "defer"}}


@pytest.mark.parametrize("lookback", list(CARRY_FORMS))
def test_carry_runs_through_thousands_of_tiles(gpu, lookback):
    """A 24 MiB line (no LF until the end): the line-end projection and the
    stars propagate through ~6,000 tiles, so every tile's carry-in is decided
    by a predecessor thousands of tiles back (the look-back reads 64 statuses
    per step).  Matches start near the front, the middle and the end.  In
    the two-launch form every tile publishes an aggregate in the sweep and
    the settle resolves the whole chain."""
    n = 24 * 2**20
    body = np.full(n, ord("x"), dtype=np.uint8)
    body[[5, n // 2, n - 300]] = ord("a")
    body[[n // 3, n - 10]] = ord("b")
    body[-1] = ord("\n")
    groups = [["a.*b", "a[^\n]*b"], ["ax*b"], ["q"], ["x{3}b"]]
    print(f"n={n} lookback={lookback}", flush=True)
    ref = assert_parity(groups, body, gpu, options=CARRY_FORMS[lookback], repeat=2, label=f"24 MiB line {lookback}")
    assert [O.popcount(r) for r in ref] == [1, 1, 0, 1]


@pytest.mark.parametrize("seed", range(6))
def test_tile_scope_replays(gpu, seed):
    """Replays that recompute the whole tile (basis loads, class header,
    byte-pair table) instead of keeping the header live across passes:
    star-heavy groups on long and short lines, merged and with filters."""
    rng = random.Random(300 + seed)
    alpha = np.frombuffer(b"abcdxyzq0123\n", dtype=np.uint8)
    p = [0.95 / 12] * 12 + [0.05] if seed % 2 else [1 / 13] * 13
    n = rng.choice([9000, 70000, 300001])
    data = np.random.default_rng(300 + seed).choice(alpha, size=n, p=p)
    groups = [rng.sample(POOL, rng.randint(1, 4)) for _ in range(rng.randint(1, 4))]
    opts = {"replay_scope": "tile", "carry_settle": "inline", "groups_per_kernel": 1 + seed % 3,
            "bigram_filter": seed % 2}
    print(f"seed={seed} n={n} opts={opts} groups={groups}", flush=True)
    assert_parity(groups, data, gpu, options=opts, repeat=2, label=f"tile scope seed {seed}")


def test_long_shifts_use_multi_word_halo(gpu):
    rng = np.random.default_rng(3)
    data = rng.choice(np.frombuffer(b"abcx\n", dtype=np.uint8), size=200_000, p=[.2, .2, .2, .39, .01])
    groups = [["a.{0,100}b"], ["a.{150}b", "c[ab]{70,90}c"], ["ab.{0,40}ba.{0,40}ab.{0,40}c"]]
    assert_parity(groups, data, gpu, label="long shifts")


@pytest.mark.parametrize("seed", range(4))
def test_waiting_lookback_matches_oracle(gpu, seed):
    """The waiting look-back (every carry op waits for its predecessor), the
    form used for programs with more than 64 carry ops, on star-heavy groups
    with long and short lines."""
    rng = random.Random(100 + seed)
    alpha = np.frombuffer(b"abcdxyzq0123\n", dtype=np.uint8)
    p = [0.9 / 12] * 12 + [0.1] if seed % 2 else [1 / 13] * 13
    data = np.random.default_rng(seed).choice(alpha, size=rng.choice([9000, 300001]), p=p)
    groups = [["a.*b", "x[^\\n]*y", "q(a|b)z*[xy]{2}"], ["ab[0-9]*c", "d.*3"]]
    print(f"seed={seed} n={data.size}", flush=True)
    assert_parity(groups, data, gpu, options={"lookback": "wait"}, label=f"wait seed {seed}")


def test_more_than_64_carry_ops(gpu):
    """A group whose program has more than 64 carry ops (stars of 70
    regexes) runs with the waiting look-back."""
    from bitgen_amd import compiler
    from bitgen_amd.codegen import GroupProgram

    regs = [f"{chr(97 + i % 26)}{i % 10}.*{chr(97 + (i * 7) % 26)}" for i in range(70)]
    _, _, post, vnm = compiler.group_programs(regs)
    assert GroupProgram(0, post, vnm, regs).n_carry > 64
    rng = np.random.default_rng(5)
    alpha = np.frombuffer(b"abcdefghij0123456789 \n", dtype=np.uint8)
    data = rng.choice(alpha, size=150_000)
    assert_parity([regs], data, gpu, label="70 stars")


@pytest.mark.parametrize("settle", ["launch", "defer"])
@pytest.mark.parametrize("seed", range(10))
def test_phased_settle(gpu, seed, settle):
    """carry_settle="launch": a sweep launch that never waits at the end of a
    tile, then a settle launch that replays only tiles whose assumed
    carry-ins were wrong or whose sweep aborted (deferred statuses);
    carry_settle="defer": the same phases in one launch, each tile settled
    after its wave's next pass.  Star- and loop-heavy groups, short and long
    lines, single, merged and filtered kernels, stream-level groups; each
    run twice (the second run's epoch must not see the first run's
    statuses)."""
    rng = random.Random(400 + seed)
    alpha = np.frombuffer(b"abcdxyzq0123\n", dtype=np.uint8)
    p = [0.96 / 12] * 12 + [0.04] if seed % 2 else [1 / 13] * 13
    n = rng.choice([70, 9000, 70000, 300001, 2_000_003])
    data = np.random.default_rng(400 + seed).choice(alpha, size=n, p=p)
    for _ in range(rng.randint(0, 6)):  # long runs that carries cross
        run = rng.choice([b"bc", b"ab", b"x", b"a"]) * rng.randint(100, 20000)
        off = rng.randrange(0, n)
        seg = np.frombuffer(run[:n - off], dtype=np.uint8)
        data[off:off + seg.size] = seg
    pool = POOL if seed % 3 == 0 else FUSED_POOL
    groups = [rng.sample(pool, rng.randint(1, 4)) for _ in range(rng.randint(1, 6))]
    opts = {"carry_settle": settle, "groups_per_kernel": 1 + seed % 4 if pool is FUSED_POOL else 1,
            "bigram_filter": seed % 2 if pool is FUSED_POOL else 0, "min_waves": 5 if seed % 4 == 3 else 0}
    print(f"seed={seed} n={n} opts={opts} groups={groups}", flush=True)
    assert_parity(groups, data, gpu, options=opts, repeat=2, label=f"{settle} seed {seed}")


def test_shared_filter_kernels_refuse_launch_without_answers(gpu):
    """A kernel compiled with shared byte-pair filters is refused (an error,
    before anything is launched) by bg_group_launch without the answer
    buffer and by bg_group_launch_static; the set's own run (pre-pass, then
    kernels) is bit-exact."""
    import torch

    from bitgen_amd import compiler
    from bitgen_amd.backend.hip import runtime as rt

    groups = [["hello", "ab[0-9]+c"], ["x[yz]q", "zz[^\\n]*yy"]]
    data = np.frombuffer(b"hello ab12c xyq zzkkyy\n" * 300, dtype=np.uint8)
    assert_parity(groups, data, gpu, options={"bigram_filter": 1, "shared_filter": 1}, repeat=2, label="shared")
    ks = compiler.compile_pattern_set(groups, {"bigram_filter": 1, "shared_filter": 1}).load(0)
    buf = torch.zeros(8 * 64, dtype=torch.uint64, device=gpu)
    status = ks.alloc_status(1, 64)
    args = rt.GroupArgs()
    args.basis = args.result = buf.data_ptr()
    args.status = status.data_ptr()
    args.ctl = ks.ctl[0].data_ptr()
    args.n_words, args.n_tiles, args.n_carry = 64, 1, ks.metas[0]["status_slots"]
    with pytest.raises(rt.HipRuntimeError, match="shared byte-pair answers"):
        rt.group_launch(ks.kernels[0], args, rt.current_stream_ptr())
    flag = torch.zeros(1, dtype=torch.int32, device=gpu)
    with pytest.raises(rt.HipRuntimeError, match="shared byte-pair filters"):
        rt.group_launch_static(ks.kernels[0], buf.data_ptr(), buf.data_ptr(), 64, 1, 1, flag.data_ptr(),
                               rt.current_stream_ptr())
    torch.cuda.synchronize()
    ks.free()
