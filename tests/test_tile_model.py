"""The generator's tile geometry against the oracle, on the CPU.

tests/tile_model.py computes each group's streams tile by tile exactly as
the generated kernel partitions them (halo words, carry entry lanes, window
edges, zero tests and byte-pair filters over the window).  Equality with the
oracle's whole-stream result shows the halo and entry lanes the generator
chose are large enough -- the property GPU parity rests on -- for inputs
built to stress tile boundaries:

* matches planted so that they end in every lane around the start and end
  of a tile's output range, with their line ends in the next tiles;
* the IDS pattern set over a 1 MiB prefix, in every autotuner shape;
* a regression case: a 21-byte literal ending in the last output lane of a
  tile, whose line-end projection carries into the next tile (its operand
  words in the successor's entry lane were not all exact).
"""
import random

import numpy as np
import pytest

from bitgen_amd import autotune, compiler, workloads
from oracle import oracle as O
from tests.tile_model import run_tiles


def _oracle(groups, data):
    basis = O.transpose(data, O.n_words_for(data.size))
    out = []
    for g in groups:
        pre, vnm, _, _ = compiler.group_programs(g)
        out.append(O.mask_to_positions(O.run(pre, vnm, basis), data.size + 1))
    return np.stack(out), basis


def _model(groups, data, options, basis):
    o = dict(compiler.DEFAULT_OPTIONS)
    o.update(options or {})
    from bitgen_amd.codegen.gen_stream import StreamGroup
    from tests.test_stream_plan import run_plan_cpu

    rows = []
    for prog in compiler.kernel_programs(groups, o):
        if isinstance(prog, StreamGroup):
            streams = [run_plan_cpu(prog, basis)]
        else:
            streams = run_tiles(prog, basis)
        rows.extend(O.mask_to_positions(s, data.size + 1) for s in streams)
    return np.stack(rows)


def _assert_model(groups, data, options=None, label=""):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    ref, basis = _oracle(groups, data)
    got = _model(groups, data, options, basis)
    for g in range(len(groups)):
        if not np.array_equal(got[g], ref[g]):
            diff = np.nonzero(got[g] != ref[g])[0]
            raise AssertionError(f"{label}: group {g} differs in {diff.size} words (first {diff[0]}); "
                                 f"model {O.popcount(got[g])} oracle {O.popcount(ref[g])}; {groups[g]}")
    return ref


def _tile_starts(prog, W):
    """First output word of every tile after tile 0."""
    s0 = 64 - prog.HR
    return [s0 + (t - 1) * prog.S for t in range(1, prog.n_tiles(W))]


LITERALS = ["]8CSJ[%?VYNP}[v Tk7';", "hello", "abcdefghijklmnopqrstuvwxyz0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZ=+",
            "qz"]


@pytest.mark.parametrize("lit", LITERALS, ids=lambda s: f"len{len(s)}")
def test_literal_ends_around_every_tile_boundary(lit):
    """A literal ending at each offset -70..+70 positions around the first
    output position of consecutive tiles, its line end 100-900 bytes later:
    the line-end projection (a ScanThru) carries across tiles."""
    regs = [compiler_escape(lit), "x[0-9]+y", "k.*m"]
    groups = [regs]
    _, _, post, vnm = compiler.group_programs(regs)
    from bitgen_amd.codegen.gen_hip import GroupProgram

    prog = GroupProgram(0, post, vnm, regs)
    n = 64 * 64 * 40
    W = O.n_words_for(n)
    rng = random.Random(len(lit))
    data = np.frombuffer(bytes(rng.choice(b"abcdefgh ") for _ in range(n)), dtype=np.uint8).copy()
    lb = np.frombuffer(lit.encode(), dtype=np.uint8)
    offsets = list(range(-70, 71, 7)) + [-1, 0, 1, 7, 8, 9, 63, 64, 65]
    starts = _tile_starts(prog, W)[1:-2]
    for i, ts in enumerate(starts):
        end = 64 * ts + offsets[i % len(offsets)]  # match-end marker position
        a = end - lb.size
        if a < 0 or end + 1000 >= n:
            continue
        data[a:end] = lb
        data[end + rng.randrange(100, 900)] = ord("\n")
    ref = _assert_model(groups, data, label=f"literal {lit!r}")
    assert O.popcount(ref) > 5


def compiler_escape(lit):
    return workloads.lit(lit.encode())


@pytest.mark.parametrize("seed", range(6))
def test_random_groups_all_shapes(seed):
    """Random groups from the GPU tests' pools, single and merged kernels,
    with and without byte-pair filters, over text with long and short lines."""
    from tests.test_gpu_parity import BIGRAM_POOL, POOL

    rng = random.Random(700 + seed)
    alpha = np.frombuffer(b"abcdxyzq0123hel \n", dtype=np.uint8)
    p = None if seed % 2 else [0.98 / 16] * 16 + [0.02]
    n = rng.choice([5000, 40000, 120000])
    data = np.random.default_rng(700 + seed).choice(alpha, size=n, p=p)
    pool = POOL + BIGRAM_POOL
    groups = [rng.sample(pool, rng.randint(1, 5)) for _ in range(rng.randint(1, 5))]
    for opts in ({}, {"groups_per_kernel": 2}, {"groups_per_kernel": 4, "bigram_filter": 1},
                 {"groups_per_kernel": 2, "bigram_filter": 1, "shared_filter": 1}):
        _assert_model(groups, data, opts, label=f"seed {seed} {opts}")


@pytest.fixture(scope="module")
def ids_prefix():
    regexes, n_groups, _ = workloads.ids(n_bytes=1 << 16)
    groups = compiler.group_regex_in_list(regexes, n_groups)
    # the ids byte mix with planted matches of the set, 1 MiB
    data = workloads.text_ids_fast(1 << 20, 41)
    workloads.plant(data, regexes, 400, random.Random(41), avoid_overlap=False)
    return groups, data


@pytest.mark.parametrize("cand", autotune.candidates(16), ids=lambda c: "gpk%d_mw%d_q%d_bg%d_bt%d" % c)
def test_ids_prefix_every_tuner_shape(ids_prefix, cand):
    groups, data = ids_prefix
    ref = _assert_model(groups, data, autotune.options_of(cand), label=f"IDS 1 MiB {cand}")
    assert O.popcount(ref) > 100
