"""Pinning the oracle (CPU): closed forms, exhaustive byte values, the
per-position twin, hand-worked vectors and the reference's known answer."""
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from oracle import np_ref
from bitgen_amd.inst import (
    BsAnd, BsOr, BsXor, BsNot, BsAssign, BsAdvance, BsMatchStar, BsScanThru,
    BsStreamLoad, BsStreamStore, BsWhile, BsIf, BsSkipZero, BsLabel, ADV_LEFT,
)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_transpose_every_byte_value_at_every_bit_offset():
    # 256 values x 64 offsets: position p holds value (p * 37 + p // 64) & 255
    n = 256 * 64 + 17
    data = np.array([(p * 37 + p // 64) & 255 for p in range(n)], dtype=np.uint8)
    basis = O.transpose(data)
    bits = np.unpackbits(data[:, None], axis=1)  # [n, 8], bit 0 = MSB
    for i in range(8):
        got = np.unpackbits(basis[i].view(np.uint8), bitorder="little")[:n]
        assert np.array_equal(got, bits[:, i]), f"basis{i}"
    # the extra position(s) past the data are zero
    W = O.n_words_for(n)
    assert basis.shape == (8, W)
    tail = np.unpackbits(basis.view(np.uint8), bitorder="little").reshape(8, -1)[:, n:]
    assert not tail.any()


def _prog(ops, vnm):
    insts = [BsStreamLoad([0], 1 + i, i) for i in range(8)]
    insts += ops
    return insts


def test_advance_closed_form():
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, size=5000, dtype=np.uint8)
    basis = O.transpose(data)
    vnm = {0: "basic_stream", **{1 + i: f"basis{i}" for i in range(8)}, 20: "y", 21: "out"}
    for n in (1, 7, 63, 64, 65, 130, 4000):
        for left in (False, True):
            adv = BsAdvance([3, n], 20, ADV_LEFT) if left else BsAdvance([3, n], 20)
            insts = _prog([adv, BsStreamStore([20], 21)], vnm)
            out = O.run(insts, vnm, basis)
            x = np.unpackbits(basis[2].view(np.uint8), bitorder="little")
            y = np.zeros_like(x)
            if left:
                y[:len(x) - n] = x[n:]
            else:
                y[n:] = x[:len(x) - n]
            assert np.array_equal(np.unpackbits(out.view(np.uint8), bitorder="little"), y), (n, left)


def test_hand_worked_matchstar_scanthru_vectors():
    cases = json.load(open(os.path.join(GOLDEN, "matchstar_paper.json")))
    vnm = {0: "basic_stream", 10: "M", 11: "C", 12: "r", 13: "out"}
    for c in cases:
        L = len(c["M"])
        m = sum(1 << i for i, ch in enumerate(c["M"]) if ch == "1")
        cc = sum(1 << i for i, ch in enumerate(c["C"]) if ch == "1")
        cls = BsMatchStar if c["op"] == "MatchStar" else BsScanThru
        # M and C enter as basis0 / basis1 words directly
        basis = np.zeros((8, 1), dtype=np.uint64)
        basis[0, 0], basis[1, 0] = m, cc
        insts = [BsStreamLoad([0], 10, 0), BsStreamLoad([0], 11, 1), cls([10, 11], 12), BsStreamStore([12], 13)]
        out = int(O.run(insts, vnm, basis)[0])
        got = "".join("1" if (out >> i) & 1 else "0" for i in range(L))
        assert got == c["expect"], (c, got)


# ---- random programs: word-level C oracle vs per-position twin -------------------
def random_program(rng, n_ops=30, with_loops=True):
    """Random IR over basis0..7 using every operation kind."""
    vnm = {0: "basic_stream", 9: "0", 10: "1"}
    insts = []
    for i in range(8):
        vnm[1 + i] = f"basis{i}"
        insts.append(BsStreamLoad([0], 1 + i, i))
    live = list(range(1, 9))
    nxt = [11]

    def new(name="t"):
        v = nxt[0]
        nxt[0] += 1
        vnm[v] = f"{name}{v}"
        return v

    label = [0]
    for _ in range(n_ops):
        k = rng.randrange(11 if with_loops else 9)
        a, b = rng.choice(live), rng.choice(live)
        r = new()
        if k == 0:
            insts.append(BsAnd([a, b], r))
        elif k == 1:
            insts.append(BsOr([a, b], r))
        elif k == 2:
            insts.append(BsXor([a, b], r))
        elif k == 3:
            insts.append(BsNot([a], r))
        elif k == 4:
            insts.append(BsAdvance([a, rng.choice([1, 2, 5, 63, 64, 65, 100])], r))
        elif k == 5:
            insts.append(BsAdvance([a, rng.choice([1, 3, 64, 70])], r, ADV_LEFT))
        elif k == 6:
            insts.append(BsMatchStar([a, b], r))
        elif k == 7:
            insts.append(BsScanThru([a, b], r))
        elif k == 8:
            # zero skip over an AND chain
            t = new()
            insts.append(BsAssign([9], t))
            insts.append(BsAnd([a, b], r))
            insts.append(BsSkipZero(r, label[0]))
            insts.append(BsAnd([r, rng.choice(live)], t))
            insts.append(BsLabel(label[0]))
            label[0] += 1
            live.append(t)
        elif k == 9:
            # canonical star loop of class b from marker a
            test, acc, t1, adv, nacc, nw = new("test"), new("accum"), new(), new(), new(), new()
            insts += [BsAssign([a], test), BsAssign([a], acc)]
            body = [BsAnd([test, b], t1), BsAdvance([t1, 1], adv), BsNot([acc], nacc), BsAnd([adv, nacc], nw),
                    BsOr([acc, adv], acc), BsAssign([nw], test)]
            insts.append(BsWhile(body, test))
            insts.append(BsAssign([acc], r))
        else:
            c = rng.choice(live)
            insts.append(BsIf([BsAnd([a, b], r)], [BsOr([a, b], r)], c))
        live.append(r)
    out = new("out")
    insts.append(BsStreamStore([live[-1]], out))
    return insts, vnm


@pytest.mark.parametrize("seed", range(40))
def test_word_oracle_matches_per_position_twin(seed):
    rng = random.Random(seed)
    insts, vnm = random_program(rng)
    L = rng.choice([7, 64, 65, 129, 200])
    data = bytes(rng.randrange(256) if rng.random() < 0.5 else rng.choice(b"ab") for _ in range(L))
    W = (L + 63) // 64
    basis = O.transpose(np.frombuffer(data, dtype=np.uint8), W)
    recs, nslots = O.flatten(insts, vnm)
    out = O.run_records(recs, nslots, basis)
    twin = np_ref.run_per_position(recs, nslots, np_ref.bytes_to_bits(data, 64 * W))
    assert np_ref.words_to_bits(out, 64 * W) == twin


def test_slot_reuse_does_not_change_results():
    rng = random.Random(99)
    insts, vnm = random_program(rng, n_ops=80)
    data = np.random.default_rng(5).integers(0, 256, size=3000, dtype=np.uint8)
    basis = O.transpose(data)
    recs, nslots = O.flatten(insts, vnm)
    # one slot per variable: rebuild without the allocator
    f = O._Flattener(vnm)
    f.body(insts)
    f.finish()
    keys = {}
    for r in f.recs:
        refs = (r[1], r[2]) if r[0] == O.OP["IF"] else (r[1], r[2], r[3])
        for k in refs + ((r[4],) if r[0] == O.OP["SEL"] else ()):
            if k != -1:
                keys.setdefault(k, len(keys))

    def slot(k):
        return keys[k] if k != -1 else -1

    plain = np.array([[r[0], slot(r[1]), slot(r[2]), r[3] if r[0] == O.OP["IF"] else slot(r[3]),
                       slot(r[4]) if r[0] == O.OP["SEL"] else r[4]] for r in f.recs], dtype=np.int32)
    assert nslots <= len(keys)
    assert np.array_equal(O.run_records(recs, nslots, basis), O.run_records(plain, len(keys), basis))


def test_while_fixed_point_equals_matchstar():
    rng = np.random.default_rng(11)
    data = rng.choice(np.frombuffer(b"aab\n", dtype=np.uint8), size=3000)
    basis = O.transpose(data)
    vnm = {0: "basic_stream", 9: "0", **{1 + i: f"basis{i}" for i in range(8)}}
    names = ["m", "c", "test", "accum", "t", "adv", "nacc", "new", "out", "ms", "o2"]
    ids = {n: 20 + i for i, n in enumerate(names)}
    vnm.update({v: n for n, v in ids.items()})
    g = ids
    pre = [BsStreamLoad([0], 1 + i, i) for i in range(8)]
    # m = bytes with bit 0 set ('a'), c = bytes with bit 3 clear (not LF)
    pre += [BsAssign([8], g["m"]), BsNot([5], g["c"])]
    loop = [BsAssign([g["m"]], g["test"]), BsAssign([g["m"]], g["accum"]),
            BsWhile([BsAnd([g["test"], g["c"]], g["t"]), BsAdvance([g["t"], 1], g["adv"]),
                     BsNot([g["accum"]], g["nacc"]), BsAnd([g["adv"], g["nacc"]], g["new"]),
                     BsOr([g["accum"], g["adv"]], g["accum"]), BsAssign([g["new"]], g["test"])], g["test"]),
            BsStreamStore([g["accum"]], g["out"])]
    star = [BsMatchStar([g["m"], g["c"]], g["ms"]), BsStreamStore([g["ms"]], g["o2"])]
    a = O.run(pre + loop, vnm, basis)
    b = O.run(pre + star, vnm, basis)
    assert np.array_equal(a, b)
    assert O.popcount(a) > 0


def test_snort_known_answer_count_16():
    from bitgen_amd import compiler
    from bitgen_amd.tool.pcre import read_regex_from_pcre_file

    kat = json.load(open(os.path.join(GOLDEN, "snort_kat.json")))
    regexes = read_regex_from_pcre_file(os.path.join(GOLDEN, kat["regex_file"]))
    data = np.fromfile(os.path.join(GOLDEN, kat["input_file"]), dtype=np.uint8)[:kat["input_size"]]
    groups = compiler.group_regex_in_list(regexes, kat["groups"])
    assert len(groups) == kat["groups"]
    basis = O.transpose(data)
    total = 0
    for g in groups:
        pre, vnm, _, _ = compiler.group_programs(g)
        total += O.popcount(O.mask_to_positions(O.run(pre, vnm, basis), data.size + 1))
    assert total == kat["count"]


def test_oracle_loop_guard_raises():
    # while on a stream that never empties: ~0 advanced stays nonzero forever
    vnm = {0: "basic_stream", 9: "1", 20: "test", 21: "x", 22: "out", **{1 + i: f"basis{i}" for i in range(8)}}
    insts = [BsStreamLoad([0], 1 + i, i) for i in range(8)]
    insts += [BsNot([1], 20), BsWhile([BsOr([20, 1], 21), BsNot([1], 20)], 20), BsStreamStore([20], 22)]
    data = np.zeros(100, dtype=np.uint8)
    with pytest.raises(RuntimeError):
        O.run(insts, vnm, O.transpose(data))


@pytest.fixture(scope="module")
def asan_oracle():
    import subprocess

    subprocess.run(["make", "-C", os.path.dirname(O.LIB_PATH), "oracle_asan"], check=True,
                   stdout=subprocess.DEVNULL)
    return os.path.join(os.path.dirname(O.LIB_PATH), "oracle_asan")


@pytest.mark.parametrize("seed", range(12))
def test_sanitizer_build_of_the_oracle(asan_oracle, tmp_path, seed):
    """The oracle's C code built with AddressSanitizer and UBSan (any invalid
    access or undefined behaviour aborts) gives the same words as the
    library build, on random programs with loops over lengths around word
    boundaries."""
    import struct
    import subprocess

    rng = random.Random(1000 + seed)
    insts, vnm = random_program(rng, n_ops=40)
    L = rng.choice([0, 1, 63, 64, 65, 1000, 4097])
    data = bytes(rng.randrange(256) if rng.random() < 0.5 else rng.choice(b"ab") for _ in range(L))
    W = O.n_words_for(L)
    recs, nslots = O.flatten(insts, vnm)
    job = tmp_path / "job.bin"
    job.write_bytes(struct.pack("<iiQQ", recs.shape[0], nslots, L, W) + recs.astype("<i4").tobytes() + data)
    out = tmp_path / "out.bin"
    r = subprocess.run([asan_oracle, str(job), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    got = np.frombuffer(out.read_bytes(), dtype=np.uint64)
    ref = O.run_records(recs, nslots, O.transpose(np.frombuffer(data, dtype=np.uint8), W))
    assert np.array_equal(got, ref)
