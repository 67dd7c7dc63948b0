"""Passes preserve semantics (CPU): the oracle gives the same match stream on
the pre-pass and the post-pass IR of the same group (positions 0..n)."""
import random

import numpy as np
import pytest

from bitgen_amd import compiler
from bitgen_amd.bitparser import BitstreamParser
from bitgen_amd.inst import BsInstType, walk
from bitgen_amd.passes import (
    PassManager, star_rewrite_pass, cc_advance_pass, remove_alias_pass, short_circuit_pass, graph_break_pass,
)
from bitgen_amd.tool.regex2pablo import regexes_to_pablo
from oracle import oracle as O

POOL = ["ab", "a[bc]+d", "x.*y", "^q", "z$", "(ab|cd){1,3}e", "b[^\\n]*c", "a(bc)*d", "c?d",
        "[0-9]{2,4}", "abcdx", "q(a|b)z*[xy]{2}", "a.{0,3}b", "(a|bb)+c", "dd[^a]*c", "^a[0-9]*$",
        "a.{70}b", "x[ab]{0,40}y"]


def run_both(regexes, data, flags, star_mode):
    text = regexes_to_pablo(regexes, star_mode=star_mode)
    a = BitstreamParser(backend="hip")
    pre = a.parse_text(text)
    b = BitstreamParser(backend="hip")
    post = b.parse_text(text)
    pm = PassManager()
    sr, cca, ra, sc, start, interval = flags
    pm.add_pass(star_rewrite_pass(), enabled=sr)
    pm.add_pass(cc_advance_pass(), enabled=cca)
    pm.add_pass(remove_alias_pass(), enabled=ra)
    pm.add_pass(short_circuit_pass(start=start, interval=interval), enabled=sc)
    pm.add_pass(graph_break_pass(-1))
    post, vnm = pm.run(post, b.var_name_map)
    basis = O.transpose(data)
    n = data.size + 1
    return (O.mask_to_positions(O.run(pre, a.var_name_map, basis), n),
            O.mask_to_positions(O.run(post, vnm, basis), n), post, vnm)


@pytest.mark.parametrize("seed", range(50))
def test_passes_preserve_match_streams(seed):
    rng = random.Random(seed)
    n = rng.choice([0, 10, 200, 700, 3000])
    data = np.frombuffer(bytes(rng.choice(b"abcdxyzq0123\n") for _ in range(n)), dtype=np.uint8)
    regs = rng.sample(POOL, rng.randint(1, 5))
    flags = (rng.random() < .8, rng.random() < .8, rng.random() < .9, rng.random() < .8,
             rng.randint(0, 2), rng.randint(1, 3))
    mode = rng.choice(["matchstar", "while"])
    ref, got, _, _ = run_both(regs, data, flags, mode)
    assert np.array_equal(ref, got), (regs, flags, mode)


def test_star_rewrite_removes_emitted_star_loops():
    text = regexes_to_pablo(["ab[0-9]*[0-5]", "x[^\\n]*y"], star_mode="while")
    p = BitstreamParser(backend="hip")
    insts = p.parse_text(text)
    assert sum(i.type == BsInstType.WHILE for i in walk(insts)) == 2
    sr = star_rewrite_pass()
    out, _ = sr.run(insts, p.var_name_map)
    assert sr.rewritten == 2
    assert not any(i.type == BsInstType.WHILE for i in walk(out))
    assert sum(i.type == BsInstType.MATCHSTAR for i in walk(out)) >= 2


def test_star_rewrite_keeps_other_loops():
    text = regexes_to_pablo(["a(bc)*d"], star_mode="while")
    p = BitstreamParser(backend="hip")
    insts = p.parse_text(text)
    out, _ = star_rewrite_pass().run(insts, p.var_name_map)
    assert any(i.type == BsInstType.WHILE for i in walk(out))


def test_cc_advance_removes_marker_advances_and_realigns_output():
    regs = ["abcdefgh", "xy[0-9]{2}z"]
    data = np.frombuffer(b"zzabcdefghzz\nxy12z\nabcdefg\n" * 40, dtype=np.uint8)
    ref, got, post, vnm = run_both(regs, data, (True, True, True, False, 0, 1), "matchstar")
    assert np.array_equal(ref, got) and O.popcount(ref) == 80
    rights = [i for i in post if i.type == BsInstType.ADVANCE and not i.is_left]
    lefts = [i for i in post if i.type == BsInstType.ADVANCE and i.is_left]
    assert len(lefts) > len(rights)
    # the stored stream is realigned by a right shift named *_realign*
    store = post[-1]
    assert store.type == BsInstType.STREAMSTORE
    assert "realign" in vnm[store.operands[0]]


def test_short_circuit_inserts_checks_and_zero_inits():
    regs = ["abcdefghij", "klmnop"]
    data = np.frombuffer(b"abcdefghij\nklmnop\nxxxx\n", dtype=np.uint8)
    ref, got, post, vnm = run_both(regs, data, (True, True, True, True, 0, 1), "matchstar")
    assert np.array_equal(ref, got)
    skips = [i for i in post if i.type == BsInstType.SKIPZERO]
    labels = [i for i in post if i.type == BsInstType.LABEL]
    assert len(skips) >= 4 and len(labels) == len(skips)
    # a check never guards a carry operation
    pos = {id(i): k for k, i in enumerate(post)}
    lab = {i.label: pos[id(i)] for i in labels}
    for s in skips:
        body = post[pos[id(s)]:lab[s.label]]
        assert not any(i.type in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU) for i in body)


def test_short_circuit_skip_is_exact_when_forced():
    # every check fires (input without the patterns) and the output is still exact
    regs = ["abcdef", "a.{2}bcd", "qq[0-9]+r"]
    data = np.frombuffer(b"zzzz\n" * 300, dtype=np.uint8)
    ref, got, _, _ = run_both(regs, data, (True, True, True, True, 0, 1), "matchstar")
    assert np.array_equal(ref, got) and O.popcount(ref) == 0


def test_remove_alias_keeps_loop_carried_and_output():
    text = regexes_to_pablo(["a(bc)*d"], star_mode="while")
    p = BitstreamParser(backend="hip")
    insts = p.parse_text(text)
    out, vnm = remove_alias_pass().run(insts, p.var_name_map)
    names = [vnm[i.ret] for i in walk(out) if i.type == BsInstType.ASSIGN]
    assert "output.matches" in names
    assert any(n.startswith("test") for n in names) and any(n.startswith("accum") for n in names)


def test_graph_break_modes_place_breaks_like_the_reference():
    """Breaks around loops/carry ops (0), also around Advances (2), before
    every instruction but loads and stores (1); -1 leaves the program alone
    (bitgen/passes/graph_break.py:373-396)."""
    from bitgen_amd.inst import BsInstType

    _, _, post, vnm = compiler.group_programs(["ab[0-9]{1,3}c", "x[^\\n]*yz"],
                                              {"pass_cc_advanced": 0, "pass_short_circuit": 0})

    def breaks_around(insts, kinds):
        for k, inst in enumerate(insts):
            if inst.type in kinds:
                assert insts[k - 1].type == BsInstType.GRAPHBREAK and insts[k + 1].type == BsInstType.GRAPHBREAK

    n = lambda insts: sum(i.type == BsInstType.GRAPHBREAK for i in insts)  # noqa: E731
    carry = (BsInstType.MATCHSTAR, BsInstType.SCANTHRU, BsInstType.WHILE)
    m0, _ = graph_break_pass(0).run(list(post), vnm)
    breaks_around(m0, carry)
    m2, _ = graph_break_pass(2).run(list(post), vnm)
    breaks_around(m2, carry + (BsInstType.ADVANCE,))
    m1, _ = graph_break_pass(1).run(list(post), vnm)
    assert n(m1) == sum(i.type not in (BsInstType.STREAMLOAD, BsInstType.STREAMSTORE) for i in post)
    assert n(m2) > n(m0) > 0
    assert graph_break_pass(-1).run(list(post), vnm)[0] == list(post)
    with pytest.raises(ValueError):
        graph_break_pass(3)


def test_group_balancing_matches_reference_rule():
    regs = ["\\x00+", "abc", "[a-z]{3}", "abcdef", "a", "xyz"]
    groups = compiler.balance_regex_groups(regs, 2)
    # sorted (stably) by custom length -- \\xHH and [...] count as one
    # character: a 1, \\x00+ 2, abc 3, xyz 3, [a-z]{3} 4, abcdef 6 -- then
    # each to the currently lightest group (scripts/run_bitgen.py:102-116)
    assert groups == [["a", "abc", "[a-z]{3}"], ["\\x00+", "xyz", "abcdef"]]


@pytest.mark.parametrize("seed", range(20))
def test_merged_program_computes_every_group(seed):
    """Several groups merged into one program (one kernel): each store of
    the merged program gives exactly its group's own stream, and the groups
    share their basis loads and class header."""
    from bitgen_amd.codegen.gen_hip import GroupProgram

    rng = random.Random(100 + seed)
    groups = [rng.sample(POOL, rng.randint(1, 4)) for _ in range(rng.randint(2, 4))]
    n = rng.choice([0, 50, 900, 5000])
    data = np.frombuffer(bytes(rng.choice(b"abcdxyzq0123\n") for _ in range(n)), dtype=np.uint8)
    parts = [compiler.group_programs(g)[2:] for g in groups]
    merged, vnm = compiler.merge_programs(parts)
    stores = [i for i, inst in enumerate(merged) if inst.type == BsInstType.STREAMSTORE]
    assert len(stores) == len(groups)
    assert sum(inst.type == BsInstType.STREAMLOAD for inst in merged) <= 8
    assert len(merged) < sum(len(p[0]) for p in parts)
    labels = [inst.label for inst in merged if inst.type == BsInstType.LABEL]
    assert len(labels) == len(set(labels))
    basis = O.transpose(data)
    for j in range(len(groups)):
        keep = [inst for i, inst in enumerate(merged) if inst.type != BsInstType.STREAMSTORE or i == stores[j]]
        got = O.mask_to_positions(O.run(keep, vnm, basis), n + 1)
        ref = O.mask_to_positions(O.run(parts[j][0], parts[j][1], basis), n + 1)
        assert np.array_equal(got, ref), (groups, j)
    if any(inst.type == BsInstType.WHILE for inst in walk(merged)):
        return  # loops other than stars have no hip form
    src = GroupProgram(0, merged, vnm).source()
    last = len(groups) - 1
    assert f"result[(int64_t){last} * W + w] = out{last};" in src
