"""Regex -> pablo emitter and pablo parser (CPU).

The emitter's programs, run by the oracle, must mark exactly the LF of every
line that Python's ``re`` finds a match in (``icgrep -c`` semantics with LF as
the only line break), brute force on small random inputs."""
import random
import re

import numpy as np
import pytest

from bitgen_amd.bitparser import BitstreamParser, PabloSyntaxError
from bitgen_amd.inst import BsInstType, walk
from bitgen_amd.tool.regex2pablo import regexes_to_pablo, parse_regex, RegexSyntaxError, cc_name
from oracle import oracle as O

POOL = ["ab", "a[bc]+d", "x.*y", "^q", "z$", "(ab|cd){1,3}e", "b[^\\n]*c", "a(bc)*d", "c?d",
        "[0-9]{2,4}", "abcdx", "q(a|b)z*[xy]{2}", "a.{0,3}b", "(a|bb)+c", "dd[^a]*c", "^a[0-9]*$",
        "(?:ab|c)x", "\\x61\\d", "[^\\s]z", "\\w{2}1", "a.*?b", "(^|\\x0a)q"]


def matched_line_ends(regexes, data):
    pats = [re.compile(r.encode("latin-1")) for r in regexes]
    out, pos = [], 0
    for line in data.split(b"\n")[:-1]:
        end = pos + len(line)
        if any(p.search(line) for p in pats):
            out.append(end)
        pos = end + 1
    return out


def oracle_line_ends(regexes, data, star_mode):
    text = regexes_to_pablo(regexes, star_mode=star_mode)
    p = BitstreamParser(backend="hip")
    insts = p.parse_text(text)
    arr = np.frombuffer(data, dtype=np.uint8)
    out = O.mask_to_positions(O.run(insts, p.var_name_map, O.transpose(arr)), len(data) + 1)
    return list(O.match_positions(out))


@pytest.mark.parametrize("seed", range(60))
def test_emitter_matches_python_re_per_line(seed):
    rng = random.Random(seed)
    alpha = b"abcdxyzq0123 \n"
    data = bytes(rng.choice(alpha) for _ in range(rng.randint(0, 120)))
    regs = rng.sample(POOL, rng.randint(1, 4))
    for mode in ("matchstar", "while"):
        assert oracle_line_ends(regs, data, mode) == matched_line_ends(regs, data), (regs, mode, data)


def test_every_class_excludes_lf_and_matches_members():
    # single-class regexes against every byte value, one byte per line
    data = b"".join(bytes([b]) + b"\n" for b in range(256) if b != 10)
    for r in ["[a-z0-9]", "\\d", "\\W", "[^ab]", ".", "\\s", "[\\x00-\\x1f]", "[^\\x80-\\xff]"]:
        assert oracle_line_ends([r], data, "matchstar") == matched_line_ends([r], data), r


def test_posix_class_equals_explicit_range():
    # Python's re has no POSIX classes: compare with the explicit class instead
    data = b"a3\n13\nx 3\n93\n"
    assert oracle_line_ends(["[[:digit:]]3"], data, "matchstar") == matched_line_ends(["[0-9]3"], data)


def test_regex_parser_errors():
    for bad in ["(ab", "a)", "[ab", "*a", "a{2,1}", "\\q", "(?i)a"]:
        with pytest.raises(RegexSyntaxError):
            parse_regex(bad)


def test_cc_names_follow_icgrep_style():
    assert cc_name((1 << 0x41) | (1 << 0x42) | (1 << 0x43) | (1 << 0x61)) == "CC__41_43_61"


# ---- pablo parser -------------------------------------------------------------
PABLO = """
# comment line
not_1 = (~basis0)
and_ = (basis1 & not_1)
or_ = (basis6 | basis7)
sel = (basis3 ? and_ : or_)
nx = Next(sel)
l_0 = pablo.Advance(nx, 1)
ms = pablo.MatchStar(l_0, and_)
st = pablo.ScanThru(ms, or_)
test = st
accum = st

while test:
  accum = test
  m1 = (test & and_)
  adv = pablo.Advance(m1, 1)

  test = adv
if accum:
  x = (accum ^ or_)
else:
  x = (accum | or_)
if x:
  x = (~x)
output.matches = x
"""


def test_parser_grammar_and_stream_mode():
    p = BitstreamParser(backend="hip")
    insts = p.parse_text(PABLO)
    types = [i.type for i in insts]
    assert types[:8] == [BsInstType.STREAMLOAD] * 8
    assert types[-1] == BsInstType.STREAMSTORE
    # select split into not/and/and/or for stream backends
    assert [i.type for i in insts[11:15]] == [BsInstType.NOT, BsInstType.AND, BsInstType.AND, BsInstType.OR]
    loops = [i for i in insts if i.type == BsInstType.WHILE]
    assert len(loops) == 1 and len(loops[0].body) == 4  # blank line inside the block kept it open
    ifs = [i for i in insts if i.type == BsInstType.IF]
    assert len(ifs) == 2 and ifs[0].body_2 is not None and ifs[1].body_2 is None
    assert any(i.type == BsInstType.MATCHSTAR for i in walk(insts))
    assert any(i.type == BsInstType.SCANTHRU for i in walk(insts))
    # torch-style parsing keeps the select
    q = BitstreamParser(backend="torch").parse_text(PABLO)
    assert any(i.type == BsInstType.SEL for i in q)


@pytest.mark.parametrize("bad", ["x = (basis0 + basis1)", "x = pablo.Foo(basis0)", "x = (y & basis0)",
                                 "x = pablo.Advance(basis0, -1)", "x = pablo.Advance(basis0, a)",
                                 "while basis0\n  x = basis0", "x = y = basis0"])
def test_parser_rejects(bad):
    with pytest.raises(PabloSyntaxError):
        BitstreamParser(backend="hip").parse_text(bad)


def test_parser_oracle_select_semantics():
    text = "sel = (basis0 ? basis1 : basis2)\noutput.matches = sel\n"
    data = np.random.default_rng(3).integers(0, 256, size=500, dtype=np.uint8)
    basis = O.transpose(data)
    a = BitstreamParser(backend="hip")
    b = BitstreamParser(backend="torch")
    ia = a.parse_text(text)
    ib = b.parse_text(text)
    from bitgen_amd.inst import BsStreamLoad, BsStreamStore

    ib = [BsStreamLoad([99], b.name_var_map[f"basis{i}"], i) for i in range(8)] + ib + \
         [BsStreamStore([b.name_var_map["output.matches"]], 100)]
    ra = O.run(ia, a.var_name_map, basis)
    rb = O.run(ib, b.var_name_map, basis)
    assert np.array_equal(ra, rb)
    x = np.unpackbits(basis.view(np.uint8), bitorder="little").reshape(8, -1)
    exp = np.where(x[0] == 1, x[1], x[2])
    assert np.array_equal(np.unpackbits(ra.view(np.uint8), bitorder="little"), exp)
