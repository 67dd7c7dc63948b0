"""Code generator analysis, kernel compilation and the C ABI (CPU only: no
compute call is made)."""
import ctypes
import os
import re
import subprocess

import pytest

from bitgen_amd import compiler, workloads
from bitgen_amd.codegen import GroupProgram, HipCodegenError
from bitgen_amd.csrc import build as nb
from bitgen_amd.inst import BsInstType

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _prog(regexes, **opt):
    _, _, post, vnm = compiler.group_programs(regexes, opt or None)
    return GroupProgram(0, post, vnm, regexes)


def test_halo_covers_total_shift_distance():
    # 'a' followed 150 positions later by 'b': the look-ahead after cc-advance
    # spans 151 positions -> 3 words of right halo
    p = _prog(["a.{150}b"])
    assert p.HR * 64 >= 150 and p.HR <= 3
    assert p.S == 64 - p.HL - p.HR
    # without cc-advance the same distance is a look-back
    q = _prog(["a.{150}b"], pass_cc_advanced=0)
    assert q.HL * 64 >= 150
    # a literal needs at most one word either side
    r = _prog(["abcdefghij"])
    assert r.HL <= 1 and r.HR <= 1


def test_carry_ops_get_exact_start_lane():
    p = _prog(["ab[^\\n]*cd", "x.*y"])
    assert p.n_carry >= 2  # stars + line projection
    for inst, L in p.carry_ops:
        assert 0 <= L <= p.HL


def test_tile_count_covers_stream():
    p = _prog(["abc"])
    for W in (1, 63, 64, 65, 1000, 123457):
        T = p.n_tiles(W)
        s0 = 64 - p.HR
        covered = s0 + (T - 1) * p.S
        assert covered >= W and (T == 1 or s0 + (T - 2) * p.S < W)


def test_general_loops_go_to_stream_level():
    """A fused kernel cannot hold a general while loop (its condition is a
    whole-stream property); the set compiler runs such a group at stream
    level instead."""
    from bitgen_amd.codegen.gen_stream import StreamGroup

    with pytest.raises(HipCodegenError):
        _prog(["a(bc)*d"])
    progs = compiler.kernel_programs([["a(bc)*d"], ["hello"]], dict(compiler.DEFAULT_OPTIONS, groups_per_kernel=2))
    assert isinstance(progs[0], StreamGroup) and not isinstance(progs[1], StreamGroup)
    assert "bg_group_0_s0" in progs[0].source()


def test_generated_source_shape():
    src = _prog(["abcd", "x[0-9]+y"]).source()
    assert 'extern "C" __global__' in src and "bg_group_0" in src
    assert "bg_matchstar" in src or "bg_scanthru" in src
    assert "goto L" in src  # zero-block skipping
    assert "__syncthreads" not in src  # waves never synchronise


@pytest.mark.parametrize("regexes", [
    ["a.{0,3}b", "[0-9]{2,4}", "a(bc){0,3}d", "(ab|cd){1,3}e"],
    ["c?d"],
    ["ab[^\\n]*cd", "x.*y", "q[0-9]+z"],
    ["hello", "w.*d", "^a[0-9]*$"],
])
def test_pass_loop_replays_exactly(regexes):
    """The speculative pass loop: every variable assigned inside the loop is
    reset at its top and none is also assigned before it, so a replay pass
    recomputes what one straight run computes (header values are read-only
    inside the loop)."""
    p = _prog(regexes)
    src = p.source()
    assert p.n_carry >= 1 and p.loop_start is not None
    before = {d for i in p.insts[:p.loop_start] for d in i.defs()}
    inside = {d for i in p.insts[p.loop_start:] for d in i.defs()}
    assert not (before & inside)
    loop = src[src.index("for (;;) {"):]
    resets = loop[:loop.index("out0 = 0;")]
    from bitgen_amd.codegen.gen_hip import _cname
    for v in inside:
        assert f"{_cname(v)} = 0;" in resets
    assert src.count("goto bg_pass_end;") == p.n_carry
    assert f"bg_settle<{p.n_carry}>(tc)" in src


def _gprog(regexes, **kw):
    _, _, post, vnm = compiler.group_programs(regexes)
    return GroupProgram(0, post, vnm, regexes, **kw)


def test_two_launch_form_has_one_pass_per_launch():
    """carry_settle="launch": no pass loop; the settle launch starts each tile
    with bg_settle_begin (before any load), the sweep ends with bg_sweep_end;
    the queue keeps its epoch after the sweep; the marker kernel tells the
    runtime to launch twice, in the plain and in the batched form."""
    from bitgen_amd.codegen.gen_hip import batch_source
    p = _gprog(["ab[^\\n]*cd", "x.*y", "q[0-9]+z"], carry_settle="launch")
    src = p.source()
    assert p.two_launch and p.loop_start is None and p.meta()["launches"] == 2
    assert "for (;;)" not in src and "bg_settle<" not in src
    begin = src.index(f"if (phase != 0 && bg_settle_begin<{p.n_carry}>(tc)) return true;")
    assert begin < src.index("bas[i] = basis[")
    assert f"const bool bg_keep = bg_sweep_end<{p.n_carry}>(tc);" in src
    assert src.index("if (!bg_keep) return false;") < src.index("result[w] = out0;")
    assert src.count("goto bg_pass_end;") == p.n_carry
    assert "bg_queue_close(Q, phase != 0);" in src
    assert 'extern "C" __global__ void bg_group_0_two_launch() {}' in src
    b = batch_source(p)
    assert "bg_batch_0_two_launch()" in b and "bg_queue_close(Q, phase != 0);" in b
    # programs that wait for every carry stay one launch
    w = _gprog(["x.*y"], carry_settle="launch", lookback="wait")
    assert not w.two_launch and "bg_queue_close(Q);" in w.source() and "_two_launch" not in w.source()
    with pytest.raises(ValueError):
        _gprog(["x.*y"], carry_settle="later")


def test_deferred_settle_form():
    """carry_settle="defer": the tile program of the two-launch form (phase 0
    a single speculative pass, phase 1 bg_settle_begin + replay) driven by
    one launch: each iteration makes the one inlined call, on a new tile
    (phase 0) or on a tile passed before (phase 1); a tile is settled after
    the wave's next pass, an aborted pass at once, the last pending tile
    after the queue runs dry; no marker kernel, one launch; the tile index
    and the pass outcome are made wave-uniform."""
    from bitgen_amd.codegen.gen_hip import batch_source
    p = _gprog(["ab[^\\n]*cd", "x.*y", "q[0-9]+z"], carry_settle="defer")
    src = p.source()
    assert p.phased and not p.two_launch and p.loop_start is None and p.meta()["launches"] == 1
    assert "for (;;)" not in src and "bg_settle<" not in src and "_two_launch" not in src
    assert src.count("bg_ok = bg_prog_0(") == 1
    assert f"if (phase != 0 && bg_settle_begin<{p.n_carry}>(tc)) return true;" in src
    loop = src[src.index("__global__"):]
    assert "if (!bg_ok) now = t;" in loop and "else { now = pend; pend = t; }" in loop
    assert loop.index("if (now != 0xFFFFFFFFu)") < loop.index("else if (more)") < loop.index("else if (pend != 0xFFFFFFFFu)")
    assert "t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);" in loop
    assert "bg_ok = __builtin_amdgcn_readfirstlane((int)bg_ok) != 0;" in loop
    assert "bg_queue_close(Q);" in loop and "bg_queue_open(a.ctl, a.n_queues, a.n_tiles, 1)" in loop
    b = batch_source(p)
    assert b.count("bg_ok = bg_prog_0(") == 1 and "bg_queue_open(a.ctl, a.n_queues, a.total_tiles, 1)" in b
    # programs without speculation stay in the plain loop
    w = _gprog(["x.*y"], carry_settle="defer", lookback="wait")
    assert not w.phased and "pend" not in w.source()


def test_carry_bounds_follow_data_flow():
    """Speculation bounds: a carry op's operand bound only involves the carry
    ops its operands are computed from."""
    p = _prog(["ab[^\\n]*cd", "x.*y"])
    p.source()
    k_last = p.n_carry - 1
    # the stars of independent regexes depend on no earlier carry op
    for k in range(k_last):
        assert p.carry_bounds[k] == {}
    # the line projection depends on every star (through the match streams)
    assert set(p.carry_bounds[k_last]) == set(range(k_last))
    # chained stars: the second star's operand comes from the first
    q = _prog(["a.*b.*c"])
    src = q.source()
    assert q.n_carry >= 3
    assert 0 in q.carry_bounds[1]
    # op 0 has no operand bound; op 1's bound is raised right after op 0
    assert "bg_matchstar<0, S, 0>(" in src or "bg_scanthru<0, S, 0>(" in src
    first = src[src.index(", S, 0>("):]
    assert ", tc, -1, bg_infl);" in first.split("\n")[0]
    assert "bg_bnd_1 = bg_imax(bg_bnd_1, bg_infl" in src
    assert ", tc, bg_bnd_1, bg_infl);" in src


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc not installed")
def test_generated_kernel_compiles_for_gfx950(tmp_path):
    src = _prog(["ab[0-9]{1,3}c", "q.*z", "a.{80}b"]).source()
    path = tmp_path / "k.hip"
    path.write_text(src)
    out = tmp_path / "k.hsaco"
    nb.compile_kernel(str(path), str(out))
    assert out.stat().st_size > 1000


def _declared_symbols():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    return sorted(set(re.findall(r"^\s*(?:int|uint32_t|const char \*)\s*(bg_\w+)\s*\(", text, re.M)))


def test_abi_library_exports_every_declared_symbol():
    from bitgen_amd.backend.hip import runtime as rt

    if not os.path.exists(nb.LIB_PATH):
        nb.build_runtime()
    lib = ctypes.CDLL(nb.LIB_PATH)
    declared = _declared_symbols()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(rt.EXPORTS)
    # bg_last_error works without a GPU
    rt.load_library().bg_last_error()


def test_missing_code_object_raises_cleanly(tmp_path):
    """A pattern set whose .hsaco is gone fails with HipRuntimeError naming
    the file, before any device call (the loader checks the path first)."""
    from bitgen_amd.backend.hip import runtime as rt

    missing = str(tmp_path / "kernel_0.hsaco")
    with pytest.raises(rt.HipRuntimeError, match="kernel_0.hsaco"):
        rt.Module(missing)


def test_missing_library_raises_cleanly(tmp_path, monkeypatch):
    """No libbitgen_hip.so: every entry point raises HipRuntimeError (there is
    no fallback path)."""
    from bitgen_amd.backend.hip import runtime as rt

    monkeypatch.setattr(rt, "_lib", None)
    monkeypatch.setenv("BITGEN_HIP_LIB", str(tmp_path / "libbitgen_hip.so"))
    with pytest.raises(rt.HipRuntimeError, match="missing"):
        rt.load_library()


def test_abi_structs_and_ctl_layout_match_header(tmp_path):
    """ctypes mirrors vs the C header: sizes and offsets of every struct and
    the control-block word indices, compiled with gcc."""
    from bitgen_amd.backend.hip import runtime as rt

    structs = {"bg_group_args_t": rt.GroupArgs, "bg_batch_args_t": rt.BatchArgs, "bg_job_t": rt.Job,
               "bg_tdesc_t": rt.TDesc, "bg_kernel_info_t": rt.KernelInfo}
    body = []
    for cname, py in structs.items():
        body.append(f'printf("%zu ", sizeof({cname}));')
        body += [f'printf("%zu ", offsetof({cname}, {f}));' for f, _ in py._fields_]
    consts = ["BG_MAX_QUEUES", "BG_CTL_STRIDE", "BG_CTL_QDONE", "BG_CTL_ERROR", "BG_CTL_EPOCH", "BG_CTL_WORDS",
              "BG_CTL_COUNTER(3)", "BG_CTL_DONE(3)"]
    body += [f'printf("%d ", (int)({c}));' for c in consts]
    src = tmp_path / "s.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "bitgen_hip.h"\n'
                   f'int main(void) {{ {"".join(body)} return 0; }}\n')
    exe = tmp_path / "s"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = []
    for py in structs.values():
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    want += [rt.MAX_QUEUES, rt.CTL_STRIDE, rt.CTL_QDONE, rt.CTL_ERROR, rt.CTL_EPOCH, rt.CTL_WORDS,
             3 * rt.CTL_STRIDE, (rt.MAX_QUEUES + 3) * rt.CTL_STRIDE]
    assert got == want


def test_transpose_core_host_build(tmp_path):
    """The device transpose's bit logic, compiled for the host with g++ and
    checked against the per-byte definition."""
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <cstdio>
#include <cstdlib>
#include "transpose_core.h"
int main() {
  srand(7); int bad = 0;
  for (int it = 0; it < 5000; it++) {
    unsigned char b[64]; for (int i = 0; i < 64; i++) b[i] = rand() & 255;
    uint64_t x[8], y[8];
    for (int j = 0; j < 8; j++) { x[j] = 0; for (int k = 0; k < 8; k++) x[j] |= (uint64_t)b[8*j+k] << (8*k); }
    bg_transpose_block(x, y);
    for (int i = 0; i < 8; i++) { uint64_t e = 0;
      for (int p = 0; p < 64; p++) if ((b[p] >> (7 - i)) & 1) e |= 1ull << p;
      bad += (e != y[i]); }
  }
  printf("%d\n", bad); return bad != 0;
}''')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-I", os.path.join(ROOT, "bitgen_amd", "csrc"), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0"


def test_untranspose_core_host_build(tmp_path):
    """The inverse block transpose (used by the bigram filter) gives back the
    bytes for every block, and bg_bigram_hash keeps b and the low 7 bits of a
    (word = (a & 3) << 8 | b, bit = bits 2..6 of a)."""
    src = tmp_path / "u.cpp"
    src.write_text(r'''
#include <cstdio>
#include <cstdlib>
#include "transpose_core.h"
int main() {
  srand(11); int bad = 0;
  for (int it = 0; it < 5000; it++) {
    uint64_t x[8], y[8], z[8];
    for (int j = 0; j < 8; j++) x[j] = ((uint64_t)rand() << 40) ^ ((uint64_t)rand() << 20) ^ (uint64_t)rand();
    bg_transpose_block(x, y);
    bg_untranspose_block(y, z);
    for (int j = 0; j < 8; j++) bad += (z[j] != x[j]);
  }
  for (unsigned a = 0; a < 256; a++)
    for (unsigned b = 0; b < 256; b++)
      bad += bg_bigram_hash(a, b) != (((((a & 3u) << 8) | b) << 5) | ((a >> 2) & 31u));
  printf("%d\n", bad); return bad != 0;
}''')
    exe = tmp_path / "u"
    subprocess.run(["g++", "-O2", "-I", os.path.join(ROOT, "bitgen_amd", "csrc"), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "0"


def _interp(prog, data_bits, n, fused, data_bytes=None):
    """Whole-stream interpreter of a generated program (Python ints as
    n-bit streams, bit p = position p).  With fused=True, bitwise
    instructions are evaluated through the generator's bitop3 trees and
    absorbed instructions are skipped, as in the emitted code.  A bigram
    test jumps when one of its clauses has none of its pair hashes among the
    hashes of the byte pairs at consecutive positions of the whole stream
    (positions past the data hold byte 0, as in the basis)."""
    from bitgen_amd.inst import BsInstType as T
    from bitgen_amd.codegen.gen_hip import _LUT_IN, _lut, _leaves, BigramTest, bigram_hash

    present = None
    if data_bytes is not None:
        padded = list(data_bytes) + [0] * max(0, n - len(data_bytes))
        present = {bigram_hash(padded[q], padded[q + 1]) for q in range(n - 1)}

    mask = (1 << n) - 1
    trees, absorbed = prog._fuse()
    env = {}

    def val(v):
        c = prog._const(v)
        if c is not None:
            return mask if c == "1" else 0
        return env.get(v, 0)

    def lut_eval(tree):
        leaves = _leaves(tree, [])
        lut = _lut(tree, leaves) if leaves else (0xFF if _lut(tree, []) & 1 else 0)
        vals = [val(v) for v in leaves] + [val(leaves[0]) if leaves else 0] * (3 - len(leaves))
        out = 0
        for idx in range(8):
            if (lut >> idx) & 1:
                a, b, c = (idx >> 2) & 1, (idx >> 1) & 1, idx & 1
                term = (vals[0] if a else ~vals[0]) & (vals[1] if b else ~vals[1]) & (vals[2] if c else ~vals[2])
                out |= term
        return out & mask

    skip_to = None
    out = None
    for k, inst in enumerate(prog.insts):
        t = inst.type
        if skip_to is not None:
            if t == T.LABEL and inst.label == skip_to:
                skip_to = None
            continue
        if t == T.STREAMLOAD:
            env[inst.ret] = data_bits[int(inst.stream_index)]
        elif t in (T.AND, T.OR, T.XOR, T.NOT):
            if fused:
                if k not in absorbed:
                    env[inst.ret] = lut_eval(trees[k])
            else:
                a = val(inst.operands[0])
                if t == T.NOT:
                    env[inst.ret] = ~a & mask
                else:
                    b = val(inst.operands[1])
                    env[inst.ret] = {T.AND: a & b, T.OR: a | b, T.XOR: a ^ b}[t]
        elif t == T.ASSIGN:
            env[inst.ret] = val(inst.operands[0])
        elif t == T.ADVANCE:
            a = val(inst.operands[0])
            env[inst.ret] = ((a >> inst.amount) if inst.is_left else (a << inst.amount)) & mask
        elif t == T.MATCHSTAR:
            m, c = val(inst.operands[0]), val(inst.operands[1])
            env[inst.ret] = ((((m & c) + c) ^ c) | m) & mask
        elif t == T.SCANTHRU:
            m, c = val(inst.operands[0]), val(inst.operands[1])
            env[inst.ret] = ((m + c) & ~c) & mask
        elif isinstance(inst, BigramTest):
            if present is not None and not all(any(h in present for h in c) for c in inst.clauses):
                skip_to = inst.label
        elif t == T.SKIPZERO:
            if val(inst.operands[0]) == 0:
                skip_to = inst.label
        elif t == T.STREAMSTORE:
            out = val(inst.operands[0])
    return out


@pytest.mark.parametrize("seed", range(10))
def test_bitop3_fusion_computes_the_same_streams(seed):
    """The generator's three-input trees (v_bitop3_b32) and the skipped
    (absorbed) instructions give the same match stream as the instructions
    evaluated one by one, on random programs and inputs."""
    import random
    import numpy as np

    from oracle import oracle as O
    from tests.test_gpu_parity import FUSED_POOL

    rng = random.Random(seed)
    regs = rng.sample(FUSED_POOL, rng.randint(1, 6))
    p = _prog(regs)
    p.source()
    from bitgen_amd.codegen.gen_hip import _leaves

    trees, absorbed = p.fused
    assert all(len(_leaves(t, [])) <= 3 for t in trees.values())
    alpha = np.frombuffer(b"abcdxyzq0123\n", dtype=np.uint8)
    data = np.random.default_rng(seed).choice(alpha, size=3000)
    W = O.n_words_for(data.size)
    basis = O.transpose(data, W)
    n = 64 * W
    bits = [int.from_bytes(basis[i].astype("<u8").tobytes(), "little") for i in range(8)]
    assert _interp(p, bits, n, fused=True) == _interp(p, bits, n, fused=False)


def test_bitop3_lut_convention():
    """LUT bit (a << 2 | b << 1 | c) holds f(a, b, c): a few known tables."""
    from bitgen_amd.codegen.gen_hip import _lut

    a, b, c = ("leaf", "a"), ("leaf", "b"), ("leaf", "c")
    L = ["a", "b", "c"]
    assert _lut(("and", ("and", a, b), c), L) == 0x80
    assert _lut(("or", ("or", a, b), c), L) == 0xFE
    assert _lut(("xor", ("xor", a, b), c), L) == 0x96
    # ~a & b: rows with a=0, b=1 -> indices 2 and 3
    assert _lut(("and", ("not", a), b), L) == 0x0C
    # a & ~b: rows with a=1, b=0 -> indices 4 and 5
    assert _lut(("and", a, ("not", b)), L) == 0x30


LITERALS = ["hello", "abcz", "q0q0", "xyzzy", "a[bc]de", "[ab]cd", "[0-9]ab", "zz[^\\n]*yy", "ca.*bd", "ab",
            "dcba(xy|zq)", "c?abc", "^abc", "xyz$", "bb[0-9]{2,4}c"]


@pytest.mark.parametrize("seed", range(40))
def test_bigram_filter_rewrite_is_exact(seed):
    """The bigram rewrite (filter tests in front of literal first stages,
    their stage instructions moved behind the filter) computes the oracle's
    stream whether or not the filters jump, with and without the bitop3
    trees; the filters' clauses come from the byte sets of the stages."""
    import random
    import numpy as np

    from oracle import oracle as O
    from bitgen_amd.tool.regex2pablo import regexes_to_pablo
    from bitgen_amd.codegen.gen_hip import BigramTest

    rng = random.Random(seed)
    regs = rng.sample(LITERALS, rng.randint(2, 6))
    n = rng.choice([40, 300, 2000])
    alpha = b"abcdxyzq0123h\nelo"
    data = bytearray(rng.choice(alpha) for _ in range(n))
    for _ in range(rng.randint(0, 4)):
        lit = rng.choice([b"hello", b"abcz", b"xyzzy", b"acde", b"q0q0", b"bb12c"])
        off = rng.randrange(0, max(1, n - len(lit)))
        data[off:off + len(lit)] = lit[:n - off]
    data = np.frombuffer(bytes(data), dtype=np.uint8)
    pre, pre_vnm, post, post_vnm = compiler.group_programs(regs)
    prog = GroupProgram(0, post, post_vnm, regs, bigram_filter=True)
    assert prog.bigram_keys and any(isinstance(i, BigramTest) for i in prog.insts)
    basis = O.transpose(data)
    npos = data.size + 1
    ref = O.mask_to_positions(O.run(pre, pre_vnm, basis), npos)
    W = basis.shape[1]
    bits = [int.from_bytes(basis[i].tobytes(), "little") for i in range(8)]
    for fused in (False, True):
        got = _interp(prog, bits, W * 64, fused, data_bytes=data.tolist())
        words = np.frombuffer(got.to_bytes(W * 8, "little"), dtype=np.uint64)
        assert np.array_equal(O.mask_to_positions(words, npos), ref), (regs, fused)


def test_bigram_filter_keeps_carry_ops_in_every_tile():
    """A filter in front of a stage that follows a star must leave the star's
    MatchStar ahead of it: a tile that skips a carry op never publishes the
    carry status its successor waits for.  The generator checks that no
    carry op sits inside a skip range; here on a star followed by a literal
    and on the merged ids programs."""
    from bitgen_amd.codegen.gen_hip import BigramTest
    from bitgen_amd.inst import BsInstType as T

    def carry_ops_in_ranges(insts):
        open_, n = set(), 0
        for i in insts:
            if i.type == T.SKIPZERO:
                open_.add(i.label)
            elif i.type == T.LABEL:
                open_.discard(i.label)
            elif i.type in (T.MATCHSTAR, T.SCANTHRU) and open_:
                n += 1
        return n

    regs = ["ca.*bd", "x[yz]q", "zz[^\\n]*yy"]
    _, _, post, vnm = compiler.group_programs(regs)
    prog = GroupProgram(0, post, vnm, regs, bigram_filter=True)
    tests = [i for i in prog.insts if isinstance(i, BigramTest)]
    assert len(tests) >= 2 and carry_ops_in_ranges(prog.insts) == 0
    stars = [k for k, i in enumerate(prog.insts) if i.type == T.MATCHSTAR]
    after_star = [k for k, i in enumerate(prog.insts) if isinstance(i, BigramTest) and k > stars[0]]
    assert after_star, "a filter after the first star"
    regexes, n_groups, _buf = workloads.ids(n_bytes=1 << 16, n_regex=1000, n_groups=16)
    groups = compiler.group_regex_in_list(regexes, n_groups)
    o = dict(compiler.DEFAULT_OPTIONS, groups_per_kernel=4, min_waves=5, bigram_filter=1)
    for p in compiler.kernel_programs(groups[:8], o):
        assert p.bigram_keys and carry_ops_in_ranges(p.insts) == 0



def test_bigram_table_slot_matches_hash():
    """The table insert takes word = bits 0..9 and bit = bits 10..14 of the
    16-bit key (first byte << 8 | second byte), for both keys of a 32-bit
    v_perm result; that is the 15-bit pair hash the filters are keyed on,
    for every byte pair (the first byte's top bit is not part of the hash),
    and it is one-to-one on the 15 bits it keeps."""
    from bitgen_amd.codegen.gen_hip import bigram_hash

    seen = set()
    for a in range(256):
        for b in range(256):
            key = (a << 8) | b
            k32 = (key << 16) | 0x1234
            h = bigram_hash(a, b)
            assert key & 0x3FF == h >> 5 and (key >> 10) & 31 == h & 31
            assert (k32 >> 16) & 0x3FF == h >> 5 and (k32 >> 26) & 31 == h & 31
            assert h < 1 << 15
            seen.add(h)
    assert len(seen) == 1 << 15


def test_inst_stats_dyn_slot_and_counting():
    """inst_stats_dyn: one more status slot per tile (index n_carry) holding
    the statistics word; every zero / byte-pair test that falls through
    counts one, each pass adds 1 << 20; the word is written by lane 0 before
    the stores, in both settle forms; without the option nothing changes."""
    regs = ["ab[^\\n]*cd", "x.*y", "hello", "q[0-9]+z"]
    for settle in ("inline", "launch", "defer"):
        p = _gprog(regs, carry_settle=settle, inst_stats_dyn=True, bigram_filter=True)
        src = p.source()
        m = p.meta()
        assert m["status_slots"] == p.n_carry + 1 and m["dyn_stats"] is True
        from bitgen_amd.codegen.gen_hip import BigramTest
        n_tests = sum(1 for i in p.insts if isinstance(i, BigramTest) or i.type == BsInstType.SKIPZERO)
        assert any(isinstance(i, BigramTest) for i in p.insts)
        assert src.count("bg_dyn++;") == n_tests > 0
        assert f"uint32_t *bg_dw = status + (size_t)t * n_carry + {p.n_carry};" in src
        assert src.index("bg_dw") < src.index("result[w] = out0;")
        if settle == "inline":
            assert "bg_dyn = (bg_dyn & 0xFFF00000u) + (1u << 20);" in src  # each pass of the loop
        else:
            assert "uint32_t bg_dyn = 1u << 20;" in src and "if (!bg_keep) return false;" in src
    q = _gprog(regs)
    assert "bg_dyn" not in q.source() and q.meta()["status_slots"] == max(1, q.n_carry)
    assert q.meta()["dyn_stats"] is False


def test_tile_model_counts_tests_that_fall_through():
    """The tile model's per-tile test count (the GPU statistics' reference):
    a tile of text without any of a literal's bytes passes no test of it."""
    import numpy as np

    from oracle import oracle as O
    from tests.tile_model import run_tiles

    p = _gprog(["zqxwvu"])
    n_tests = sum(1 for i in p.insts if i.type == BsInstType.SKIPZERO)
    data = np.full(20000, ord("a"), dtype=np.uint8)
    data[15000:15006] = np.frombuffer(b"zqxwvu", dtype=np.uint8)
    stats = []
    run_tiles(p, O.transpose(data, O.n_words_for(data.size)), stats)
    assert len(stats) == p.n_tiles(O.n_words_for(data.size))
    # every test of the literal passes in the tile holding it, none elsewhere
    assert sorted(stats) == [0] * (len(stats) - 1) + [n_tests] and n_tests > 0


def test_shared_filter_geometry_and_sources():
    """shared_filter: the kernels of a set take the widest halo of any of
    them (carry entry lanes move with HL), their answer words sit one after
    another in a tile's row, each kernel reads its words (no table build in
    the kernel) and the set's filter pre-pass builds the table once and
    answers every kernel's filters with that kernel's own keys."""
    from bitgen_amd.codegen.gen_hip import filter_source

    groups = [["ab[0-9]+c", "hello"], ["x.{0,40}y", "[ab]cd"], ["zz[^\\n]*yy", "q0q0"]]
    o = dict(compiler.DEFAULT_OPTIONS, groups_per_kernel=1, bigram_filter=1)
    own = compiler.kernel_programs(groups, o)
    shared = compiler.kernel_programs(groups, dict(o, shared_filter=1))
    HL = max(p.HL for p in own)
    HR = max(p.HR for p in own)
    assert {(p.HL, p.HR, p.S) for p in shared} == {(HL, HR, 64 - HL - HR)}
    off = 0
    for a, b in zip(own, shared):
        assert b.shared_filter and b.filter_off == off and b.filter_words == a.filter_words
        off += b.filter_words
        assert [L - a.HL for _, L in a.carry_ops] == [L - b.HL for _, L in b.carry_ops]
        src = b.source()
        assert "bg_filter_words<" in src and "bg_bigram_build" not in src and "a.filter +" in src
        # the marker that makes the runtime refuse a launch without answers
        assert f"void {b.kernel_name}_shared_filter()" in src and "_shared_filter()" not in a.source()
        assert "bg_bigram_build" in a.source() and "bg_filter_words<" not in a.source()
    assert all(p.filter_row == off for p in shared)
    # lane 0 stores a chunk's two answer words as one 8-byte store: rows and
    # every kernel's offset are even
    assert off % 2 == 0 and all(p.filter_off % 2 == 0 for p in shared)
    fsrc = filter_source(shared)
    # one table build per tile in each of the two pre-pass kernels (one
    # buffer, a batch's global tiles)
    assert fsrc.count("bg_bigram_build(") == 2
    assert 'void bg_filter(' in fsrc and 'void bg_filter_batch(' in fsrc
    for p in shared:
        # lane f holds filter f's clauses; one ballot per 64 filters, two
        # answer words per ballot at the kernel's offset
        nchunks = (len(p.bigram_keys) + 63) // 64
        for c in range(nchunks):
            if p.bigram_pairs == 1:
                # single-pair clauses: lanes hold table slots, not hashes
                assert f"bg_fkey_test(bg_tables, kr{p.gid}[{c}][0])" in fsrc
            else:
                assert f"bg_bigram_clause<{p.bigram_pairs}>(bgt, kr{p.gid}[{c}][0])" in fsrc
            # lane 0 stores the chunk's two answer words at the kernel's offset
            assert f"if (lane == 0) *reinterpret_cast<uint64_t *>(frow + {p.filter_off + 2 * c}) = m;" in fsrc
        slot = "bg_fkey_slot(wbase, bg_fkeys_" if p.bigram_pairs == 1 else "bg_fkeys_"
        assert f"kr{p.gid}[c][j] = c * 64 + lane < {len(p.bigram_keys)} ? {slot}{p.gid}[" in fsrc
    # the batch kernels read the batched pre-pass's rows by global tile
    from bitgen_amd.codegen.gen_hip import batch_source

    bsrc = batch_source(shared[1])
    assert "bg_bigram_build" not in bsrc and "bg_filter_words<" in bsrc
    assert f"a.filter + (size_t)t * {shared[1].filter_row}u + {shared[1].filter_off}u" in bsrc
    assert "bg_bigram_build" in batch_source(own[1])


def test_shared_filter_set_compiles_with_filter_code_object():
    cs = compiler.compile_pattern_set([["hello", "ab[0-9]+c"], ["x[yz]q"]],
                                      {"bigram_filter": 1, "shared_filter": 1, "groups_per_kernel": 1})
    rows = {m["filter_row"] for m in cs.metas}
    assert len(rows) == 1 and rows.pop() > 0
    f = cs.metas[0]["filter"]
    assert f["kernel"] == "bg_filter" and os.path.exists(f["hsaco"])
    # cached: the filter code object comes back from the index
    again = compiler.compile_pattern_set([["hello", "ab[0-9]+c"], ["x[yz]q"]],
                                         {"bigram_filter": 1, "shared_filter": 1, "groups_per_kernel": 1})
    assert again.metas[0]["filter"]["hsaco"] == f["hsaco"]


def test_bigram_key_words_are_canonical():
    """Every clause word holds two pair hashes (a one-pair clause repeats its
    pair) and every filter holds bigram_clauses words (a filter with fewer
    clauses repeats its first): a clause word is true iff one of its pairs
    is present, exactly the filter's clause, and ANDing a repeated clause
    changes nothing -- so the branch-free test gives the filter's answer for
    every table."""
    import random as _r

    from bitgen_amd.codegen.gen_hip import GroupProgram

    regs = ["hello", "ab[0-9]+c", "x[yz]q", "q0q0", "[ab]cd", "zz[^\\n]*yy", "a.{0,3}b", "dcba(xy|zq)"]
    _, _, post, vnm = compiler.group_programs(regs)
    prog = GroupProgram(0, post, vnm, regs, bigram_filter=True)
    words = prog.bigram_key_words()
    ncl = prog.bigram_clauses
    # bigram_pairs: the most pairs of any clause; with 1 both halves of every
    # word are equal, so the device reads only the low half
    assert prog.bigram_pairs == max(len(c) for k in prog.bigram_keys for c in k)
    if prog.bigram_pairs == 1:
        assert all(w & 0xFFFF == w >> 16 for w in words)
    assert len(words) == ncl * len(prog.bigram_keys)
    rng = _r.Random(5)
    for _ in range(200):
        present = {rng.randrange(1 << 15) for _ in range(rng.randrange(0, 3000))}
        for f, clauses in enumerate(prog.bigram_keys):
            # a few of the filter's own pairs planted at random
            if rng.random() < 0.5:
                present |= {h for c in clauses for h in c if rng.random() < 0.5}
            want = all(any(h in present for h in c) for c in clauses)
            got = all(((w & 0xFFFF) in present) or ((w >> 16) in present)
                      for w in words[f * ncl:(f + 1) * ncl])
            assert got == want, (f, clauses)


def test_batch_set_with_stream_level_groups_compiles():
    """A set with a general loop: batch kernels for its tile-parallel groups,
    none for the stream-level group (its meta keeps the plain form's code
    object and its plan), the plain form compiled beside it; the second
    construction comes from the cache."""
    from bitgen_amd.batch import BatchPatternSet

    groups = [["a(bc)*d", "xy"], ["hello", "w.*d"]]
    bs = BatchPatternSet(groups)
    kinds = [m.get("kind") for m in bs.metas]
    assert kinds[0] == "stream" and kinds[1] != "stream"
    assert bs.has_stream_groups and bs.hsacos[0] is None and os.path.exists(bs.hsacos[1])
    assert bs.plain is not None and bs.plain.metas[0].get("kind") == "stream"
    again = BatchPatternSet(groups)
    assert again.hsacos == bs.hsacos and again.plain is not None
