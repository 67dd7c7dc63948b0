"""Code generator analysis, kernel compilation and the C ABI (CPU only: no
compute call is made)."""
import ctypes
import os
import re
import subprocess

import pytest

from bitgen_amd import compiler
from bitgen_amd.codegen import GroupProgram, HipCodegenError
from bitgen_amd.csrc import build as nb

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


def test_general_loops_are_rejected():
    with pytest.raises(HipCodegenError):
        _prog(["a(bc)*d"])


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
    resets = loop[:loop.index("out = 0;")]
    from bitgen_amd.codegen.gen_hip import _cname
    for v in inside:
        assert f"{_cname(v)} = 0;" in resets
    assert src.count("goto bg_pass_end;") == p.n_carry
    assert f"bg_settle<{p.n_carry}>(tc)" in src


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
    q.source()
    assert q.n_carry >= 3
    assert 0 in q.carry_bounds[1]
    assert "bg_infl_0" in q._bound_expr(1)
    assert q._bound_expr(0) == "-1"


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc not installed")
def test_generated_kernel_compiles_for_gfx950(tmp_path):
    src = _prog(["ab[0-9]{1,3}c", "q.*z", "a.{80}b"]).source()
    path = tmp_path / "k.hip"
    path.write_text(src)
    out = tmp_path / "k.hsaco"
    nb.compile_kernel(str(path), str(out))
    assert out.stat().st_size > 1000


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "bitgen_hip.h")).read()
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
