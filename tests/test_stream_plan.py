"""Stream-level plans (graph-break modes 0/1/2, general while loops) on the
CPU: every kernel step runs through the tile model exactly as the generated
kernel partitions it, copies / zeroing / loops run as the host runner does,
and the group's stream equals the oracle's whole-stream result."""
import random

import numpy as np
import pytest

from bitgen_amd import compiler
from bitgen_amd.codegen.gen_stream import CopyStep, KernelStep, LoopStep, StreamGroup, ZeroStep, needs_stream_plan
from oracle import oracle as O
from tests.tile_model import run_tiles

LOOPS = ["a(bc)*d", "(ab)+c", "(a|bb)+c", "x(yz|q)*w", "b(a(bc)*d)+x", "q(ab|c)*z[0-9]"]
PLAIN = ["ab[0-9]{1,3}c", "x[^\\n]*yz", "hello", "c?d", "a.{0,3}b", "z(a|b)z"]


def run_plan_cpu(sg, basis):
    W = basis.shape[1]
    ws = np.zeros((sg.n_slots, W), dtype=np.uint64)
    ws[:8] = basis

    def execute(steps):
        for s in steps:
            if isinstance(s, KernelStep):
                outs = run_tiles(s.program, ws)
                ws[s.out_base:s.out_base + len(s.outputs)] = outs
            elif isinstance(s, CopyStep):
                ws[s.dst] = ws[s.src]
            elif isinstance(s, ZeroStep):
                ws[s.slot] = 0
            elif isinstance(s, LoopStep):
                n = 0
                while ws[s.cond_slot].any():
                    n += 1
                    assert n < 64 * W + 16, "loop did not end"
                    execute(s.body)
            else:
                raise TypeError(s)

    execute(sg.steps)
    return ws[sg.result_slot]


def _check(regs, data, options):
    data = np.ascontiguousarray(data, dtype=np.uint8)
    basis = O.transpose(data, O.n_words_for(data.size))
    pre, pvnm, _, _ = compiler.group_programs(regs)
    ref = O.mask_to_positions(O.run(pre, pvnm, basis), data.size + 1)
    o = dict(compiler.DEFAULT_OPTIONS)
    o.update(options)
    progs = compiler.kernel_programs([regs], o)
    assert len(progs) == 1
    sg = progs[0]
    assert isinstance(sg, StreamGroup), "expected a stream-level group"
    got = O.mask_to_positions(run_plan_cpu(sg, basis), data.size + 1)
    assert np.array_equal(got, ref), (regs, options, O.popcount(got), O.popcount(ref))
    return ref


def _text(rng, n, alpha=b"abcdxyzqw0123 \n"):
    return np.frombuffer(bytes(rng.choice(alpha) for _ in range(n)), dtype=np.uint8).copy()


@pytest.mark.parametrize("seed", range(8))
def test_general_loops_fused_mode(seed):
    """While loops other than single-class stars: the loop runs over whole
    streams, the rest of the group fused around it."""
    rng = random.Random(seed)
    regs = rng.sample(LOOPS, rng.randint(1, 3)) + rng.sample(PLAIN, rng.randint(0, 2))
    data = _text(rng, rng.choice([200, 5000, 30000]))
    for _ in range(rng.randint(2, 10)):  # long repetitions that cross tiles
        rep = bytes(rng.choice([b"bc", b"ab", b"bb", b"yz", b"q"]) * rng.randint(50, 3000))
        off = rng.randrange(0, data.size)
        seg = np.frombuffer(rep[:data.size - off], dtype=np.uint8)
        data[off:off + seg.size] = seg
    _check(regs, data, {})


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("seed", range(3))
def test_graph_break_modes(mode, seed):
    """The reference's ablation modes: subgraphs exchange whole streams."""
    rng = random.Random(100 + seed)
    pool = PLAIN + LOOPS[:3]
    regs = rng.sample(pool, rng.randint(1, 3))
    data = _text(rng, rng.choice([300, 9000]))
    _check(regs, data, {"pass_graph_break": mode, "pass_cc_advanced": 0, "pass_short_circuit": 0})


def test_graph_break_modes_with_passes():
    """Modes 0 and 2 behind cc-advance and short-circuit: the skips are
    dropped at stream level (zero-block skipping never changes results)."""
    rng = random.Random(7)
    data = _text(rng, 20000)
    for mode in (0, 2):
        _check(["ab[0-9]{1,3}c", "x[^\\n]*yz", "a(bc)*d"], data, {"pass_graph_break": mode})


def test_fused_groups_stay_fused():
    pre, pvnm, post, vnm = compiler.group_programs(["ab[0-9]{1,3}c", "x[^\\n]*yz"])
    assert not needs_stream_plan(post)
    _, _, post2, _ = compiler.group_programs(["a(bc)*d"])
    assert needs_stream_plan(post2)


def test_generator_plugin_routes_loops_to_stream_level():
    """The reference-facing generator (create_generator("hip")) builds a
    StreamGroup for a group with a general loop (the --xml program among
    them) and a fused GroupProgram otherwise."""
    from bitgen_amd.codegen import create_generator
    from bitgen_amd.codegen.gen_hip import GroupProgram
    from bitgen_amd.tool.xml_lexical import xml_lexical_pablo

    xml, xvnm = compiler.parse_pablo(xml_lexical_pablo())
    xml, xvnm = compiler.apply_passes(xml, xvnm, {})
    _, _, plain, pvnm = compiler.group_programs(["hello", "x.*y"])
    gen = create_generator("hip")([xml, plain], [xvnm, pvnm], regex_list=[["xml_parser"], ["hello", "x.*y"]])
    progs = gen.build_programs()
    assert isinstance(progs[0], StreamGroup) and isinstance(progs[1], GroupProgram)
    m = progs[0].meta()
    assert m["kind"] == "stream" and m["plan"][1][0] == "loop"
    assert all(f"bg_group_0_s{i}" in progs[0].source() for i in range(len(m["step_kernels"])))


def test_generated_module_of_a_stream_group_is_valid_python(tmp_path):
    """lower() writes kernels/kernel_0.py with the kernel metas as a Python
    literal; a stream-level group's meta holds None fields (no single
    kernel), which must survive as None (compiled for gfx950 here, nothing
    launched)."""
    import ast
    import os

    from bitgen_amd import config as cfg
    from bitgen_amd.codegen import create_generator
    from bitgen_amd.tool.xml_lexical import xml_lexical_pablo

    xml, xvnm = compiler.parse_pablo(xml_lexical_pablo())
    xml, xvnm = compiler.apply_passes(xml, xvnm, {})
    saved = dict(cfg.get_config_dict())
    cwd = os.getcwd()
    try:
        os.chdir(tmp_path)
        cfg.set_config("__cached_module_path", "code")
        cfg.set_config("use_cached_code", False)
        path = create_generator("hip")([xml], [xvnm], regex_list=[["xml_parser"]]).lower()
        src = open(path).read()
    finally:
        os.chdir(cwd)
        cfg.set_from_dict(saved)
    compile(src, path, "exec")
    groups = ast.literal_eval(src.split("GROUPS = ", 1)[1].split("\nHERE = ", 1)[0])
    assert groups[0]["kind"] == "stream" and groups[0]["kernel"] is None
    assert groups[0]["hsaco"] == "kernel_0.hsaco"
