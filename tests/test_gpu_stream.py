"""Groups run at stream level on the GPU, bit-exact against the oracle:
general while loops (repetitions of sub-expressions, whose condition is a
whole-stream property), the reference's unfused graph-break modes 0/1/2,
and the XML lexical program of --xml."""
import random

import numpy as np
import pytest

from bitgen_amd import compiler
from oracle import oracle as O
from tests.gpu_helpers import assert_parity

pytestmark = pytest.mark.gpu

LOOPS = ["a(bc)*d", "(ab)+c", "(a|bb)+c", "x(yz|q)*w", "b(a(bc)*d)+x", "q(ab|c)*z[0-9]"]
PLAIN = ["ab[0-9]{1,3}c", "x[^\\n]*yz", "hello", "c?d", "a.{0,3}b", "z(a|b)z"]


def _text(rng, n, alpha=b"abcdxyzqw0123 \n"):
    g = np.random.default_rng(rng.randrange(1 << 30))
    return g.choice(np.frombuffer(alpha, dtype=np.uint8), size=n)


@pytest.mark.parametrize("seed", range(6))
def test_general_loops(gpu, seed):
    rng = random.Random(seed)
    groups = [rng.sample(LOOPS, rng.randint(1, 3)) + rng.sample(PLAIN, rng.randint(0, 2))
              for _ in range(rng.randint(1, 3))]
    data = _text(rng, rng.choice([300, 9000, 70000, 300001]))
    for _ in range(rng.randint(2, 12)):  # long repetitions that cross many tiles
        rep = np.frombuffer(rng.choice([b"bc", b"ab", b"bb", b"yz", b"q"]) * rng.randint(50, 6000), dtype=np.uint8)
        off = rng.randrange(0, data.size)
        data[off:off + rep.size] = rep[:data.size - off]
    print(f"seed={seed} n={data.size} groups={groups}", flush=True)
    assert_parity(groups, data, gpu, repeat=2, label=f"loops seed {seed}")


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_graph_break_modes(gpu, mode):
    rng = random.Random(40 + mode)
    groups = [["ab[0-9]{1,3}c", "x[^\\n]*yz"], ["hello", "a(bc)*d"], ["q(a|b)z*[xy]{2}"]]
    data = _text(rng, 120_000)
    opts = {"pass_graph_break": mode, "pass_cc_advanced": 0, "pass_short_circuit": 0}
    print(f"mode={mode}", flush=True)
    assert_parity(groups, data, gpu, options=opts, repeat=2, label=f"graph break {mode}")


def test_graph_break_modes_behind_other_passes(gpu):
    rng = random.Random(50)
    data = _text(rng, 60_000)
    for mode in (0, 2):
        assert_parity([["ab[0-9]{1,3}c", "x[^\\n]*yz", "(ab)+c"]], data, gpu, options={"pass_graph_break": mode},
                      label=f"graph break {mode} with passes")


def test_xml_lexical_program(gpu, tmp_path):
    """--xml: the XML lexical program through the generator plugin and the
    generated module (from_icgrep.compile_xml -> init_kernel -> run)."""
    import torch

    from bitgen_amd import config as cfg
    from bitgen_amd import from_icgrep
    from bitgen_amd.bitstream import transpose_bytestream_to_bitstream
    from bitgen_amd.tool.xml_lexical import xml_lexical_pablo

    rng = random.Random(3)
    parts = []
    for _ in range(3000):
        name = "".join(rng.choice("abcdefghijXYZ") for _ in range(rng.randint(1, 12)))
        parts.append(rng.choice([f"<{name}>", f"</{name}>", f"<?{name} v='1'?>", f"<!{name}>", f"< {name}>",
                                 "text ", "\n"]))
    data = "".join(parts).encode()
    saved = dict(cfg.get_config_dict())
    cfg.reset_config()
    cfg.apply_hip_defaults()
    cfg.set_config("__cached_path", str(tmp_path / "bitgen_code"))
    try:
        x = torch.frombuffer(bytearray(data), dtype=torch.uint8).reshape(1, 1, -1)
        basis = transpose_bytestream_to_bitstream(x)
        kernel_path = from_icgrep.compile_xml("xml_parser", None, x, basis)
        result, _, handle, run, _ = from_icgrep.init_kernel(1, kernel_path, x, basis)
        result.fill_(0x5A)
        run(basis)
        got = O.mask_to_positions(result.cpu().numpy()[0, 0, 0], len(data) + 1)
        handle.free()
    finally:
        cfg.set_from_dict(saved)
    pre, vnm = compiler.parse_pablo(xml_lexical_pablo())
    arr = np.frombuffer(data, dtype=np.uint8)
    ref = O.mask_to_positions(O.run(pre, vnm, O.transpose(arr, O.n_words_for(arr.size))), arr.size + 1)
    assert O.popcount(ref) > 100
    assert np.array_equal(got, ref)
