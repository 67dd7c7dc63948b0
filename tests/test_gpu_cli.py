"""The reference-facing path end to end on the GPU: the run_bitgen command line
with the quick-start flags of the reference's README (8 Snort regexes, 4
groups, the first 10,000 bytes of snort_10KB.input, cc-advance,
short-circuit, graph-break -1, --check 1), which drives
create_generator("hip") -> HipGenerator.lower() -> the generated
kernels/kernel_0.py -> init_kernel -> run -> Bitstream -> validation.

Known answer: "Validation Passed: count 16" (README.md:124-145 of the
reference).  Beyond the count, the positions of the generated module's
result equal the oracle's, the init_kernel tuple has the contract's shapes
and types, and the reference-layout conversion round-trips."""
import json
import os

import numpy as np
import pytest

from bitgen_amd import compiler
from bitgen_amd import config as cfg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KAT = json.load(open(os.path.join(GOLDEN, "snort_kat.json")))


def quick_start_argv():
    return ["-f", os.path.join(GOLDEN, KAT["regex_file"]), "-i", os.path.join(GOLDEN, KAT["input_file"]),
            f"--input-size={KAT['input_size']}", f"--regex_group_number={KAT['groups']}",
            "--pass-graph-break=-1", "--pass-cc-advanced=1", "--pass-short-circuit=1",
            "--pass-short-circuit-start=0", "--pass-short-circuit-interval=1", "--check", "1",
            "--warmup_iters", "1", "--exec_iters", "2"]


@pytest.fixture
def fresh_config(tmp_path):
    saved = dict(cfg.get_config_dict())
    cfg.reset_config()
    cfg.set_config("__cached_path", str(tmp_path / "bitgen_code"))
    yield tmp_path
    cfg.set_from_dict(saved)


def test_run_bitgen_quick_start_known_answer(gpu, fresh_config, capsys):
    from bitgen_amd import run_bitgen

    ok = run_bitgen.main(quick_start_argv())
    out = capsys.readouterr().out
    print(out[-2000:])
    assert ok is True
    assert f"Validation Passed: count {KAT['count']}" in out
    # the generated module was written where the generator plugin puts it
    kernels = list((fresh_config / "bitgen_code").glob("code*/kernels/kernel_0.py"))
    assert len(kernels) == 1


def test_generated_module_contract_and_positions(gpu, fresh_config):
    import torch

    from bitgen_amd import from_icgrep, run_bitgen
    from bitgen_amd.backend.hip import runtime as rt
    from bitgen_amd.bitstream import Bitstream, transpose_bytestream_to_bitstream
    from bitgen_amd.tool.pcre import read_regex_from_pcre_file

    run_bitgen.configure(run_bitgen.build_parser().parse_args(quick_start_argv()))
    data = open(os.path.join(GOLDEN, KAT["input_file"]), "rb").read()[:KAT["input_size"]]
    x = torch.frombuffer(bytearray(data), dtype=torch.uint8).repeat(1, 1).unsqueeze(0)  # [S=1, M=1, N]
    basis = transpose_bytestream_to_bitstream(x)
    W = O.n_words_for(len(data))
    assert basis.dtype == torch.uint64 and tuple(basis.shape) == (1, 1, 8, W)
    groups = compiler.group_regex_in_list(read_regex_from_pcre_file(os.path.join(GOLDEN, KAT["regex_file"])),
                                          KAT["groups"])
    kernel_path = from_icgrep.compile_regex_group_list(groups, None, x, basis)
    assert kernel_path.endswith(os.path.join("kernels", "kernel_0.py"))
    result, tmp_streams, handle, run, dyn_stats = from_icgrep.init_kernel(len(groups), kernel_path, x, basis)
    assert result.dtype == torch.uint64 and tuple(result.shape) == (1, 1, len(groups), W)
    assert tmp_streams.dtype == torch.int32 and tmp_streams.dim() == 3 and tmp_streams.size(0) == 1
    assert dyn_stats is None
    result.fill_(0x5A)
    run(basis)
    got = result.cpu().numpy()[0, 0]
    ob = O.transpose(np.frombuffer(data, dtype=np.uint8), W)
    total = 0
    for g, regs in enumerate(groups):
        pre, vnm, _, _ = compiler.group_programs(regs)
        ref = O.mask_to_positions(O.run(pre, vnm, ob), len(data) + 1)
        assert np.array_equal(O.mask_to_positions(got[g], len(data) + 1), ref), f"group {g}"
        total += O.popcount(ref)
    assert total == KAT["count"]
    # the reference's u32 layout (position p = bit 31 - p%32 of word p/32)
    ref32 = torch.empty((1, 1, len(groups), 2 * W), dtype=torch.int32, device=result.device)
    rt.to_reference_layout(result.data_ptr(), ref32.data_ptr(), W, len(groups), rt.current_stream_ptr())
    torch.cuda.synchronize()
    r32 = ref32.cpu().numpy().view(np.uint32)
    a = Bitstream(r32, n_positions=len(data) + 1).data
    b = Bitstream(got[None, None], n_positions=len(data) + 1).data
    assert np.array_equal(a, b)
    assert Bitstream(r32, n_positions=len(data) + 1).get_count() == KAT["count"]
    handle.free()


def test_run_bitgen_inst_stats_dyn(gpu, fresh_config, capsys):
    """--pass-inst-stats-dyn 1 (the reference's dynamic statistics,
    from_icgrep.py:619-640): same known answer, the statistics report is
    printed and the per-tile statistics cover every tile once at least."""
    from bitgen_amd import run_bitgen
    from bitgen_amd.tool.timer import global_timer

    global_timer.inst_result = {}
    ok = run_bitgen.main(quick_start_argv() + ["--pass-inst-stats-dyn", "1"])
    out = capsys.readouterr().out
    print(out[-2000:])
    assert ok is True and f"Validation Passed: count {KAT['count']}" in out
    assert "iter_num:" in out and "dyn_offset_num:" in out
    res = global_timer.inst_result
    # 4 groups over 10,000 bytes: one tile each, at least one pass per tile
    assert res["iter_num_total"] >= KAT["groups"]
