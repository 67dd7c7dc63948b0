"""Command line front end with the reference's flags
(``scripts/run_bitgen.py:293-593``) and ``--backend=hip`` (the default and
only backend of this package).

    python -m bitgen_amd.run_bitgen -f regexes.regex -i input.txt \\
        --input-size=10000 --regex_group_number=4 --check 1 \\
        --pass-graph-break=-1 --pass-cc-advanced=1 --pass-short-circuit=1
"""
import argparse
import os

from . import config as cfg
from .compiler import group_regex_in_list
from .from_icgrep import (
    run_regex_kernel, compile_regex, compile_regex_group_list, compile_xml, load_input_stream,
)
from .log import MyLogger
from .tool.pcre import read_regex_from_pcre_file, select_regex
from .tool.timer import global_timer


def split_input(input_stream_tensor, split_input_num):
    if split_input_num < 1:
        return input_stream_tensor
    if split_input_num == 1:
        return input_stream_tensor.unsqueeze(0)
    raise NotImplementedError("--split-input > 1 is not supported (same as the reference)")


def run_bitgen(input_path, regex, regex_from_file=False, is_xml=False):
    import torch

    from .bitstream import transpose_bytestream_to_bitstream

    data = load_input_stream(input_path)
    input_size = cfg.get_config("input_size")
    if input_size and input_size > 0:
        MyLogger.info(f"[Inputstream]: Limit input stream length to {input_size} bytes")
        data = data[:input_size]
    x = torch.frombuffer(bytearray(data), dtype=torch.uint8) if data else torch.zeros(0, dtype=torch.uint8)
    if cfg.get_config("repeat_input", 1) > 1:
        x = x.repeat(cfg.get_config("repeat_input"))
    x = x.repeat(max(1, cfg.get_config("multi_input", 1)), 1)
    x = split_input(x, cfg.get_config("split_input", 1))
    basis = transpose_bytestream_to_bitstream(x)
    if is_xml:
        kernel_path = compile_xml("xml_parser", input_path, x, basis)
        return run_regex_kernel(kernel_path, "xml_parser", input_path, x, basis)
    if regex_from_file:
        regexes = select_regex(read_regex_from_pcre_file(regex), cfg.get_config("regex_num", -1))
        groups = group_regex_in_list(regexes, cfg.get_config("regex_group_number", 1))
        if cfg.get_config("regex_group_serial"):
            ok = True
            for gid, g in enumerate(groups):
                global_timer.set_group(f"regex_group_{gid}")
                kernel_path = compile_regex_group_list([g], input_path, x, basis)
                if not cfg.get_config("dry_run"):
                    ok = run_regex_kernel(kernel_path, [g], input_path, x, basis) is not False and ok
            return ok
        kernel_path = compile_regex_group_list(groups, input_path, x, basis)
        if not cfg.get_config("dry_run"):
            return run_regex_kernel(kernel_path, groups, input_path, x, basis)
        return None
    kernel_path = compile_regex(regex, input_path, x, basis)
    if not cfg.get_config("dry_run"):
        return run_regex_kernel(kernel_path, regex, input_path, x, basis)
    return None


def build_parser():
    ap = argparse.ArgumentParser(description="BitGen (hip backend)",
                                 formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    ap.add_argument("--backend", default="hip", choices=["hip", "cuda", "torch"],
                    help="Backend (only hip is provided by this package).")
    ap.add_argument("--warmup_iters", default=4, type=int)
    ap.add_argument("--exec_iters", default=8, type=int)
    ap.add_argument("--dry_run", default=0, type=int)
    grp = ap.add_mutually_exclusive_group(required=True)
    grp.add_argument("-e", "--regex", default="", type=str)
    grp.add_argument("-f", "--regex_file", default="", type=str)
    grp.add_argument("--xml", default=False, action="store_true")
    ap.add_argument("--regex_num_from_file", default=-1, type=int)
    ap.add_argument("--regex_group_number", default=1, type=int)
    ap.add_argument("--regex_group_serial", default=0, type=int)
    ap.add_argument("-i", "--input_file", default="", type=str, required=True)
    ap.add_argument("--input-size", default=-1, type=int)
    ap.add_argument("--repeat-input", default=1, type=int)
    ap.add_argument("--multi-input", default=1, type=int)
    ap.add_argument("--split-input", default=1, type=int)
    ap.add_argument("--parallel_compile_parabix", default=1, type=int)
    ap.add_argument("--parallel_compile_cuda", default=1, type=int, help="accepted; hip kernels always compile in parallel")
    ap.add_argument("--parallel_compile_cuda_lto", default=0, type=int, help="accepted, no effect")
    ap.add_argument("--pass-cc-advanced", default=0, type=int)
    ap.add_argument("--pass-cc-advanced-max", default=1, type=int, help="accepted, no effect on the hip backend")
    ap.add_argument("--pass-cc-advanced-merge-cc", default=1, type=int)
    ap.add_argument("--pass-short-circuit", default=0, type=int)
    ap.add_argument("--pass-short-circuit-start", default=2, type=int)
    ap.add_argument("--pass-short-circuit-interval", default=2, type=int)
    ap.add_argument("--pass-short-circuit-interval-disable", default=0, type=int)
    ap.add_argument("--pass-short-circuit-syncpoint", default=1, type=int, help="accepted, no effect")
    ap.add_argument("--pass-graph-break", default=0, choices=[-1, 0, 1, 2], type=int)
    ap.add_argument("--pass-print-graph", default=0, type=int)
    ap.add_argument("--pass-inst-stats", default=0, type=int)
    ap.add_argument("--pass-inst-stats-dyn", default=0, type=int)
    ap.add_argument("--pass-star-rewrite", default=1, type=int, help="rewrite star loops into MatchStar")
    ap.add_argument("--star-mode", default="matchstar", choices=["matchstar", "while"],
                    help="how the IR producer writes single-class stars")
    ap.add_argument("--cuda-register-num", default=128, type=int, help="accepted, no effect")
    ap.add_argument("--cuda_block_size", default=256, type=int, help="threads per block of the group kernels")
    ap.add_argument("-c", "--check", default=0, type=int)
    ap.add_argument("-r", "--reuse_cache", default=False, action="store_true")
    ap.add_argument("--torch-compile", default=0, type=int)
    return ap


def configure(args):
    """Set the process-global config from parsed flags (run_bitgen.py:536-586
    of the reference)."""
    if args.backend != "hip":
        raise SystemExit(f"--backend={args.backend}: this package provides the hip backend only")
    c = cfg.set_config
    c("warmup_iters", args.warmup_iters)
    c("exec_iters", args.exec_iters)
    c("dry_run", args.dry_run)
    c("regex_num", args.regex_num_from_file)
    c("regex_group_number", args.regex_group_number)
    c("regex_group_serial", args.regex_group_serial)
    c("input_size", args.input_size)
    c("repeat_input", args.repeat_input)
    c("multi_input", args.multi_input)
    c("split_input", args.split_input)
    c("pass_cc_advanced", args.pass_cc_advanced)
    c("pass_cc_advanced_max", args.pass_cc_advanced_max)
    c("pass_cc_advanced_merge_cc", args.pass_cc_advanced_merge_cc)
    c("pass_short_circuit", args.pass_short_circuit)
    c("pass_short_circuit_start", args.pass_short_circuit_start)
    c("pass_short_circuit_interval", args.pass_short_circuit_interval)
    c("pass_short_circuit_interval_disable", args.pass_short_circuit_interval_disable)
    c("pass_short_circuit_syncpoint", args.pass_short_circuit_syncpoint)
    c("pass_graph_break", args.pass_graph_break)
    c("pass_print_graph", args.pass_print_graph)
    c("pass_inst_stats", args.pass_inst_stats)
    c("pass_inst_stats_dyn", args.pass_inst_stats_dyn)
    c("pass_star_rewrite", args.pass_star_rewrite)
    c("star_mode", args.star_mode)
    c("parallel_compile_parabix", args.parallel_compile_parabix)
    c("backend", args.backend)
    c("block_size", (args.cuda_block_size, 1, 1))
    c("validation", args.check)
    c("use_cached_code", args.reuse_cache)


def main(argv=None):
    args = build_parser().parse_args(argv)
    configure(args)
    regex = args.regex if args.regex else args.regex_file
    try:
        global_timer.reset_timings()
        return run_bitgen(args.input_file, regex, regex_from_file=bool(args.regex_file), is_xml=args.xml)
    finally:
        global_timer.display_timings()
        root = os.environ.get("BITGEN_ROOT")
        if root:
            global_timer.save_to_file(os.path.join(root, "raw_results/ac"))


if __name__ == "__main__":
    main()
