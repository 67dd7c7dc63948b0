"""Pipeline orchestrator with the reference's entry points
(``bitgen/from_icgrep.py``): IR production per group, parsing, passes,
code generation, kernel initialisation, the warm-up/timed run loop and the
count check.

Differences: the IR producer is the in-process regex emitter (no icgrep
subprocess); ``validation`` counts matching lines with Python's ``re`` per
line, the same quantity ``icgrep -c`` reports (from_icgrep.py:59-121).
"""
import concurrent.futures
import importlib
import os
import re
import shutil
import sys
from datetime import datetime

from . import config as cfg
from .bitparser import BitstreamParser
from .bitstream import Bitstream
from .codegen import create_generator
from .log import MyLogger
from .passes import (
    PassManager, star_rewrite_pass, cc_advance_pass, remove_alias_pass,
    short_circuit_pass, graph_break_pass, inst_stats_pass,
)
from .tool import global_timer
from .tool.pcre2pablo import pcre2pablo, pcrefile2pablo


def load_input_stream(input_path):
    with open(input_path, "rb") as f:
        return f.read()


# ---------------------------------------------------------------------------
def run_passes(insts, var_name_map):
    """Pass order of from_icgrep.py:124-166 with the hip additions
    (star rewrite first; graph break -1 only)."""
    pm = PassManager()
    pm.add_pass(star_rewrite_pass(), enabled=bool(cfg.get_config("pass_star_rewrite", 1)))
    pm.add_pass(cc_advance_pass(), enabled=bool(cfg.get_config("pass_cc_advanced", 0)))
    pm.add_pass(remove_alias_pass(), enabled=True)
    pm.add_pass(short_circuit_pass(), enabled=bool(cfg.get_config("pass_short_circuit", 0)))
    pm.add_pass(graph_break_pass(graph_break_type=cfg.get_config("pass_graph_break", -1)), enabled=True)
    pm.add_pass(inst_stats_pass(), enabled=bool(cfg.get_config("pass_inst_stats", 0)))
    insts, var_name_map = pm.run(insts, var_name_map)
    return insts


def parse_pablo_file(path):
    parser = BitstreamParser(backend=cfg.get_config("backend", "hip"))
    insts = parser.parse_file(path)
    return insts, parser.var_name_map


# ---------------------------------------------------------------------------
def init_cached_path():
    if not cfg.get_config("__cached_path"):
        root = cfg.get_config("cache_dir", ".cache_bitgen")
        if cfg.get_config("use_cached_code"):
            # -r: the latest run's code (the cache root also holds code
            # objects, indexes and tuning results)
            runs = sorted(n for n in os.listdir(root) if n.startswith("bitgen_code_")) if os.path.isdir(root) else []
            if not runs:
                raise FileNotFoundError(f"-r: no cached bitgen_code_* run under {root}")
            name = runs[-1]
        else:
            name = f"bitgen_code_{datetime.now().strftime('%Y_%m%d_%H%M%S_%f')}"
        cfg.set_config("__cached_path", os.path.join(root, name))


def init_cached_module_path():
    base = os.path.join(cfg.get_config("__cached_path"), "code")
    if cfg.get_config("use_cached_code"):
        assert os.path.exists(os.path.join(os.getcwd(), base))
    elif os.path.exists(os.path.join(os.getcwd(), base)):
        i = 1
        while os.path.exists(os.path.join(os.getcwd(), f"{base}_{i}")):
            i += 1
        base = f"{base}_{i}"
    cfg.set_config("__cached_module_path", base)


def _pablo_dir():
    d = os.path.join(os.getcwd(), cfg.get_config("__cached_module_path"), "pablo_src")
    os.makedirs(d, exist_ok=True)
    return d


def compile_regex_group(regex_group_id, regex_group):
    d = _pablo_dir()
    pablo_path = os.path.join(d, f"code_{regex_group_id}.pablo")
    group_file = os.path.join(d, f"regex_group_{regex_group_id}.txt")
    with open(group_file, "w", encoding="latin-1") as f:
        f.write("\n".join(regex_group))
    out = pcrefile2pablo(group_file, pablo_path)
    if out.returncode != 0:
        raise RuntimeError(f"Failed to compile regex group {regex_group_id}: {out.stderr}")
    return pablo_path


def pablo_list2kernel(pablo_regexes, pablo_paths, input_path=None, input_stream_tensor=None,
                      basic_stream_tensor=None):
    insts_list, vnm_list = [], []
    if not cfg.get_config("use_cached_code"):
        for gid, path in enumerate(pablo_paths):
            try:
                insts, vnm = parse_pablo_file(path)
                insts = run_passes(insts, vnm)
            except Exception:
                MyLogger.error(f"Failed to parse Pablo: {gid}: {path}")
                raise
            insts_list.append(insts)
            vnm_list.append(vnm)
    Gen = create_generator(cfg.get_config("backend", "hip"))
    gen = Gen(insts_list, vnm_list, regex_list=pablo_regexes)
    kernel_path = gen.lower(input_path=input_path)
    gen.close()
    return kernel_path


def compile_regex_group_list(regex_group_list, input_path=None, input_stream_tensor=None,
                             basic_stream_tensor=None):
    init_cached_path()
    init_cached_module_path()
    pablo_paths = [None] * len(regex_group_list)
    if not cfg.get_config("use_cached_code"):
        workers = min(len(regex_group_list), os.cpu_count() or 4, 16) or 1
        if cfg.get_config("parallel_compile_parabix", 1) and len(regex_group_list) > 1:
            with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
                futs = {ex.submit(compile_regex_group, i, g): i for i, g in enumerate(regex_group_list)}
                for fut in concurrent.futures.as_completed(futs):
                    pablo_paths[futs[fut]] = fut.result()
        else:
            for i, g in enumerate(regex_group_list):
                pablo_paths[i] = compile_regex_group(i, g)
    return pablo_list2kernel(regex_group_list, pablo_paths, input_path, input_stream_tensor,
                             basic_stream_tensor)


def compile_regex(regex_str, input_path=None, input_stream_tensor=None, basic_stream_tensor=None):
    init_cached_path()
    init_cached_module_path()
    pablo_path = os.path.join(_pablo_dir(), "code.pablo")
    pcre2pablo(regex_str, pablo_path)
    return pablo_list2kernel([[regex_str]], [pablo_path], input_path)


def compile_xml(regex_str, input_path, input_stream_tensor=None, basic_stream_tensor=None):
    """``--xml`` (from_icgrep.py:390-425): the XML lexical program (restated
    in tool/xml_lexical.py) compiled as one group; its pablo text is kept
    under ``kernels/pablo/`` of the module, as the reference keeps its copy.
    Its while loop runs at stream level (a whole-stream condition)."""
    from .tool.xml_lexical import XML_NAME, xml_lexical_pablo

    init_cached_path()
    init_cached_module_path()
    d = os.path.join(os.getcwd(), cfg.get_config("__cached_module_path"), "kernels", "pablo")
    os.makedirs(d, exist_ok=True)
    pablo_path = os.path.join(d, XML_NAME)
    with open(pablo_path, "w") as f:
        f.write(xml_lexical_pablo())
    return pablo_list2kernel([regex_str], [pablo_path], input_path, input_stream_tensor, basic_stream_tensor)


# ---------------------------------------------------------------------------
def init_kernel(regex_func_num, kernel_path, input_stream_tensor, basic_stream_tensor):
    module_root = os.path.dirname(os.path.dirname(kernel_path))
    sys.path.insert(0, module_root)
    try:
        for name in [m for m in sys.modules if m == "kernels" or m.startswith("kernels.")]:
            del sys.modules[name]
        kernels = importlib.import_module("kernels")
        mod = getattr(kernels, os.path.basename(kernel_path).split(".")[0])
    finally:
        sys.path.remove(module_root)
    return mod.init_kernel(regex_func_num, input_stream_tensor, basic_stream_tensor, cfg.get_config_dict())


# ---------------------------------------------------------------------------
def count_matching_lines(regexes, data, multi_input=1):
    """Lines (LF-terminated) containing a match of any regex: ``icgrep -c``."""
    pats = [re.compile(r.encode("latin-1")) for r in regexes]
    n = 0
    for line in data.split(b"\n")[:-1]:
        if any(p.search(line) for p in pats):
            n += 1
    return n * multi_input


def validation(regex, input_data, result, as_one_group=False):
    """Compare the popcount of ``result`` with the matching-line count.

    As in the reference, ``as_one_group`` counts each line once for the whole
    regex list while the GPU result counts it once per group that matches it
    (from_icgrep.py:83-121, README 'Note on Match Counts')."""
    multi = cfg.get_config("multi_input", 1)
    if isinstance(regex, list):
        groups = regex
        if as_one_group:
            flat = [r for g in groups for r in g]
            ref = count_matching_lines(flat, input_data, multi)
        else:
            ref = sum(count_matching_lines(g, input_data, multi) for g in groups)
    else:
        ref = count_matching_lines([regex], input_data, multi)
    global_timer.append_data("run_regex", "ref_count", ref)
    got = result.get_count()
    if got != ref:
        print(f"Validation Failed: bs_count: {got}, reference line count: {ref}")
        return False
    print(f"Validation Passed: count {got}")
    return True


def run_regex_kernel(kernel_path, regex, input_path, input_stream_tensor, basic_stream_tensor=None):
    import torch

    regex_func_num = len(regex) if isinstance(regex, list) else 1
    result, tmp_streams, kernel_handle, kernel, dyn_stats = init_kernel(
        regex_func_num, kernel_path, input_stream_tensor, basic_stream_tensor)
    warmup = cfg.get_config("warmup_iters", 4)
    iters = cfg.get_config("exec_iters", 8)
    with torch.no_grad():
        for it in range(warmup + iters):
            if it < warmup:
                kernel(basic_stream_tensor)
            else:
                with global_timer.time("run_regex"):
                    kernel(basic_stream_tensor)
    n_bytes = input_stream_tensor.size(-1)
    size_mb = input_stream_tensor.numel() / 1024 / 1024
    global_timer.append_data("run_regex", "input_size", size_mb)
    avg = global_timer.get_value("run_regex", "avg_duration")
    global_timer.append_data("run_regex", "throughput", size_mb / avg * 1024 if avg > 0 else float("nan"))
    ok = None
    if cfg.get_config("validation"):
        kernel(basic_stream_tensor)
        bs = Bitstream(result.cpu().numpy(), n_positions=n_bytes + 1)
        global_timer.append_data("run_regex", "count", bs.get_count())
        for gid, c in enumerate(bs.get_count_array()):
            print(f"Count array {gid}: {c}")
        data = input_stream_tensor.reshape(-1, n_bytes)[0].cpu().numpy().tobytes()
        ok = validation(regex if isinstance(regex, list) else regex, data, bs,
                        as_one_group=cfg.get_config("validation_as_one_group", False))
        global_timer.append_data("run_regex", "check", ok)
    if cfg.get_config("pass_inst_stats_dyn") and dyn_stats is not None:
        # the reference's report of its dynamic statistics (from_icgrep.py:619-640):
        # here iterations = passes over tiles, offsets = tests passed per tile
        iter_num = dyn_stats[..., :1]
        dyn_offset_num = dyn_stats[..., 1:]
        print("iter_num:", tuple(iter_num.shape), iter_num.sum().item())
        print("dyn_offset_num:", tuple(dyn_offset_num.shape), dyn_offset_num.sum().item())
        res = global_timer.inst_result
        res["iter_num_total"] = res.get("iter_num_total", 0) + iter_num.sum().item()
        res["dyn_offset_num_max"] = res.get("dyn_offset_num_max", 0) + dyn_offset_num.to(torch.float32).max().item()
        res["dyn_offset_num_std"] = dyn_offset_num.to(torch.float32).std().item()
        res["dyn_offset_num_total"] = res.get("dyn_offset_num_total", 0) + dyn_stats.sum().item()
        print(res)
    torch.cuda.synchronize()
    if kernel_handle is not None:
        kernel_handle.free()
    return ok
