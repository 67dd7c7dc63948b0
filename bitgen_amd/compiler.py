"""Compile facade: regex groups -> loaded gfx950 group kernels.

The reference's ``bitgen/compiler.py`` is a dead NFA path; its live compile
API is ``from_icgrep.compile_regex_group_list`` / ``run_regex_kernel``, which
``bitgen_amd.from_icgrep`` mirrors.  This module offers the same pipeline as
plain functions (no process-global cache path, no module import) for
library callers, the batch runner and the serving loop.
"""
import hashlib
import json
import os
import re
from dataclasses import dataclass, field
from typing import List

from . import config as cfg
from .bitparser import BitstreamParser
from .codegen.gen_hip import GroupProgram, compile_sources
from .passes import (
    PassManager, star_rewrite_pass, cc_advance_pass, remove_alias_pass,
    short_circuit_pass, graph_break_pass,
)
from .tool.regex2pablo import regexes_to_pablo

DEFAULT_OPTIONS = {
    "star_mode": "matchstar",
    "pass_star_rewrite": 1,
    "pass_cc_advanced": 1,
    "pass_short_circuit": 1,
    "pass_short_circuit_start": 0,
    "pass_short_circuit_interval": 1,
    "pass_graph_break": -1,
    # carry look-back of the hip kernels: "speculate" (assume carry-in 0,
    # resolve once per tile, replay when wrong) or "wait" (every carry op
    # waits for its predecessor tile)
    "lookback": "speculate",
    # instruction scheduling of the generated group programs: zero tests
    # marked likely taken (skipped code laid out of line), and each regex's
    # first stage computed ahead of the previous regex's first zero test
    "sched_expect": 1,
    "sched_pipeline": 0,
}


# ---- grouping (scripts/run_bitgen.py:71-116) ---------------------------------
def regex_custom_length(regex):
    regex = re.sub(r"\\x[0-9A-Fa-f]{2}", "X", regex)
    regex = re.sub(r"\[[^\]]+\]", "Y", regex)
    return len(regex)


def balance_regex_groups(regex_list, regex_group_number):
    """Greedy: regexes sorted by custom length, each to the lightest group."""
    regex_list = sorted(regex_list, key=regex_custom_length)
    groups = [[] for _ in range(regex_group_number)]
    lengths = [0] * regex_group_number
    for r in regex_list:
        i = lengths.index(min(lengths))
        groups[i].append(r)
        lengths[i] += regex_custom_length(r)
    return groups


def group_regex_in_list(regex_list, regex_group_number):
    if regex_group_number >= 1:
        regex_group_number = min(regex_group_number, len(regex_list))
        groups = balance_regex_groups(list(regex_list), regex_group_number)
    else:
        groups = [[r] for r in regex_list]
    return [g for g in groups if g]


# ---- pipeline ----------------------------------------------------------------
def parse_pablo(text, backend="hip"):
    p = BitstreamParser(backend=backend)
    insts = p.parse_text(text)
    return insts, p.var_name_map


def apply_passes(insts, vnm, options=None):
    o = dict(DEFAULT_OPTIONS)
    o.update(options or {})
    pm = PassManager()
    pm.add_pass(star_rewrite_pass(), enabled=bool(o["pass_star_rewrite"]))
    pm.add_pass(cc_advance_pass(), enabled=bool(o["pass_cc_advanced"]))
    pm.add_pass(remove_alias_pass())
    pm.add_pass(short_circuit_pass(start=o["pass_short_circuit_start"], interval=o["pass_short_circuit_interval"]),
                enabled=bool(o["pass_short_circuit"]))
    pm.add_pass(graph_break_pass(o["pass_graph_break"]))
    return pm.run(insts, vnm)


def group_pablo(regexes, options=None):
    o = dict(DEFAULT_OPTIONS)
    o.update(options or {})
    return regexes_to_pablo(list(regexes), star_mode=o["star_mode"])


def group_programs(regexes, options=None):
    """(pre-pass insts, vnm, post-pass insts, vnm) of one group."""
    text = group_pablo(regexes, options)
    pre, pre_vnm = parse_pablo(text)
    post, post_vnm = parse_pablo(text)
    post, post_vnm = apply_passes(post, post_vnm, options)
    return pre, pre_vnm, post, post_vnm


def program_for(gid, insts, vnm, regexes, options):
    """GroupProgram of one group's post-pass IR under the compile options."""
    return GroupProgram(gid, insts, vnm, regexes, lookback=options["lookback"],
                        expect_zero=bool(options["sched_expect"]), pipeline=bool(options["sched_pipeline"]))


@dataclass
class CompiledPatternSet:
    groups: List[List[str]]
    metas: List[dict]
    base_dir: str
    options: dict = field(default_factory=dict)

    def load(self, device=0):
        from .backend.hip.template import HipGroupSet

        return HipGroupSet(self.metas, self.base_dir, device)


_PKG = os.path.dirname(os.path.abspath(__file__))
_code_digest = None


def code_digest():
    """Digest of every source that shapes a compiled pattern set: the regex
    front end, parser, IR, passes, code generator, device headers and build
    flags.  A cached set is reused only under the same digest."""
    global _code_digest
    if _code_digest is None:
        h = hashlib.sha256()
        paths = [os.path.join(_PKG, "compiler.py"), os.path.join(_PKG, "csrc", "build.py"),
                 os.path.join(_PKG, "template", "hip", "bitgen_device.h"),
                 os.path.join(os.path.dirname(_PKG), "include", "bitgen_hip.h")]
        for sub in ("tool", "bitparser", "inst", "passes", "codegen"):
            d = os.path.join(_PKG, sub)
            paths += [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith(".py")]
        for path in paths:
            h.update(path.encode())
            with open(path, "rb") as f:
                h.update(f.read())
        _code_digest = h.hexdigest()[:24]
    return _code_digest


def _set_dir(groups, options):
    h = hashlib.sha256(json.dumps([groups, sorted(options.items())]).encode()).hexdigest()[:20]
    root = cfg.get_config("cache_dir", ".cache_bitgen")
    d = os.path.join(root, "sets", h)
    os.makedirs(d, exist_ok=True)
    return d


def load_cached_set(d, name):
    """Metas of a set compiled earlier under the same code digest whose code
    objects are all still present, else None."""
    from .codegen.gen_hip import hsaco_cache_dir

    path = os.path.join(d, name)
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("digest") != code_digest():
        return None
    metas = []
    for m in rec["metas"]:
        m = dict(m, hsaco=os.path.abspath(os.path.join(hsaco_cache_dir(), m["hsaco"])))
        if not os.path.exists(m["hsaco"]):
            return None
        metas.append(m)
    return metas


def store_cached_set(d, name, metas):
    rec = {"digest": code_digest(), "metas": [dict(m, hsaco=os.path.basename(m["hsaco"])) for m in metas]}
    tmp = os.path.join(d, name + ".tmp")
    with open(tmp, "w") as f:
        json.dump(rec, f)
    os.replace(tmp, os.path.join(d, name))


def compile_pattern_set(groups, options=None, workers=None):
    """Compile regex groups (list of lists) into gfx950 code objects."""
    o = dict(DEFAULT_OPTIONS)
    o.update(options or {})
    groups = [list(g) for g in groups]
    d = _set_dir(groups, o)
    cached = load_cached_set(d, "plain.json")
    if cached is not None:
        return CompiledPatternSet(groups, cached, d, o)
    programs, sources = [], {}
    for gid, regs in enumerate(groups):
        _, _, post, post_vnm = group_programs(regs, o)
        prog = program_for(gid, post, post_vnm, regs, o)
        programs.append(prog)
        path = os.path.join(d, f"kernel_{gid}.hip")
        src = prog.source()
        with open(path, "w") as f:
            f.write(src)
        sources[path] = src
    hsacos = compile_sources(sources, workers=workers)
    metas = []
    for prog in programs:
        m = prog.meta()
        m["hsaco"] = os.path.abspath(hsacos[os.path.join(d, f"kernel_{prog.gid}.hip")])
        metas.append(m)
    store_cached_set(d, "plain.json", metas)
    return CompiledPatternSet(groups, metas, d, o)
