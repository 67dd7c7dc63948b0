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
from .codegen.gen_hip import FILTER_KERNEL, GroupProgram, compile_sources, filter_source, plan_shared_filter
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
    # waves per SIMD the register allocator must leave room for (0: its own)
    "min_waves": 0,
    # groups merged into one kernel (one launch computes all their streams
    # and shares their basis loads, class header, tile queue and look-back)
    "groups_per_kernel": 1,
    # answer literal first-stage zero tests from a per-tile byte-pair table
    "bigram_filter": 0,
    "replay_scope": "stage",
    # threads per block of the plain kernels (4 waves of one tile each)
    "block_threads": 256,
    # speculative carry-ins settled "inline" (each tile at the end of its
    # pass, waiting for its predecessor), by a second "launch" of the kernel
    # (a sweep that never waits, then a settle that replays only the tiles
    # that need it), or "defer"red in one launch (a wave settles each tile
    # after its next tile's pass, when the predecessor has long published).
    # IDS workload, 4 kernels per 100 MiB on MI355X: inline 0.85 ms,
    # launch 1.05 ms, defer 0.70 ms
    "carry_settle": "defer",
    # per-tile statistics word (passes, tests passed) after the carry
    # statuses (the reference's --pass-inst-stats-dyn)
    "inst_stats_dyn": 0,
    # byte-pair answers of all kernels of a set from one filter pre-pass
    # (plain form; the kernels share one tile geometry)
    "shared_filter": 0,
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


_PURE = None


def merge_programs(parts):
    """One program that computes several groups' match streams.

    ``parts``: [(post-pass insts, vnm)] of the groups, in order.  Variable ids
    and skip labels of part j are moved past those of the parts before it;
    the basis loads of every part become those of part 0; and a pure,
    single-definition instruction before its part's first zero test that
    computes the same function of the same inputs as an earlier one is
    dropped, its result renamed to the earlier one (the groups share their
    character-class header).  Each part keeps its STREAMSTORE, in part
    order: store j writes group j's stream.  Returns (insts, vnm)."""
    from .inst import BsInstType, walk

    global _PURE
    if _PURE is None:
        _PURE = {BsInstType.AND, BsInstType.OR, BsInstType.XOR, BsInstType.NOT, BsInstType.ADVANCE,
                 BsInstType.ASSIGN}
    commutative = {BsInstType.AND, BsInstType.OR, BsInstType.XOR}
    merged, vnm = [], {}
    next_id, next_label = 0, 0
    known = {}   # expression key -> canonical variable id
    canon = {}   # renamed id -> canonical id
    for insts, pvnm in parts:
        ids = set(pvnm)
        labels = set()
        ndefs = {}
        for inst in walk(insts):
            ids.update(inst.defs())
            ids.update(inst.uses())
            for d in inst.defs():
                ndefs[d] = ndefs.get(d, 0) + 1
            if inst.type in (BsInstType.SKIPZERO, BsInstType.LABEL):
                labels.add(inst.label)
        base = next_id
        next_id = base + (max(ids) + 1 if ids else 0)
        lbase = next_label
        next_label = lbase + (max(labels) + 1 if labels else 0)
        for v, name in pvnm.items():
            vnm[base + v] = name
        first_skip = next((k for k, i in enumerate(insts) if i.type == BsInstType.SKIPZERO), len(insts))

        def operand_key(v):
            name = vnm.get(v)
            return ("c", name) if name in ("0", "1") else ("v", canon.get(v, v))

        def moved(inst):
            """Copy of inst (and of any nested loop or branch body) with this
            part's ids, labels and the shared-header renames applied."""
            new = inst.deepcopy()
            new.replace_uses({u: canon.get(base + u, base + u) for u in inst.uses()})
            if new.ret is not None:
                new.ret = base + new.ret
            if new.type in (BsInstType.SKIPZERO, BsInstType.LABEL):
                new.label = lbase + new.label
            if new.type == BsInstType.WHILE:
                new.body = [moved(i) for i in inst.body]
            elif new.type == BsInstType.IF:
                new.body_1 = [moved(i) for i in inst.body_1]
                new.body_2 = [moved(i) for i in inst.body_2] if inst.body_2 else inst.body_2
            return new

        for k, inst in enumerate(insts):
            new = moved(inst)
            key = None
            if new.type == BsInstType.STREAMLOAD:
                key = ("load", int(new.stream_index))
            elif new.type in _PURE and k < first_skip and ndefs.get(inst.ret) == 1:
                ops = [operand_key(u) for u in new.uses()]
                if new.type in commutative:
                    ops.sort()
                extra = (new.amount, new.is_left) if new.type == BsInstType.ADVANCE else ()
                key = (new.type.value, tuple(ops), extra)
            if key is not None:
                if key in known:
                    canon[new.ret] = known[key]
                    continue
                known[key] = new.ret
            merged.append(new)
    return merged, vnm


def program_for(gid, insts, vnm, regexes, options):
    """GroupProgram of one group's post-pass IR under the compile options."""
    return GroupProgram(gid, insts, vnm, regexes, lookback=options["lookback"],
                        expect_zero=bool(options["sched_expect"]), pipeline=bool(options["sched_pipeline"]),
                        min_waves=int(options["min_waves"]), bigram_filter=bool(options["bigram_filter"]),
                        replay_scope=options["replay_scope"], block_threads=int(options.get("block_threads", 256)),
                        carry_settle=options.get("carry_settle", "defer"),
                        inst_stats_dyn=bool(options.get("inst_stats_dyn", 0)),
                        shared_filter=bool(options.get("shared_filter", 0)))


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
                 os.path.join(_PKG, "csrc", "transpose_core.h"),
                 os.path.join(_PKG, "template", "hip", "bitgen_device.h"),
                 os.path.join(os.path.dirname(_PKG), "include", "bitgen_hip.h")]
        for sub in ("tool", "bitparser", "inst", "passes", "codegen"):
            d = os.path.join(_PKG, sub)
            paths += [os.path.join(d, f) for f in sorted(os.listdir(d)) if f.endswith(".py")]
        root = os.path.dirname(_PKG)
        for path in paths:
            # relative names: the digest must not depend on where the tree lives
            h.update(os.path.relpath(path, root).encode())
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


def _index_path(d, name):
    """Cached metas of a set live apart from its generated sources (a small
    index file per set and form), so the sources can be dropped without
    losing the cache."""
    root = cfg.get_config("cache_dir", ".cache_bitgen")
    os.makedirs(os.path.join(root, "index"), exist_ok=True)
    return os.path.join(root, "index", f"{os.path.basename(os.path.normpath(d))}-{name}")


def load_cached_set(d, name):
    """Metas of a set compiled earlier under the same code digest whose code
    objects are all still present, else None."""
    from .codegen.gen_hip import hsaco_cache_dir

    path = _index_path(d, name)
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
        if m.get("filter"):
            m["filter"] = dict(m["filter"], hsaco=os.path.abspath(os.path.join(hsaco_cache_dir(),
                                                                               m["filter"]["hsaco"])))
            if not os.path.exists(m["filter"]["hsaco"]):
                return None
        metas.append(m)
    return metas


def store_cached_set(d, name, metas):
    def rel(m):
        m = dict(m, hsaco=os.path.basename(m["hsaco"]))
        if m.get("filter"):
            m["filter"] = dict(m["filter"], hsaco=os.path.basename(m["filter"]["hsaco"]))
        return m

    rec = {"digest": code_digest(), "metas": [rel(m) for m in metas]}
    path = _index_path(d, name)
    with open(path + ".tmp", "w") as f:
        json.dump(rec, f)
    os.replace(path + ".tmp", path)


def kernel_groups(n_groups, groups_per_kernel):
    """Groups of each kernel: consecutive runs of ``groups_per_kernel``."""
    k = max(1, int(groups_per_kernel))
    return [list(range(i, min(i + k, n_groups))) for i in range(0, n_groups, k)]


def kernel_programs(groups, o):
    """One GroupProgram per kernel: the kernel's groups merged into one
    program (their streams are the program's stores, in group order).  A
    group whose program needs whole streams between its steps (graph breaks,
    while loops) becomes a StreamGroup instead, and then no group of the set
    is merged."""
    from .codegen.gen_stream import StreamGroup, needs_stream_plan

    posts = [group_programs(g, o)[2:] for g in groups]
    if any(needs_stream_plan(insts) for insts, _ in posts):
        programs = []
        for g, (insts, vnm) in enumerate(posts):
            if needs_stream_plan(insts):
                step_opts = dict(o, bigram_filter=0, inst_stats_dyn=0)

                def make(gid, step_insts, step_vnm, regs=groups[g], so=step_opts):
                    return program_for(gid, step_insts, step_vnm, regs, so)

                prog = StreamGroup(g, insts, vnm, groups[g], make)
            else:
                prog = program_for(g, insts, vnm, groups[g], dict(o, shared_filter=0))
            prog.groups = [g]
            programs.append(prog)
        return programs
    programs = []
    for kid, members in enumerate(kernel_groups(len(groups), o["groups_per_kernel"])):
        parts = []
        for g in members:
            parts.append(posts[g])
        insts, vnm = parts[0] if len(parts) == 1 else merge_programs(parts)
        prog = program_for(kid, insts, vnm, [r for g in members for r in groups[g]], o)
        prog.groups = members
        programs.append(prog)
    plan_shared_filter(programs)
    return programs


def kernel_meta(prog):
    m = prog.meta()
    m["groups"] = list(prog.groups)
    m["first_group"] = prog.groups[0]
    return m


def compile_pattern_set(groups, options=None, workers=None):
    """Compile regex groups (list of lists) into gfx950 code objects: one
    kernel per ``groups_per_kernel`` consecutive groups."""
    o = dict(DEFAULT_OPTIONS)
    o.update(options or {})
    groups = [list(g) for g in groups]
    d = _set_dir(groups, o)
    cached = load_cached_set(d, "plain.json")
    if cached is not None:
        return CompiledPatternSet(groups, cached, d, o)
    programs = kernel_programs(groups, o)
    sources = {}
    for prog in programs:
        path = os.path.join(d, f"kernel_{prog.gid}.hip")
        src = prog.source()
        with open(path, "w") as f:
            f.write(src)
        sources[path] = src
    row = max((getattr(p, "filter_row", 0) for p in programs if getattr(p, "shared_filter", False)), default=0)
    fpath = os.path.join(d, "filter.hip")
    if row:
        src = filter_source(programs)
        with open(fpath, "w") as f:
            f.write(src)
        sources[fpath] = src
    hsacos = compile_sources(sources, workers=workers)
    metas = []
    for prog in programs:
        m = kernel_meta(prog)
        m["hsaco"] = os.path.abspath(hsacos[os.path.join(d, f"kernel_{prog.gid}.hip")])
        if row:
            # the set's filter pre-pass (one code object for the set)
            m["filter"] = {"hsaco": os.path.abspath(hsacos[fpath]), "kernel": FILTER_KERNEL, "row": row}
        metas.append(m)
    store_cached_set(d, "plain.json", metas)
    return CompiledPatternSet(groups, metas, d, o)
