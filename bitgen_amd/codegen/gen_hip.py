"""HIP (gfx950) code generator for the bitstream IR.

Drop-in for the reference's ``CudaGenerator`` (``bitgen/codegen/gen_cuda.py:15-497``):
same constructor ``(insts_list, var_name_map_list, regex_list)`` and
``lower(input_path=None) -> path of kernels/kernel_0.py``.  It writes, under
``cfg["__cached_module_path"]/kernels/``:

* ``kernel_<g>.hip`` -- one ``extern "C" __global__ bg_group_<g>`` per regex
  group (the reference emits one ``__device__ regex_<g>`` per group and an
  if-chain dispatcher on blockIdx.x, gen_cuda.py:335-473);
* ``kernel_<g>.hsaco`` -- the gfx950 code object (``hipcc --genco``, compiled
  in parallel and cached by content hash);
* ``kernel_0.py`` + ``__init__.py`` -- the generated-module contract
  (``init_kernel`` returning ``(result, tmp_streams, kernel_handle, run,
  dyn_stats)``, as bitgen/backend/cuda/template/template.py:124-149).

Execution model of a group kernel (see template/hip/bitgen_device.h): a
wavefront is a tile of 64 words; lane l holds word w0 + l of every stream.
The generator sizes the halo from the program: ``need[v] = (back, ahead)`` is
how many positions before/after the tile's output range v must be exact,
propagated backwards from the store (Advance by n adds n to ``back`` of its
operand, a look-ahead by n adds n to ``ahead``).  HL = ceil(max back / 64),
HR = ceil(max ahead / 64) halo words; every carry operation k gets L_k, the
first lane where its result must be exact, which is where the look-back
carry enters.
"""
import concurrent.futures
import json
import math
import os
import pprint

from .gen import Generator
from ..inst import BsInst, BsInstType
from .. import config as cfg
from ..csrc import build as native_build

WAVE = 64


class HipCodegenError(NotImplementedError):
    pass


_BITWISE = (BsInstType.AND, BsInstType.OR, BsInstType.XOR, BsInstType.NOT)
# truth-table patterns of the three bitop3 inputs (v_bitop3_b32 convention:
# result bit = LUT[(a << 2) | (b << 1) | c])
_LUT_IN = (0xF0, 0xCC, 0xAA)


def _lut(expr, leaves):
    """Truth table (8 bits) of an expression tree over up to three leaves."""
    kind = expr[0]
    if kind == "leaf":
        return _LUT_IN[leaves.index(expr[1])]
    if kind == "const":
        return 0xFF if expr[1] else 0x00
    if kind == "not":
        return ~_lut(expr[1], leaves) & 0xFF
    a, b = _lut(expr[1], leaves), _lut(expr[2], leaves)
    return {"and": a & b, "or": a | b, "xor": a ^ b}[kind]


def _leaves(expr, out):
    if expr[0] == "leaf":
        if expr[1] not in out:
            out.append(expr[1])
    elif expr[0] != "const":
        for e in expr[1:]:
            _leaves(e, out)
    return out


def _nops(expr):
    if expr[0] in ("leaf", "const"):
        return 0
    return 1 + sum(_nops(e) for e in expr[1:])


def _cname(v):
    return f"v{v}"


class BigramTest(BsInst):
    """Zero test answered by the tile's byte-pair table (bg_bigram_build in
    the device header): jump to ``label`` unless every clause of filter
    ``slot`` may hold in the tile (a clause holds when one of its one or two
    byte-pair hashes occurs at consecutive positions).  It is placed in front
    of a first-stage zero test that stays in place behind it, so a tile
    whose pairs are present (or collide) still reaches the exact tests."""

    def __init__(self, slot, label, clauses):
        super().__init__(BsInstType.SKIPZERO, [], None)
        self.slot = slot
        self.label = label
        self.clauses = clauses

    def uses(self):
        return []

    def __repr__(self):
        return f"BigramTest({self.slot}:{self.clauses}->L{self.label})"


_ALL_BYTES = (1 << 256) - 1
_MOVABLE = (BsInstType.AND, BsInstType.OR, BsInstType.XOR, BsInstType.NOT, BsInstType.ASSIGN, BsInstType.ADVANCE)


def _basis_set(i):
    """Byte values whose basis-i bit (bit 7 - i of the byte) is set."""
    m = 0
    for b in range(256):
        if (b >> (7 - i)) & 1:
            m |= 1 << b
    return m


def _members(m):
    return [b for b in range(256) if (m >> b) & 1]


def bigram_hash(a, b):
    """Same function as bg_bigram_hash in the device header: table word
    (hash >> 5) = the low 2 bits of a and all of b, bit in the word (hash &
    31) = bits 2..6 of a.  The word's low bits, which pick the LDS bank, are
    b's low bits (spread evenly in text; its top bits take a few values)."""
    return ((((a & 3) << 8) | (b & 0xFF)) << 5) | ((a >> 2) & 31)


class GroupProgram:
    """Analysis + HIP source of one group program."""

    def __init__(self, gid, insts, vnm, regexes=None, lookback="speculate", expect_zero=False, pipeline=False,
                 min_waves=0, bigram_filter=False, replay_scope="stage", block_threads=256, carry_settle="inline",
                 inst_stats_dyn=False, shared_filter=False):
        """lookback: "speculate" (carry-ins assumed 0, resolved once per tile,
        replay when wrong) or "wait" (every carry op waits for its
        predecessor's status).  expect_zero: mark every zero test as likely
        taken, so the compiler lays the skipped code out of line and the
        common path falls through.  pipeline: compute the first stage of the
        next regex (the part before its first zero test) ahead of the
        current regex's first zero test, so its instructions fill the
        latencies of the test and of the lane moves.  min_waves: waves per
        SIMD the register allocator must leave room for (0: its own choice;
        a higher count may spill).  bigram_filter: answer first-stage zero
        tests of the form "S1 at p and S2 at p+1" (at most two byte pairs)
        from a per-tile table of the byte pairs present, computing the stage
        only when one of its pairs may occur.  replay_scope: what a replay
        pass recomputes -- "stage": from the first zero test or carry op on
        (the basis and the class header stay in registers across passes),
        "tile": the whole tile from its basis loads (nothing stays live
        across the pass loop's back edge but the tile's carry state).
        block_threads: threads per block of the plain kernel (waves per
        block x 64; a wave is a tile either way, the runtime launches the
        kernel with the block size its launch bounds name).  carry_settle
        (speculative programs): "inline" -- each tile settles its carry-ins
        at the end of its pass, waiting for its predecessor, and replays in
        a loop; "launch" -- the kernel runs twice, a sweep that never waits
        and a settle that replays only the tiles that need it (one pass per
        tile and launch, so no pass loop); "defer" -- one launch in which a
        wave settles each tile after its next tile's pass, by when the
        tile's predecessor has almost always published its statuses (a pass
        that aborted is settled at once).  inst_stats_dyn: each tile also
        writes one statistics word into an extra status slot (index
        n_carry): passes run (bits 20..31) and zero / byte-pair tests that
        fell through in its exact pass (bits 0..19).  shared_filter (with
        bigram_filter): the tile's byte-pair answers are read from a buffer
        that one filter pre-pass of the whole pattern set fills (the
        kernels of a set then share one tile geometry, see filter_source)
        instead of each kernel building the tile's pair table itself."""
        if block_threads % WAVE or not 64 <= block_threads <= 1024:
            raise ValueError(f"block_threads must be a multiple of 64 in 64..1024, not {block_threads}")
        self.block_threads = int(block_threads)
        if replay_scope not in ("stage", "tile"):
            raise ValueError(f"replay_scope must be 'stage' or 'tile', not {replay_scope!r}")
        self.replay_scope = replay_scope
        if lookback not in ("speculate", "wait"):
            raise ValueError(f"lookback must be 'speculate' or 'wait', not {lookback!r}")
        if carry_settle not in ("inline", "launch", "defer"):
            raise ValueError(f"carry_settle must be 'inline', 'launch' or 'defer', not {carry_settle!r}")
        self.carry_settle = carry_settle
        self.inst_stats_dyn = bool(inst_stats_dyn)
        self.gid = gid
        self.insts = insts
        self.vnm = vnm
        self.regexes = regexes or []
        self.lookback = lookback
        self.expect_zero = bool(expect_zero)
        self.pipeline = bool(pipeline)
        self.min_waves = int(min_waves)
        self.kernel_name = f"bg_group_{gid}"
        self._check()
        # streams the program loads: the 8 basis streams, or (a stream-level
        # step) any workspace slots
        used = sorted({int(i.stream_index) for i in insts if i.type == BsInstType.STREAMLOAD})
        self.in_slots = list(range(8)) if not used or used[-1] < 8 else used
        if bigram_filter and self.in_slots != list(range(8)):
            raise HipCodegenError(f"group {gid}: byte-pair filters need the basis streams as inputs")
        self.bigram_keys = []
        self.bigram_clauses = 0
        self.bigram_pairs = 2
        if bigram_filter:
            self._bigram_rewrite()
        self._check_regions()
        self._halo()
        self.shared_filter = bool(shared_filter) and bool(self.bigram_keys)
        # word offset of this program's answers in a tile's row of the
        # set's filter buffer, and the row length (set by plan_shared_filter)
        self.filter_off = 0
        self.filter_row = 0

    def _check_regions(self):
        """Carry ops never sit inside a skip range: every tile runs them, so
        every tile publishes the carry status its successor waits for."""
        open_ = set()
        for inst in self.insts:
            if inst.type == BsInstType.SKIPZERO:
                open_.add(inst.label)
            elif inst.type == BsInstType.LABEL:
                open_.discard(inst.label)
            elif inst.type in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU) and open_:
                raise HipCodegenError(f"group {self.gid}: carry operation {inst} inside a skip range")

    def _check(self):
        stores = 0
        for inst in self.insts:
            t = inst.type
            if t in (BsInstType.WHILE, BsInstType.IF):
                raise HipCodegenError(
                    f"group {self.gid}: {t.name.lower()} blocks are not supported by the hip backend "
                    "(star loops are rewritten to MatchStar; other loops have no tile-parallel form)")
            if t in (BsInstType.SEL, BsInstType.STR, BsInstType.ADD):
                raise HipCodegenError(f"group {self.gid}: unsupported instruction {inst}")
            if t == BsInstType.STREAMSTORE:
                stores += 1
        if stores < 1:
            raise HipCodegenError(f"group {self.gid}: expected a stream store, found none")
        # one store per group: a program merged from several groups writes
        # their streams in order (result + j * W for the j-th store)
        self.n_out = stores

    def _bigram_rewrite(self, max_pairs=2, max_clauses=3):
        """Put a BigramTest in front of every top-level zero test whose skip
        range has pair clauses, and move the instructions that only feed that
        test's operand (and the skipped range) behind the BigramTest.

        Position patterns: a single-definition value that is an AND of
        position-local values (bitwise functions of the basis, each the set
        of byte values for which it is 1) and of their advances or
        look-aheads is nonzero at p only if the byte at p + o lies in S_o for
        every offset o of its pattern {o: S_o}.  Values shifted in at the
        window's edges are zero, so a nonzero value needs the bytes at p + o
        and p + o + 1 (for consecutive offsets) inside the tile: the pair
        clause (S_o x S_o+1) then holds in the tile's table.  The range of a
        top-level zero test may be skipped when any of its tests would fail:
        the test itself, and every nested test whose label closes together
        with the range's and in front of which nothing defines a value used
        after the range.  Each clause comes from one of those tests; up to
        ``max_clauses`` clauses of at most ``max_pairs`` pairs each, most
        selective first."""
        insts = self.insts
        ndefs, def_at, uses_at, labels, defs_of = {}, {}, {}, {}, {}
        for k, inst in enumerate(insts):
            for d in inst.defs():
                ndefs[d] = ndefs.get(d, 0) + 1
                def_at[d] = k
                defs_of.setdefault(d, []).append(k)
            for u in inst.uses():
                uses_at.setdefault(u, []).append(k)
            if inst.type == BsInstType.LABEL:
                labels[inst.label] = k
        sets, pats = {}, {}

        def local(v):
            c = self._const(v)
            if c is not None:
                return _ALL_BYTES if c == "1" else 0
            return sets.get(v)

        def pattern(v):
            c = self._const(v)
            if c is not None:
                return {} if c == "1" else {0: 0}
            return pats.get(v)

        def inst_pattern(inst):
            """Pattern of the value an AND / advance / assign computes."""
            t = inst.type
            if t == BsInstType.AND:
                merged = {}
                for u in inst.uses():
                    for o, m in (pattern(u) or {}).items():
                        merged[o] = merged.get(o, _ALL_BYTES) & m
                return merged or None
            if t == BsInstType.ADVANCE:
                p = pattern(inst.operands[0])
                if p:
                    d = inst.amount if inst.is_left else -inst.amount
                    return {o + d: m for o, m in p.items()}
            elif t == BsInstType.ASSIGN:
                p = pattern(inst.operands[0])
                return dict(p) if p else None
            return None

        for inst in insts:
            v = inst.ret
            if v is None or ndefs.get(v) != 1:
                continue
            t = inst.type
            ops = [local(u) for u in inst.uses()]
            if t == BsInstType.STREAMLOAD:
                sets[v] = _basis_set(int(inst.stream_index))
            elif t in (BsInstType.AND, BsInstType.OR, BsInstType.XOR) and None not in ops:
                a, b = ops
                sets[v] = {BsInstType.AND: a & b, BsInstType.OR: a | b, BsInstType.XOR: a ^ b}[t]
            elif t == BsInstType.NOT and None not in ops:
                sets[v] = _ALL_BYTES ^ ops[0]
            elif t == BsInstType.ASSIGN and None not in ops:
                sets[v] = ops[0]
            if v in sets:
                pats[v] = {0: sets[v]}
            else:
                p = inst_pattern(inst)
                if p:
                    pats[v] = p

        def clauses_of(v, p=None):
            out = []
            p = (pattern(v) if p is None else p) or {}
            for o in sorted(p):
                if o + 1 not in p:
                    continue
                a_set, b_set = _members(p[o]), _members(p[o + 1])
                if not a_set or not b_set or len(a_set) * len(b_set) > max_pairs:
                    continue
                out.append(tuple(sorted({bigram_hash(a, b) for a in a_set for b in b_set})))
            return out

        plan = []   # (skip index, clauses, moved instruction indices)
        open_ = set()
        seg_start = 0
        for k, inst in enumerate(insts):
            if inst.type == BsInstType.LABEL:
                open_.discard(inst.label)
                if not open_:
                    seg_start = k + 1
                continue
            if inst.type != BsInstType.SKIPZERO:
                continue
            top = not open_
            open_.add(inst.label)
            v = inst.operands[0]
            if not top or inst.label not in labels:
                continue
            end = labels[inst.label]
            found = list(clauses_of(v))
            for n in range(k + 1, end):
                nt = insts[n]
                if nt.type != BsInstType.SKIPZERO or nt.label not in labels:
                    continue
                if any(insts[j].type != BsInstType.LABEL for j in range(labels[nt.label], end)):
                    continue
                # nothing in front of the nested test defines a value read
                # after the range
                if any(u >= end for j in range(k + 1, n) for d in insts[j].defs() for u in uses_at.get(d, ())):
                    continue
                found.extend(clauses_of(nt.operands[0]))
            # the range's one result: zero before the range, computed once
            # inside it; a clause of that computation that fails makes the
            # result zero again, the value skipping leaves
            outs = {d for j in range(k + 1, end) for d in insts[j].defs()
                    if any(u > end for u in uses_at.get(d, ()))}
            if len(outs) == 1:
                x = outs.pop()
                xdefs = [j for j, i2 in enumerate(insts) if x in i2.defs()]
                if (len(xdefs) == 2 and seg_start <= xdefs[0] < k < xdefs[1] < end
                        and insts[xdefs[0]].type == BsInstType.ASSIGN
                        and self._const(insts[xdefs[0]].operands[0]) == "0"):
                    found.extend(clauses_of(None, inst_pattern(insts[xdefs[1]])))
            chosen = sorted(set(found), key=lambda c: (len(c), c))[:max_clauses]
            if not chosen:
                continue
            # bitwise instructions and shifts that only feed v: move them
            # behind the filter (carry ops stay: every tile publishes their
            # carry status)
            feed, stack = set(), [v]
            while stack:
                u = stack.pop()
                j = def_at.get(u) if ndefs.get(u) == 1 else None
                if j is None or j < seg_start or j >= k or j in feed or insts[j].type not in _MOVABLE:
                    continue
                # its inputs must hold the same values at the new place
                if any(k2 > j for w in insts[j].uses() if ndefs.get(w, 0) > 1 for k2 in defs_of[w] if k2 < k):
                    continue
                feed.add(j)
                stack.extend(insts[j].uses())
            changed = True
            while changed:
                changed = False
                for j in sorted(feed):
                    ok = all(c in feed or c == k or k < c < end for c in uses_at.get(insts[j].ret, []))
                    if not ok:
                        feed.discard(j)
                        changed = True
            plan.append((k, chosen, sorted(feed)))
        if not plan:
            return
        self.bigram_clauses = max(len(c) for _, c, _ in plan)
        # pairs per clause (1: every clause is a single pair; the device
        # test then reads one table word per clause)
        self.bigram_pairs = max(len(cl) for _, c, _ in plan for cl in c)
        moved = {j for _, _, f in plan for j in f}
        at = {k: (c, f) for k, c, f in plan}
        out = []
        for k, inst in enumerate(insts):
            if k in moved:
                continue
            if k in at:
                clauses, f = at[k]
                out.append(BigramTest(len(self.bigram_keys), inst.label, clauses))
                self.bigram_keys.append(clauses)
                out.extend(insts[j] for j in f)
            out.append(inst)
        self.insts = out

    def bigram_key_words(self):
        """Filter f's clause c as one word: two 15-bit pair hashes (low and
        high half), one of which must be in the tile's table;
        bigram_clauses words per filter.  A clause of one pair repeats it in
        both halves (a or a = a), and a filter with fewer clauses repeats
        its first clause (the clauses are ANDed), so every word is tested
        the same way, with no branch (bg_bigram_clause)."""
        words = []
        for clauses in self.bigram_keys:
            for c in range(self.bigram_clauses):
                h = clauses[c] if c < len(clauses) else clauses[0]
                words.append(h[0] | ((h[1] if len(h) > 1 else h[0]) << 16))
        return words

    def _const(self, v):
        n = self.vnm.get(v)
        return n if n in ("0", "1") else None

    def _halo(self):
        need = {}

        def bump(v, back, ahead):
            if self._const(v) is not None:
                return
            b, a = need.get(v, (0, 0))
            need[v] = (max(b, back), max(a, ahead))

        for inst in reversed(self.insts):
            t = inst.type
            if t == BsInstType.STREAMSTORE:
                bump(inst.operands[0], 0, 0)
                continue
            if t in (BsInstType.LABEL, BsInstType.GRAPHBREAK):
                continue
            if t == BsInstType.SKIPZERO:
                # the zero test is exact whatever the halo holds
                continue
            r = need.get(inst.ret, (0, 0)) if inst.ret is not None else (0, 0)
            if t == BsInstType.ADVANCE:
                n = inst.amount
                if inst.is_left:
                    bump(inst.operands[0], max(0, r[0] - n), r[1] + n)
                else:
                    bump(inst.operands[0], r[0] + n, max(0, r[1] - n))
            elif t in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU):
                # the wave adder works on whole words: the result is exact
                # from lane L = HL - ceil(back / 64) on only if the operands
                # are exact over all of lane L, i.e. from a word boundary
                back = WAVE * math.ceil(r[0] / WAVE)
                for o in inst.uses():
                    bump(o, back, r[1])
            else:
                for o in inst.uses():
                    bump(o, r[0], r[1])
        self.need = need
        max_back = max((b for b, _ in need.values()), default=0)
        max_ahead = max((a for _, a in need.values()), default=0)
        self.HL = math.ceil(max_back / WAVE)
        self.HR = math.ceil(max_ahead / WAVE)
        self.S = WAVE - self.HL - self.HR
        if self.S < 1:
            raise HipCodegenError(
                f"group {self.gid}: shift distances need {self.HL}+{self.HR} halo words, more than a 64-word tile")
        self.carry_ops = []
        for inst in self.insts:
            if inst.type in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU):
                back = need.get(inst.ret, (0, 0))[0]
                L = self.HL - math.ceil(back / WAVE)
                self.carry_ops.append((inst, L))
        self.n_carry = len(self.carry_ops)

    def set_geometry(self, HL, HR):
        """Widen the tile's halo to HL / HR words (at least the program's own):
        the kernels of a set that share a filter pre-pass cover the same
        windows.  Each carry op's entry lane moves with HL (its operands stay
        exact from the same distance behind it)."""
        if HL < self.HL or HR < self.HR:
            raise ValueError(f"group {self.gid}: halo {HL}/{HR} is narrower than the program's {self.HL}/{self.HR}")
        if WAVE - HL - HR < 1:
            raise HipCodegenError(f"group {self.gid}: a {HL}+{HR}-word halo leaves no tile")
        self.HL, self.HR = HL, HR
        self.S = WAVE - HL - HR
        self.carry_ops = [(inst, HL - math.ceil(self.need.get(inst.ret, (0, 0))[0] / WAVE))
                          for inst, _ in self.carry_ops]

    @property
    def filter_words(self):
        """32-bit answer words of this program's byte-pair filters per tile."""
        return 2 * ((len(self.bigram_keys) + 63) // 64)

    @property
    def phased(self):
        """The tile program runs in phases: phase 0 a single speculative pass
        that stores its values without settling its carry-ins, phase 1 the
        settle (and the replay when one is needed).  Both the sweep + settle
        launches and the deferred settle run the program this way."""
        return 0 < self.n_carry <= 64 and self.lookback == "speculate" and self.carry_settle in ("launch", "defer")

    @property
    def two_launch(self):
        """The kernel runs as a sweep launch and a settle launch."""
        return self.phased and self.carry_settle == "launch"

    def n_tiles(self, n_words):
        s0 = WAVE - self.HR
        if n_words <= s0:
            return 1
        return 1 + (n_words - s0 + self.S - 1) // self.S

    def val(self, v):
        c = self._const(v)
        if c == "0":
            return "0ull"
        if c == "1":
            return "~0ull"
        return _cname(v)

    def device_function(self, shared=None):
        """The group program for one tile as a device function (shared by the
        plain per-group kernel and the batched kernel).  shared: read the
        byte-pair answers from the filter pre-pass (default: the program's
        shared_filter)."""
        shared = self.shared_filter if shared is None else bool(shared)
        lines = []
        add = lines.append
        regex_note = json.dumps(self.regexes)[:2000].replace("*/", "* /")
        add(f"/* group {self.gid}: {regex_note} */")
        nf = len(self.bigram_keys)
        if nf and not shared:
            words = self.bigram_key_words()
            keys = ", ".join(f"0x{k:08X}u" for k in words)
            add(f"static __constant__ uint32_t bg_keys_{self.gid}[{len(words)}] = {{{keys}}};")
        add(f"static __device__ __forceinline__ bool bg_prog_{self.gid}(const uint64_t *__restrict__ basis, "
            "uint64_t *__restrict__ result, uint32_t *status, int64_t W, uint32_t t, uint32_t epoch, "
            "uint32_t n_carry, uint32_t *err, uint32_t *bgt, const uint32_t *__restrict__ bgf, uint32_t phase) {")
        add(f"  constexpr int HL = {self.HL}, HR = {self.HR}, S = {self.S};")
        add("  const int lane = bg_lane();")
        add("  const int64_t w0 = (t == 0) ? 0 : (int64_t)(64 - HR) + (int64_t)(t - 1) * S - HL;")
        add("  const int64_t w = w0 + lane;")
        add("  const bool inb = (w >= 0) && (w < W);")
        add("  bg_tile tc = bg_tile_make(t, w0, W, status, n_carry, epoch, err);")
        two = self.phased
        if two:
            # settle phase: most tiles are exact from their speculative pass
            # and stop here, before any load
            add("  tc.sweep = phase == 0;")
            add(f"  if (phase != 0 && bg_settle_begin<{self.n_carry}>(tc)) return true;")
        else:
            add("  (void)phase;")
        if self.n_carry > 64 or self.lookback == "wait":
            # every op waits for its predecessor (speculation state is kept in
            # 64-bit masks, so programs with more carry ops always wait)
            add("  tc.mode = BG_REPLAY;")
        add("  (void)tc; (void)bgt;")
        outs = [f"out{j}" for j in range(self.n_out)]
        tile_scope = 0 < self.n_carry <= 64 and self.lookback == "speculate" and self.replay_scope == "tile" and not two
        # all eight loads in flight at once: clamp the index instead of
        # predicating each load, then clear the lanes outside the stream
        add("  const int64_t wc = w < 0 ? 0 : (w >= W ? W - 1 : w);")
        if tile_scope:
            # the pass loop covers the whole tile: every value of the tile is
            # declared (and so zeroed) inside it, and the load index passes
            # through a fence so that nothing computed from the basis is
            # hoisted out of the loop (and kept live across it)
            add("  uint64_t " + ", ".join(f"{o} = 0" for o in outs) + ";")
            add("  for (;;) {")
            add("  " + " ".join(f"{o} = 0;" for o in outs))
            add("  uint64_t wl = (uint64_t)wc;")
            add("  bg_fence(wl);")
        else:
            add("  const uint64_t wl = (uint64_t)wc;")
        slots = self.in_slots
        if slots == list(range(8)):
            add("  uint64_t bas[8];")
            add("#pragma unroll")
            add("  for (int i = 0; i < 8; i++) bas[i] = basis[(int64_t)i * W + (int64_t)wl];")
        else:
            # a stream-level step: its inputs are workspace slots (stream s at
            # basis + s * W)
            add(f"  uint64_t bas[{len(slots)}];")
            for i, sl in enumerate(slots):
                add(f"  bas[{i}] = basis[(int64_t){sl} * W + (int64_t)wl];")
        add("#pragma unroll")
        add(f"  for (int i = 0; i < {len(slots)}; i++) bas[i] = inb ? bas[i] : 0ull;")
        if nf and shared:
            # the answers of every filter test, from the set's filter pre-pass
            # (bgf: this tile's words; wave-uniform)
            add("  (void)bgt;")
            add(f"  uint32_t bgm[{self.filter_words}];")
            add("  bg_filter_words<" + str(self.filter_words) + ">(bgf, bgm);")
        elif nf:
            # byte-pair table of the tile and the answers of every filter test
            add("  (void)bgf;")
            add("  bg_bigram_build(bas, bgt);")
            add(f"  uint32_t bgm[{self.filter_words}];")
            add(f"  bg_bigram_lookup<{nf}, {self.bigram_clauses}, {self.bigram_pairs}>(bgt, bg_keys_{self.gid}, bgm);")
        else:
            add("  (void)bgf;")
        # declarations (goto must not cross initialisations)
        vars_ = set()
        for inst in self.insts:
            for d in inst.defs():
                vars_.add(d)
            for u in inst.uses():
                if self._const(u) is None:
                    vars_.add(u)
        vars_ = sorted(vars_)
        for i in range(0, len(vars_), 16):
            chunk = ", ".join(f"{_cname(v)} = 0" for v in vars_[i:i + 16])
            add(f"  uint64_t {chunk};")
        if not tile_scope:
            add("  uint64_t " + ", ".join(f"{o} = 0" for o in outs) + ";")
        store_index = {}
        for inst in self.insts:
            if inst.type == BsInstType.STREAMSTORE:
                store_index[id(inst)] = len(store_index)
        carry_index = {id(inst): (k, L) for k, (inst, L) in enumerate(self.carry_ops)}
        # With carry ops the program after its header runs in a loop of
        # passes: a speculative pass and, when bg_settle finds an assumed
        # carry-in was wrong, a replay.  The loop starts before the first skip
        # test or carry op, and earlier if a variable defined before that
        # point is assigned again after it; variables defined in the loop are
        # reset at the top of a pass, so a replay computes exactly what a
        # single straight run would, and the header stays outside.
        speculative = 0 < self.n_carry <= 64 and self.lookback == "speculate"
        # the two-launch form runs one pass per tile and launch: no pass loop
        looped = speculative and not two
        body_start = next((k for k, i in enumerate(self.insts)
                           if i.type in (BsInstType.SKIPZERO, BsInstType.MATCHSTAR, BsInstType.SCANTHRU)),
                          len(self.insts))
        while body_start > 0:
            later = {d for i in self.insts[body_start:] for d in i.defs()}
            first = next((k for k, i in enumerate(self.insts[:body_start]) if later & set(i.defs())), None)
            if first is None:
                break
            body_start = first
        body_vars = list(dict.fromkeys(d for i in self.insts[body_start:] for d in i.defs()))
        if tile_scope:
            body_start = 0
        self.loop_start = body_start if looped else None
        self.carry_bounds = self._carry_bounds(carry_index)
        # operand bound of carry op k = max over the carry ops j it depends on
        # of (j's result bound + lanes of right advance from j to k), kept as
        # one running maximum per dependent op
        consumers = {k: d for k, d in self.carry_bounds.items() if d}
        feeds = {}
        for k, d in consumers.items():
            for j, bits in d.items():
                feeds.setdefault(j, []).append((k, math.ceil(bits / WAVE)))
        if speculative:
            add("  int bg_infl = -1; uint32_t bg_vseen = 0;" + "".join(f" int bg_bnd_{k} = -1;" for k in sorted(consumers)))
        stats = self.inst_stats_dyn
        if stats:
            # passes in bits 20..31, tests passed in this pass in bits 0..19
            add("  uint32_t bg_dyn = " + ("0u;" if looped else "1u << 20;"))
        # three-input bitwise trees -> v_bitop3_b32
        trees, absorbed = self._fuse()
        self.fused = (trees, absorbed)

        def eff_uses(k, inst):
            if k in trees:
                return _leaves(trees[k], [])
            return inst.uses()

        # values shared by every regex (the basis and the group header defined
        # before the first zero test) are fenced at regex boundaries
        first_skip = next((k for k, i in enumerate(self.insts) if i.type == BsInstType.SKIPZERO), len(self.insts))
        header = [d for k, i in enumerate(self.insts[:first_skip]) if k not in absorbed for d in i.defs()]
        later = {u for k, i in enumerate(self.insts) if k >= first_skip and k not in absorbed for u in eff_uses(k, i)}
        fence_vars = [v for v in dict.fromkeys(header) if v in later]
        hoist_at, hoisted = self._pipeline_plan(trees, absorbed) if self.pipeline else ({}, set())
        chunks = [[] for _ in self.insts]
        body_lines = lines
        open_regions = set()
        for idx, inst in enumerate(self.insts):
            lines = chunks[idx]
            add = lines.append
            if looped and idx == body_start:
                if not tile_scope:
                    add("  for (;;) {")
                if self.bigram_keys:
                    add("  " + " ".join(f"bg_sfence(bgm[{c}]);" for c in range(2 * ((len(self.bigram_keys) + 63) // 64))))
                if stats:
                    add("  bg_dyn = (bg_dyn & 0xFFF00000u) + (1u << 20);")
                if not tile_scope:
                    zeros = [f"{_cname(v)} = 0;" for v in body_vars]
                    for i in range(0, len(zeros), 16):
                        add("  " + " ".join(zeros[i:i + 16]))
                    add("  " + " ".join(f"{o} = 0;" for o in outs)
                        + "".join(f" bg_bnd_{k} = -1;" for k in sorted(consumers)))
            t = inst.type
            if t == BsInstType.STREAMLOAD:
                add(f"  {_cname(inst.ret)} = bas[{self.in_slots.index(int(inst.stream_index))}];")
            elif t in _BITWISE:
                if idx not in absorbed:
                    add(f"  {_cname(inst.ret)} = {self._emit_tree(trees[idx])};")
            elif t == BsInstType.ASSIGN:
                add(f"  {_cname(inst.ret)} = {self.val(inst.operands[0])};")
            elif t == BsInstType.ADVANCE:
                fn = "bg_advl" if inst.is_left else "bg_advr"
                add(f"  {_cname(inst.ret)} = {fn}<{inst.amount}>({self.val(inst.operands[0])});")
            elif t in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU):
                k, L = carry_index[id(inst)]
                fn = "bg_matchstar" if t == BsInstType.MATCHSTAR else "bg_scanthru"
                if speculative:
                    bnd = f"bg_bnd_{k}" if k in consumers else "-1"
                    add(f"  {_cname(inst.ret)} = {fn}<{L}, S, {k}>({self.val(inst.operands[0])}, "
                        f"{self.val(inst.operands[1])}, tc, {bnd}, bg_infl);")
                    add("  if (tc.abort) goto bg_pass_end;")
                    later = [c for c in sorted(consumers) if c > k]
                    if k in consumers and later:
                        # the op verified the earlier ops' assumed carry-ins:
                        # their influence on later operands is void
                        add("  if (tc.vfy != bg_vseen) { bg_vseen = tc.vfy;"
                            + "".join(f" bg_bnd_{c} = -1;" for c in later) + " }")
                    for c, lanes in feeds.get(k, []):
                        add(f"  bg_bnd_{c} = bg_imax(bg_bnd_{c}, bg_infl" + (f" + {lanes});" if lanes else ");"))
                else:
                    add(f"  {{ int bg_unused = -1; {_cname(inst.ret)} = {fn}<{L}, S, {k}>("
                        f"{self.val(inst.operands[0])}, {self.val(inst.operands[1])}, tc, -1, bg_unused); }}")
            elif isinstance(inst, BigramTest):
                test = f"!(bgm[{inst.slot // 32}] & 0x{1 << (inst.slot % 32):X}u)"
                if self.expect_zero:
                    test = f"__builtin_expect({test}, 1)"
                add(f"  if ({test}) goto L{inst.label};")
                if stats:
                    add("  bg_dyn++;")
                open_regions.add(inst.label)
            elif t == BsInstType.SKIPZERO:
                test = f"bg_wave_zero({self.val(inst.operands[0])})"
                if self.expect_zero:
                    test = f"__builtin_expect({test}, 1)"
                add(f"  if ({test}) goto L{inst.label};")
                if stats:
                    add("  bg_dyn++;")
                open_regions.add(inst.label)
            elif t == BsInstType.LABEL:
                add(f"L{inst.label}:;")
                open_regions.discard(inst.label)
                if not open_regions:
                    # end of a regex's skip regions: later classes are
                    # recomputed instead of kept live from here
                    add("  " + " ".join(f"bg_fence({_cname(v)});" for v in fence_vars))
            elif t == BsInstType.STREAMSTORE:
                add(f"  out{store_index[id(inst)]} = {self.val(inst.operands[0])};")
            elif t == BsInstType.GRAPHBREAK:
                pass
            else:
                raise HipCodegenError(f"unsupported instruction {inst}")
        # assemble: hoisted first stages go right before the zero test they
        # were scheduled against (hoisting never crosses a pass-loop header:
        # every target is a zero test, and the loop starts at or before the
        # first one)
        lines = body_lines
        add = lines.append
        for idx in range(len(self.insts)):
            if idx in hoist_at:
                head = chunks[idx]
                # the pass-loop header, if it opens here, stays in front
                lead = []
                while head and not head[0].lstrip().startswith("if ("):
                    lead.append(head.pop(0))
                lines.extend(lead)
                for h in hoist_at[idx]:
                    lines.extend(chunks[h])
                lines.extend(head)
            elif idx not in hoisted:
                lines.extend(chunks[idx])
        if looped:
            add("bg_pass_end:;")
            add(f"  if (bg_settle<{self.n_carry}>(tc)) break;")
            add("  }")
        elif two:
            add("bg_pass_end:;")
            add(f"  const bool bg_keep = bg_sweep_end<{self.n_carry}>(tc);")
        if stats:
            # one word per tile after its carry statuses; a settle replay
            # keeps the sweep's pass count and replaces the test count
            add("  if (lane == 0) {")
            add(f"    uint32_t *bg_dw = status + (size_t)t * n_carry + {self.n_carry};")
            if two:
                add("    *bg_dw = phase != 0 ? (*bg_dw & 0xFFF00000u) + bg_dyn : (bg_keep ? bg_dyn : bg_dyn & 0xFFF00000u);")
            else:
                add("    *bg_dw = bg_dyn;")
            add("  }")
        if two:
            add("  if (!bg_keep) return false;")
        add("  const int lo = (t == 0) ? 0 : HL;")
        add("  const int hi = (t == 0) ? 64 - HR : HL + S;")
        add("  if (inb && lane >= lo && lane < hi) {")
        for j, o in enumerate(outs):
            add(f"    result[{f'(int64_t){j} * W + ' if j else ''}w] = {o};")
        add("  }")
        add("  return true;")
        add("}")
        return "\n".join(lines) + "\n"

    def _pipeline_plan(self, trees, absorbed):
        """Which first-stage instructions move up one regex.

        Regexes are the stretches between labels that close every open skip
        region.  For consecutive regexes j and j+1, the pure instructions of
        j+1 before its first zero test (bitwise trees, shifts) move to just
        before j's first zero test when every input is already defined there:
        a constant, a single-definition value defined before that test, or
        another moved instruction.  A moved value is defined once, so it is
        the same wherever it is computed.  Returns ({zero-test index: [moved
        indices in program order]}, set of moved indices)."""
        insts = self.insts
        ndefs = {}
        def_at = {}
        for k, inst in enumerate(insts):
            for d in inst.defs():
                ndefs[d] = ndefs.get(d, 0) + 1
                def_at[d] = k
        # segments: (start index, first top-level zero test or None, end index)
        segs, start, open_, first = [], 0, set(), None
        for k, inst in enumerate(insts):
            if inst.type == BsInstType.SKIPZERO:
                if not open_ and first is None:
                    first = k
                open_.add(inst.label)
            elif inst.type == BsInstType.LABEL:
                open_.discard(inst.label)
                if not open_:
                    segs.append((start, first, k))
                    start, first = k + 1, None
        segs.append((start, first, len(insts) - 1))

        def uses(k):
            if k in trees:
                return _leaves(trees[k], [])
            return insts[k].uses()

        hoist_at, hoisted = {}, set()
        for (s0, f0, _e0), (s1, f1, _e1) in zip(segs, segs[1:]):
            if f0 is None or f1 is None:
                continue
            moved = []
            moved_defs = set()
            for k in range(s1, f1):
                inst = insts[k]
                if k in absorbed:
                    continue
                pure = (inst.type in _BITWISE and k in trees) or inst.type == BsInstType.ADVANCE
                if not pure or inst.ret is None or ndefs.get(inst.ret) != 1:
                    continue
                ok = True
                for u in uses(k):
                    if self._const(u) is not None or u in moved_defs:
                        continue
                    if ndefs.get(u) != 1 or def_at[u] >= f0:
                        ok = False
                        break
                if ok:
                    moved.append(k)
                    moved_defs.add(inst.ret)
            if moved:
                hoist_at[f0] = moved
                hoisted.update(moved)
        return hoist_at, hoisted

    def _fuse(self):
        """Bitwise expression trees for v_bitop3_b32 (three-input lookup
        table): an AND/OR/XOR/NOT result that feeds exactly one other bitwise
        op, and any NOT, is folded into its consumer while the consumer's tree
        has at most three distinct inputs.  Only single-definition values are
        folded, and a folded value's own inputs must not be reassigned between
        its definition and the consumer, so moving the computation to the
        consumer computes the same value.  Returns ({index: tree} for emitted
        bitwise instructions, set of absorbed instruction indices)."""
        insts = self.insts
        ndefs = {}
        def_at = {}
        uses = {}
        for k, inst in enumerate(insts):
            for d in inst.defs():
                ndefs[d] = ndefs.get(d, 0) + 1
                def_at[d] = k
            for u in inst.uses():
                uses.setdefault(u, []).append(k)
        defs_of = {}
        for k, inst in enumerate(insts):
            for d in inst.defs():
                defs_of.setdefault(d, []).append(k)
        # skip regions open at each instruction: a value defined inside a
        # region is only folded into consumers inside that region
        regions, open_ = [], set()
        for inst in insts:
            if inst.type == BsInstType.LABEL:
                open_.discard(inst.label)
            regions.append(frozenset(open_))
            if inst.type == BsInstType.SKIPZERO:
                open_.add(inst.label)

        def stable(v, lo, hi):
            # v keeps its value from instruction lo to instruction hi
            return all(not (lo < k < hi) for k in defs_of.get(v, ()))

        def foldable(v, consumer):
            if self._const(v) is not None or ndefs.get(v) != 1:
                return False
            k = def_at[v]
            inst = insts[k]
            if k >= consumer or inst.type not in _BITWISE or not regions[k] <= regions[consumer]:
                return False
            if inst.type != BsInstType.NOT and len(uses.get(v, ())) != 1:
                return False
            return all(self._const(o) is not None or stable(o, k, consumer) for o in inst.operands)

        absorbed_uses = {}
        trees = {}
        absorbed = set()

        def leaf(v):
            c = self._const(v)
            return ("const", c == "1") if c is not None else ("leaf", v)

        def expand(v, consumer, budget):
            """Tree for operand v of `consumer`, folding while the consumer's
            distinct inputs stay within three."""
            if not foldable(v, consumer):
                return leaf(v), []
            inst = insts[def_at[v]]
            kind = {BsInstType.AND: "and", BsInstType.OR: "or", BsInstType.XOR: "xor", BsInstType.NOT: "not"}[inst.type]
            subs, folded = [], [v]
            inner = list(budget)
            ops = list(inst.operands)
            for i, o in enumerate(ops):
                reserve = list(inner)
                for later_op in ops[i + 1:]:
                    reserve = _leaves(leaf(later_op), reserve)
                e, f = expand(o, consumer, reserve)
                subs.append(e)
                folded += f
                inner = _leaves(e, inner)
            tree = (kind, *subs)
            if len(_leaves(tree, list(budget))) > 3:
                return leaf(v), []
            return tree, folded

        for k in range(len(insts) - 1, -1, -1):
            inst = insts[k]
            if inst.type not in _BITWISE or k in absorbed:
                continue
            kind = {BsInstType.AND: "and", BsInstType.OR: "or", BsInstType.XOR: "xor", BsInstType.NOT: "not"}[inst.type]
            subs, folded_all, budget = [], [], []
            ops = list(inst.operands)
            for i, o in enumerate(ops):
                # keep room for the operands still to come (each at least a leaf)
                reserve = list(budget)
                for later_op in ops[i + 1:]:
                    reserve = _leaves(leaf(later_op), reserve)
                e, f = expand(o, k, reserve)
                subs.append(e)
                folded_all += f
                budget = _leaves(e, budget)
            tree = (kind, *subs)
            assert len(_leaves(tree, [])) <= 3, tree
            trees[k] = tree
            for v in folded_all:
                absorbed_uses[v] = absorbed_uses.get(v, 0) + 1
                if absorbed_uses[v] == len(uses.get(v, ())):
                    absorbed.add(def_at[v])
        return trees, absorbed

    def _emit_tree(self, tree):
        """C expression of a fused tree: plain C for a single two-input op,
        else one bg_op3 lookup (two v_bitop3_b32)."""
        kind = tree[0]
        if _nops(tree) <= 1 and all(e[0] in ("leaf", "const") for e in tree[1:]):
            vals = [("~0ull" if e[1] else "0ull") if e[0] == "const" else _cname(e[1]) for e in tree[1:]]
            if kind == "not":
                return f"~{vals[0]}"
            return f"{vals[0]} {dict(andop='&', orop='|', xorop='^')[kind + 'op']} {vals[1]}"
        leaves = _leaves(tree, [])
        if not leaves:
            return "~0ull" if _lut(tree, []) & 1 else "0ull"
        args = [_cname(v) for v in leaves] + [_cname(leaves[0])] * (3 - len(leaves))
        return f"bg_op3<0x{_lut(tree, leaves):02X}>({', '.join(args)})"

    def _carry_bounds(self, carry_index):
        """For each carry op k: {earlier carry op j: lanes of right advance
        between j's result and k's operands}, from a program-order walk.  A
        variable's map only grows (reassignments and skip-region joins merge),
        and a carry op's result depends on itself alone: its reported bound
        already covers its operands'."""
        deps = {}
        bounds = {}

        def merge(into, other, shift=0):
            for j, b in other.items():
                into[j] = max(into.get(j, 0), b + shift)

        for inst in self.insts:
            t = inst.type
            if t in (BsInstType.SKIPZERO, BsInstType.LABEL, BsInstType.STREAMSTORE, BsInstType.GRAPHBREAK):
                continue
            if t == BsInstType.STREAMLOAD:
                deps.setdefault(inst.ret, {})
                continue
            ops = {}
            for u in inst.uses():
                merge(ops, deps.get(u, {}))
            if t in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU):
                k = carry_index[id(inst)][0]
                bounds[k] = ops
                res = {k: 0}
            elif t == BsInstType.ADVANCE and not inst.is_left:
                res = {}
                merge(res, ops, inst.amount)
            else:
                res = ops
            cur = deps.setdefault(inst.ret, {})
            merge(cur, res)
        return bounds

    def launch_bounds(self, block_threads=256):
        if self.min_waves:
            return f"__launch_bounds__({block_threads}, {self.min_waves})"
        return f"__launch_bounds__({block_threads})"

    def table_decl(self, block_threads, shared=None):
        """The wave's byte-pair table in LDS (4 KiB per wave), or none."""
        shared = self.shared_filter if shared is None else bool(shared)
        if not self.bigram_keys or shared:
            return ["  uint32_t *bgt = nullptr;"]
        return [f"  __shared__ uint32_t bg_tables[{block_threads // WAVE}][BG_BIGRAM_WORDS];",
                "  uint32_t *bgt = bg_tables[threadIdx.x / BG_WAVE];"]

    def source(self):
        """Plain form: one kernel per group, one launch per (group, buffer)."""
        bgf = (f"a.filter + (size_t)t * {self.filter_row}u + {self.filter_off}u" if self.shared_filter
               else "nullptr")
        call = (f"bg_ok = bg_prog_{self.gid}(a.basis, a.result, a.status, (int64_t)a.n_words, t, Q.epoch, "
                f"a.n_carry, a.ctl + BG_CTL_ERROR, bgt, {bgf}, phase);")
        chunk = [
            "const uint32_t tl = g + (uint32_t)bg_lane();",
            "uint32_t *row = tl < a.n_tiles ? a.status + (size_t)tl * a.n_carry : nullptr;",
            "bg_tile tc0 = bg_tile_make(g, 0, (int64_t)a.n_words, a.status, a.n_carry, Q.epoch, a.ctl + BG_CTL_ERROR);",
            "const uint32_t *prev = a.status + (size_t)(g ? g - 1 : 0) * a.n_carry;",
            f"need = bg_settle_chunk<{self.n_carry}>(row, tl == 0, prev, tc0);",
        ]
        return "\n".join([
            '#include "bitgen_device.h"',
            self.device_function(),
            f'extern "C" __global__ {self.launch_bounds(self.block_threads)} void {self.kernel_name}(bg_group_args_t a, '
            'uint32_t phase) {',
            *self.table_decl(self.block_threads),
            *self.work_loop("a.n_tiles", [call], chunk),
            "}",
            *self.two_launch_marker(self.kernel_name),
            *self.shared_filter_marker(self.kernel_name),
            "",
        ])

    def work_loop(self, n_tiles, run, chunk):
        """Kernel body: persistent waves take tiles from the control block's
        queues (the next take issued before the current tile runs) and run
        each (`run`: statements for tile t, phase `phase`, that set `bg_ok`
        -- false when a phase-0 pass aborted; one call site, as the tile
        program is inlined).  Two-launch form, settle launch (phase 1): waves
        take chunks of 64 consecutive tiles from 8 queues, resolve a chunk's
        carry-ins at once (`chunk`: statements setting `need` for tiles g ..
        g + 63) and run only the tiles in `need`, in order.  Deferred
        settle: each iteration runs either a new tile's speculative pass
        (phase 0) or the settle of a tile passed earlier (phase 1): a tile is
        settled after the wave's next pass, or at once when its own pass
        aborted (its later statuses are not published until it replays, so
        it must not wait behind another pass), and the last one after the
        queue runs dry."""
        ind = lambda xs, n: [" " * n + x for x in xs]
        if self.phased and self.carry_settle == "defer":
            none = "0xFFFFFFFFu"
            return [
                f"  const bg_queue Q = bg_queue_open(a.ctl, a.n_queues, {n_tiles}, 1);",
                "  (void)phase;",
                "  uint32_t cur, next = 0;",
                "  bool more = bg_queue_take(Q, cur);",
                f"  uint32_t pend = {none}, now = {none};  // passed, not yet settled; to settle next",
                "  while (true) {",
                "    uint32_t t, ph;",
                f"    if (now != {none}) {{ t = now; ph = 1; now = {none}; }}",
                "    else if (more) { t = cur; ph = 0; next = bg_queue_issue(Q); }",
                f"    else if (pend != {none}) {{ t = pend; ph = 1; pend = {none}; }}",
                "    else break;",
                "    // the tile and the pass outcome are wave-uniform: said explicitly, as",
                "    // the compiler cannot see it through the loop (it would otherwise keep",
                "    // the tile's addresses in vector registers)",
                "    t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);",
                "    bool bg_ok;",
                "    {",
                "      const uint32_t phase = ph;",
                *ind(run, 6),
                "      bg_ok = __builtin_amdgcn_readfirstlane((int)bg_ok) != 0;",
                "    }",
                "    if (ph == 0) {",
                "      more = bg_queue_finish(Q, next, cur);",
                "      if (!bg_ok) now = t;",
                "      else { now = pend; pend = t; }",
                "    }",
                "  }",
                "  bg_queue_close(Q);",
            ]
        run = ["{", "  bool bg_ok;", *ind(run, 2), "  (void)bg_ok;", "}"]
        if not self.two_launch:
            return [
                "  // persistent: each wave takes tiles from its block's queue until the",
                "  // queue is exhausted",
                f"  const bg_queue Q = bg_queue_open(a.ctl, a.n_queues, {n_tiles}, a.take);",
                "  uint32_t first;",
                "  bool more = bg_queue_take(Q, first);",
                "  while (more) {",
                "    const uint32_t next = bg_queue_issue(Q);  // next tiles, taken ahead",
                "    for (uint32_t i = 0, t = first; i < Q.take && t < Q.n; i++, t += Q.nq) {",
                *ind(run, 6),
                "    }",
                "    more = bg_queue_finish(Q, next, first);",
                "  }",
                "  bg_queue_close(Q);",
            ]
        return [
            "  // sweep (phase 0): one tile per take from the launch's queues;",
            "  // settle (phase 1): chunks of 64 tiles from 8 queues",
            f"  const bg_queue Q = phase == 0 ? bg_queue_open(a.ctl, a.n_queues, {n_tiles}, 1)",
            f"                                : bg_queue_open(a.ctl, 8, ({n_tiles} + 63u) / 64u, 1);",
            "  uint32_t cur;",
            "  bool more = bg_queue_take(Q, cur);",
            "  uint64_t need = 0;",
            "  uint32_t g = 0, next = 0;",
            "  while (true) {",
            "    uint32_t t;",
            "    if (phase == 0) {",
            "      if (!more) break;",
            "      t = cur;",
            "      next = bg_queue_issue(Q);  // next tile, taken ahead",
            "    } else {",
            "      while (need == 0 && more) {",
            "        g = cur * 64u;",
            "        next = bg_queue_issue(Q);",
            *ind(chunk, 8),
            "        more = bg_queue_finish(Q, next, cur);",
            "      }",
            "      if (need == 0) break;",
            "      t = g + (uint32_t)__builtin_ctzll(need);",
            "      need &= need - 1;",
            "    }",
            *ind(run, 4),
            "    if (phase == 0) more = bg_queue_finish(Q, next, cur);",
            "  }",
            "  bg_queue_close(Q, phase != 0);",
        ]

    def two_launch_marker(self, kernel):
        """An empty kernel named <kernel>_two_launch: its presence in the code
        object tells the runtime to launch `kernel` twice (phase 0, then 1)."""
        if not self.two_launch:
            return []
        return [f'extern "C" __global__ void {kernel}_two_launch() {{}}']

    def shared_filter_marker(self, kernel):
        """An empty kernel named <kernel>_shared_filter: its presence tells
        the runtime that `kernel` reads its byte-pair answers from a filter
        pre-pass (a launch without them is refused)."""
        if not self.shared_filter:
            return []
        return [f'extern "C" __global__ void {kernel}_shared_filter() {{}}']

    def meta(self):
        return {
            "gid": self.gid,
            "kernel": self.kernel_name,
            "HL": self.HL,
            "HR": self.HR,
            "S": self.S,
            "n_carry": self.n_carry,
            # look-back status words per tile: one per carry op (>= 1)
            "status_slots": self.n_carry + 1 if self.inst_stats_dyn else max(1, self.n_carry),
            # statistics word of each tile at status slot n_carry
            "dyn_stats": self.inst_stats_dyn,
            # launches per run: 2 for the sweep + settle form
            "launches": 2 if self.two_launch else 1,
            "n_out": self.n_out,
            "n_bigram_tests": len(self.bigram_keys),
            # byte-pair answers from the set's filter pre-pass: this kernel's
            # word offset in a tile's row and the row length (0: none)
            "filter_off": self.filter_off if self.shared_filter else 0,
            "filter_row": self.filter_row if self.shared_filter else 0,
            "n_insts": len(self.insts),
            "regexes": list(self.regexes),
        }


def batch_kernel_name(gid):
    return f"bg_batch_{gid}"


def batch_source(program, block_threads=256):
    """Batched form of one group: a persistent kernel whose waves take tiles
    from a device work queue over many buffers (global tile index -> job
    through ``tile_job``; each job's tiles are contiguous and increasing, so a
    tile's predecessor always has a smaller index)."""
    g = program.gid
    # shared filters: the set's batched filter pre-pass wrote each global
    # tile's answer row (the kernels of a set have the same tiles per job)
    bgf = (f"a.filter + (size_t)t * {program.filter_row}u + {program.filter_off}u" if program.shared_filter
           else "nullptr")
    run = [
        "const uint32_t j = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.tile_job[t]);",
        "const bg_job_t J = a.jobs[j];",
        f"bg_ok = bg_prog_{g}(J.basis, J.result, J.status, (int64_t)J.n_words, t - J.tile_start, Q.epoch, "
        f"J.n_carry, a.ctl + BG_CTL_ERROR, bgt, {bgf}, phase);",
    ]
    chunk = [
        "// lane i: global tile g + i, its job and row; lane 0's job for the",
        "// look-back before the chunk",
        "const uint32_t tl = g + (uint32_t)bg_lane();",
        "uint32_t *row = nullptr;",
        "bool first = false;",
        "if (tl < a.total_tiles) {",
        "  const bg_job_t &Jl = a.jobs[a.tile_job[tl]];",
        "  row = Jl.status + (size_t)(tl - Jl.tile_start) * Jl.n_carry;",
        "  first = tl == Jl.tile_start;",
        "}",
        "const bg_job_t J0 = a.jobs[(uint32_t)__builtin_amdgcn_readfirstlane((int)a.tile_job[g])];",
        "const uint32_t t0 = g - J0.tile_start;",
        "bg_tile tc0 = bg_tile_make(t0, 0, (int64_t)J0.n_words, J0.status, J0.n_carry, Q.epoch, a.ctl + BG_CTL_ERROR);",
        "const uint32_t *prev = J0.status + (size_t)(t0 ? t0 - 1 : 0) * J0.n_carry;",
        f"need = bg_settle_chunk<{program.n_carry}>(row, first, prev, tc0);",
    ] if program.two_launch else []
    return "\n".join([
        '#include "bitgen_device.h"',
        program.device_function(),
        f'extern "C" __global__ {program.launch_bounds(block_threads)} void {batch_kernel_name(g)}(bg_batch_args_t a, '
        'uint32_t phase) {',
        *program.table_decl(block_threads),
        *program.work_loop("a.total_tiles", run, chunk),
        "}",
        *program.two_launch_marker(batch_kernel_name(g)),
        *program.shared_filter_marker(batch_kernel_name(g)),
        "",
    ])


FILTER_KERNEL = "bg_filter"
FILTER_BLOCK_THREADS = 256


def plan_shared_filter(programs):
    """Give the kernels of a set that read shared byte-pair answers one tile
    geometry (the widest halo of any of them) and their word offsets in a
    tile's answer row.  Returns the row length (0 when no kernel shares).
    A set with a single kernel that has filters shares nothing (the
    pre-pass would only add a launch): that kernel builds its own table."""
    progs = [p for p in programs if getattr(p, "shared_filter", False)]
    if len(progs) < 2:
        for p in progs:
            p.shared_filter = False
        return 0
    HL = max(p.HL for p in progs)
    HR = max(p.HR for p in progs)
    off = 0
    for p in progs:
        p.set_geometry(HL, HR)
        p.filter_off = off
        off += p.filter_words
    for p in progs:
        p.filter_row = off
    return off


FILTER_BATCH_KERNEL = "bg_filter_batch"


def filter_source(programs):
    """The set's filter pre-pass, in two kernels: ``bg_filter`` (one buffer)
    and ``bg_filter_batch`` (the global tiles of a batch, each mapped to its
    job and the job's own tile index through ``tile_job``, as the batch
    kernels map them).  Persistent waves take tiles (wave i of the grid:
    tiles i, i + waves, ...); for each, load the tile's window of the basis
    in the set's shared geometry, build its byte-pair table (the same
    bg_bigram_build the kernels would run) and answer the filters of every
    sharing kernel with that kernel's keys, then store the tile's answer
    row: word filter_off + i of kernel k's answers is its bgm[i].  Lane f
    holds the clauses of filter f (of each 64-filter chunk) in registers for
    the whole launch.  The answers are those each kernel computes for itself
    from the same window (bg_bigram_lookup's), so sharing changes no result."""
    progs = [p for p in programs if getattr(p, "shared_filter", False)]
    row = progs[0].filter_row
    HL, HR, S = progs[0].HL, progs[0].HR, progs[0].S
    wpb = FILTER_BLOCK_THREADS // WAVE
    lines = ['#include "bitgen_device.h"']
    for p in progs:
        keys = ", ".join(f"0x{k:08X}u" for k in p.bigram_key_words())
        lines.append(f"static __constant__ uint32_t bg_fkeys_{p.gid}[{len(p.bigram_key_words())}] = {{{keys}}};")

    # the tile's answers (t: the tile within its buffer; basis, W: the buffer)
    tile = [
        "    const int64_t w0 = (t == 0) ? 0 : (int64_t)(64 - HR) + (int64_t)(t - 1) * S - HL;",
        "    const int64_t w = w0 + lane;",
        "    const bool inb = (w >= 0) && (w < W);",
        "    const int64_t wc = w < 0 ? 0 : (w >= W ? W - 1 : w);",
        "    uint64_t bas[8];",
        "#pragma unroll",
        "    for (int i = 0; i < 8; i++) bas[i] = basis[(int64_t)i * W + wc];",
        "#pragma unroll",
        "    for (int i = 0; i < 8; i++) bas[i] = inb ? bas[i] : 0ull;",
        "    bg_bigram_build(bas, bgt);",
        f"    uint32_t *__restrict__ frow = filter + (size_t)g * {row}u;",
    ]
    for p in progs:
        nf, ncl = len(p.bigram_keys), p.bigram_clauses
        for c in range((nf + WAVE - 1) // WAVE):
            # every clause evaluated (no short-circuit: no branch per clause)
            if p.bigram_pairs == 1:
                # a lane's clause register holds its table word's LDS byte
                # address above its bit index (bg_fkey_slot): one shift for
                # the address, one for the bit
                test = " & ".join(f"bg_fkey_test(bg_tables, kr{p.gid}[{c}][{j}])" for j in range(ncl))
            else:
                test = " & ".join(f"bg_bigram_clause<{p.bigram_pairs}>(bgt, kr{p.gid}[{c}][{j}])"
                                  for j in range(ncl))
            live = min(WAVE, nf - c * WAVE)
            # lanes past the last filter: their answers dropped from the ballot
            keep = "" if live == WAVE else f" & 0x{(1 << live) - 1:016X}ull"
            # the fence keeps this chunk's table reads from being hoisted
            # ahead of the earlier chunks' (each would hold registers)
            tile += [
                "    {",
                "      " + " ".join(f"bg_fence(kr{p.gid}[{c}][{j}]);" for j in range(ncl)),
                f"      const uint64_t m = __builtin_amdgcn_ballot_w64((({test}) & 1u) != 0u){keep};",
            ]
            # the chunk's two answer words, stored by lane 0 as it goes (no
            # ballot held across chunks); rows and offsets are even
            j = p.filter_off + 2 * c
            tile += [f"      if (lane == 0) *reinterpret_cast<uint64_t *>(frow + {j}) = m;", "    }"]
    keys = ["  // every lane's clauses, once per wave"]
    for p in progs:
        nf, ncl = len(p.bigram_keys), p.bigram_clauses
        nc = (nf + WAVE - 1) // WAVE
        slot = "bg_fkey_slot(wbase, {})" if p.bigram_pairs == 1 else "{}"
        key = f"bg_fkeys_{p.gid}[(c * 64 + lane) * {ncl} + j]"
        keys += [
            f"  uint32_t kr{p.gid}[{nc}][{ncl}];",
            "#pragma unroll",
            f"  for (int c = 0; c < {nc}; c++)",
            "#pragma unroll",
            f"    for (int j = 0; j < {ncl}; j++)",
            f"      kr{p.gid}[c][j] = c * 64 + lane < {nf} ? {slot.format(key)} : 0u;",
        ]
    head = [
        f"  __shared__ uint32_t bg_tables[{wpb}][BG_BIGRAM_WORDS];",
        "  uint32_t *bgt = bg_tables[threadIdx.x / BG_WAVE];",
        "  const uint32_t wbase = (threadIdx.x / BG_WAVE) * (BG_BIGRAM_WORDS * 4u);",
        f"  constexpr int HL = {HL}, HR = {HR}, S = {S};",
        "  const int lane = bg_lane();",
        *keys,
        f"  const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * {wpb}u + threadIdx.x / BG_WAVE));",
        f"  const uint32_t stride = gridDim.x * {wpb}u;",
    ]
    lines += [
        f'extern "C" __global__ __launch_bounds__({FILTER_BLOCK_THREADS}) void {FILTER_KERNEL}('
        "const uint64_t *__restrict__ basis, uint64_t n_words, uint32_t n_tiles, uint32_t *__restrict__ filter) {",
        *head,
        "  const int64_t W = (int64_t)n_words;",
        "  for (uint32_t g = first; g < n_tiles; g += stride) {",
        "    const uint32_t t = g;",
        *tile,
        "  }",
        "}",
        "",
        f'extern "C" __global__ __launch_bounds__({FILTER_BLOCK_THREADS}) void {FILTER_BATCH_KERNEL}('
        "const bg_job_t *__restrict__ jobs, const uint32_t *__restrict__ tile_job, uint32_t total_tiles, "
        "uint32_t *__restrict__ filter) {",
        *head,
        "  for (uint32_t g = first; g < total_tiles; g += stride) {",
        "    const bg_job_t J = jobs[(uint32_t)__builtin_amdgcn_readfirstlane((int)tile_job[g])];",
        "    const uint64_t *__restrict__ basis = J.basis;",
        "    const int64_t W = (int64_t)J.n_words;",
        "    const uint32_t t = g - J.tile_start;",
        *tile,
        "  }",
        "}",
        "",
    ]
    return "\n".join(lines)


def hsaco_cache_dir():
    d = cfg.get_config("hsaco_cache_dir") or os.path.join(cfg.get_config("cache_dir", ".cache_bitgen"), "hsaco")
    os.makedirs(d, exist_ok=True)
    return d


def compile_sources(sources, workers=None):
    """Compile {path: source} to cached hsaco files; returns {path: hsaco}."""
    cache = hsaco_cache_dir()
    jobs = {}
    out = {}
    for path, src in sources.items():
        key = native_build.source_key(src)
        hs = os.path.join(cache, f"{key}.hsaco")
        out[path] = hs
        if not os.path.exists(hs):
            jobs[path] = hs
    if jobs:
        workers = workers or min(len(jobs), int(os.environ.get("BITGEN_COMPILE_JOBS", "0")) or os.cpu_count() or 4, 16)
        with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
            futs = [ex.submit(native_build.compile_kernel, p, hs) for p, hs in jobs.items()]
            for f in futs:
                f.result()
    return out


class HipGenerator(Generator):
    def __init__(self, insts_list, var_name_map_list, regex_list=None):
        super().__init__(insts_list, var_name_map_list, regex_list)
        self.programs = []

    def build_programs(self):
        """One program per group: a fused GroupProgram, or a StreamGroup when
        the group needs whole streams between its steps (graph breaks,
        while loops)."""
        from .. import compiler
        from .gen_stream import StreamGroup, needs_stream_plan

        # the library's launch options; the reference's --pass-inst-stats-dyn
        # asks for per-tile statistics (gen_cuda.py:152-172)
        opts = dict(compiler.DEFAULT_OPTIONS, inst_stats_dyn=int(bool(cfg.get_config("pass_inst_stats_dyn"))))
        step_opts = dict(opts, inst_stats_dyn=0, bigram_filter=0)
        self.programs = []
        for g, (insts, vnm) in enumerate(zip(self.insts_list, self.var_name_map_list)):
            regs = self.regex_list[g] if self.regex_list else None
            if isinstance(regs, str):
                regs = [regs]
            if needs_stream_plan(insts):
                self.programs.append(StreamGroup(
                    g, insts, vnm, regs, lambda gid, i, v, r=regs: compiler.program_for(gid, i, v, r or [], step_opts)))
            else:
                self.programs.append(compiler.program_for(g, insts, vnm, regs or [], opts))
        return self.programs

    def lower(self, input_path=None):
        wrapper = self.get_kernel_path("kernel_0.py")
        if cfg.get_config("use_cached_code") and os.path.exists(wrapper):
            print(f"Use cached code: {wrapper}")
            return wrapper
        self.build_programs()
        sources = {}
        for prog in self.programs:
            path = self.get_kernel_path(f"kernel_{prog.gid}.hip")
            src = prog.source()
            with open(path, "w") as f:
                f.write(src)
            sources[path] = src
        hsacos = compile_sources(sources)
        metas = []
        for prog in self.programs:
            m = prog.meta()
            src_path = self.get_kernel_path(f"kernel_{prog.gid}.hip")
            local = self.get_kernel_path(f"kernel_{prog.gid}.hsaco")
            with open(hsacos[src_path], "rb") as fi, open(local, "wb") as fo:
                fo.write(fi.read())
            m["hsaco"] = os.path.basename(local)
            metas.append(m)
        with open(wrapper, "w") as f:
            f.write('"""Generated by bitgen_amd.codegen.gen_hip (hip backend module)."""\n')
            f.write("import os\n")
            f.write("from bitgen_amd.backend.hip.template import init_kernel as _init_kernel\n\n")
            # a Python literal (metas hold None for stream-level groups)
            f.write(f"GROUPS = {pprint.pformat(metas, width=110)}\n")
            f.write("HERE = os.path.dirname(os.path.abspath(__file__))\n")
            f.write(f"grid_size = ({len(metas)}, {cfg.get_config('multi_input', 1)}, 1)\n")
            f.write(f"block_size = {tuple(cfg.get_config('block_size', (256, 1, 1)))}\n\n\n")
            f.write("def init_kernel(regex_func_num, input_stream_tensor, basic_stream_tensor=None, cfg={}):\n")
            f.write("    return _init_kernel(GROUPS, HERE, regex_func_num, input_stream_tensor, basic_stream_tensor, cfg)\n")
        self.gen_kernel_module_file(["kernel_0.py"])
        return wrapper
