"""Stream-level (unfused) execution of a group program.

The fused form (``GroupProgram``) runs a whole group program per tile.  Two
cases need whole streams between steps instead:

* the reference's ablation modes ``--pass-graph-break`` 0, 1 and 2
  (``bitgen/passes/graph_break.py:368-407``), which split a program into
  subgraphs that hand streams to each other through global memory
  (``tmp_streams``, ``bitgen/codegen/gen_cuda.py:191-195``);
* ``while`` loops other than the single-class stars that become MatchStar:
  the loop repeats while its condition stream has a set bit anywhere in the
  stream (``bitgen/inst/inst_block_while.py:50-70``), a whole-stream
  property that no tile can decide alone.

A plan is a list of steps over a per-input workspace of streams: slots 0-7
hold the basis, later slots hold temporaries.  A ``KernelStep`` is a
straight-line subgraph compiled as a tile-parallel group kernel (same halo,
carry look-back and geometry as the fused form) that loads its inputs from
workspace slots and stores each output to a fresh slot; a ``LoopStep`` runs
its body steps while its condition slot holds a set bit; ``CopyStep`` and
``ZeroStep`` move and clear slots.  A kernel never reads a slot it writes
(outputs always get fresh slots), so tiles never see a neighbour's new
values through their halo; loop-carried values are copied back to their
entry slots at the end of each iteration, so the body's kernels (compiled
once) read the same slots in every iteration.
"""
from dataclasses import dataclass, field
from typing import List

from ..inst import BsInstType, BsStreamLoad, BsStreamStore

N_BASIS = 8


class StreamPlanError(NotImplementedError):
    pass


@dataclass
class KernelStep:
    insts: list           # straight-line instructions (loads/stores added)
    outputs: List[int]    # variables stored, in store order
    out_base: int = 0     # first output slot (outputs are consecutive)
    program: object = None


@dataclass
class LoopStep:
    cond_slot: int
    body: list = field(default_factory=list)


@dataclass
class CopyStep:
    src: int
    dst: int


@dataclass
class ZeroStep:
    slot: int


_STRAIGHT = (BsInstType.AND, BsInstType.OR, BsInstType.XOR, BsInstType.NOT, BsInstType.ASSIGN,
             BsInstType.ADVANCE, BsInstType.MATCHSTAR, BsInstType.SCANTHRU, BsInstType.STREAMLOAD)


def _walk_uses(insts):
    for inst in insts:
        yield from inst.uses()
        if inst.type == BsInstType.WHILE:
            yield from _walk_uses(inst.body)
        elif inst.type == BsInstType.IF:
            yield from _walk_uses(inst.body_1)
            yield from _walk_uses(inst.body_2 or [])


def _walk_defs(insts):
    for inst in insts:
        yield from inst.defs()
        if inst.type == BsInstType.WHILE:
            yield from _walk_defs(inst.body)


def _exposed(insts):
    """Variables an instruction list reads before defining them (in program
    order; a nested loop's reads count as exposed unless defined before it)."""
    defined, exposed = set(), set()
    for inst in insts:
        if inst.type == BsInstType.WHILE:
            exposed.update(u for u in _walk_uses([inst]) if u not in defined)
            defined.update(_walk_defs(inst.body))
            continue
        exposed.update(u for u in inst.uses() if u not in defined)
        defined.update(inst.defs())
    return exposed


def _segments(insts):
    """Top-level pieces: lists of straight-line instructions (split at graph
    breaks) and single WHILE instructions."""
    out, cur = [], []
    for inst in insts:
        t = inst.type
        if t in (BsInstType.SKIPZERO, BsInstType.LABEL):
            continue  # zero-block skipping is exact either way; dropped here
        if t == BsInstType.GRAPHBREAK:
            if cur:
                out.append(cur)
            cur = []
        elif t == BsInstType.WHILE:
            if cur:
                out.append(cur)
            cur = []
            out.append(inst)
        elif t == BsInstType.IF:
            raise StreamPlanError("if blocks are not supported by the hip backend")
        elif t in _STRAIGHT or t == BsInstType.STREAMSTORE:
            cur.append(inst)
        else:
            raise StreamPlanError(f"unsupported instruction {inst}")
    if cur:
        out.append(cur)
    return out


class StreamPlanner:
    """Builds the steps of one group program.  ``vnm``: variable names (the
    constants "0"/"1" are not variables)."""

    def __init__(self, insts, vnm):
        self.vnm = vnm
        self.n_slots = N_BASIS
        self.slot = {}       # variable -> slot holding its current value
        self.result_slot = None
        self.steps = self._plan(insts, outer_uses=set())

    def _const(self, v):
        n = self.vnm.get(v)
        return n if n in ("0", "1") else None

    def _fresh(self, n=1):
        base = self.n_slots
        self.n_slots += n
        return base

    def _plan(self, insts, outer_uses):
        pieces = _segments(insts)
        steps = []
        for i, piece in enumerate(pieces):
            later = set(outer_uses)
            for p in pieces[i + 1:]:
                later.update(_walk_uses(p if isinstance(p, list) else [p]))
            if isinstance(piece, list):
                steps.extend(self._kernel(piece, later))
            else:
                steps.extend(self._loop(piece, later))
        return steps

    def _kernel(self, piece, later):
        """One straight-line subgraph: inputs are loaded from their slots,
        values used later (or the group's stored stream) go to fresh slots."""
        body, defined, outputs, store = [], set(), [], None
        loads = {}
        for inst in piece:
            if inst.type == BsInstType.STREAMLOAD:
                # basis streams sit in workspace slots 0-7
                self.slot[inst.ret] = int(inst.stream_index)
                continue
            if inst.type == BsInstType.STREAMSTORE:
                store = inst.operands[0]
                continue
            for u in inst.uses():
                if self._const(u) is None and u not in defined and u not in loads:
                    if u not in self.slot:
                        raise StreamPlanError(f"variable {u} ({self.vnm.get(u)}) is read before any definition")
                    loads[u] = self.slot[u]
            body.append(inst.deepcopy())
            defined.update(inst.defs())
        head = [BsStreamLoad([0], v, stream_index=s) for v, s in loads.items()]
        outputs = [v for v in dict.fromkeys(d for inst in body for d in inst.defs()) if v in later]
        if store is not None and store not in outputs:
            if self._const(store) is not None:
                raise StreamPlanError("a constant group result is not supported")
            if store not in defined and store in self.slot:
                # the stored stream was computed by an earlier step
                self.result_slot = self.slot[store]
                store = None
            else:
                outputs.append(store)
        if not outputs:
            return []
        base = self._fresh(len(outputs))
        insts = head + body + [BsStreamStore([v], None) for v in outputs]
        for j, v in enumerate(outputs):
            self.slot[v] = base + j
        if store is not None:
            self.result_slot = self.slot[store]
        return [KernelStep(insts, outputs, base)]

    def _loop(self, loop, later):
        cond = loop.condition
        if self._const(cond) is not None:
            raise StreamPlanError("while on a constant condition")
        body_defs = set(_walk_defs(loop.body))
        # read by the next iteration (before the body redefines them), by the
        # condition, or after the loop
        live = _exposed(loop.body) | {cond}
        # values the loop changes that are read after it or by a later
        # iteration (or its condition) keep one slot across iterations
        carried = sorted(v for v in body_defs if v in later or v in live)
        steps = []
        entry = {}
        for v in carried:
            if v in self.slot:
                entry[v] = self.slot[v]
            else:
                # not defined before the loop: 0 if the loop never runs
                entry[v] = self._fresh()
                steps.append(ZeroStep(entry[v]))
                self.slot[v] = entry[v]
        if cond not in self.slot:
            raise StreamPlanError(f"loop condition {cond} is not defined before the loop")
        cond_slot = self.slot[cond]
        body_steps = self._plan(loop.body, outer_uses=later | live)
        for v in carried:
            if self.slot[v] != entry[v]:
                body_steps.append(CopyStep(self.slot[v], entry[v]))
            self.slot[v] = entry[v]
        steps.append(LoopStep(cond_slot, body_steps))
        return steps


def plan_program(insts, vnm):
    """Steps of one group program, the workspace slot count and the slot of
    the group's result stream."""
    p = StreamPlanner(insts, vnm)
    if p.result_slot is None:
        raise StreamPlanError("the program stores no stream")
    return p.steps, p.n_slots, p.result_slot


def kernel_steps(steps):
    """Every KernelStep of a plan, in execution order of first appearance."""
    for s in steps:
        if isinstance(s, KernelStep):
            yield s
        elif isinstance(s, LoopStep):
            yield from kernel_steps(s.body)


def needs_stream_plan(insts):
    """True when the fused form cannot run the program: graph breaks (the
    reference's unfused ablation modes) or while loops."""
    from ..inst import walk

    return any(i.type in (BsInstType.GRAPHBREAK, BsInstType.WHILE) for i in walk(insts))


def _serialize(steps):
    out = []
    for s in steps:
        if isinstance(s, KernelStep):
            out.append(["kernel", s.index, s.out_base])
        elif isinstance(s, CopyStep):
            out.append(["copy", s.src, s.dst])
        elif isinstance(s, ZeroStep):
            out.append(["zero", s.slot])
        else:
            out.append(["loop", s.cond_slot, _serialize(s.body)])
    return out


class StreamGroup:
    """A group program run at stream level: every KernelStep of its plan
    compiled as a group kernel ``bg_group_<gid>_s<i>`` (one code object for
    all of them), plus the plan the host runs.  ``make_program(gid, insts,
    vnm)`` builds the GroupProgram of one step under the set's options."""

    def __init__(self, gid, insts, vnm, regexes, make_program):
        self.gid = gid
        self.regexes = list(regexes or [])
        self.steps, self.n_slots, self.result_slot = plan_program(insts, vnm)
        self.programs = []
        for i, ks in enumerate(kernel_steps(self.steps)):
            ks.index = i
            ks.program = make_program(f"{gid}_s{i}", ks.insts, vnm)
            self.programs.append(ks.program)
        if not self.programs:
            raise StreamPlanError(f"group {gid}: no kernel step")
        self.n_out = 1
        self.groups = [gid]

    @property
    def kernel_name(self):
        return f"bg_stream_{self.gid}"

    def source(self):
        parts = [p.source() for p in self.programs]
        return "\n".join(parts)

    def meta(self):
        steps = [p.meta() for p in self.programs]
        return {
            "kind": "stream",
            "gid": self.gid,
            "kernel": None,
            "n_out": 1,
            "n_slots": self.n_slots,
            "result_slot": self.result_slot,
            "plan": _serialize(self.steps),
            "step_kernels": steps,
            "status_slots": max(m["status_slots"] for m in steps),
            "n_carry": max(m["n_carry"] for m in steps),
            "HL": max(m["HL"] for m in steps),
            "HR": max(m["HR"] for m in steps),
            "S": min(m["S"] for m in steps),
            "n_bigram_tests": 0,
            "n_insts": sum(m["n_insts"] for m in steps),
            "regexes": self.regexes,
        }
