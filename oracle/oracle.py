"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(),
bench.py cpu_baseline).  The product never imports this module.

Python side of the CPU oracle:

* ``flatten(insts, var_name_map)`` serialises the bitstream IR
  (``bitgen_amd.inst``) into the record format of ``pablo_oracle.c``, assigning
  storage slots by live range so large inputs fit in memory.
* ``run(insts, var_name_map, basis, n_words)`` executes it whole-stream in C.
* ``transpose(data, n_words)`` is the byte -> basis transpose
  (``bitgen/template/template.py:135-145``).
* ``match_positions`` / ``popcount`` decode result words like the reference's
  ``Bitstream`` (``bitgen/bitstream.py:30-67``) but with LSB-first words.

``run_per_position`` (np_ref.py) re-executes the same records with one Python
bool per position; tests use it to pin the word-level C code.
"""
import ctypes
import os
import subprocess
import threading

import numpy as np

from bitgen_amd.inst import BsInstType

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OP = dict(NOP=0, LOAD=1, ZERO=2, ONES=3, COPY=4, AND=5, OR=6, XOR=7, NOT=8,
          ADVR=9, ADVL=10, MSTAR=11, SCAN=12, SEL=13, WHILE=14, IF=15,
          SKIPZ=16, STORE=17, LABEL=18, ADD=19)

_lib = None
_lib_lock = threading.Lock()


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                build()
            L = ctypes.CDLL(LIB_PATH)
            L.bg_oracle_run.argtypes = [
                ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
            ]
            L.bg_oracle_run.restype = ctypes.c_int
            L.bg_oracle_transpose.argtypes = [
                ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
            ]
            L.bg_oracle_transpose.restype = ctypes.c_int
            _lib = L
    return _lib


def n_words_for(n_bytes):
    """Words per stream: positions 0..n_bytes inclusive (a match may end at n)."""
    return (n_bytes + 1 + 63) // 64


def transpose(data, n_words=None):
    data = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = int(data.size)
    W = n_words_for(n) if n_words is None else int(n_words)
    basis = np.zeros((8, W), dtype=np.uint64)
    rc = lib().bg_oracle_transpose(data.ctypes.data, n, basis.ctypes.data, W)
    if rc != 0:
        raise ValueError(f"oracle transpose failed ({rc})")
    return basis


# ---------------------------------------------------------------------------
# IR -> records
class _Flattener:
    def __init__(self, var_name_map):
        self.names = var_name_map
        self.recs = []
        self.labels = {}      # label -> record index of its BsLabel
        self.skips = []       # (record index, label)
        self.order = []       # (var, record index) for every def/use, in order
        self.loops = []       # (start, end) record ranges of loop bodies

    def _const(self, v):
        name = self.names.get(v, None)
        return name if name in ("0", "1") else None

    def rec(self, op, dst=-1, a=-1, b=-1, imm=0):
        self.recs.append([OP[op], dst, a, b, imm])
        return len(self.recs) - 1

    def val(self, v):
        """Operand: constants become fresh ZERO/ONES temporaries."""
        c = self._const(v)
        if c is not None:
            t = ("const", len(self.recs))
            self.rec("ZERO" if c == "0" else "ONES", dst=t)
            return t
        return v

    def body(self, insts):
        for inst in insts:
            self.one(inst)

    def one(self, inst):
        t = inst.type
        if t == BsInstType.STREAMLOAD:
            self.rec("LOAD", dst=inst.ret, imm=int(inst.stream_index))
        elif t in (BsInstType.AND, BsInstType.OR, BsInstType.XOR):
            a, b = self.val(inst.operands[0]), self.val(inst.operands[1])
            self.rec({BsInstType.AND: "AND", BsInstType.OR: "OR", BsInstType.XOR: "XOR"}[t],
                     dst=inst.ret, a=a, b=b)
        elif t == BsInstType.NOT:
            self.rec("NOT", dst=inst.ret, a=self.val(inst.operands[0]))
        elif t == BsInstType.ASSIGN:
            c = self._const(inst.operands[0])
            if c is not None:
                self.rec("ZERO" if c == "0" else "ONES", dst=inst.ret)
            else:
                self.rec("COPY", dst=inst.ret, a=inst.operands[0])
        elif t == BsInstType.ADVANCE:
            op = "ADVL" if inst.is_left else "ADVR"
            self.rec(op, dst=inst.ret, a=self.val(inst.operands[0]), imm=inst.amount)
        elif t == BsInstType.MATCHSTAR:
            self.rec("MSTAR", dst=inst.ret, a=self.val(inst.operands[0]), b=self.val(inst.operands[1]))
        elif t == BsInstType.SCANTHRU:
            self.rec("SCAN", dst=inst.ret, a=self.val(inst.operands[0]), b=self.val(inst.operands[1]))
        elif t == BsInstType.SEL:
            a, b = self.val(inst.operands[0]), self.val(inst.operands[1])
            self.rec("SEL", dst=inst.ret, a=a, b=b, imm=self.val(inst.condition))
        elif t == BsInstType.WHILE:
            cond = self.val(inst.condition)
            w = self.rec("WHILE", a=cond)
            start = len(self.recs)
            self.body(inst.body)
            if self._const(inst.condition) is not None:
                raise ValueError("while on a constant condition")
            self.recs[w][4] = len(self.recs) - start
            self.loops.append((w, len(self.recs)))
        elif t == BsInstType.IF:
            cond = self.val(inst.condition)
            i = self.rec("IF", a=cond)
            s1 = len(self.recs)
            self.body(inst.body_1)
            s2 = len(self.recs)
            self.body(inst.body_2 or [])
            self.recs[i][3] = s2 - s1
            self.recs[i][4] = len(self.recs) - s2
        elif t == BsInstType.SKIPZERO:
            i = self.rec("SKIPZ", a=self.val(inst.operands[0]))
            self.skips.append((i, inst.label))
        elif t == BsInstType.LABEL:
            self.labels[inst.label] = self.rec("LABEL")
        elif t == BsInstType.STREAMSTORE:
            self.rec("STORE", a=self.val(inst.operands[0]))
        elif t in (BsInstType.GRAPHBREAK, BsInstType.STR):
            pass
        else:
            raise NotImplementedError(f"oracle: {inst}")

    def finish(self):
        for i, label in self.skips:
            j = self.labels[label]
            if j <= i:
                raise ValueError("zero-skip label before its check")
            self.recs[i][4] = j - i - 1
        return self


def _assign_slots(recs, loops):
    """Map variable keys to storage slots by live range (loops keep every
    value they touch alive across the whole loop)."""
    first, last = {}, {}
    for idx, r in enumerate(recs):
        op = r[0]
        keys = [r[1], r[2]] if op == OP["IF"] else [r[1], r[2], r[3]]  # IF: b is a length
        if op == OP["SEL"]:
            keys.append(r[4])
        for k in keys:
            if k == -1 or k is None:
                continue
            first.setdefault(k, idx)
            last[k] = idx
    for (s, e) in loops:
        for k in list(first):
            if first[k] <= e and last[k] >= s:
                first[k] = min(first[k], s)
                last[k] = max(last[k], e)
    events = sorted(first, key=lambda k: first[k])
    free, active, slot_of = [], [], {}
    nslots = 0
    for k in events:
        start = first[k]
        still = []
        for (end, kk) in active:
            if end < start:
                free.append(slot_of[kk])
            else:
                still.append((end, kk))
        active = still
        if free:
            slot_of[k] = free.pop()
        else:
            slot_of[k] = nslots
            nslots += 1
        active.append((last[k], k))
    return slot_of, nslots


def flatten(insts, var_name_map):
    """Return (records int32[n,5], nslots)."""
    f = _Flattener(var_name_map)
    f.body(insts)
    f.finish()
    slot_of, nslots = _assign_slots(f.recs, f.loops)
    out = np.zeros((len(f.recs), 5), dtype=np.int32)
    for i, r in enumerate(f.recs):
        op, d, a, b, imm = r
        if op == OP["SEL"]:
            imm = slot_of[imm]
        if op != OP["IF"]:  # IF: b is the length of the then-body
            b = slot_of.get(b, -1) if b != -1 else -1
        out[i] = [op, slot_of.get(d, -1) if d != -1 else -1,
                  slot_of.get(a, -1) if a != -1 else -1, b, imm]
    return out, nslots


def run_records(recs, nslots, basis):
    basis = np.ascontiguousarray(basis, dtype=np.uint64)
    W = basis.shape[1]
    out = np.zeros(W, dtype=np.uint64)
    recs = np.ascontiguousarray(recs, dtype=np.int32)
    rc = lib().bg_oracle_run(recs.ctypes.data, recs.shape[0], nslots,
                             basis.ctypes.data, W, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle run failed ({rc})")
    return out


def run(insts, var_name_map, basis):
    recs, nslots = flatten(insts, var_name_map)
    return run_records(recs, nslots, basis)


def mask_to_positions(words, n_positions):
    """Clear every bit at or beyond n_positions."""
    words = np.array(words, dtype=np.uint64, copy=True)
    W = words.shape[-1]
    full = int(n_positions) // 64  # Python ints: 1 << 63 must not wrap
    if full < W:
        rem = int(n_positions) % 64
        words[..., full] &= np.uint64((1 << rem) - 1) if rem else np.uint64(0)
        words[..., full + 1:] = 0
    return words


def popcount(words):
    w = np.ascontiguousarray(words, dtype=np.uint64)
    return int(np.unpackbits(w.view(np.uint8)).sum())


def match_positions(words):
    w = np.ascontiguousarray(words, dtype=np.uint64).reshape(-1)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")
    return np.nonzero(bits)[0]
