"""Tile model of the generated group kernels -- test infrastructure (CPU).

A Python restatement of what one wavefront of ``bg_group_<k>`` computes for
one tile, from the generator's own geometry (``GroupProgram``: HL, HR, S,
the entry lane L of every carry op) and final instruction list:

* the tile is a window of 64 words (4,096 positions, a Python int here)
  starting at word w0 = 0 for tile 0 and S0 + (t-1)S - HL after it
  (S0 = 64 - HR); words outside the stream are zero, and so is everything
  shifted in across the window's edges;
* Advance / look-ahead are shifts of the window with zero fill;
* MatchStar / ScanThru add from lane L upward with the carry the previous
  tile hands over (its carry into lane L + S); lanes below L add per lane
  without carries between them, as the wave adder does;
* a zero test skips to its label when the value is zero over the window
  (values defined in the skipped range stay zero); a byte-pair filter test
  skips when a clause has none of its pairs among the window's byte pairs;
* the stores write lanes [0, 64 - HR) of tile 0 and [HL, HL + S) after it.

Tiles run in order with exact carry-ins (the speculation and replays of the
device only decide *when* a tile learns its carry-in, not the values).  If
the generator's halo or entry lanes are too small, the model's streams
differ from the oracle's: that is what the CPU tests check, independently
of the GPU.
"""
import numpy as np

from bitgen_amd.codegen.gen_hip import BigramTest, bigram_hash
from bitgen_amd.inst import BsInstType

WAVE = 64
BITS = WAVE * 64
MASK = (1 << BITS) - 1
LANE = (1 << 64) - 1


def _window(words, w0, W):
    """Python int of words [w0, w0 + 64) of a stream (zero outside [0, W))."""
    lo, hi = max(w0, 0), min(w0 + WAVE, W)
    if lo >= hi:
        return 0
    chunk = np.ascontiguousarray(words[lo:hi], dtype="<u8").tobytes()
    return int.from_bytes(chunk, "little") << (64 * (lo - w0))


def _window_bytes(basis_ints):
    """Byte value at each of the window's 4,096 positions (basis i = bit 7 - i)."""
    out = np.zeros(BITS, dtype=np.uint8)
    for i, x in enumerate(basis_ints):
        bits = np.unpackbits(np.frombuffer(x.to_bytes(BITS // 8, "little"), dtype=np.uint8), bitorder="little")
        out |= (bits << (7 - i)).astype(np.uint8)
    return out


def _wave_add(a, b, l0, cin, pass_lane):
    """Sum of the wave adder and its carry into lane ``pass_lane``."""
    base = 64 * l0
    lo = 0
    for lane in range(l0):
        s = (((a >> (64 * lane)) & LANE) + ((b >> (64 * lane)) & LANE)) & LANE
        lo |= s << (64 * lane)
    A, B = a >> base, b >> base
    full = A + B + cin
    s = lo | ((full << base) & MASK)
    if pass_lane >= WAVE:
        c = (full >> (BITS - base)) & 1
    else:
        c = ((full ^ A ^ B) >> (64 * (pass_lane - l0))) & 1
    return s, c


def run_tiles(prog, basis, stats=None):
    """Streams [n_out, W] computed tile by tile as the generated kernel does.
    ``prog``: a GroupProgram (its final instruction list and geometry);
    ``basis``: uint64 [streams, W] -- the 8 basis streams, or the workspace
    of a stream-level step.  ``stats``: a list that receives, per tile, the
    number of zero / byte-pair tests that fell through (what the kernel's
    statistics word counts in the tile's exact pass)."""
    W = basis.shape[1]
    HL, HR, S = prog.HL, prog.HR, prog.S
    insts = prog.insts
    labels = {inst.label: k for k, inst in enumerate(insts) if inst.type == BsInstType.LABEL}
    carry_index = {id(inst): (k, L) for k, (inst, L) in enumerate(prog.carry_ops)}
    stores = [k for k, inst in enumerate(insts) if inst.type == BsInstType.STREAMSTORE]
    out = np.zeros((len(stores), W), dtype=np.uint64)
    cin = [0] * prog.n_carry
    for t in range(prog.n_tiles(W)):
        w0 = 0 if t == 0 else (WAVE - HR) + (t - 1) * S - HL
        # the streams the program loads: the basis, or workspace slots of a
        # stream-level step (``basis`` is then the whole workspace)
        bas = {i: _window(basis[i], w0, W) for i in prog.in_slots}
        pairs = None
        env = {}
        outs = [0] * len(stores)
        cout = [0] * prog.n_carry

        def val(v):
            c = prog._const(v)
            if c is not None:
                return MASK if c == "1" else 0
            return env.get(v, 0)

        passed = 0
        pc = 0
        while pc < len(insts):
            inst = insts[pc]
            ty = inst.type
            pc += 1
            if ty == BsInstType.STREAMLOAD:
                env[inst.ret] = bas[int(inst.stream_index)]
            elif ty == BsInstType.AND:
                env[inst.ret] = val(inst.operands[0]) & val(inst.operands[1])
            elif ty == BsInstType.OR:
                env[inst.ret] = val(inst.operands[0]) | val(inst.operands[1])
            elif ty == BsInstType.XOR:
                env[inst.ret] = val(inst.operands[0]) ^ val(inst.operands[1])
            elif ty == BsInstType.NOT:
                env[inst.ret] = ~val(inst.operands[0]) & MASK
            elif ty == BsInstType.ASSIGN:
                env[inst.ret] = val(inst.operands[0])
            elif ty == BsInstType.ADVANCE:
                x = val(inst.operands[0])
                env[inst.ret] = (x >> inst.amount) if inst.is_left else ((x << inst.amount) & MASK)
            elif ty in (BsInstType.MATCHSTAR, BsInstType.SCANTHRU):
                k, L = carry_index[id(inst)]
                l0 = 0 if t == 0 else L
                m, c = val(inst.operands[0]), val(inst.operands[1])
                if ty == BsInstType.MATCHSTAR:
                    s, cout[k] = _wave_add(m & c, c, l0, cin[k] if t else 0, L + S)
                    env[inst.ret] = ((s ^ c) | m) & MASK
                else:
                    s, cout[k] = _wave_add(m, c, l0, cin[k] if t else 0, L + S)
                    env[inst.ret] = s & ~c & MASK
            elif isinstance(inst, BigramTest):
                if pairs is None:
                    b = _window_bytes([bas[i] for i in range(8)])
                    nxt = np.append(b[1:], 0)
                    a64, n64 = b.astype(np.int64), nxt.astype(np.int64)
                    pairs = set(((((a64 & 3) << 8 | n64) << 5) | ((a64 >> 2) & 31)).tolist())
                if not all(any(h in pairs for h in clause) for clause in inst.clauses):
                    pc = labels[inst.label] + 1
                else:
                    passed += 1
            elif ty == BsInstType.SKIPZERO:
                if val(inst.operands[0]) == 0:
                    pc = labels[inst.label] + 1
                else:
                    passed += 1
            elif ty == BsInstType.STREAMSTORE:
                outs[stores.index(pc - 1)] = val(inst.operands[0])
            elif ty in (BsInstType.LABEL, BsInstType.GRAPHBREAK):
                pass
            else:
                raise NotImplementedError(f"tile model: {inst}")
        cin = cout
        if stats is not None:
            stats.append(passed)
        lo, hi = (0, WAVE - HR) if t == 0 else (HL, HL + S)
        for j, x in enumerate(outs):
            for lane in range(lo, hi):
                w = w0 + lane
                if 0 <= w < W:
                    out[j, w] = (x >> (64 * lane)) & LANE
    return out
