"""ORACLE -- test infrastructure only.

Per-position twin of pablo_oracle.c: executes the same records with one Python
bool per position and the textbook per-position definitions, so the word-level
C code (shifts across words, carries across words) is checked independently:

* Advance(x, n)[p] = x[p - n]  (bitgen/template/template.py:63-70)
* MatchStar(M, C)[p] = 1 iff some q <= p has M[q] and C[q..p-1] all set
  (the closure definition; the formula in bitgen/template/template.py:96-102
  is what the C code evaluates)
* ScanThru / add: bit-serial full adder (bitgen/template/template.py:73-85)
  followed by & ~C (bitgen/template/template_torch.py:384-389)

Only for tiny inputs (pure Python loops).
"""
from .oracle import OP


def _any(x):
    return any(x)


def run_per_position(recs, nslots, basis_bits):
    """basis_bits: list of 8 lists of bools (length L). Returns list of bools."""
    L = len(basis_bits[0])
    slots = [[False] * L for _ in range(nslots)]
    out = [False] * L

    def adder(a, b):
        s, carry = [False] * L, False
        for p in range(L):
            s[p] = a[p] ^ b[p] ^ carry
            carry = (a[p] and b[p]) or (carry and (a[p] ^ b[p]))
        return s

    def run(pc, end):
        nonlocal out
        while pc < end:
            op, d, a, b, imm = [int(v) for v in recs[pc]]
            pc += 1
            A = slots[a] if a >= 0 else None
            B = slots[b] if b >= 0 else None
            if op in (OP["NOP"], OP["LABEL"]):
                continue
            if op == OP["LOAD"]:
                slots[d] = list(basis_bits[imm])
            elif op == OP["ZERO"]:
                slots[d] = [False] * L
            elif op == OP["ONES"]:
                slots[d] = [True] * L
            elif op == OP["COPY"]:
                slots[d] = list(A)
            elif op == OP["AND"]:
                slots[d] = [x and y for x, y in zip(A, B)]
            elif op == OP["OR"]:
                slots[d] = [x or y for x, y in zip(A, B)]
            elif op == OP["XOR"]:
                slots[d] = [x != y for x, y in zip(A, B)]
            elif op == OP["NOT"]:
                slots[d] = [not x for x in A]
            elif op == OP["ADVR"]:
                slots[d] = [A[p - imm] if p - imm >= 0 else False for p in range(L)]
            elif op == OP["ADVL"]:
                slots[d] = [A[p + imm] if p + imm < L else False for p in range(L)]
            elif op == OP["MSTAR"]:
                r = [False] * L
                for q in range(L):
                    if A[q]:
                        r[q] = True
                        p = q
                        while p < L and B[p]:
                            p += 1
                            if p < L:
                                r[p] = True
                slots[d] = r
            elif op == OP["SCAN"]:
                s = adder(A, B)
                slots[d] = [x and not c for x, c in zip(s, B)]
            elif op == OP["ADD"]:
                slots[d] = adder(A, B)
            elif op == OP["SEL"]:
                C = slots[imm]
                slots[d] = [(c and x) or ((not c) and y) for c, x, y in zip(C, A, B)]
            elif op == OP["WHILE"]:
                body_end = pc + imm
                guard = 0
                while _any(slots[a]):
                    guard += 1
                    if guard > 64 * L + 2:
                        raise RuntimeError("loop did not converge")
                    run(pc, body_end)
                pc = body_end
            elif op == OP["IF"]:
                then_end = pc + b
                if_end = then_end + imm
                if _any(A):
                    run(pc, then_end)
                else:
                    run(then_end, if_end)
                pc = if_end
            elif op == OP["SKIPZ"]:
                if not _any(A):
                    pc += imm
            elif op == OP["STORE"]:
                out = list(A)
            else:
                raise ValueError(f"bad op {op}")

    run(0, len(recs))
    return out


def bytes_to_bits(data, L):
    """basis_bits[i][p] = bit (7 - i) of byte p; zero beyond the data."""
    bits = [[False] * L for _ in range(8)]
    for p, c in enumerate(bytes(data)):
        for i in range(8):
            bits[i][p] = bool((c >> (7 - i)) & 1)
    return bits


def words_to_bits(words, L):
    return [bool((int(words[p // 64]) >> (p % 64)) & 1) for p in range(L)]
