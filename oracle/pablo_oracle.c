/*
 * ORACLE -- test infrastructure only.  Not part of the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this.
 *
 * CPU restatement of the reference's bitstream semantics, executed over whole
 * streams of 64-bit words (position p = bit p%64 of word p/64, LSB-first):
 *
 *   transpose       bitgen/template/template.py:135-145 (basis0 = MSB of the
 *                   byte) and bitgen/template/cuda/kernel.cu:60-100
 *   and/or/xor/not  bitgen/backend/cuda/template/template.cu:122-159
 *   Advance(x,n)    bitgen/template/template.py:63-70 (toward higher positions)
 *   left Advance    bitgen/backend/cuda/template/template.cu:296-337
 *                   (look-ahead form used by the cc-advance pass)
 *   MatchStar       bitgen/template/template.py:96-102:
 *                   (((M & C) + C) ^ C) | M
 *   ScanThru        bitgen/template/template_torch.py:377-389: (M + C) & ~C
 *                   (template.py:88-93 is not followed: it mutates M and
 *                   masks with ~M)
 *   add             bitgen/template/template.py:73-85 (full adder chain,
 *                   carries toward higher positions)
 *   while           bitgen/inst/inst_block_while.py:61-68 (repeat while the
 *                   condition has any set bit)
 *   if              bitgen/inst/inst_block_if.py:366-376
 *   select          bitgen/inst/inst_sel.py:42-57
 *   zero skip       bitgen/passes/short_circuit.py:318-350 (jump when the
 *                   value is all zero)
 *
 * Program format: records of 5 int32 {op, dst, a, b, imm}; see enum below.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
  OP_NOP = 0,
  OP_LOAD = 1,   /* dst = basis[imm]                                  */
  OP_ZERO = 2,   /* dst = 0                                           */
  OP_ONES = 3,   /* dst = all ones                                    */
  OP_COPY = 4,   /* dst = a                                           */
  OP_AND = 5,
  OP_OR = 6,
  OP_XOR = 7,
  OP_NOT = 8,    /* dst = ~a                                          */
  OP_ADVR = 9,   /* dst[p] = a[p - imm]                               */
  OP_ADVL = 10,  /* dst[p] = a[p + imm]                               */
  OP_MSTAR = 11, /* dst = MatchStar(a, b)                             */
  OP_SCAN = 12,  /* dst = ScanThru(a, b)                              */
  OP_SEL = 13,   /* dst = (slot[imm] & a) | (~slot[imm] & b)          */
  OP_WHILE = 14, /* while any(a): run next imm records                */
  OP_IF = 15,    /* if any(a): next b records else following imm      */
  OP_SKIPZ = 16, /* if !any(a): skip next imm records                 */
  OP_STORE = 17, /* out = a                                           */
  OP_LABEL = 18,
  OP_ADD = 19    /* dst = a + b                                       */
};

typedef struct {
  const int32_t *prog;
  int32_t nrec;
  uint64_t **slot;
  const uint64_t *basis; /* [8][W] */
  uint64_t W;
  uint64_t *out;
  uint64_t *tmp;
  uint64_t max_iter;
  int err;
} vm_t;

static int any_set(const uint64_t *x, uint64_t W) {
  for (uint64_t i = 0; i < W; i++)
    if (x[i]) return 1;
  return 0;
}

static void adv_right(uint64_t *y, const uint64_t *x, uint64_t W, uint64_t n, uint64_t *tmp) {
  uint64_t q = n / 64, r = n % 64;
  for (uint64_t w = 0; w < W; w++) {
    uint64_t hi = (w >= q) ? x[w - q] : 0;
    uint64_t lo = (w >= q + 1) ? x[w - q - 1] : 0;
    tmp[w] = r ? ((hi << r) | (lo >> (64 - r))) : hi;
  }
  memcpy(y, tmp, W * sizeof(uint64_t));
}

static void adv_left(uint64_t *y, const uint64_t *x, uint64_t W, uint64_t n, uint64_t *tmp) {
  uint64_t q = n / 64, r = n % 64;
  for (uint64_t w = 0; w < W; w++) {
    uint64_t lo = (w + q < W) ? x[w + q] : 0;
    uint64_t hi = (w + q + 1 < W) ? x[w + q + 1] : 0;
    tmp[w] = r ? ((lo >> r) | (hi << (64 - r))) : lo;
  }
  memcpy(y, tmp, W * sizeof(uint64_t));
}

/* tmp = a + b over the whole stream, carry toward higher words. */
static void add_stream(uint64_t *tmp, const uint64_t *a, const uint64_t *b, uint64_t W) {
  uint64_t carry = 0;
  for (uint64_t w = 0; w < W; w++) {
    uint64_t s = a[w] + b[w];
    uint64_t c1 = s < a[w];
    uint64_t s2 = s + carry;
    uint64_t c2 = s2 < s;
    tmp[w] = s2;
    carry = c1 | c2;
  }
}

static int run_range(vm_t *vm, int32_t pc, int32_t end);

static int step(vm_t *vm, int32_t *ppc) {
  const int32_t *r = vm->prog + 5 * (*ppc);
  int32_t op = r[0], d = r[1], a = r[2], b = r[3], imm = r[4];
  uint64_t W = vm->W;
  uint64_t *D = (d >= 0) ? vm->slot[d] : NULL;
  const uint64_t *A = (a >= 0) ? vm->slot[a] : NULL;
  const uint64_t *B = (b >= 0) ? vm->slot[b] : NULL;
  uint64_t i;
  *ppc += 1;
  switch (op) {
  case OP_NOP:
  case OP_LABEL:
    return 0;
  case OP_LOAD:
    if (imm < 0 || imm > 7) return -2;
    memcpy(D, vm->basis + (uint64_t)imm * W, W * sizeof(uint64_t));
    return 0;
  case OP_ZERO:
    memset(D, 0, W * sizeof(uint64_t));
    return 0;
  case OP_ONES:
    memset(D, 0xff, W * sizeof(uint64_t));
    return 0;
  case OP_COPY:
    if (D != A) memmove(D, A, W * sizeof(uint64_t));
    return 0;
  case OP_AND:
    for (i = 0; i < W; i++) D[i] = A[i] & B[i];
    return 0;
  case OP_OR:
    for (i = 0; i < W; i++) D[i] = A[i] | B[i];
    return 0;
  case OP_XOR:
    for (i = 0; i < W; i++) D[i] = A[i] ^ B[i];
    return 0;
  case OP_NOT:
    for (i = 0; i < W; i++) D[i] = ~A[i];
    return 0;
  case OP_ADVR:
    adv_right(D, A, W, (uint64_t)imm, vm->tmp);
    return 0;
  case OP_ADVL:
    adv_left(D, A, W, (uint64_t)imm, vm->tmp);
    return 0;
  case OP_MSTAR: {
    /* t = (M & C) + C ; out = (t ^ C) | M */
    uint64_t *t = vm->tmp, carry = 0;
    for (i = 0; i < W; i++) {
      uint64_t m = A[i] & B[i];
      uint64_t s = m + B[i];
      uint64_t c1 = s < m;
      uint64_t s2 = s + carry;
      uint64_t c2 = s2 < s;
      t[i] = s2;
      carry = c1 | c2;
    }
    for (i = 0; i < W; i++) t[i] = (t[i] ^ B[i]) | A[i];
    memcpy(D, t, W * sizeof(uint64_t));
    return 0;
  }
  case OP_SCAN:
    add_stream(vm->tmp, A, B, W);
    for (i = 0; i < W; i++) vm->tmp[i] &= ~B[i];
    memcpy(D, vm->tmp, W * sizeof(uint64_t));
    return 0;
  case OP_ADD:
    add_stream(vm->tmp, A, B, W);
    memcpy(D, vm->tmp, W * sizeof(uint64_t));
    return 0;
  case OP_SEL: {
    const uint64_t *C = vm->slot[imm];
    for (i = 0; i < W; i++) vm->tmp[i] = (C[i] & A[i]) | (~C[i] & B[i]);
    memcpy(D, vm->tmp, W * sizeof(uint64_t));
    return 0;
  }
  case OP_WHILE: {
    int32_t body = *ppc, body_end = *ppc + imm;
    uint64_t it = 0;
    while (any_set(A, W)) {
      if (++it > vm->max_iter) return -3;
      int rc = run_range(vm, body, body_end);
      if (rc) return rc;
      A = vm->slot[a];
    }
    *ppc = body_end;
    return 0;
  }
  case OP_IF: {
    int32_t then_start = *ppc, else_start = *ppc + b, end = else_start + imm;
    int rc = any_set(A, W) ? run_range(vm, then_start, else_start)
                           : run_range(vm, else_start, end);
    *ppc = end;
    return rc;
  }
  case OP_SKIPZ:
    if (!any_set(A, W)) *ppc += imm;
    return 0;
  case OP_STORE:
    memcpy(vm->out, A, W * sizeof(uint64_t));
    return 0;
  default:
    return -1;
  }
}

static int run_range(vm_t *vm, int32_t pc, int32_t end) {
  while (pc < end) {
    int rc = step(vm, &pc);
    if (rc) return rc;
  }
  return 0;
}

/* Execute a flat program.  basis: [8][W] words, out: [W] words.
 * Returns 0 on success, -1 bad opcode, -2 bad operand, -3 loop did not
 * converge within max_iter iterations, -4 out of memory. */
int bg_oracle_run(const int32_t *prog, int32_t nrec, int32_t nslots,
                  const uint64_t *basis, uint64_t W, uint64_t *out) {
  vm_t vm;
  uint64_t *mem;
  int32_t s;
  int rc;
  if (W == 0) return 0;
  mem = (uint64_t *)calloc((size_t)(nslots + 1) * W, sizeof(uint64_t));
  vm.slot = (uint64_t **)calloc((size_t)nslots + 1, sizeof(uint64_t *));
  if (!mem || !vm.slot) {
    free(mem);
    free(vm.slot);
    return -4;
  }
  for (s = 0; s < nslots; s++) vm.slot[s] = mem + (uint64_t)s * W;
  vm.tmp = mem + (uint64_t)nslots * W;
  vm.prog = prog;
  vm.nrec = nrec;
  vm.basis = basis;
  vm.W = W;
  vm.out = out;
  vm.max_iter = 64 * W + 2;
  vm.err = 0;
  memset(out, 0, W * sizeof(uint64_t));
  rc = run_range(&vm, 0, nrec);
  free(mem);
  free(vm.slot);
  return rc;
}

/* bytes -> basis[8][W] (LSB-first positions, basis0 = bit 7 of the byte).
 * Positions >= n_bytes are zero.  W must be >= ceil(n_bytes / 64). */
int bg_oracle_transpose(const uint8_t *bytes, uint64_t n_bytes, uint64_t *basis, uint64_t W) {
  uint64_t p;
  int i;
  if (W * 64 < n_bytes) return -2;
  memset(basis, 0, 8 * W * sizeof(uint64_t));
  for (p = 0; p < n_bytes; p++) {
    uint8_t c = bytes[p];
    for (i = 0; i < 8; i++)
      if ((c >> (7 - i)) & 1) basis[(uint64_t)i * W + p / 64] |= (uint64_t)1 << (p % 64);
  }
  return 0;
}
