// Byte -> basis bit transpose of one 64-byte block, shared by the gfx950
// transpose kernel and the host unit test that checks it with g++.
//
// In:  x[j] = bytes 8j..8j+7 of the block, little-endian (byte b of x[j] at
//      bits 8b..8b+7).
// Out: y[i] = basis word i: bit p set iff bit (7 - i) of byte p is set
//      (basis0 = most significant bit, bitgen/template/template.py:135-145).
//
// Two stages, all register-resident: an 8x8 bit transpose inside every word
// (3 delta swaps) turns byte b of x[j] into "bit b of bytes 8j..8j+7", then an
// 8x8 byte transpose across the eight words (3 swap stages) gathers byte b of
// every word into y[7 - b].
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define BG_HD __host__ __device__ __forceinline__
// fully unrolled: an indexed loop over the register arrays would move its
// words through relative register addressing
#define BG_UNROLL8 _Pragma("unroll")
#else
#define BG_HD static inline
#define BG_UNROLL8
#endif

BG_HD uint64_t bg_bit_transpose8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

// swap the `mask` lanes of a with the (mask << s) lanes of b
#define BG_SWAP(a, b, s, mask)                                  \
  do {                                                          \
    uint64_t _t = (((a) >> (s)) ^ (b)) & (mask);                \
    (b) ^= _t;                                                  \
    (a) ^= _t << (s);                                           \
  } while (0)

BG_HD void bg_transpose_block(const uint64_t x_in[8], uint64_t y[8]) {
  uint64_t x[8];
  BG_UNROLL8
  for (int j = 0; j < 8; j++) x[j] = bg_bit_transpose8(x_in[j]);
  // byte matrix M[j][b] = byte b of x[j]; want T[b][j].
  BG_SWAP(x[0], x[4], 32, 0x00000000FFFFFFFFull);
  BG_SWAP(x[1], x[5], 32, 0x00000000FFFFFFFFull);
  BG_SWAP(x[2], x[6], 32, 0x00000000FFFFFFFFull);
  BG_SWAP(x[3], x[7], 32, 0x00000000FFFFFFFFull);
  BG_SWAP(x[0], x[2], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[1], x[3], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[4], x[6], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[5], x[7], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[0], x[1], 8, 0x00FF00FF00FF00FFull);
  BG_SWAP(x[2], x[3], 8, 0x00FF00FF00FF00FFull);
  BG_SWAP(x[4], x[5], 8, 0x00FF00FF00FF00FFull);
  BG_SWAP(x[6], x[7], 8, 0x00FF00FF00FF00FFull);
  BG_UNROLL8
  for (int b = 0; b < 8; b++) y[7 - b] = x[b];
}

// Inverse of bg_transpose_block: basis words y[0..7] of a 64-position block
// back to its bytes, x[j] = bytes 8j..8j+7 little-endian.  The byte
// transpose and the in-word bit transpose are each their own inverse, so the
// stages of bg_transpose_block run again in reverse order.
BG_HD void bg_untranspose_block(const uint64_t y[8], uint64_t x_out[8]) {
  uint64_t x[8];
  BG_UNROLL8
  for (int b = 0; b < 8; b++) x[b] = y[7 - b];
  BG_SWAP(x[0], x[1], 8, 0x00FF00FF00FF00FFull);
  BG_SWAP(x[2], x[3], 8, 0x00FF00FF00FF00FFull);
  BG_SWAP(x[4], x[5], 8, 0x00FF00FF00FF00FFull);
  BG_SWAP(x[6], x[7], 8, 0x00FF00FF00FF00FFull);
  BG_SWAP(x[0], x[2], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[1], x[3], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[4], x[6], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[5], x[7], 16, 0x0000FFFF0000FFFFull);
  BG_SWAP(x[0], x[4], 32, 0x00000000FFFFFFFFull);
  BG_SWAP(x[1], x[5], 32, 0x00000000FFFFFFFFull);
  BG_SWAP(x[2], x[6], 32, 0x00000000FFFFFFFFull);
  BG_SWAP(x[3], x[7], 32, 0x00000000FFFFFFFFull);
  BG_UNROLL8
  for (int j = 0; j < 8; j++) x_out[j] = bg_bit_transpose8(x[j]);
}

// Bigram filter hash: 15 bits of the byte pair (a, b) at positions (p, p+1):
// the low 7 bits of a and all of b.  Pairs that differ only in bit 7 of a
// share a slot (a superset filter, never a missed pair).  Table word (hash >>
// 5) = (a & 3) << 8 | b, bit (hash & 31) = bits 2..6 of a: the word's low
// bits select the LDS bank, and b's low bits spread evenly in text.
BG_HD uint32_t bg_bigram_hash(uint32_t a, uint32_t b) {
  return ((((a & 3u) << 8) | (b & 0xFFu)) << 5) | ((a >> 2) & 31u);
}

