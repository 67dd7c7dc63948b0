// Device primitives for generated bitstream group kernels (gfx950, wave64).
//
// Replaces the reference's CUDA template (bitgen/backend/cuda/template/
// template.cu:21-386).  Execution model, MI355X-first:
//
//  * one wavefront = one tile; lane l owns one 64-bit word (64 consecutive
//    positions, LSB-first) of every stream, so a tile spans 64 words = 4096
//    positions and waves never synchronise with each other;
//  * Advance / look-ahead shifts move words between lanes with DPP
//    (wave_shr:1 / wave_shl:1) or ds_bpermute, never through LDS or barriers
//    (the reference spends two __syncthreads per shift, template.cu:241-255);
//  * a tile recomputes HL halo words before and HR after its output range,
//    sized by the generator from the program's total shift distances, so
//    every output word is exact without inter-tile data exchange;
//  * MatchStar / ScanThru additions carry across lanes with wavefront
//    ballots and across tiles with a decoupled look-back over per-(tile, op)
//    status words;
//  * zero-block skipping is a wave-uniform branch on a ballot (the reference
//    spends three __syncthreads and an atomicOr per test, template.cu:356-374).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bitgen_hip.h"
#include "transpose_core.h"

#define BG_WAVE 64

// ---- cross-lane moves (32-bit halves) ----------------------------------------
__device__ __forceinline__ int bg_lane() { return (int)__lane_id(); }

// Value of lane (l - K); zero where l - K < 0.
template <int K>
__device__ __forceinline__ uint32_t bg_prev32(uint32_t x) {
  if constexpr (K == 0) {
    return x;
  } else if constexpr (K == 1) {
    // bound_ctrl: lane 0 (no source lane) reads 0, no initialised destination needed
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x138 /*wave_shr:1*/, 0xF, 0xF, true);
  } else if constexpr (K >= BG_WAVE) {
    return 0u;
  } else {
    const int l = bg_lane();
    const uint32_t v = (uint32_t)__builtin_amdgcn_ds_bpermute(((l - K) & (BG_WAVE - 1)) << 2, (int)x);
    return (l >= K) ? v : 0u;
  }
}

// Value of lane (l + K); zero where l + K > 63.
template <int K>
__device__ __forceinline__ uint32_t bg_next32(uint32_t x) {
  if constexpr (K == 0) {
    return x;
  } else if constexpr (K == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x130 /*wave_shl:1*/, 0xF, 0xF, true);
  } else if constexpr (K >= BG_WAVE) {
    return 0u;
  } else {
    const int l = bg_lane();
    const uint32_t v = (uint32_t)__builtin_amdgcn_ds_bpermute(((l + K) & (BG_WAVE - 1)) << 2, (int)x);
    return (l + K < BG_WAVE) ? v : 0u;
  }
}

__device__ __forceinline__ uint32_t bg_lo(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t bg_hi(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ uint64_t bg_cat(uint32_t hi, uint32_t lo) { return ((uint64_t)hi << 32) | lo; }

// ((a:b) >> s)[31:0] for s in 0..31 (v_alignbit_b32)
__device__ __forceinline__ uint32_t bg_align(uint32_t a, uint32_t b, uint32_t s) {
  return __builtin_amdgcn_alignbit(a, b, s);
}

// ---- shifts ------------------------------------------------------------------
// Advance by N (toward higher positions): y[p] = x[p - N].  Words come from
// lanes l - Q and l - Q - 1 (Q = N / 64); only the 32-bit halves that reach
// the result are moved, and the funnel is two v_alignbit_b32.
template <int N>
__device__ __forceinline__ uint64_t bg_advr(uint64_t x) {
  constexpr int Q = N / 64, R = N % 64;
  const uint32_t xl = bg_lo(x), xh = bg_hi(x);
  if constexpr (R == 0) {
    return bg_cat(bg_prev32<Q>(xh), bg_prev32<Q>(xl));
  } else if constexpr (R < 32) {
    const uint32_t cl = bg_prev32<Q>(xl), ch = bg_prev32<Q>(xh);
    const uint32_t ph = bg_prev32<Q + 1>(xh);
    return bg_cat(bg_align(ch, cl, 32 - R), bg_align(cl, ph, 32 - R));
  } else if constexpr (R == 32) {
    return bg_cat(bg_prev32<Q>(xl), bg_prev32<Q + 1>(xh));
  } else {
    constexpr int r = R - 32;
    const uint32_t cl = bg_prev32<Q>(xl);
    const uint32_t pl = bg_prev32<Q + 1>(xl), ph = bg_prev32<Q + 1>(xh);
    return bg_cat(bg_align(cl, ph, 32 - r), bg_align(ph, pl, 32 - r));
  }
}

// Look-ahead by N (toward lower positions): y[p] = x[p + N].
template <int N>
__device__ __forceinline__ uint64_t bg_advl(uint64_t x) {
  constexpr int Q = N / 64, R = N % 64;
  const uint32_t xl = bg_lo(x), xh = bg_hi(x);
  if constexpr (R == 0) {
    return bg_cat(bg_next32<Q>(xh), bg_next32<Q>(xl));
  } else if constexpr (R < 32) {
    const uint32_t cl = bg_next32<Q>(xl), ch = bg_next32<Q>(xh);
    const uint32_t nl = bg_next32<Q + 1>(xl);
    return bg_cat(bg_align(nl, ch, R), bg_align(ch, cl, R));
  } else if constexpr (R == 32) {
    return bg_cat(bg_next32<Q + 1>(xl), bg_next32<Q>(xh));
  } else {
    constexpr int r = R - 32;
    const uint32_t ch = bg_next32<Q>(xh);
    const uint32_t nl = bg_next32<Q + 1>(xl), nh = bg_next32<Q + 1>(xh);
    return bg_cat(bg_align(nh, nl, r), bg_align(nl, ch, r));
  }
}

// Keeps the compiler from sharing (and keeping live) values computed from x
// before this point with values computed after it: issued at regex
// boundaries so every regex recomputes its classes next to their use.
__device__ __forceinline__ void bg_fence(uint64_t &x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void bg_fence(uint32_t &x) { asm volatile("" : "+v"(x)); }

// Scalar counterpart, for wave-uniform values: issued at the top of every
// pass so that tests derived from x are not all hoisted in front of the
// pass loop (where each would hold its own scalar registers).
__device__ __forceinline__ void bg_sfence(uint64_t &x) { asm volatile("" : "+s"(x)); }
__device__ __forceinline__ void bg_sfence(uint32_t &x) { asm volatile("" : "+s"(x)); }

// ---- zero tests ------------------------------------------------------------
__device__ __forceinline__ bool bg_wave_zero(uint64_t x) {
  return __builtin_amdgcn_ballot_w64(x != 0ull) == 0ull;
}

// ---- bigram filter -------------------------------------------------------------
// A regex whose first zero test is "class S1 at p and class S2 at p+1" (a
// literal's first two bytes) is nonzero in a tile only if one of its byte
// pairs occurs in the tile.  Each wave builds, once per tile, a table of the
// byte pairs of its window (15-bit hash, bg_bigram_hash: 32,768 bits = 4 KiB
// of LDS per wave) and answers every such test with one table lookup, lane
// f looking up test f; the answers become wave-uniform bit masks and a test
// whose pairs are all absent jumps over the regex without computing it.  A
// present pair (or a hash collision) falls through to the exact zero test,
// so the table only has to hold every pair of the window (a superset is
// fine), never exactly them.
#define BG_BIGRAM_WORDS 1024

// Table of the byte pairs (p, p+1) of the wave's window (bas = the tile's
// basis words after the out-of-stream lanes are cleared).  Lane 63's last
// position pairs with 0 (no lane above): one extra entry, harmless.
__device__ __forceinline__ void bg_bigram_build(const uint64_t bas[8], uint32_t *tab) {
  const int lane = bg_lane();
  uint4 *t4 = reinterpret_cast<uint4 *>(tab);
#pragma unroll
  for (int i = 0; i < BG_BIGRAM_WORDS / 4 / BG_WAVE; i++) t4[lane + BG_WAVE * i] = make_uint4(0u, 0u, 0u, 0u);
  uint64_t x[8];
  bg_untranspose_block(bas, x);
  // the byte after this lane's last position: the next lane's first byte
  const uint32_t nxt = bg_next32<1>((uint32_t)x[0]);
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const uint32_t lo = (uint32_t)x[j], hi = (uint32_t)(x[j] >> 32);
    const uint32_t after = (j < 7) ? (uint32_t)x[j + 1] : nxt;
    // v_perm_b32 bytes: selector 0-3 = lo, 4-7 = hi operand.  Each 32-bit
    // result holds two 16-bit pair keys (second byte low, first byte high):
    // (b0,b1) (b1,b2); (b2,b3) (b3,b4); (b4,b5) (b5,b6); (b6,b7) (b7,b8).
    const uint32_t kk[4] = {__builtin_amdgcn_perm(hi, lo, 0x01020001u), __builtin_amdgcn_perm(hi, lo, 0x03040203u),
                            __builtin_amdgcn_perm(after, hi, 0x01020001u),
                            __builtin_amdgcn_perm(after, hi, 0x03040203u)};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      // word = hash >> 5 = bits 0..9 of the 16-bit key (b and the low two
      // bits of a), bit = hash & 31 = bits 10..14 (a's bits 2..6; the first
      // byte's top bit, bit 15, is not part of the hash)
      const uint32_t k = kk[q];
      __hip_atomic_fetch_or(tab + (k & 0x3FFu), 1u << ((k >> 10) & 31u), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_WAVEFRONT);
      __hip_atomic_fetch_or(tab + __builtin_amdgcn_ubfe(k, 16u, 10u), 1u << ((k >> 26) & 31u), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  }
}

// One clause of a filter against the tile's table: two 15-bit pair hashes
// (low and high half; a one-pair clause repeats its pair, a filter with
// fewer clauses repeats its first), one of which must be present.  No
// branch.  NP = 1: every clause of the program has one pair (both halves
// equal), so only the low half is read.
template <int NP>
__device__ __forceinline__ uint32_t bg_bigram_clause(const uint32_t *tab, uint32_t k) {
  const uint32_t a = k & 0xFFFFu;
  if (NP == 1) return (tab[a >> 5] >> (a & 31u)) & 1u;
  const uint32_t b = k >> 16;
  return ((tab[a >> 5] >> (a & 31u)) | (tab[b >> 5] >> (b & 31u))) & 1u;
}

// A single-pair clause as a lane of the filter pre-pass holds it for a whole
// launch: the LDS byte address of its table word (wbase: the byte offset of
// the wave's table in the block's tables) shifted up by 5, over the bit index
// in that word (k: the clause's key word; its low half is the pair hash).
__device__ __forceinline__ uint32_t bg_fkey_slot(uint32_t wbase, uint32_t k) {
  const uint32_t a = k & 0xFFFFu;
  return ((wbase + (a >> 5) * 4u) << 5) | (a & 31u);
}

// The clause's table bit (bit 0 of the result; higher bits are garbage),
// tabs: the block's tables.
template <class T>
__device__ __forceinline__ uint32_t bg_fkey_test(const T &tabs, uint32_t s) {
  const uint32_t w = *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(&tabs[0][0]) + (s >> 5));
  return w >> (s & 31u);
}

// Wave-uniform answers of NF filters of NCL clauses each: bit f % 32 of
// masks[f / 32] is set when every clause of filter f may hold in the tile.
// keys[f * NCL + c] holds clause c (bg_bigram_clause: two 15-bit pair
// hashes, one of which must be in the table; NP pairs per clause).  32-bit masks keep each test a
// single scalar AND (its result sets SCC for the branch).
template <int NF, int NCL, int NP>
__device__ __forceinline__ void bg_bigram_lookup(const uint32_t *tab, const uint32_t *keys, uint32_t *masks) {
  const int lane = bg_lane();
#pragma unroll
  for (int c = 0; c < (NF + 63) / 64; c++) {
    const int f = c * 64 + lane;
    const int fc = f < NF ? f : NF - 1;  // lanes past the last filter: any key, answer dropped
    uint32_t hit = f < NF;
#pragma unroll
    for (int j = 0; j < NCL; j++) hit &= bg_bigram_clause<NP>(tab, keys[fc * NCL + j]);
    const uint64_t m = __builtin_amdgcn_ballot_w64(hit != 0u);
    masks[2 * c] = (uint32_t)m;
    masks[2 * c + 1] = (uint32_t)(m >> 32);
  }
}

// A tile's byte-pair answers from the pattern set's filter pre-pass: n words
// at a wave-uniform address, written by an earlier launch, read through the
// constant address space so that they arrive as scalar loads.
template <int N>
__device__ __forceinline__ void bg_filter_words(const uint32_t *src, uint32_t *masks) {
  const __attribute__((address_space(4))) uint32_t *p = (const __attribute__((address_space(4))) uint32_t *)src;
#pragma unroll
  for (int i = 0; i < N; i++) masks[i] = p[i];
}

// ---- three-input bitwise functions ---------------------------------------------
// f(a, b, c) for the 8-bit truth table LUT (bit (a << 2 | b << 1 | c) holds
// f's value; a = 0xF0, b = 0xCC, c = 0xAA give the table), one v_bitop3_b32
// per 32-bit half.
template <unsigned LUT>
__device__ __forceinline__ uint64_t bg_op3(uint64_t a, uint64_t b, uint64_t c) {
  const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, (uint32_t)c, LUT);
  const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32), LUT);
  return ((uint64_t)hi << 32) | lo;
}

// ---- tile context ----------------------------------------------------------
// Geometry (words):  tile 0 covers words [0, 64) and outputs lanes
// [0, 64 - HR);  tile t >= 1 covers [S0 + (t-1)*S - HL, +64) and outputs lanes
// [HL, HL + S), where S0 = 64 - HR and S = 64 - HL - HR.
struct bg_tile {
  uint32_t t;         // tile index within this (group, buffer)
  int64_t w0;         // stream word held by lane 0
  int64_t W;          // words in the stream
  uint32_t *status;   // status words of this (group, buffer): [n_tiles][n_carry]
  uint32_t n_carry;   // carry operations of the program
  uint32_t epoch;     // launch epoch (status words of older launches are ignored)
  uint32_t *error;    // set to nonzero when a look-back wait gives up
  // Carry-in handling, wave-uniform (see bg_wave_add and bg_settle).
  uint32_t sweep;     // 1: first launch of a two-launch kernel (bg_sweep_end)
  uint32_t mode;      // BG_SPECULATE or BG_REPLAY
  uint32_t k_exact;   // replay: carry-ins of ops < k_exact are in `cin`
  uint32_t abort;     // speculate: an op could not publish; the pass stops
  uint32_t vfy;       // speculate: carry-ins of ops < vfy are known to be 0
  uint64_t cin;       // replay: carry-in bits of ops < k_exact
  uint64_t agg;       // speculate: ops that published an aggregate
  uint64_t out1;      // speculate: their carry-out for carry-in 1
};

#define BG_SPECULATE 0u
#define BG_REPLAY 1u

__device__ __forceinline__ bg_tile bg_tile_make(uint32_t t, int64_t w0, int64_t W, uint32_t *status,
                                                uint32_t n_carry, uint32_t epoch, uint32_t *error) {
  bg_tile tc;
  tc.t = t;
  tc.w0 = w0;
  tc.W = W;
  tc.status = status;
  tc.n_carry = n_carry;
  tc.epoch = epoch;
  tc.error = error;
  tc.sweep = 0;
  tc.mode = BG_SPECULATE;
  tc.k_exact = 0;
  tc.abort = 0;
  tc.vfy = 0;
  tc.cin = 0;
  tc.agg = 0;
  tc.out1 = 0;
  return tc;
}

// Upper bound on look-back polls (each sleeps ~64 cycles): far above any
// legitimate wait, low enough that a fault never hangs the device.
#ifndef BG_SPIN_LIMIT
#define BG_SPIN_LIMIT (1u << 21)
#endif

// Status word: bits 0..1 value, bits 2..3 kind, bits 4..31 epoch.
#define BG_ST_AGG 1u   // aggregate of a propagating segment (value unused)
#define BG_ST_INC 2u   // inclusive carry into the successor (value = carry)
#define BG_ST_DEF 3u   // deferred: the tile's sweep stopped before this op; its
                       // settle launch publishes the status
#define BG_LB_UNKNOWN 2u  // bg_lookback in a sweep: a deferred status decides

__device__ __forceinline__ uint32_t bg_st_pack(uint32_t epoch, uint32_t kind, uint32_t v) {
  return (epoch << 4) | (kind << 2) | v;
}

__device__ __forceinline__ void bg_st_store(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t bg_st_load(const uint32_t *p) {
  return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// First lane >= x whose bit in `notp` is set (64 if none); x in [0, 64].
__device__ __forceinline__ int bg_imax(int a, int b) { return a > b ? a : b; }

__device__ __forceinline__ int bg_first_from(uint64_t notp, int x) {
  if (x >= 64) return 64;
  const uint64_t m = notp & (~0ull << x);
  return m ? (int)__builtin_ctzll(m) : 64;
}

// Look-back for the carry into op K of tile t: the nearest predecessor whose
// status is not an aggregate decides it (aggregates pass the carry on
// unchanged; before tile 0 the carry is 0).  The predecessor is read (and
// waited for) with one scalar load, as almost always it decides.  Behind an
// aggregate the next 64 statuses are read at once (lane i reads tile j-1-i),
// so a run of propagating tiles -- a long line under a line-end projection,
// a star over a long run -- costs one load per 64 tiles; the aggregates
// passed in the last step before an inclusive status get that inclusive
// value written back (their own carry-out, the value they publish
// themselves), which shortens the walks of later tiles.  st is tile t - 1's
// status word if already loaded (else 0).  Waits are bounded so that a
// missing predecessor shows up as an error flag instead of a hung wave.
// A deferred status is waited for like a missing one, except in a sweep
// launch, where only the next launch publishes it: the result is then
// BG_LB_UNKNOWN.
__device__ __forceinline__ uint32_t bg_lookback(const bg_tile &tc, uint32_t K, uint32_t st) {
  const uint32_t ep = tc.epoch & 0x0FFFFFFFu;
  int64_t j = (int64_t)tc.t - 1;
  uint32_t spins = 0;
  bool have = st != 0;
  while (true) {
    // every lane loads the same word; readfirstlane keeps the value (and
    // everything decided from it) in scalar registers
    if (!have) st = (uint32_t)__builtin_amdgcn_readfirstlane((int)bg_st_load(tc.status + (size_t)j * tc.n_carry + K));
    have = false;
    const uint32_t kind = (st >> 2) & 3u;
    if ((st >> 4) == ep && kind == BG_ST_DEF && tc.sweep) return BG_LB_UNKNOWN;
    if ((st >> 4) != ep || kind == 0 || kind == BG_ST_DEF) {
      if (++spins > BG_SPIN_LIMIT) {
        if (bg_lane() == 0) __hip_atomic_fetch_or(tc.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    if (kind == BG_ST_INC) return st & 1u;
    // tile j passes its carry-in on: read the 64 tiles before it
    const int lane = bg_lane();
    const int64_t jj = j - 1 - lane;
    uint32_t w = bg_st_pack(ep, BG_ST_INC, 0u);
    if (jj >= 0) w = bg_st_load(tc.status + (size_t)jj * tc.n_carry + K);
    const bool agg = (w >> 4) == ep && ((w >> 2) & 3u) == BG_ST_AGG;
    const uint64_t stop = ~__builtin_amdgcn_ballot_w64(agg);
    if (stop == 0) {
      j -= 64;  // an aggregate too: read on from it
      have = true;
      continue;
    }
    const int f = (int)__builtin_ctzll(stop);
    st = (uint32_t)__builtin_amdgcn_readlane((int)w, f);
    j -= 1 + f;
    have = true;
    if ((st >> 4) == ep && ((st >> 2) & 3u) == BG_ST_INC && lane < f)
      bg_st_store(tc.status + (size_t)jj * tc.n_carry + K, bg_st_pack(ep, BG_ST_INC, st & 1u));
  }
}

// Speculative pass, op K: checks that the carry-ins assumed 0 for ops
// [tc.vfy, K) are 0, from the predecessor's statuses (one load, lane k reads
// op k; the predecessor published these ops earlier in its own pass, so they
// are usually there; the waiting look-back only where one is not inclusive
// yet).  True: every op < K computed exact values, and tc.vfy = K.  False: an
// assumption was wrong (the pass must replay).
__device__ __forceinline__ bool bg_verify(bg_tile &tc, uint32_t K) {
  const uint32_t ep = tc.epoch & 0x0FFFFFFFu;
  const int lane = bg_lane();
  const uint32_t *prev = tc.status + (size_t)(tc.t - 1) * tc.n_carry;
  const bool mine = (uint32_t)lane >= tc.vfy && (uint32_t)lane < K;
  uint32_t w = 0;
  if (mine) w = bg_st_load(prev + lane);
  const bool inc = mine && (w >> 4) == ep && ((w >> 2) & 3u) == BG_ST_INC;
  if (__builtin_amdgcn_ballot_w64(inc && (w & 1u))) return false;
  const uint64_t span = (K >= 64 ? ~0ull : ((1ull << K) - 1ull)) & (tc.vfy >= 64 ? 0ull : (~0ull << tc.vfy));
  uint64_t rest = span & ~__builtin_amdgcn_ballot_w64(inc);
  while (rest) {
    const int k = __builtin_ctzll(rest);
    rest &= rest - 1;
    const uint32_t wk = (uint32_t)__builtin_amdgcn_readlane((int)w, k);
    if (bg_lookback(tc, (uint32_t)k, wk)) return false;
  }
  tc.vfy = K;
  return true;
}

// Wave-wide addition a + b of the tile's 64-word number with the carry into
// lane L taken from the predecessor tile (decoupled look-back); returns this
// lane's sum word.  K = index of the carry operation in the program.
//
// Per lane: s = a + b, g = carry out without carry-in, p = s is all ones.  With
// G, P the wave ballots of g, p (lanes below L cleared) the carry into lane l
// is bit l of ((G | P) + G + (cin << L)) ^ P: a ripple adder whose per-lane
// generate is G and propagate is P.  The carry handed to the next tile is the
// carry into lane L + S.
//
// Three modes (wave-uniform):
//  * tile 0: carry-in 0, exact; publishes at once.
//  * BG_SPECULATE (first pass of tile t >= 1): the carry-in is assumed 0 and
//    nothing waits.  `bnd` bounds the lanes of this op's operands that
//    earlier assumptions may have changed (-1: none; the generator takes the
//    maximum over the carry ops the operands depend on, plus the lanes right
//    advances can move values up).  The carry into lane L + S cannot depend
//    on any assumption when a non-propagating lane lies between bnd and
//    L + S (or bnd < L); the status is then published at once.  Otherwise
//    the assumed carry-ins of the earlier ops are checked (bg_verify): if
//    they hold, the operands are exact and the op publishes; if not, the
//    pass aborts at this op.  `infl` returns the bound for this op's
//    result.  bg_settle then resolves every carry-in with one batch of
//    look-back loads and replays the tile if an assumption was wrong or the
//    pass aborted.
//  * BG_REPLAY: ops < k_exact take their resolved carry-in and re-publish the
//    same status; ops >= k_exact wait for their predecessor (publish, look
//    back, publish the inclusive value).
template <int L, int S, int K>
__device__ __forceinline__ uint64_t bg_wave_add(uint64_t a, uint64_t b, bg_tile &tc, int bnd, int &infl) {
  const int lane = bg_lane();
  const int l0 = (tc.t == 0) ? 0 : L;  // tile 0 is exact from lane 0
  const uint64_t s = a + b;
  const bool g = s < a;
  const bool p = s == ~0ull;
  const uint64_t keep = (l0 == 0) ? ~0ull : (~0ull << l0);
  const uint64_t G = __builtin_amdgcn_ballot_w64(g) & keep;
  const uint64_t P = __builtin_amdgcn_ballot_w64(p) & keep;
  const uint64_t GP = G | P;
  const int pass = L + S;  // lane receiving the successor's carry point (<= 64)

  // carry into lane `pass` for a given carry-in
  auto carry_out = [&](uint64_t cin) -> uint32_t {
    const uint64_t c = cin << l0;
    const uint64_t t1 = GP + G;
    const uint64_t o1 = t1 < GP;
    const uint64_t t2 = t1 + c;
    const uint64_t o2 = t2 < t1;
    if (pass >= 64) return (uint32_t)(o1 | o2);
    return (uint32_t)(((t2 ^ P) >> pass) & 1ull);
  };

  uint32_t *my = tc.status + (size_t)tc.t * tc.n_carry + K;
  uint32_t cin = 0;
  constexpr bool kSpeculate = K < 64;  // speculation state is kept in 64-bit masks
  if (tc.t == 0) {
    const uint32_t out = carry_out(0);
    if (lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_INC, out));
  } else if (!kSpeculate || tc.mode == BG_REPLAY) {
    if (kSpeculate && (uint32_t)K < tc.k_exact) {
      cin = (uint32_t)(tc.cin >> (K & 63)) & 1u;
      if (lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_INC, carry_out(cin)));
    } else {
      const uint32_t out0 = carry_out(0), out1 = carry_out(1);
      if (lane == 0) bg_st_store(my, out0 == out1 ? bg_st_pack(tc.epoch, BG_ST_INC, out0)
                                                  : bg_st_pack(tc.epoch, BG_ST_AGG, 0));
      cin = bg_lookback(tc, K, 0);
      if (out0 != out1 && lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_INC, cin));
    }
  } else {
    // speculate: operand lanes <= b may differ from their exact values
    int b = bnd > 63 ? 63 : bnd;
    const uint64_t notp = ~P & keep;
    if (b >= l0 && bg_first_from(notp, b + 1) >= pass) {
      // the carry-out may depend on the carry-ins assumed for earlier ops:
      // check them now; if they hold, the operands are exact
      if (!bg_verify(tc, (uint32_t)K)) {
        tc.abort = 1;
        tc.k_exact = K;
        return 0;
      }
      b = -1;
    }
    const uint32_t out0 = carry_out(0), out1 = carry_out(1);
    if (out0 == out1) {
      if (lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_INC, out0));
    } else {
      if (lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_AGG, 0));
      tc.agg |= 1ull << (K & 63);
      tc.out1 |= (uint64_t)out1 << (K & 63);
    }
    // lanes of this op's result that differ if an assumption was wrong: a
    // carry-in of 1 changes lanes L through the first non-propagating lane,
    // and changed operand lanes <= b carry on through the lanes that
    // propagate
    int hi = bg_first_from(notp, l0);
    if (b >= l0) {
      const int zb = bg_first_from(notp, b + 1);
      hi = hi > zb ? hi : zb;
    }
    hi = hi > 63 ? 63 : hi;
    infl = hi > b ? hi : b;
  }
  // cm is wave-uniform: most tiles generate no carry between lanes at all,
  // and then the per-lane carry add is skipped
  const uint64_t cm = (GP + G + ((uint64_t)cin << l0)) ^ P;
  if (cm == 0) return s;
  const uint64_t carry = (cm >> lane) & 1ull;
  return s + carry;
}

// End of a pass.  Returns true when the pass's values are exact.  After a
// speculative pass: resolves the carry-ins of the ops the pass reached (one
// batch of loads, then the waiting look-back only where a predecessor had not
// published yet), publishes the inclusive status of the ops that published an
// aggregate, and switches to a replay if an assumed carry-in was 1 or the
// pass aborted.  NC = carry operations of the program (<= 64).
template <int NC>
__device__ __forceinline__ bool bg_settle(bg_tile &tc) {
  static_assert(NC >= 1 && NC <= 64, "programs with more than 64 carry operations start in replay mode");
  if (tc.t == 0 || tc.mode == BG_REPLAY) return true;
  const uint32_t n = tc.abort ? tc.k_exact : (uint32_t)NC;
  // one load instruction: lane k reads the predecessor's status of op k
  const int lane = bg_lane();
  const uint32_t *prev = tc.status + (size_t)(tc.t - 1) * tc.n_carry;
  uint32_t w = 0;
  if ((uint32_t)lane < n) w = bg_st_load(prev + lane);
  const bool inc = (uint32_t)lane < n && (w >> 4) == (tc.epoch & 0x0FFFFFFFu) && ((w >> 2) & 3u) == BG_ST_INC;
  const uint64_t inc_ops = __builtin_amdgcn_ballot_w64(inc);
  uint64_t cin = __builtin_amdgcn_ballot_w64(inc && (w & 1u));
  // predecessors that had not published yet, or published an aggregate:
  // the waiting look-back, op by op (rare)
  uint64_t rest = ((n >= 64) ? ~0ull : ((1ull << n) - 1ull)) & ~inc_ops;
  while (rest) {
    const int k = __builtin_ctzll(rest);
    rest &= rest - 1;
    const uint32_t wk = (uint32_t)__builtin_amdgcn_readlane((int)w, k);
    cin |= (uint64_t)bg_lookback(tc, (uint32_t)k, wk) << k;
  }
  if (tc.agg) {
    uint64_t a = tc.agg;
    while (a) {
      const int k = __builtin_ctzll(a);
      a &= a - 1;
      const uint32_t c = (uint32_t)(cin >> k) & 1u;
      const uint32_t o1 = (uint32_t)(tc.out1 >> k) & 1u;
      if (bg_lane() == 0)
        bg_st_store(tc.status + (size_t)tc.t * tc.n_carry + k, bg_st_pack(tc.epoch, BG_ST_INC, c ? o1 : 1u - o1));
    }
  }
  if (!tc.abort && cin == 0) return true;
  tc.mode = BG_REPLAY;
  tc.k_exact = n;
  tc.cin = cin;
  tc.abort = 0;
  return false;
}

// ---- two-launch form ---------------------------------------------------------
// The same kernel runs twice over the same tiles and epoch.  The sweep
// (launch 1) runs every tile's speculative pass once and never waits at its
// end: it stores the tile's values, which are exact when every assumed
// carry-in turns out 0, and leaves the carry-ins to the settle (launch 2).
// By then every tile has published every op's status, or marked the ops
// from its abort point on as deferred, so the settle reads each tile's
// carry-ins with no wait except behind deferred tiles, and replays only
// tiles whose assumption was wrong or whose sweep aborted.  A tile runs one
// pass per launch: nothing stays live across passes.

// End of a sweep pass.  True: the pass reached its end and its values are
// stored (exact if every carry-in is 0; otherwise the settle replays the
// tile and stores again).  False: it aborted at op k_exact; ops k_exact..NC-1
// are marked deferred (one store per lane) and nothing is stored.
template <int NC>
__device__ __forceinline__ bool bg_sweep_end(const bg_tile &tc) {
  static_assert(NC >= 1 && NC <= 64, "the two-launch form needs 1..64 carry operations");
  if (!tc.abort) return true;
  const uint32_t lane = (uint32_t)bg_lane();
  if (lane >= tc.k_exact && lane < (uint32_t)NC)
    bg_st_store(tc.status + (size_t)tc.t * tc.n_carry + lane, bg_st_pack(tc.epoch, BG_ST_DEF, 0u));
  return false;
}

// Start of a tile in the settle launch.  Reads the tile's own statuses and
// its predecessor's (one load each, lane k = op k): ops before the first
// deferred op k_exact take their carry-ins from the predecessor (the
// waiting look-back only where the predecessor's status is an aggregate or
// deferred), and ops that published an aggregate publish their inclusive
// value (an aggregate's carry-out is its carry-in).  True: the sweep's values
// are exact, the tile is done.  False: the tile replays (mode BG_REPLAY with
// the resolved carry-ins of ops < k_exact; later ops wait for theirs).
template <int NC>
__device__ __forceinline__ bool bg_settle_begin(bg_tile &tc) {
  static_assert(NC >= 1 && NC <= 64, "the two-launch form needs 1..64 carry operations");
  if (tc.t == 0) return true;  // tile 0 is exact from its sweep
  const uint32_t ep = tc.epoch & 0x0FFFFFFFu;
  const int lane = bg_lane();
  const bool op = lane < NC;
  uint32_t *own = tc.status + (size_t)tc.t * tc.n_carry;
  const uint32_t *prev = own - tc.n_carry;
  uint32_t w = 0, o = 0;
  if (op) {
    o = bg_st_load(own + lane);
    w = bg_st_load(prev + lane);
  }
  const uint32_t okind = (o >> 2) & 3u;
  const bool opub = op && (o >> 4) == ep && (okind == BG_ST_AGG || okind == BG_ST_INC);
  const uint64_t unpub = ~__builtin_amdgcn_ballot_w64(opub);
  const uint32_t n = unpub ? (uint32_t)__builtin_ctzll(unpub) : 64u;  // <= NC
  const bool in = (uint32_t)lane < n;
  const bool inc = in && (w >> 4) == ep && ((w >> 2) & 3u) == BG_ST_INC;
  uint64_t cin = __builtin_amdgcn_ballot_w64(inc && (w & 1u));
  uint64_t rest = ((n >= 64) ? ~0ull : ((1ull << n) - 1ull)) & ~__builtin_amdgcn_ballot_w64(inc);
  while (rest) {
    const int k = __builtin_ctzll(rest);
    rest &= rest - 1;
    const uint32_t wk = (uint32_t)__builtin_amdgcn_readlane((int)w, k);
    cin |= (uint64_t)(bg_lookback(tc, (uint32_t)k, wk) & 1u) << k;
  }
  if (in && okind == BG_ST_AGG) bg_st_store(own + lane, bg_st_pack(tc.epoch, BG_ST_INC, (uint32_t)(cin >> lane) & 1u));
  if (n >= (uint32_t)NC && cin == 0) return true;
  tc.mode = BG_REPLAY;
  tc.k_exact = n;
  tc.cin = cin;
  return false;
}

// Settle launch, one chunk of 64 consecutive tiles at once (lane i = tile
// g + i; a chunk may span buffers in the batched form).  row: this lane's
// status row (nullptr: no tile); first: the tile is its buffer's tile 0
// (carry-in 0); prev: the status row of lane 0's predecessor (unused when
// lane 0's tile is first); tc0: lane 0's tile (for the look-back behind an
// aggregate before the chunk).
//
// The carry into a tile is the carry-out of the tile below it, and an
// aggregate's carry-out is its own carry-in; so the carry into lane i is the
// value of the nearest lane j <= i whose predecessor status is inclusive
// (or its buffer's start: 0), or the chunk's incoming carry when every
// predecessor down to lane 0 is an aggregate.  Per op: one load per lane
// (all NC issued together), three ballots and a bit scan per lane; the
// aggregates of the chunk publish their inclusive value.  Returns the tiles
// that still need the per-tile settle (bg_settle_begin): a carry-in of 1,
// an aborted sweep, or a carry-in decided by an aborted predecessor.
template <int NC>
__device__ __forceinline__ uint64_t bg_settle_chunk(uint32_t *row, bool first, const uint32_t *prev, bg_tile &tc0) {
  static_assert(NC >= 1 && NC <= 64, "the two-launch form needs 1..64 carry operations");
  const uint32_t ep = tc0.epoch & 0x0FFFFFFFu;
  const uint32_t inc0 = bg_st_pack(tc0.epoch, BG_ST_INC, 0u);
  const int lane = bg_lane();
  const bool valid = row != nullptr;
  const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const bool first0 = __builtin_amdgcn_readfirstlane((int)first) != 0;
  uint32_t pw = inc0;  // lane k: lane 0's predecessor status of op k
  if (!first0 && lane < NC) pw = bg_st_load(prev + lane);
  uint32_t o[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) o[k] = valid ? bg_st_load(row + k) : inc0;
  bool need = false;
#pragma unroll
  for (int k = 0; k < NC; k++) {
    // predecessor's status of op k: lane i - 1's own word (lane 0: pw)
    uint32_t p = (uint32_t)__builtin_amdgcn_mov_dpp((int)o[k], 0x138 /*wave_shr:1*/, 0xF, 0xF, true);
    const uint32_t p0 = (uint32_t)__builtin_amdgcn_readlane((int)pw, k);
    if (lane == 0) p = p0;
    if (first) p = inc0;
    const bool pok = (p >> 4) == ep;
    const uint32_t pk = (p >> 2) & 3u;
    const uint64_t NA = ~__builtin_amdgcn_ballot_w64(pok && pk == BG_ST_AGG);
    const uint64_t ONE = __builtin_amdgcn_ballot_w64(pok && pk == BG_ST_INC && (p & 1u));
    const uint64_t UNK = __builtin_amdgcn_ballot_w64(!(pok && (pk == BG_ST_AGG || pk == BG_ST_INC)));
    // the chunk's incoming carry, needed only behind an aggregate in lane 0
    uint32_t cin_chunk = 0;
    if (!(NA & 1ull)) cin_chunk = bg_lookback(tc0, (uint32_t)k, p0) & 1u;
    const uint64_t m = NA & upto;
    bool cin, unk;
    if (m) {
      const int j = 63 - __builtin_clzll(m);
      cin = (ONE >> j) & 1ull;
      unk = (UNK >> j) & 1ull;
    } else {
      cin = cin_chunk;
      unk = false;
    }
    const bool ook = (o[k] >> 4) == ep;
    const uint32_t okind = (o[k] >> 2) & 3u;
    const bool odef = !(ook && (okind == BG_ST_AGG || okind == BG_ST_INC));
    if (valid && ook && okind == BG_ST_AGG && !unk) bg_st_store(row + k, bg_st_pack(tc0.epoch, BG_ST_INC, cin ? 1u : 0u));
    need = need || (valid && (odef || unk || cin));
  }
  return __builtin_amdgcn_ballot_w64(need);
}

// MatchStar(M, C) = (((M & C) + C) ^ C) | M
template <int L, int S, int K>
__device__ __forceinline__ uint64_t bg_matchstar(uint64_t m, uint64_t c, bg_tile &tc, int bnd, int &infl) {
  const uint64_t sum = bg_wave_add<L, S, K>(m & c, c, tc, bnd, infl);
  return (sum ^ c) | m;
}

// ScanThru(M, C) = (M + C) & ~C
template <int L, int S, int K>
__device__ __forceinline__ uint64_t bg_scanthru(uint64_t m, uint64_t c, bg_tile &tc, int bnd, int &infl) {
  const uint64_t sum = bg_wave_add<L, S, K>(m, c, tc, bnd, infl);
  return sum & ~c;
}

// ---- tile queues -------------------------------------------------------------
// A launch's tiles are dealt round-robin over nq queues (tile t belongs to
// queue t % nq); block b serves queue b % nq, one atomic add per tile on the
// queue's own counter line.  Waves of a queue take its tiles in increasing
// order, and the look-back only waits for smaller tiles, so every wait ends
// once the first nq blocks are resident.  The control block layout
// (BG_CTL_*) is in bitgen_hip.h.
struct bg_queue_lds;

struct bg_queue {
  bg_queue_lds *lds;  // per-block state (LDS)
  uint32_t *ctl;
  uint32_t q;       // this block's queue
  uint32_t nq;      // queues of the launch
  uint32_t take;    // tiles per take (consecutive tiles of the queue)
  uint32_t n;       // tiles of the launch
  uint32_t epoch;   // status tag of the launch
};

__device__ __forceinline__ uint32_t bg_ctl_load(uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-block state of the queue protocol (LDS).
struct bg_queue_lds {
  uint32_t epoch;
  uint32_t exits;
};

// Device-scope accesses to one shared word serialise, so each block makes
// one of each: a single epoch read at the start (shared through LDS) and a
// single done-increment when its last wave leaves.
__device__ __forceinline__ bg_queue bg_queue_open(uint32_t *ctl, uint32_t nq, uint32_t n_tiles, uint32_t take) {
  __shared__ bg_queue_lds lds;
  bg_queue Q;
  Q.lds = &lds;
  Q.ctl = ctl;
  // the launch functions clamp nq to 1 .. gridDim.x; clamp again so that every
  // queue is served by at least one block whatever the caller passed
  Q.nq = (nq == 0) ? 1u : (nq < gridDim.x ? nq : gridDim.x);
  Q.q = blockIdx.x % Q.nq;
  Q.take = (take == 0) ? 1u : take;
  Q.n = n_tiles;
  if (threadIdx.x == 0) {
    lds.epoch = bg_ctl_load(ctl + BG_CTL_EPOCH) & 0x0FFFFFFFu;
    lds.exits = 0;
  }
  __syncthreads();
  Q.epoch = (uint32_t)__builtin_amdgcn_readfirstlane((int)lds.epoch);
  return Q;
}

// One take gives a wave `take` consecutive tiles of its queue: tiles
// g, g + nq, ..., g + (take-1)*nq with g = k*take*nq + q for the counter
// value k.  Taking is split in two so that a wave can ask for its next tiles
// before running the current ones (the atomic's latency then overlaps the
// tiles' own loads): bg_queue_issue sends the atomic add (the old value lands
// in lane 0), bg_queue_finish turns it into the first tile g and returns
// false when the queue is exhausted.  Each wave makes exactly one failing
// take.  A wave runs the tiles it holds in increasing order, so the smallest
// unfinished tile is always being run and every look-back wait ends.
__device__ __forceinline__ uint32_t bg_queue_issue(const bg_queue &Q) {
  uint32_t k = 0;
  if (bg_lane() == 0)
    k = __hip_atomic_fetch_add(Q.ctl + BG_CTL_COUNTER(Q.q), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return k;
}

__device__ __forceinline__ bool bg_queue_finish(const bg_queue &Q, uint32_t k, uint32_t &t) {
  k = (uint32_t)__builtin_amdgcn_readfirstlane((int)k);
  const uint64_t tt = (uint64_t)k * Q.take * Q.nq + Q.q;
  t = (uint32_t)tt;
  return tt < Q.n;
}

__device__ __forceinline__ bool bg_queue_take(const bg_queue &Q, uint32_t &t) {
  return bg_queue_finish(Q, bg_queue_issue(Q), t);
}

// Called once by every wave after its last take.  The last wave of a block
// reports the block; the last block of a queue resets the queue's counters;
// the last queue to finish advances the epoch.  Every block read the epoch
// before it reported, so the advance is seen only by the next launch.
//
// Every access is a relaxed device-scope atomic, with no fence: an
// acquire/release at device scope makes the compiler write back and
// invalidate the whole L2 of the block's XCD (buffer_wbl2 / buffer_inv),
// which, done by every block at its end, cost more than all the tiles'
// work.  The order the protocol needs comes from the waits instead: a wave
// consumes the result of its failing take before it reports, the reporting
// atomic returns before the last block resets, and device-scope atomics are
// performed at the device's coherence point in that order.  The control
// words carry no other data, and the next launch on the stream starts after
// this one's end-of-kernel release.  advance = false keeps the epoch for a
// following launch over the same status words (the settle after a sweep).
__device__ __forceinline__ void bg_queue_close(const bg_queue &Q, bool advance = true) {
  if (bg_lane() != 0) return;
  const uint32_t wpb = blockDim.x / BG_WAVE;
  const uint32_t e = __hip_atomic_fetch_add(&Q.lds->exits, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (e + 1 != wpb) return;
  const uint32_t blocks_q = (gridDim.x - Q.q + Q.nq - 1) / Q.nq;
  const uint32_t d = __hip_atomic_fetch_add(Q.ctl + BG_CTL_DONE(Q.q), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (d + 1 != blocks_q) return;
  __hip_atomic_store(Q.ctl + BG_CTL_COUNTER(Q.q), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(Q.ctl + BG_CTL_DONE(Q.q), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t qd = __hip_atomic_fetch_add(Q.ctl + BG_CTL_QDONE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (qd + 1 != Q.nq) return;
  __hip_atomic_store(Q.ctl + BG_CTL_QDONE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (advance)
    __hip_atomic_store(Q.ctl + BG_CTL_EPOCH, (Q.epoch + 1u) & 0x0FFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
