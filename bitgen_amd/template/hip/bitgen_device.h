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

// ---- zero tests ------------------------------------------------------------
__device__ __forceinline__ bool bg_wave_zero(uint64_t x) {
  return __builtin_amdgcn_ballot_w64(x != 0ull) == 0ull;
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
};

// Upper bound on look-back polls (each sleeps ~64 cycles): far above any
// legitimate wait, low enough that a fault never hangs the device.
#ifndef BG_SPIN_LIMIT
#define BG_SPIN_LIMIT (1u << 24)
#endif

// Status word: bits 0..1 value, bits 2..3 kind, bits 4..31 epoch.
#define BG_ST_AGG 1u   // aggregate of a propagating segment (value unused)
#define BG_ST_INC 2u   // inclusive carry into the successor (value = carry)

__device__ __forceinline__ uint32_t bg_st_pack(uint32_t epoch, uint32_t kind, uint32_t v) {
  return (epoch << 4) | (kind << 2) | v;
}

__device__ __forceinline__ void bg_st_store(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t bg_st_load(const uint32_t *p) {
  return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
template <int L, int S, int K>
__device__ __forceinline__ uint64_t bg_wave_add(uint64_t a, uint64_t b, const bg_tile &tc) {
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
  if (tc.t == 0) {
    const uint32_t out = carry_out(0);
    if (lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_INC, out));
  } else {
    const uint32_t out0 = carry_out(0), out1 = carry_out(1);
    if (out0 == out1) {
      if (lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_INC, out0));
    } else if (lane == 0) {
      bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_AGG, 0));
    }
    // look back for the carry into lane L; the wait is bounded so that a
    // missing predecessor shows up as an error flag instead of a hung wave
    int64_t j = (int64_t)tc.t - 1;
    uint32_t spins = 0;
    while (true) {
      const uint32_t st = bg_st_load(tc.status + (size_t)j * tc.n_carry + K);
      const uint32_t kind = (st >> 2) & 3u;
      if ((st >> 4) != (tc.epoch & 0x0FFFFFFFu) || kind == 0) {
        if (++spins > BG_SPIN_LIMIT) {
          if (lane == 0) __hip_atomic_fetch_or(tc.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          cin = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      if (kind == BG_ST_INC) {
        cin = st & 1u;
        break;
      }
      // aggregate: the whole predecessor segment propagates; look further back
      j -= 1;
      if (j < 0) {
        cin = 0;
        break;
      }
    }
    if (out0 != out1 && lane == 0) bg_st_store(my, bg_st_pack(tc.epoch, BG_ST_INC, cin));
  }
  const uint64_t cm = (GP + G + ((uint64_t)cin << l0)) ^ P;
  const uint64_t carry = (cm >> lane) & 1ull;
  return s + carry;
}

// MatchStar(M, C) = (((M & C) + C) ^ C) | M
template <int L, int S, int K>
__device__ __forceinline__ uint64_t bg_matchstar(uint64_t m, uint64_t c, const bg_tile &tc) {
  const uint64_t sum = bg_wave_add<L, S, K>(m & c, c, tc);
  return (sum ^ c) | m;
}

// ScanThru(M, C) = (M + C) & ~C
template <int L, int S, int K>
__device__ __forceinline__ uint64_t bg_scanthru(uint64_t m, uint64_t c, const bg_tile &tc) {
  const uint64_t sum = bg_wave_add<L, S, K>(m, c, tc);
  return sum & ~c;
}

// ---- tile acquisition --------------------------------------------------------
// Every wave takes the next tile index from a monotonically increasing counter
// (counter[0]); args.tile_base is the counter value at launch, so waves run
// tiles in acquisition order and a tile's predecessors always started earlier
// (required by the look-back spin).
__device__ __forceinline__ uint32_t bg_acquire_tile(uint32_t *counter, uint32_t base) {
  uint32_t t = 0;
  if (bg_lane() == 0) t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
  return t - base;
}
