// libbitgen_hip.so -- C ABI of the MI355X "hip" backend (see include/bitgen_hip.h).
//
// Kernels owned by the runtime (the group kernels are generated per pattern
// set and loaded as code objects):
//   bg_transpose_kernel  bytes -> 8 basis streams (LSB-first u64 words)
//   bg_msb32_kernel      u64 LSB-first -> the reference's u32 MSB-first layout
//   bg_popcount_kernel   match counting
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "bitgen_hip.h"
#include "bitgen_hip_static.h"
#include "bitgen_hip_serve.h"
#include "transpose_core.h"

namespace {

thread_local char g_err[512] = "";
int g_cu_count = 256;  // compute units of the current device (set by bg_init)

int set_err(int code, const char *what) {
  snprintf(g_err, sizeof(g_err), "%s", what);
  return code;
}

int hip_err(hipError_t e, const char *what) {
  snprintf(g_err, sizeof(g_err), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
  return (int)e;
}

#define BG_HIP(call, what)                \
  do {                                    \
    hipError_t _e = (call);               \
    if (_e != hipSuccess) return hip_err(_e, what); \
  } while (0)

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

struct bg_module_s {
  hipModule_t mod;
};

struct bg_kernel_s {
  hipFunction_t fn;
  uint32_t occ_threads;  // block size of the cached occupancy (0 = none)
  int32_t occ_blocks;    // blocks per CU at that block size
  uint32_t threads;      // block size of a plain launch: the kernel's launch bounds (0 = not read yet)
  uint32_t launches;     // 2: a sweep and a settle launch per run (phase argument 0, then 1)
  uint32_t needs_filter; // 1: reads its byte-pair answers from a filter pre-pass (args filter)
};

struct bg_graph_s {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

// Scratch of bg_group_launch_static: a control block (BG_CTL_WORDS words,
// zero when the code object loads; every launch leaves its queue counters at
// zero and advances the epoch) followed by BG_STATIC_STATUS_WORDS look-back
// status words (epoch-tagged, never cleared).  Each device holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for that device (g_scratch).
static __device__ __align__(256) uint32_t scratch[BG_CTL_WORDS + BG_STATIC_STATUS_WORDS];

namespace {
#define BG_MAX_DEVICES 64
uint32_t *g_scratch[BG_MAX_DEVICES] = {};
}  // namespace

// ---------------------------------------------------------------------------
// Transpose: thread = one 64-byte block (one word of each basis stream) of one
// input.  A wave whose 64 blocks (4 KiB) all hold data reads them with four
// coalesced 16-byte loads per lane (1 KiB per instruction) and hands each
// lane its own block through the wave's LDS area (rows padded to 80 bytes);
// other full blocks are read as four 16-byte loads per lane, the tail block
// byte by byte with bounds checks; blocks past the data write zero words.
__global__ __launch_bounds__(256) void bg_transpose_kernel(const uint8_t *__restrict__ bytes,
                                                           uint64_t n_bytes, uint64_t in_stride,
                                                           uint64_t *__restrict__ basis,
                                                           uint64_t n_words) {
  __shared__ uint4 stage[256 / 64][64 * 5];
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t input = blockIdx.y;
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint8_t *src = bytes + (uint64_t)input * in_stride;
  uint64_t *dst = basis + (uint64_t)input * 8 * n_words;
  const uint64_t off = w * 64;
  const uint64_t wave_off = (w - lane) * 64;  // the wave's first byte (wave-uniform)
  uint64_t x[8], y[8];
  if (wave_off + 4096 <= n_bytes) {
    // every lane's block is in range (and so w < n_words)
    const uint4 *p = reinterpret_cast<const uint4 *>(src + wave_off);
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = p[k * 64 + lane];
    // 16-byte piece e = k * 64 + lane belongs to lane e / 4, part e % 4
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t e = k * 64 + lane;
      stage[wv][(e >> 2) * 5 + (e & 3)] = v[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 r = stage[wv][lane * 5 + k];
      x[2 * k] = ((uint64_t)r.y << 32) | r.x;
      x[2 * k + 1] = ((uint64_t)r.w << 32) | r.z;
    }
  } else if (w >= n_words) {
    return;
  } else if (off + 64 <= n_bytes) {
    const uint4 *p = reinterpret_cast<const uint4 *>(src + off);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 v = p[k];
      x[2 * k] = ((uint64_t)v.y << 32) | v.x;
      x[2 * k + 1] = ((uint64_t)v.w << 32) | v.z;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      uint64_t v = 0;
      for (int b = 0; b < 8; b++) {
        const uint64_t pos = off + 8 * j + b;
        if (pos < n_bytes) v |= (uint64_t)src[pos] << (8 * b);
      }
      x[j] = v;
    }
  }
  bg_transpose_block(x, y);
#pragma unroll
  for (int i = 0; i < 8; i++) dst[(uint64_t)i * n_words + w] = y[i];
}

// Batched transpose: global word index -> buffer by binary search over the
// descriptors' word_start, then the same per-block transpose.
__global__ __launch_bounds__(256) void bg_transpose_batch_kernel(const uint8_t *__restrict__ bytes,
                                                                 uint64_t *__restrict__ basis,
                                                                 const bg_tdesc_t *__restrict__ descs,
                                                                 uint32_t n_desc, uint64_t total_words) {
  const uint64_t gw = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gw >= total_words) return;
  uint32_t lo = 0, hi = n_desc - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (descs[mid].word_start <= gw) lo = mid; else hi = mid - 1;
  }
  const bg_tdesc_t d = descs[lo];
  const uint64_t w = gw - d.word_start;
  const uint8_t *src = bytes + d.byte_off;
  uint64_t *dst = basis + d.basis_off;
  const uint64_t off = w * 64;
  uint64_t x[8], y[8];
  if (off + 64 <= d.n_bytes) {
    const uint4 *p = reinterpret_cast<const uint4 *>(src + off);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 v = p[k];
      x[2 * k] = ((uint64_t)v.y << 32) | v.x;
      x[2 * k + 1] = ((uint64_t)v.w << 32) | v.z;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      uint64_t v = 0;
      for (int b = 0; b < 8; b++) {
        const uint64_t pos = off + 8 * j + b;
        if (pos < d.n_bytes) v |= (uint64_t)src[pos] << (8 * b);
      }
      x[j] = v;
    }
  }
  bg_transpose_block(x, y);
#pragma unroll
  for (int i = 0; i < 8; i++) dst[(uint64_t)i * d.n_words + w] = y[i];
}

// dst word 2w   = positions 64w..64w+31 with position p at bit 31 - p%32,
// dst word 2w+1 = positions 64w+32..64w+63.
__global__ __launch_bounds__(256) void bg_msb32_kernel(const uint64_t *__restrict__ src,
                                                       uint32_t *__restrict__ dst, uint64_t total) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint64_t v = src[i];
  dst[2 * i] = __builtin_bitreverse32((uint32_t)v);
  dst[2 * i + 1] = __builtin_bitreverse32((uint32_t)(v >> 32));
}

__global__ __launch_bounds__(256) void bg_popcount_kernel(const uint64_t *__restrict__ src, uint64_t n,
                                                          unsigned long long *count) {
  uint64_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    acc += __popcll(src[i]);
  // wave reduction through DPP-free shuffles
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(count, (unsigned long long)acc);
}

// Moves the sticky error word of a control block to the caller's flag and
// clears it (one lane, vector atomics).
__global__ void bg_take_error_kernel(uint32_t *ctl_error, uint32_t *flag) {
  if (threadIdx.x != 0) return;
  const uint32_t e = __hip_atomic_load(ctl_error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(flag, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(ctl_error, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
extern "C" {

int bg_init(int device) {
  BG_HIP(hipSetDevice(device), "hipSetDevice");
  BG_HIP(hipFree(nullptr), "runtime init");
  if (device >= 0 && device < BG_MAX_DEVICES && !g_scratch[device]) {
    void *p = nullptr;
    BG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch[device] = static_cast<uint32_t *>(p);
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    g_cu_count = cus;
  g_err[0] = 0;
  return 0;
}

const char *bg_last_error(void) { return g_err; }

int bg_transpose(const uint8_t *bytes, uint64_t n_bytes, uint64_t in_stride, uint64_t *basis,
                 uint64_t n_words, uint32_t n_inputs, void *stream) {
  if (!basis || (n_bytes && !bytes)) return set_err(BG_EINVAL, "bg_transpose: null pointer");
  if (n_words * 64 < n_bytes) return set_err(BG_EINVAL, "bg_transpose: n_words * 64 < n_bytes");
  if (n_inputs == 0 || n_words == 0) return 0;
  if (n_inputs > 1 && in_stride < n_bytes) return set_err(BG_EINVAL, "bg_transpose: in_stride < n_bytes");
  if ((reinterpret_cast<uintptr_t>(bytes) & 15) || (n_inputs > 1 && (in_stride & 15)))
    return set_err(BG_EINVAL, "bg_transpose: input must be 16-byte aligned");
  if (n_inputs > 65535) return set_err(BG_EINVAL, "bg_transpose: too many inputs");
  const uint64_t blocks = (n_words + 255) / 256;
  if (blocks > 0x7fffffffull) return set_err(BG_EINVAL, "bg_transpose: input too large");
  hipLaunchKernelGGL(bg_transpose_kernel, dim3((uint32_t)blocks, n_inputs), dim3(256), 0, S(stream),
                     bytes, n_bytes, in_stride, basis, n_words);
  BG_HIP(hipGetLastError(), "bg_transpose launch");
  return 0;
}

int bg_transpose_batch(const uint8_t *bytes, uint64_t *basis, const bg_tdesc_t *descs, uint32_t n_desc,
                       uint64_t total_words, void *stream) {
  if (total_words == 0 || n_desc == 0) return 0;
  if (!bytes || !basis || !descs) return set_err(BG_EINVAL, "bg_transpose_batch: null pointer");
  if (reinterpret_cast<uintptr_t>(bytes) & 15)
    return set_err(BG_EINVAL, "bg_transpose_batch: byte array must be 16-byte aligned");
  const uint64_t blocks = (total_words + 255) / 256;
  if (blocks > 0x7fffffffull) return set_err(BG_EINVAL, "bg_transpose_batch: too large");
  hipLaunchKernelGGL(bg_transpose_batch_kernel, dim3((uint32_t)blocks), dim3(256), 0, S(stream), bytes, basis,
                     descs, n_desc, total_words);
  BG_HIP(hipGetLastError(), "bg_transpose_batch launch");
  return 0;
}

int bg_module_load(const char *path, bg_module_t *out) {
  if (!path || !out) return set_err(BG_EINVAL, "bg_module_load: null argument");
  hipModule_t m;
  BG_HIP(hipModuleLoad(&m, path), "hipModuleLoad");
  *out = new bg_module_s{m};
  return 0;
}

int bg_module_load_data(const void *image, bg_module_t *out) {
  if (!image || !out) return set_err(BG_EINVAL, "bg_module_load_data: null argument");
  hipModule_t m;
  BG_HIP(hipModuleLoadData(&m, image), "hipModuleLoadData");
  *out = new bg_module_s{m};
  return 0;
}

int bg_module_unload(bg_module_t module) {
  if (!module) return 0;
  hipError_t e = hipModuleUnload(module->mod);
  delete module;
  if (e != hipSuccess) return hip_err(e, "hipModuleUnload");
  return 0;
}

int bg_kernel_get(bg_module_t module, const char *name, bg_kernel_t *out) {
  if (!module || !name || !out) return set_err(BG_EINVAL, "bg_kernel_get: null argument");
  hipFunction_t f;
  hipError_t e = hipModuleGetFunction(&f, module->mod, name);
  if (e != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "kernel %s not found: %s", name, hipGetErrorString(e));
    return BG_ENOTFOUND;
  }
  // a generated kernel in the sweep + settle form comes with an empty
  // marker kernel <name>_two_launch
  uint32_t launches = 1;
  char marker[512];
  if (snprintf(marker, sizeof(marker), "%s_two_launch", name) < (int)sizeof(marker)) {
    hipFunction_t m;
    if (hipModuleGetFunction(&m, module->mod, marker) == hipSuccess) launches = 2;
    (void)hipGetLastError();  // a missing marker is not an error
  }
  // a kernel compiled with shared byte-pair filters comes with an empty
  // marker kernel <name>_shared_filter: it must be given its answers
  uint32_t needs_filter = 0;
  if (snprintf(marker, sizeof(marker), "%s_shared_filter", name) < (int)sizeof(marker)) {
    hipFunction_t m;
    if (hipModuleGetFunction(&m, module->mod, marker) == hipSuccess) needs_filter = 1;
    (void)hipGetLastError();
  }
  *out = new bg_kernel_s{f, 0, 0, 0, launches, needs_filter};
  return 0;
}

int bg_kernel_info(bg_kernel_t k, bg_kernel_info_t *info) {
  if (!k || !info) return set_err(BG_EINVAL, "bg_kernel_info: null argument");
  int v = -1;
  info->vgprs = -1;
  info->sgprs = -1;
  if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, k->fn) == hipSuccess) info->vgprs = v;
  v = 0;
  BG_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, k->fn), "shared size");
  info->lds_bytes = v;
  BG_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, k->fn), "local size");
  info->scratch_bytes = v;
  BG_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, k->fn), "max threads");
  info->max_threads = v;
  return 0;
}

namespace {

uint32_t clamp_take(uint32_t take) {
  if (take == 0) take = 1;
  return take > 8 ? 8 : take;
}

// Blocks of `threads` threads that fit on one CU (cached per kernel).
int occupancy(bg_kernel_t k, uint32_t threads, int32_t *blocks_per_cu) {
  if (k->occ_threads != threads) {
    int n = 0;
    BG_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, k->fn, (int)threads, 0),
           "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    k->occ_blocks = n > 0 ? n : 1;
    k->occ_threads = threads;
  }
  *blocks_per_cu = k->occ_blocks;
  return 0;
}

// Persistent grid: the blocks that fit on the device at once (or the
// caller's count), never more than one wave per tile.
int persistent_grid(bg_kernel_t k, uint32_t threads, uint32_t n_tiles, uint32_t take, uint32_t requested,
                    uint32_t *blocks) {
  const uint32_t wpb = threads / 64 * take;  // tiles per block and round
  uint32_t b = requested;
  if (b == 0) {
    int32_t per_cu = 1;
    int rc = occupancy(k, threads, &per_cu);
    if (rc) return rc;
    b = (uint32_t)per_cu * (uint32_t)g_cu_count;
  }
  const uint32_t need = (n_tiles + wpb - 1) / wpb;
  if (b > need) b = need;
  *blocks = b ? b : 1;
  return 0;
}


uint32_t clamp_queues(uint32_t nq, uint32_t blocks) {
  if (nq == 0) nq = 64;
  if (nq > BG_MAX_QUEUES) nq = BG_MAX_QUEUES;
  if (nq > blocks) nq = blocks;
  return nq ? nq : 1;
}

}  // namespace

int bg_kernel_occupancy(bg_kernel_t k, uint32_t block_threads, int32_t *blocks_per_cu, int32_t *cu_count) {
  if (!k || !blocks_per_cu) return set_err(BG_EINVAL, "bg_kernel_occupancy: null argument");
  if (block_threads == 0 || block_threads % 64 || block_threads > 1024)
    return set_err(BG_EINVAL, "bg_kernel_occupancy: block_threads must be a multiple of 64, at most 1024");
  int rc = occupancy(k, block_threads, blocks_per_cu);
  if (rc) return rc;
  if (cu_count) *cu_count = g_cu_count;
  return 0;
}

uint32_t bg_ctl_words(void) { return BG_CTL_WORDS; }

int bg_group_launch(bg_kernel_t k, const bg_group_args_t *args, uint32_t grid_blocks, void *stream) {
  if (!k || !args) return set_err(BG_EINVAL, "bg_group_launch: null argument");
  if (!args->basis || !args->result || !args->status || !args->ctl)
    return set_err(BG_EINVAL, "bg_group_launch: null buffer");
  if (k->needs_filter && !args->filter)
    return set_err(BG_EINVAL, "bg_group_launch: the kernel reads shared byte-pair answers; pass the buffer "
                              "bg_filter_launch filled as args->filter");
  if (args->n_tiles == 0) return 0;
  if (k->threads == 0) {
    // the block size the kernel was compiled for (its launch bounds): one
    // wave per tile, as many waves per block as the generator chose
    int v = 0;
    BG_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, k->fn), "max threads");
    k->threads = (v >= 64 && v <= 1024 && v % 64 == 0) ? (uint32_t)v : 256u;
  }
  const uint32_t threads = k->threads;
  uint32_t blocks = 0;
  int rc = persistent_grid(k, threads, args->n_tiles, clamp_take(args->take), grid_blocks, &blocks);
  if (rc) return rc;
  bg_group_args_t a = *args;
  a.n_queues = clamp_queues(a.n_queues, blocks);
  a.take = clamp_take(a.take);
  for (uint32_t phase = 0; phase < k->launches; phase++) {
    void *params[] = {&a, &phase};
    BG_HIP(hipModuleLaunchKernel(k->fn, blocks, 1, 1, threads, 1, 1, 0, S(stream), params, nullptr),
           "hipModuleLaunchKernel");
  }
  return 0;
}

int bg_group_launch_static(bg_kernel_t k, const uint64_t *basis, uint64_t *result, uint64_t n_words,
                           uint32_t n_tiles, uint32_t n_carry, uint32_t *error_flag, void *stream) {
  int dev = -1;
  BG_HIP(hipGetDevice(&dev), "hipGetDevice");
  uint32_t *s = (dev >= 0 && dev < BG_MAX_DEVICES) ? g_scratch[dev] : nullptr;
  if (!s) return set_err(BG_EINVAL, "bg_group_launch_static: call bg_init for this device first");
  if (!k || !basis || !result || !error_flag) return set_err(BG_EINVAL, "bg_group_launch_static: null argument");
  if (k->needs_filter)
    return set_err(BG_EINVAL, "bg_group_launch_static: kernels compiled with shared byte-pair filters need "
                              "their filter pre-pass; use bg_group_launch");
  if (n_carry == 0) return set_err(BG_EINVAL, "bg_group_launch_static: n_carry must be at least 1");
  if ((uint64_t)n_tiles * n_carry > BG_STATIC_STATUS_WORDS) {
    snprintf(g_err, sizeof(g_err),
             "bg_group_launch_static: %llu status words needed (%u tiles x %u), the static scratch holds %u",
             (unsigned long long)n_tiles * n_carry, n_tiles, n_carry, (unsigned)BG_STATIC_STATUS_WORDS);
    return BG_EINVAL;
  }
  bg_group_args_t a;
  a.basis = basis;
  a.result = result;
  a.status = s + BG_CTL_WORDS;
  a.ctl = s;
  a.n_words = n_words;
  a.n_tiles = n_tiles;
  a.n_queues = 0;
  a.n_carry = n_carry;
  a.take = 0;
  a.filter = nullptr;
  int rc = bg_group_launch(k, &a, 0, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(bg_take_error_kernel, dim3(1), dim3(64), 0, S(stream), s + BG_CTL_ERROR, error_flag);
  BG_HIP(hipGetLastError(), "bg_group_launch_static error hand-off");
  return 0;
}

int bg_filter_launch(bg_kernel_t k, const uint64_t *basis, uint64_t n_words, uint32_t n_tiles, uint32_t *filter,
                     uint32_t grid_blocks, void *stream) {
  if (!k || !basis || !filter) return set_err(BG_EINVAL, "bg_filter_launch: null argument");
  if (n_tiles == 0) return 0;
  if (k->threads == 0) {
    int v = 0;
    BG_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, k->fn), "max threads");
    k->threads = (v >= 64 && v <= 1024 && v % 64 == 0) ? (uint32_t)v : 256u;
  }
  // persistent: the blocks that fit at once (never more than one wave per
  // tile); wave i of the grid takes tiles i, i + waves, ...
  uint32_t blocks = 0;
  int rc = persistent_grid(k, k->threads, n_tiles, 1, grid_blocks, &blocks);
  if (rc) return rc;
  void *params[] = {&basis, &n_words, &n_tiles, &filter};
  BG_HIP(hipModuleLaunchKernel(k->fn, blocks, 1, 1, k->threads, 1, 1, 0, S(stream), params, nullptr),
         "hipModuleLaunchKernel(filter)");
  return 0;
}

int bg_filter_batch_launch(bg_kernel_t k, const bg_job_t *jobs, const uint32_t *tile_job, uint32_t total_tiles,
                           uint32_t *filter, uint32_t grid_blocks, void *stream) {
  if (!k || !jobs || !tile_job || !filter) return set_err(BG_EINVAL, "bg_filter_batch_launch: null argument");
  if (total_tiles == 0) return 0;
  if (k->threads == 0) {
    int v = 0;
    BG_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, k->fn), "max threads");
    k->threads = (v >= 64 && v <= 1024 && v % 64 == 0) ? (uint32_t)v : 256u;
  }
  uint32_t blocks = 0;
  int rc = persistent_grid(k, k->threads, total_tiles, 1, grid_blocks, &blocks);
  if (rc) return rc;
  void *params[] = {&jobs, &tile_job, &total_tiles, &filter};
  BG_HIP(hipModuleLaunchKernel(k->fn, blocks, 1, 1, k->threads, 1, 1, 0, S(stream), params, nullptr),
         "hipModuleLaunchKernel(filter batch)");
  return 0;
}

int bg_to_reference_layout(const uint64_t *src, uint32_t *dst, uint64_t n_words, uint64_t n_streams,
                           void *stream) {
  if (!src || !dst) return set_err(BG_EINVAL, "bg_to_reference_layout: null pointer");
  const uint64_t total = n_words * n_streams;
  if (total == 0) return 0;
  const uint64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffull) return set_err(BG_EINVAL, "bg_to_reference_layout: too large");
  hipLaunchKernelGGL(bg_msb32_kernel, dim3((uint32_t)blocks), dim3(256), 0, S(stream), src, dst, total);
  BG_HIP(hipGetLastError(), "bg_to_reference_layout launch");
  return 0;
}

int bg_popcount(const uint64_t *src, uint64_t n_words, uint64_t *count, void *stream) {
  if (!src || !count) return set_err(BG_EINVAL, "bg_popcount: null pointer");
  if (n_words == 0) return 0;
  uint64_t blocks = (n_words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(bg_popcount_kernel, dim3((uint32_t)blocks), dim3(256), 0, S(stream), src, n_words,
                     reinterpret_cast<unsigned long long *>(count));
  BG_HIP(hipGetLastError(), "bg_popcount launch");
  return 0;
}

int bg_batch_launch(bg_kernel_t k, const bg_batch_args_t *args, uint32_t grid_blocks, uint32_t block_threads,
                    void *stream) {
  if (!k || !args || !args->jobs || !args->tile_job || !args->ctl)
    return set_err(BG_EINVAL, "bg_batch_launch: null argument");
  if (k->needs_filter && !args->filter)
    return set_err(BG_EINVAL, "bg_batch_launch: the kernel reads shared byte-pair answers; pass the buffer "
                              "bg_filter_batch_launch filled as args->filter");
  if (block_threads == 0 || block_threads % 64 || block_threads > 1024)
    return set_err(BG_EINVAL, "bg_batch_launch: block_threads must be a multiple of 64, at most 1024");
  if (args->total_tiles == 0) return 0;
  uint32_t blocks = 0;
  int rc = persistent_grid(k, block_threads, args->total_tiles, clamp_take(args->take), grid_blocks, &blocks);
  if (rc) return rc;
  bg_batch_args_t a = *args;
  a.n_queues = clamp_queues(a.n_queues, blocks);
  a.take = clamp_take(a.take);
  for (uint32_t phase = 0; phase < k->launches; phase++) {
    void *params[] = {&a, &phase};
    BG_HIP(hipModuleLaunchKernel(k->fn, blocks, 1, 1, block_threads, 1, 1, 0, S(stream), params, nullptr),
           "hipModuleLaunchKernel(batch)");
  }
  return 0;
}

int bg_graph_begin(void *stream) {
  BG_HIP(hipStreamBeginCapture(S(stream), hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  return 0;
}

int bg_graph_end(void *stream, bg_graph_t *out) {
  if (!out) return set_err(BG_EINVAL, "bg_graph_end: null argument");
  hipGraph_t g;
  BG_HIP(hipStreamEndCapture(S(stream), &g), "hipStreamEndCapture");
  hipGraphExec_t e;
  hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  if (err != hipSuccess) {
    (void)hipGraphDestroy(g);  // the instantiation error is the one reported
    return hip_err(err, "hipGraphInstantiate");
  }
  *out = new bg_graph_s{g, e};
  return 0;
}

int bg_graph_launch(bg_graph_t graph, void *stream) {
  if (!graph) return set_err(BG_EINVAL, "bg_graph_launch: null graph");
  BG_HIP(hipGraphLaunch(graph->exec, S(stream)), "hipGraphLaunch");
  return 0;
}

int bg_graph_destroy(bg_graph_t graph) {
  if (!graph) return 0;
  const hipError_t e1 = hipGraphExecDestroy(graph->exec);
  const hipError_t e2 = hipGraphDestroy(graph->graph);
  delete graph;
  BG_HIP(e1, "hipGraphExecDestroy");
  BG_HIP(e2, "hipGraphDestroy");
  return 0;
}

int bg_stream_sync(void *stream) {
  BG_HIP(hipStreamSynchronize(S(stream)), "hipStreamSynchronize");
  return 0;
}

}  // extern "C"

#include "bitgen_serve.inc"
