/*
 * bitgen_hip.h -- C ABI of libbitgen_hip.so, the MI355X (gfx950) runtime of
 * the BitGen "hip" backend.  Plain pointers and sizes only; device buffers
 * are owned by the caller (PyTorch tensors on the Python side).  Every call
 * returns 0 on success or a nonzero status (a hipError_t value, or one of the
 * BG_E* codes below); bg_last_error() gives the message of the calling
 * thread's last failure.  Streams are passed as `void *` (a hipStream_t;
 * NULL = the default stream).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef BITGEN_HIP_H
#define BITGEN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BG_EINVAL 1001   /* bad argument (shape, alignment, null pointer) */
#define BG_ENOTFOUND 1002 /* kernel name not in module */

typedef struct bg_module_s *bg_module_t;
typedef struct bg_kernel_s *bg_kernel_t;

/* Tile queues.  A launch's tiles are dealt round-robin over n_queues
 * queues (tile t belongs to queue t % n_queues); block b takes tiles from
 * queue b % n_queues, each wave `take` consecutive tiles of its queue per
 * atomic add on the queue's counter.  Within a queue tiles are taken in
 * increasing order, and a wave only ever waits (carry look-back) for smaller
 * tiles, so no wave waits on a tile that nobody can take once the first
 * n_queues blocks are resident.  One counter per queue spreads the atomics
 * over n_queues addresses (a single device-wide counter serialises them).
 *
 * Control block ("ctl"): BG_CTL_WORDS u32 words, zeroed once before first
 * use and then owned by one launch at a time (launches that share a ctl must
 * be ordered on one stream).  Each word below sits on its own 256-byte line.
 * The last wave of each queue resets that queue's counters; the last queue to
 * finish advances the epoch, so consecutive launches (and graph replays)
 * need no host-side reset. */
#define BG_MAX_QUEUES 256
#define BG_CTL_STRIDE 64                                   /* u32 words   */
#define BG_CTL_COUNTER(q) ((q) * BG_CTL_STRIDE)            /* tile counter */
#define BG_CTL_DONE(q) ((BG_MAX_QUEUES + (q)) * BG_CTL_STRIDE) /* waves out */
#define BG_CTL_QDONE (2 * BG_MAX_QUEUES * BG_CTL_STRIDE)   /* queues out   */
#define BG_CTL_ERROR (BG_CTL_QDONE + BG_CTL_STRIDE)        /* sticky error */
#define BG_CTL_EPOCH (BG_CTL_ERROR + BG_CTL_STRIDE)        /* launch epoch */
#define BG_CTL_WORDS (BG_CTL_EPOCH + BG_CTL_STRIDE)

/* Launch descriptor of a generated group kernel (one group over one buffer).
 * Field order is part of the ABI: generated kernels take it by value. */
typedef struct bg_group_args {
  const uint64_t *basis; /* [8][n_words] basis streams, LSB-first words      */
  uint64_t *result;      /* [n_out][n_words] match-end streams of the
                            kernel's groups (one group, or several merged
                            consecutive groups: stream j at result + j*n_words) */
  uint32_t *status;      /* [n_tiles][n_carry] carry look-back status words,
                            zeroed before the first launch with this ctl     */
  uint32_t *ctl;         /* BG_CTL_WORDS words; ctl[BG_CTL_ERROR] is set
                            nonzero if a carry look-back wait gives up
                            (results are then invalid; the caller clears it) */
  uint64_t n_words;      /* words per stream                                 */
  uint32_t n_tiles;      /* tiles covering n_words                           */
  uint32_t n_queues;     /* 0 = default (64); clamped to 1 .. min(BG_MAX_QUEUES,
                            blocks)                                          */
  uint32_t n_carry;      /* carry operations in the program (>= 1 slot)      */
  uint32_t take;         /* tiles per atomic take, 0 = default; 1 .. 8       */
  const uint32_t *filter;/* [n_tiles][row] byte-pair answers of the pattern
                            set's filter pre-pass (bg_filter_launch), read by
                            kernels compiled with shared filters; else unused
                            (may be null)                                    */
} bg_group_args_t;

/* One (kernel, buffer) job of a batched launch.  Field order is ABI. */
typedef struct bg_job {
  const uint64_t *basis; /* [8][n_words] basis of the buffer               */
  uint64_t *result;      /* [n_out][n_words] match streams of the kernel's
                            groups over the buffer                         */
  uint32_t *status;      /* [n_tiles][n_carry] look-back status words       */
  uint64_t n_words;
  uint32_t group;        /* first group of the kernel (informational)       */
  uint32_t n_carry;      /* >= 1                                            */
  uint32_t tile_start;   /* first global tile index of this job             */
  uint32_t n_tiles;
} bg_job_t;

/* Launch descriptor of a generated batch kernel (one kernel of a pattern set
 * over many buffers).  Tiles are taken from a device queue: global tile index
 * g belongs to job tile_job[g]; jobs are ordered so each job's tiles are
 * contiguous and increasing. */
typedef struct bg_batch_args {
  const bg_job_t *jobs;
  const uint32_t *tile_job; /* [total_tiles] job index of each global tile */
  uint32_t *ctl;            /* BG_CTL_WORDS words (see above)              */
  uint32_t n_jobs;
  uint32_t total_tiles;
  uint32_t n_queues;        /* 0 = default; clamped like bg_group_args     */
  uint32_t take;            /* tiles per atomic take, as in bg_group_args  */
  const uint32_t *filter;   /* [total_tiles][row] byte-pair answers of the
                               set's batched filter pre-pass
                               (bg_filter_batch_launch), read by kernels
                               compiled with shared filters; else unused    */
} bg_batch_args_t;

typedef struct bg_graph_s *bg_graph_t;

/* One buffer of a batched transpose.  word_start = sum of n_words of the
 * buffers before it (the descriptors are sorted by it). */
typedef struct bg_tdesc {
  uint64_t byte_off;   /* offset of the buffer in the byte array (16-aligned) */
  uint64_t n_bytes;
  uint64_t basis_off;  /* word offset of the buffer's [8][n_words] basis      */
  uint64_t n_words;
  uint64_t word_start;
} bg_tdesc_t;

typedef struct bg_kernel_info {
  int32_t vgprs;        /* architected VGPRs per lane (-1 if unknown)   */
  int32_t sgprs;
  int32_t lds_bytes;    /* static LDS                                    */
  int32_t scratch_bytes;/* private (spill) bytes per lane                */
  int32_t max_threads;  /* max threads per block                         */
} bg_kernel_info_t;

/* Select the device and warm up the runtime.  Replaces the implicit CUDA
 * context set-up of CUDA_Kernel (bitgen/backend/cuda/cuda_kernel.py:50-63). */
int bg_init(int device);

/* Message of this thread's last failed call ("" if none). */
const char *bg_last_error(void);

/* Byte -> basis transpose.  Replaces bitgen.cuda.byte_stream_transpose
 * (bitgen/template/cuda/pybind.cpp:9, bitgen/template/cuda/kernel.cu:102-124).
 * bytes: [n_inputs][in_stride] device bytes, the first n_bytes of each used;
 * basis: [n_inputs][8][n_words] words, position p = bit p%64 of word p/64,
 * basis i = bit (7 - i) of each byte; positions >= n_bytes are zero.
 * Requires n_words * 64 >= n_bytes.  Asynchronous on `stream` (the reference
 * synchronises the device on every call, kernel.cu:123). */
int bg_transpose(const uint8_t *bytes, uint64_t n_bytes, uint64_t in_stride,
                 uint64_t *basis, uint64_t n_words, uint32_t n_inputs,
                 void *stream);

/* Transpose of many buffers of uneven length in one launch (no counterpart
 * in the reference, whose inputs are copies of one buffer).  descs: device
 * array of n_desc descriptors; total_words = sum of their n_words. */
int bg_transpose_batch(const uint8_t *bytes, uint64_t *basis, const bg_tdesc_t *descs, uint32_t n_desc,
                       uint64_t total_words, void *stream);

/* Load a gfx950 code object (.hsaco).  Replaces the NVRTC/nvJitLink build and
 * cuModuleLoadData of CUDA_Kernel.init (cuda_kernel.py:64-288). */
int bg_module_load(const char *hsaco_path, bg_module_t *out);
int bg_module_load_data(const void *image, bg_module_t *out);
int bg_module_unload(bg_module_t module);

/* Look up a kernel by name (cuModuleGetFunction, cuda_kernel.py:153-171). */
int bg_kernel_get(bg_module_t module, const char *name, bg_kernel_t *out);

/* Resource usage (cuda_kernel.py:296-398 prints the CUDA equivalents). */
int bg_kernel_info(bg_kernel_t kernel, bg_kernel_info_t *info);

/* Blocks of block_threads threads of `kernel` that fit on one CU at once
 * (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor), and the device's CU
 * count; the default persistent grid is their product. */
int bg_kernel_occupancy(bg_kernel_t kernel, uint32_t block_threads, int32_t *blocks_per_cu,
                        int32_t *cu_count);

/* Words of a control block (BG_CTL_WORDS), for callers without this header. */
uint32_t bg_ctl_words(void);

/* Launch one generated group kernel over one buffer: persistent blocks (of
 * the size the kernel's launch bounds name) that take tiles from the control
 * block's queues until none is left (grid_blocks = 0: as many blocks as fit
 * on the device at once, never more than one wave per tile).  A kernel
 * generated in the sweep + settle form (it comes with an empty marker
 * kernel <name>_two_launch) is launched twice, with phase argument 0 and 1;
 * others once.  Replaces CUDA_Kernel.run (cuda_kernel.py:411-427),
 * asynchronously: the caller synchronises. */
int bg_group_launch(bg_kernel_t kernel, const bg_group_args_t *args, uint32_t grid_blocks,
                    void *stream);

/* Filter pre-pass of a pattern set whose kernels share their byte-pair
 * answers (no counterpart in the reference: the byte-pair filters are this
 * backend's own).  One wave per tile of the set's shared tile geometry
 * builds the tile's pair table once and writes the answers of every
 * kernel's filters to filter[n_tiles][row] (row: the set's answer words per
 * tile); the set's group kernels then read their words instead of each
 * building the table.  Launch it on the same stream before the group
 * kernels of the same buffer. */
int bg_filter_launch(bg_kernel_t kernel, const uint64_t *basis, uint64_t n_words, uint32_t n_tiles,
                     uint32_t *filter, uint32_t grid_blocks, void *stream);

/* The same pre-pass over the global tiles of a batch (kernel
 * bg_filter_batch): tile g belongs to job tile_job[g] of the jobs table of
 * any kernel of the set (the kernels of a set that share filters have the
 * same tiles per job), and its answer row lands at filter[g][row].  A
 * kernel node only, so it can be captured in a graph.  grid_blocks (both
 * entries): 0 = the blocks that fit on the device at once, each wave taking
 * every (waves)-th tile; else that many blocks (at most one wave per tile). */
int bg_filter_batch_launch(bg_kernel_t kernel, const bg_job_t *jobs, const uint32_t *tile_job,
                           uint32_t total_tiles, uint32_t *filter, uint32_t grid_blocks, void *stream);

/* LSB-first u64 words -> the reference's u32 layout (position p = bit
 * 31 - p%32 of word p/32, bitgen/bitstream.py:30-52).  dst holds 2*n_words
 * u32 words per stream; n_streams streams back to back. */
int bg_to_reference_layout(const uint64_t *src, uint32_t *dst, uint64_t n_words,
                           uint64_t n_streams, void *stream);

/* Popcount of n_words words, accumulated into *count (device u64). */
int bg_popcount(const uint64_t *src, uint64_t n_words, uint64_t *count, void *stream);

/* Batched launch (no counterpart in the reference, which launches one grid
 * of (groups x inputs) blocks per pattern set): grid_blocks blocks (0 = as
 * many as fit at once) of block_threads threads of the persistent batch
 * kernel (two, sweep then settle, for a kernel in that form, as in
 * bg_group_launch).  Kernel nodes only, so it can be captured in a graph and
 * replayed (the control block resets itself at the end of every launch). */
int bg_batch_launch(bg_kernel_t kernel, const bg_batch_args_t *args, uint32_t grid_blocks,
                    uint32_t block_threads, void *stream);

/* Stream capture into a replayable graph (hipStreamBeginCapture /
 * hipStreamEndCapture / hipGraphInstantiate / hipGraphLaunch).  Work captured
 * must not contain memsets. */
int bg_graph_begin(void *stream);
int bg_graph_end(void *stream, bg_graph_t *out);
int bg_graph_launch(bg_graph_t graph, void *stream);
int bg_graph_destroy(bg_graph_t graph);

int bg_stream_sync(void *stream);

#ifdef __cplusplus
}
#endif
#endif
