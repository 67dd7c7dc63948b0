/*
 * bitgen_hip_serve.h -- staging and launch of one served batch in the
 * native runtime (libbitgen_hip.so; same conventions as bitgen_hip.h: 0 or
 * a status code, bg_last_error() for the message, streams as void *).
 *
 * No counterpart in the reference, which has no serving component.  A batch
 * is a list of jobs (a host byte buffer and the pattern set it is matched
 * against); every pattern set is a list of resident batch kernels (one per
 * kernel of the set, bg_batch_launch form).  bg_serve_pack lays the batch out
 * on the host -- input bytes, transpose descriptors, per-(set, kernel) job
 * tables and tile-to-job maps -- in one staging buffer (pinned memory) and
 * reports the layout; bg_serve_enqueue issues the whole batch on a stream:
 * one copy of the staging buffer to the device, one batched transpose, one
 * persistent launch per (set, kernel), the error flag hand-off and one copy
 * of every match stream back.  All launches of a slot share one control block,
 * so the epochs of its status words never repeat and they are never cleared.
 */
#ifndef BITGEN_HIP_SERVE_H
#define BITGEN_HIP_SERVE_H

#include <stdint.h>

#include "bitgen_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BG_ENOSPACE 1003 /* a buffer is too small: the layout holds the sizes needed */

/* One resident batch kernel: one (pattern set, kernel) pair. */
typedef struct bg_serve_kernel {
  bg_kernel_t kernel;     /* generated batch kernel                        */
  uint32_t hr;            /* halo words after a tile's output range        */
  uint32_t s;             /* output words per tile after tile 0            */
  uint32_t status_slots;  /* status words per tile (>= 1)                  */
  uint32_t n_out;         /* groups the kernel computes                    */
  uint32_t block_threads; /* threads per block (multiple of 64)            */
  uint32_t grid_cap;      /* most persistent blocks one launch may take    */
} bg_serve_kernel_t;

/* A pattern set: kernels [first_kernel, first_kernel + n_kernels). */
typedef struct bg_serve_set {
  uint32_t first_kernel;
  uint32_t n_kernels;
} bg_serve_set_t;

/* One job: n_bytes host bytes matched against pattern set `set`. */
typedef struct bg_serve_job {
  const uint8_t *data;
  uint64_t n_bytes;
  uint32_t set;
  uint32_t reserved;
} bg_serve_job_t;

/* Layout of a packed batch.  The *_words / stage_bytes fields are the sizes
 * the batch needs (set also when bg_serve_pack returns BG_ENOSPACE). */
typedef struct bg_serve_layout {
  uint64_t stage_bytes;  /* staging buffer (host and device copy)          */
  uint64_t basis_words;  /* u64 basis words of all jobs                    */
  uint64_t result_words; /* u64 match-stream words of all (job, kernel)    */
  uint64_t status_words; /* u32 look-back status words                     */
  uint64_t tdesc_off;    /* byte offset of the transpose descriptors       */
  uint32_t n_jobs;
  uint32_t n_queues;     /* persistent launches: one per (set, kernel)     */
} bg_serve_layout_t;

/* One persistent launch of a packed batch. */
typedef struct bg_serve_queue {
  uint32_t kernel;       /* index into the kernel table                    */
  uint32_t n_jobs;
  uint32_t total_tiles;
  uint32_t blocks;       /* persistent blocks of the launch                */
  uint64_t jobs_off;     /* byte offsets in the staging buffer             */
  uint64_t tile_job_off;
} bg_serve_queue_t;

/* Lay out a batch (host only, no device call).  stage: pinned host buffer of
 * stage_cap bytes; dev_basis / dev_result / dev_status: device addresses of
 * the slot's basis (basis_cap u64 words), result (result_cap u64 words) and
 * status (status_cap u32 words) buffers, written into the job tables.
 * queues: room for queue_cap launches.  job_result[j * max_kernels + k]
 * receives the u64 word offset, in the result buffer, of the streams of job
 * j's k-th set kernel (n_out streams of W_j words each).  Returns BG_ENOSPACE
 * (with the needed sizes in *layout) when a buffer is too small. */
int bg_serve_pack(const bg_serve_set_t *sets, uint32_t n_sets, const bg_serve_kernel_t *kernels,
                  uint32_t n_kernels, const bg_serve_job_t *jobs, uint32_t n_jobs, uint8_t *stage,
                  uint64_t stage_cap, uint64_t dev_basis, uint64_t basis_cap, uint64_t dev_result,
                  uint64_t result_cap, uint64_t dev_status, uint64_t status_cap, bg_serve_layout_t *layout,
                  bg_serve_queue_t *queues, uint32_t queue_cap, uint64_t *job_result, uint32_t max_kernels);

/* Issue a packed batch on `stream`: copy the staging buffer to dev_stage,
 * transpose every job, launch every queue with the slot's control block ctl,
 * move the launches' look-back error word to *dev_flag (cleared in ctl), and
 * copy the result words, then the flag, to host_out (pinned, result_words + 1
 * u64 words; the flag lands in the low 32 bits of the last word). */
int bg_serve_enqueue(const bg_serve_layout_t *layout, const bg_serve_queue_t *queues,
                     const bg_serve_kernel_t *kernels, const uint8_t *stage, uint8_t *dev_stage,
                     uint64_t *dev_basis, uint64_t *dev_result, uint32_t *ctl, uint32_t *dev_flag,
                     uint64_t *host_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
