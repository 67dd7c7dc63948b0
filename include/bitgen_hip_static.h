/*
 * bitgen_hip_static.h -- the workspace-free launch of libbitgen_hip.so (C ABI,
 * same conventions as bitgen_hip.h: 0 or a status code, bg_last_error() for
 * the message, streams as void *).
 */
#ifndef BITGEN_HIP_STATIC_H
#define BITGEN_HIP_STATIC_H

#include <stdint.h>

#include "bitgen_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Words of look-back status the static scratch of bg_group_launch_static holds
 * (16 MiB: e.g. a 100 MiB buffer with up to 160 carry operations per program). */
#define BG_STATIC_STATUS_WORDS (4u << 20)

/* bg_group_launch without a workspace: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under stream capture,
 * serving loops that must not call hipMalloc, harnesses that forbid it).  The
 * control block and status words live in a file-scope device array of
 * BG_CTL_WORDS + BG_STATIC_STATUS_WORDS u32 words, one per device, held from
 * bg_init (which resolves its address) until the process ends; nothing is
 * allocated.  Launches one group kernel over one buffer like bg_group_launch
 * (n_tiles tiles, n_carry >= 1 status words per tile, at most
 * BG_STATIC_STATUS_WORDS in total, else BG_EINVAL), then moves the launch's
 * look-back error word into *error_flag (device u32, 0 = results valid) and
 * clears it.  Every call on a device shares that device's scratch: calls
 * that may overlap on one device (two streams, two threads) race and must be
 * ordered by the caller, e.g. on one stream.  Kernels compiled with shared
 * byte-pair filters (they need their set's filter pre-pass and its buffer)
 * are refused with BG_EINVAL. */
int bg_group_launch_static(bg_kernel_t kernel, const uint64_t *basis, uint64_t *result, uint64_t n_words,
                           uint32_t n_tiles, uint32_t n_carry, uint32_t *error_flag, void *stream);

#ifdef __cplusplus
}
#endif
#endif
