/*
 * ORACLE -- test infrastructure only.  Driver of the sanitizer build of the
 * oracle (pablo_oracle.c compiled with AddressSanitizer and UBSan into one
 * host executable, so no sanitizer runtime has to be preloaded into Python):
 *
 *   oracle_asan <job file> <output file>
 *
 * Job file (little-endian): int32 nrec, int32 nslots, uint64 n_bytes,
 * uint64 W, then nrec*5 int32 records, then n_bytes input bytes.  The driver
 * transposes the bytes into W-word basis streams, runs the program, and
 * writes the W result words.  Exit status: 0 ok, 2 bad job file, 3 oracle
 * error.  Any memory or undefined-behaviour error aborts with the
 * sanitizer's report.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int bg_oracle_run(const int32_t *prog, int32_t nrec, int32_t nslots, const uint64_t *basis, uint64_t W,
                  uint64_t *out);
int bg_oracle_transpose(const uint8_t *bytes, uint64_t n_bytes, uint64_t *basis, uint64_t W);

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[2];
  uint64_t sizes[2];
  if (fread(hdr, sizeof hdr, 1, f) != 1 || fread(sizes, sizeof sizes, 1, f) != 1) return 2;
  const int32_t nrec = hdr[0], nslots = hdr[1];
  const uint64_t n_bytes = sizes[0], W = sizes[1];
  if (nrec < 0 || nslots < 0 || W == 0 || W * 64 < n_bytes) return 2;
  int32_t *prog = malloc((size_t)nrec * 5 * sizeof(int32_t) + 1);
  uint8_t *bytes = malloc(n_bytes + 1);
  uint64_t *basis = calloc(8 * W, sizeof(uint64_t));
  uint64_t *out = calloc(W, sizeof(uint64_t));
  if (!prog || !bytes || !basis || !out) return 2;
  if (fread(prog, sizeof(int32_t), (size_t)nrec * 5, f) != (size_t)nrec * 5) return 2;
  if (fread(bytes, 1, n_bytes, f) != n_bytes) return 2;
  fclose(f);
  if (bg_oracle_transpose(bytes, n_bytes, basis, W) != 0) return 3;
  if (bg_oracle_run(prog, nrec, nslots, basis, W, out) != 0) return 3;
  FILE *o = fopen(argv[2], "wb");
  if (!o || fwrite(out, sizeof(uint64_t), W, o) != W) return 2;
  fclose(o);
  free(prog);
  free(bytes);
  free(basis);
  free(out);
  return 0;
}
