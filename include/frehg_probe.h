/* Achievable-bandwidth probe of libfrehg_gw.so.
 *
 * No counterpart in the reference: a measurement utility.  It times plain
 * FP64 streaming kernels with HIP events on the current device, so a
 * benchmark can report the stencil kernels' bandwidth against the rate a
 * streaming kernel reaches on the same part, beside the published HBM peak.
 */
#ifndef FREHG_PROBE_H
#define FREHG_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FREHG_PROBE_OK 0
#define FREHG_PROBE_EINVAL 1
#define FREHG_PROBE_EDEVICE 5

/* Allocates eight FP64 arrays of n elements and times, as the mean over
 * `iterations` launches after one warm-up, a copy (1 read + 1 write) and a
 * stream of 6 reads + 2 writes.  Returns the algorithmic GB/s of each
 * (bytes = arrays x n x 8 / mean launch time).  Frees everything. */
int frehg_hbm_probe(int64_t n, int iterations, double* copy_gbs, double* stream_gbs);
/* One configuration of frehg_hbm_probe_sweep and its rate. */
typedef struct frehg_hbm_probe_result {
  int reads, writes;        /* FP64 arrays read and written per element */
  int bytes_per_lane;       /* 8 (one double a lane) or 16 (double2) */
  int nontemporal;          /* 1: non-temporal loads and stores */
  int blocks_per_cu;        /* grid = CUs x this, 256 threads a block */
  double mean_ms;           /* mean launch time */
  double gbs;               /* (reads + writes) x n x 8 B / mean_ms */
} frehg_hbm_probe_result;

/* Times every read/write mix of the step's kernels over eight FP64 arrays of
 * n elements: copy (1r1w), read-only (1r), write-only (1w), 6r2w (the PCG
 * direction kernel's compulsory arrays) and 5r1w (an SpMV row), each at 8 and
 * 16 bytes a lane, with ordinary and non-temporal loads/stores, and at
 * 1/2/4/8/16/32 blocks of 256 threads per CU: 120 configurations, the mean of
 * `iterations` launches after one warm-up each.  Writes up to `capacity`
 * results and their number to *count. */
int frehg_hbm_probe_sweep(int64_t n, int iterations, frehg_hbm_probe_result* out, int capacity,
                          int* count);
/* Message of the last failed probe call on this thread. */
const char* frehg_probe_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
