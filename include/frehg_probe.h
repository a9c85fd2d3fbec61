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
/* Message of the last failed probe call on this thread. */
const char* frehg_probe_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
