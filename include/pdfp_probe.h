/* libpdfp_probe: a measurement probe of the benchmark, built beside
 * libpdfp_hip but not part of it (bench.py loads it for its roofline).
 * No counterpart in the reference. */
#ifndef PDFP_PROBE_H
#define PDFP_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The random-gather ceiling.  Mean microseconds of one pass of
 * n_gathers gathers of record_bytes-byte records (4, 8, 16, 32 or 64) at
 * uniformly random rows of a table_bytes-byte device table, with the 4-byte
 * row indices streamed from HBM: the floor of a CSR pass whose columns are
 * scattered at random, for the benchmark's roofline.  Allocates and frees
 * its own buffers on `device`. */
int pdfp_gather_ceiling(int64_t table_bytes, int32_t record_bytes, int64_t n_gathers,
                        int32_t device, double* us_per_pass);
const char* pdfp_probe_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PDFP_PROBE_H */
