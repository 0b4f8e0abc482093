/*
 * tt_testing.h -- test hooks of the trace-transform library (no reference
 * counterpart; not part of the path's interface).  The test suite uses them
 * to reach code the public entry points cannot isolate.
 */
#ifndef TT_TESTING_H
#define TT_TESTING_H

#include "tt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Lets the next `allowed` workspace allocations of every context succeed and
 * makes the ones after fail as out of memory, so that TT_ERR_NOMEM can be
 * exercised; a negative value removes the limit. */
TT_API tt_status tt_testing_fail_allocations(long long allowed);

/* Compares the line kernel's short square root (v_sqrt_f32 and one residual
 * test per neighbour) with the correctly rounded sqrtf on the GPU for the
 * float bit patterns [lo, lo + count), skipping its documented exceptions
 * (0, 2^-96) and the negative denormals: the number of patterns whose
 * results differ (NaN results compare equal) and the smallest of them
 * (0xFFFFFFFF if none). */
TT_API tt_status tt_testing_short_sqrt(tt_context *ctx, uint32_t lo, uint64_t count,
                                       uint64_t *mismatches, uint32_t *first_bad);

#ifdef __cplusplus
}
#endif
#endif
