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

/* Runs the FP64 matrix-core products of the nearest-orthonormal iteration
 * on a strided batch of `count` column-major matrices (device pointers,
 * matrices stored back to back): C = A^T B (trans_a: A is k x m) or A B
 * (A is m x k; when k is a multiple of 4, B must be symmetric: it is read
 * transposed), B k x n, C m x n.  sym: m == n and the product is known to
 * be symmetric; only the blocks on and below the diagonal are computed and
 * each value is written to (i, j) and (j, i).  With coef (5 doubles,
 * sym, A = B = S of k x k, k a multiple of 4) the product is the polynomial
 * step's S S, whose kernel also writes T = c3 S + c4 S^2 to d_t and
 * Q = c0 I + c1 S + c2 S^2 to d_q.  Returns after the context's stream has
 * finished. */
TT_API tt_status tt_testing_nos_product(tt_context *ctx, int trans_a, int sym, int count,
                                        int m, int n, int k, const double *d_a,
                                        const double *d_b, double *d_c, const double *coef,
                                        double *d_t, double *d_q);

#ifdef __cplusplus
}
#endif
#endif
