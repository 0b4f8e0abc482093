/*
 * tt.h -- C ABI of the MI355X trace transform (sinogram + circus path).
 *
 * Plain C: pointers and sizes only, no C++ or torch types.  Every entry
 * point returns a tt_status; no exception crosses the ABI.  Buffers are
 * caller-owned.  Device-pointer entry points (d_*) take memory already in
 * HBM and run on the context's stream; *_host entry points take host memory
 * and block until the results are back on the host.
 *
 * Layouts follow the reference's Eigen column-major MatrixXf:
 *   padded image     n x n,  element (row, col) at row + col*n
 *   sinogram (one T) n x A,  element (col, angle) at col + angle*n, i.e. one
 *                    angle's n values are contiguous
 *   sinograms        nt consecutive sinograms: [nt][A][n]
 *   circus functions [n_sino][np][A]
 *   index outputs    [A][n] (index angle*n + col)
 *
 * Functional codes are the reference's enum ordinals:
 *   TFunctional Radon,T1..T7 = 0..7      (c++/src/sinogram.hpp:22-31)
 *   PFunctional Hermite=0, P1=1, P2=2, P3=3  (c++/src/circus.hpp:27-32)
 */
#ifndef TT_H
#define TT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TT_API __attribute__((visibility("default")))

typedef enum tt_status {
    TT_OK = 0,
    TT_ERR_INVALID = 1,     /* bad argument; tt_last_error() says which     */
    TT_ERR_NOMEM = 2,       /* device or host allocation failed             */
    TT_ERR_DEVICE = 3,      /* HIP runtime error                            */
    TT_ERR_UNSUPPORTED = 4  /* valid request this build does not implement  */
} tt_status;

enum {
    TT_T_RADON = 0, TT_T1 = 1, TT_T2 = 2, TT_T3 = 3,
    TT_T4 = 4, TT_T5 = 5, TT_T6 = 6, TT_T7 = 7
};
enum { TT_P_HERMITE = 0, TT_P1 = 1, TT_P2 = 2, TT_P3 = 3 };

/* One P-functional: the reference's PFunctionalWrapper with its optional
 * Hermite arguments made explicit (c++/src/circus.hpp:34-57).  order and
 * center are read only when code == TT_P_HERMITE; center < 0 means "unset",
 * which the reference leaves undefined and this ABI rejects. */
typedef struct tt_pfunc {
    int32_t code;
    int32_t order;
    int32_t center;
} tt_pfunc;

/* Optional index outputs of tt_sinograms, each [A][n] int32 in device
 * memory, or NULL.  These are the integer decisions the reference makes per
 * column; they are exported so parity can be checked bit for bit. */
typedef struct tt_sino_aux {
    int32_t *median_x;    /* findWeightedMedian(x)            (T1,T2,T7)    */
    int32_t *median_sqrt; /* findWeightedMedian(sqrt(x))      (T3,T4,T5,T6) */
    int32_t *t6_rank;     /* T6: position k* in the sorted weighted tail    */
    int32_t *t6_src;      /* T6: tail offset r of that element (idx[k*])    */
    int32_t *t7_rank;     /* T7: position k* in the sorted tail             */
    int32_t *t7_src;      /* T7: row m + idx[k*] of that element, idx the
                             stable order of the tail (the adjoint's source) */
} tt_sino_aux;

typedef struct tt_context tt_context;

/* ---- context ----------------------------------------------------------- */

/* One context per host thread and device: owns a HIP stream, the device
 * workspaces (trig and T3-T5 tables, transposed image, Hermite weights, the
 * line kernel's median-walk checkpoints: 8 * ceil((n - 1) / 64) * A * n
 * bytes, ~100 MB at n = 1449 and 360 angles) and the cached graph of the
 * batched path.  No reference counterpart: the
 * reference keeps no state between getSinograms calls. */
TT_API tt_status tt_context_create(int device, tt_context **out);
TT_API tt_status tt_context_destroy(tt_context *ctx);
/* Run subsequent work on a caller stream (hipStream_t as void*); NULL is
 * HIP's default (null) stream.  A new context runs on a private stream. */
TT_API tt_status tt_context_set_stream(tt_context *ctx, void *stream);
TT_API tt_status tt_context_synchronize(tt_context *ctx);
/* Message of the last failure on this host thread ("" if none). */
TT_API const char *tt_last_error(void);

/* ---- geometry ---------------------------------------------------------- */

/* Replaces the size computation of pad (c++/src/auxiliary.cpp:257-275):
 * side n of the padded square and the offset (dfx, dfy) of the image in it. */
TT_API tt_status tt_pad_geometry(int rows, int cols, int *n, int *dfx,
                                 int *dfy);

/* Replaces pad (c++/src/auxiliary.cpp:257-275) on the device: rows x cols
 * column-major image in, n x n zero-padded column-major image out. */
TT_API tt_status tt_pad(tt_context *ctx, const float *d_image, int rows,
                        int cols, float *d_padded);

/* Number of angles for a step, as c++/src/sinogram.cpp:77:
 * floor(360 / step).  Fractional steps are a superset of the reference's
 * unsigned steps. */
TT_API int tt_angle_count(float angle_step_deg);

/* ---- sinograms (c++/src/sinogram.hpp:57-59, getSinograms) -------------- */

/* d_image: n x n padded image.  Angles k*step for k < n_angles (the
 * reference: step unsigned, n_angles = floor(360/step)).  tcodes: nt codes in
 * 0..7 (host memory).  d_sino: nt x A x n.  aux may be NULL. */
TT_API tt_status tt_sinograms(tt_context *ctx, const float *d_image, int n,
                              int n_angles, float angle_step_deg,
                              const int32_t *tcodes, int nt, float *d_sino,
                              const tt_sino_aux *aux);

/* Host-memory variant of tt_sinograms (aux unsupported). */
TT_API tt_status tt_sinograms_host(tt_context *ctx, const float *image, int n,
                                   int n_angles, float angle_step_deg,
                                   const int32_t *tcodes, int nt, float *sino);

/* ---- circus functions (c++/src/circus.hpp:69-71, getCircusFunctions) --- */

/* d_sino: n_sino sinograms, each n x n_cols (n_cols = number of angles).
 * pf: np P-functionals (host memory).  d_circus: n_sino x np x n_cols.
 * d_p2_rank (optional, n_sino x np x n_cols): P2's weighted-median position
 * in the sorted column. */
TT_API tt_status tt_circus(tt_context *ctx, const float *d_sino, int n,
                           int n_cols, int n_sino, const tt_pfunc *pf, int np,
                           float *d_circus, int32_t *d_p2_rank);

/* Host-memory variant of tt_circus for one sinogram. */
TT_API tt_status tt_circus_host(tt_context *ctx, const float *sino, int n,
                                int n_cols, const tt_pfunc *pf, int np,
                                float *circus);

/* tt_circus without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  No reference
 * counterpart.  P1, P2 and P3 only (Hermite needs allocated weight tables:
 * TT_ERR_UNSUPPORTED).  P3's spectra go to a library-owned device array of
 * 8 Mi floats (32 MiB) per GPU, so n_sino * n_cols * 2 * (n/2 + 1) may not
 * exceed 8388608 (TT_ERR_INVALID).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (different streams or contexts) race.  No
 * per-kernel timing.  Results equal tt_circus's bit for bit. */
TT_API tt_status tt_circus_noalloc(tt_context *ctx, const float *d_sino, int n,
                                   int n_cols, int n_sino, const tt_pfunc *pf,
                                   int np, float *d_circus, int32_t *d_p2_rank);

/* ---- steps either side of the path -------------------------------------- */

/* Replaces resize (c++/src/auxiliary.cpp:198-221), used by the orthonormal
 * mode (c++/src/transform.cpp:35-41): in_rows x in_cols -> rows x cols,
 * column-major, one-pixel zero border. */
TT_API tt_status tt_resize(tt_context *ctx, const float *d_in, int in_rows,
                           int in_cols, float *d_out, int rows, int cols);

/* Replaces zscore (c++/src/auxiliary.cpp:305-318) on `count` consecutive
 * vectors of length n (e.g. the circus functions [nt][np][A] -> the
 * signature columns of c++/src/transform.cpp:93-102). */
TT_API tt_status tt_zscore(tt_context *ctx, const float *d_in, int n, int count,
                           float *d_out);

/* Replaces nearest_orthonormal_sinogram (c++/src/circus.cpp:68-102): column
 * medians, alignment to the sinogram centre, then the polar factor
 * U[:, :A] V^T of the aligned (n + |max| + |min|) x A matrix, computed in
 * double (a polynomial polar iteration on FP64 matrix-core products;
 * rocSOLVER dsyevd / dgesvdj for matrices beyond its reach: condition
 * numbers above about 1e8, rank deficiency) and stored as float.  d_out holds
 * out_capacity_rows x A floats; *out_rows receives the aligned height and
 * *center the Hermite centre (c++/src/transform.cpp:78-84).  The aligned
 * height must be at least A (the reference's U block needs it). */
TT_API tt_status tt_nearest_orthonormal_sinogram(tt_context *ctx,
                                                 const float *d_sino, int n,
                                                 int n_cols, float *d_out,
                                                 int out_capacity_rows,
                                                 int *out_rows, int *center);

/* nearest_orthonormal_sinogram (c++/src/circus.cpp:68-102) of n_sino
 * sinograms (d_sino: n_sino x n_cols x n, as tt_sinograms writes them) in
 * one batched solve, as getTransform applies it to each T-functional's
 * sinogram (c++/src/transform.cpp:75-84).  Sinogram t's factor is an
 * out_rows[t] x n_cols column-major matrix at d_out + t * out_capacity_rows
 * * n_cols; centers[t] is its Hermite centre (host arrays of n_sino). */
TT_API tt_status tt_nearest_orthonormal_sinograms(tt_context *ctx,
                                                  const float *d_sino, int n,
                                                  int n_cols, int n_sino, float *d_out,
                                                  int out_capacity_rows,
                                                  int *out_rows, int *centers);

/* ---- batched forward (no reference counterpart) ------------------------- */

/* One image of a batch; all pointers are device memory. */
typedef struct tt_image {
    const float *d_image;  /* n x n padded image, column-major              */
    int32_t n;             /* padded side, 3..2730                           */
    float *d_sino;         /* nt x A x n sinograms                           */
    float *d_circus;       /* nt x np x A circus functions (NULL if np == 0) */
    void *d_saved;         /* optional, tt_saved_bytes(n, A, nt) bytes: the
                              line data (medians, T3..T5 integrals, P3
                              spectra, P2 sources) a later tt_backward_batched
                              of this batch reads instead of recomputing the
                              forward; NULL keeps nothing                    */
} tt_image;

/* Bytes of one image's d_saved block (0 for invalid sizes).  No reference
 * counterpart: the reference has no adjoint. */
TT_API size_t tt_saved_bytes(int n, int n_angles, int nt);
/* Byte offsets of the parts of a d_saved block: the weighted medians
 * findWeightedMedian(x) (med_x, [A][n] int32, written when T1/T2 are
 * requested) and findWeightedMedian(sqrt(x)) (med_s, with T3..T5), the
 * complex T3..T5 integrals ([3][A][n] float pairs), P3's half spectra
 * ([nt][A][2 (n/2 + 1)] float) and P2's source rows ([nt][A] int32). */
TT_API tt_status tt_saved_offsets(int n, int n_angles, int nt, size_t *med_x,
                                  size_t *med_s, size_t *z345, size_t *spec,
                                  size_t *p2src);

/* Sinograms (T0..T5) and circus functions (P1..P3) of every image of a batch
 * in one persistent launch per stage, the whole pass captured into a HIP graph
 * on the first call for a batch and replayed when the same batch (pointers,
 * sizes, functionals) comes again.  The reference processes images one after
 * another (c++/src/demo.cpp:179); results equal tt_sinograms/tt_circus per
 * image (P3 to float rounding of its twiddles).  The launch shape is the
 * default or, with tt_context_set_autotune, the fastest of the candidate
 * shapes timed on the batch's first call. */
TT_API tt_status tt_forward_batched(tt_context *ctx, const tt_image *images,
                                    int nimg, int n_angles, float angle_step_deg,
                                    const int32_t *tcodes, int nt,
                                    const tt_pfunc *pf, int np);
/* Autotuning of the batched paths: on a batch's first call the forward
 * times its line kernel with every candidate line tile and then every launch
 * shape, the backward its transposed rotation with every candidate gather
 * tile, and each keeps the fastest. */
TT_API tt_status tt_context_set_autotune(tt_context *ctx, int enable);
/* Launch shape chosen for the current batch: waves per workgroup and
 * persistent workgroups of the line kernel. */
TT_API tt_status tt_batched_launch_shape(tt_context *ctx, int *waves_per_block,
                                         int *blocks);

/* Work tiles of the batched paths; they tile the reference's loops over
 * angles and columns (c++/src/sinogram.cpp:110-150) and over the rotated
 * image's pixels (c++/src/auxiliary.cpp:223-255, transposed).
 *   line_angles x line_columns: one wave item of the batched T0..T5 line
 *     kernel: 64, 32 or 16 consecutive columns of line_angles consecutive
 *     angles, a multiple of 64 / line_columns up to 16 (candidates 1x64,
 *     2x64, 4x64, 2x32, 4x32, 4x16, 8x16);
 *   gather_tile x gather_tile pixels (32 or 16), gather_waves waves per
 *     workgroup (1..16), of the batched backward's transposed rotation.
 * A line's results and a pixel's gradient are the same bits whatever the
 * tiles.  Zero fields mean the default (1x64, 32 pixels, the most waves that
 * fit the LDS) or, with autotuning, the tuned choice. */
typedef struct tt_tiles {
    int32_t line_angles;
    int32_t line_columns;
    int32_t gather_tile;
    int32_t gather_waves;
} tt_tiles;
/* Forces tiles for the batched calls that follow (NULL: none forced).
 * Invalid tiles: TT_ERR_INVALID. */
TT_API tt_status tt_context_set_tiles(tt_context *ctx, const tt_tiles *tiles);
/* Tiles of the last planned batched forward (line fields) and batched
 * backward (gather fields, zero before any). */
TT_API tt_status tt_batched_tiles(tt_context *ctx, tt_tiles *tiles);

/* ---- adjoint (no reference counterpart) --------------------------------- */

/* Vector-Jacobian products of the forward path with the weighted medians
 * and P2's rank held at their forward values (they are piecewise constant in
 * the input) and the sample positions of the forward rotation.  The forward
 * quantities are recomputed from d_image; rotated images are never stored.
 * T6 and T7 hold their sort (hence their source row) fixed as well: with it,
 * T6 = q x[sm + q] and T7 = x[src] are linear in one sample each.
 * Supported: T0..T7 (the batched form T0..T5), P1, P2, P3, Hermite; padded
 * sides up to 1800. */

/* d_dsino (n_sino x n_cols x n) = J_circus^T d_dcircus (n_sino x np x n_cols)
 * for the sinograms d_sino. */
TT_API tt_status tt_circus_backward(tt_context *ctx, const float *d_sino, int n,
                                    int n_cols, int n_sino, const tt_pfunc *pf,
                                    int np, const float *d_dcircus,
                                    float *d_dsino);
/* d_dimage (n x n) = J_sinograms^T d_dsino (nt x A x n). */
TT_API tt_status tt_sinograms_backward(tt_context *ctx, const float *d_image,
                                       int n, int n_angles, float angle_step_deg,
                                       const int32_t *tcodes, int nt,
                                       const float *d_dsino, float *d_dimage);
/* d_dimage (n x n) = J^T d_dcircus (nt x np x A) of the whole path
 * image -> sinograms -> circus functions. */
TT_API tt_status tt_backward(tt_context *ctx, const float *d_image, int n,
                             int n_angles, float angle_step_deg,
                             const int32_t *tcodes, int nt, const tt_pfunc *pf,
                             int np, const float *d_dcircus, float *d_dimage);

/* One image of a batched backward; all pointers are device memory. */
typedef struct tt_image_grad {
    const float *d_image;    /* n x n padded image, column-major            */
    int32_t n;               /* padded side, 3..1800                         */
    const float *d_dcircus;  /* nt x np x A upstream gradient                */
    float *d_dimage;         /* n x n gradient out                           */
    const float *d_sino;     /* with d_saved: the forward's sinograms        */
    const void *d_saved;     /* optional: the block tt_forward_batched filled
                                for this image (tt_image.d_saved), with the
                                same angles and functionals and the image
                                unchanged since; NULL recomputes the forward */
} tt_image_grad;

/* tt_backward for every image of a batch, each stage one launch over the
 * batch (forward recompute unless d_saved, circus VJP, line records,
 * transposed-rotation gather), captured into a HIP graph on the first call
 * for a batch and replayed for the same batch.  T0..T5, P1..P3.  d_saved
 * must be set for every image or for none; a block not written by a
 * tt_forward_batched call on this context with the image's size, sinograms,
 * angles and functionals is rejected (TT_ERR_INVALID).  The T6/T7 adjoints
 * stay single-image only (tt_backward): the batched forward computes
 * T0..T5. */
TT_API tt_status tt_backward_batched(tt_context *ctx, const tt_image_grad *images,
                                     int nimg, int n_angles, float angle_step_deg,
                                     const int32_t *tcodes, int nt,
                                     const tt_pfunc *pf, int np);

/* ---- kernel timing ------------------------------------------------------ */

/* No reference counterpart (the reference times whole getTransform calls,
 * c++/src/demo.cpp:222-248).  When enabled, the context brackets each of its
 * kernel launches with HIP events on the stream the kernel runs on (the
 * batched entry points then launch their stages directly instead of
 * replaying their graphs, so that each stage is bracketed);
 * tt_profile_read synchronizes, returns the summed kernel time (ms) and
 * launch count of one kernel kind since the last read, and resets them. */
enum {
    TT_KERNEL_SINO_T05 = 0,   /* T0..T5 line kernel                         */
    TT_KERNEL_SINO_SORT = 1,  /* T6/T7 line kernel                          */
    TT_KERNEL_CIRCUS = 2,     /* P-functional kernels (P2 wave sort incl.)  */
    TT_KERNEL_AUX = 3,        /* transpose/pad/bounding-box helpers         */
    TT_KERNEL_BATCH = 4,      /* (unused: batched calls report per stage)   */
    TT_KERNEL_GATHER = 5,     /* adjoint: transposed rotation               */
    TT_KERNEL_CIRCUS_VJP = 6, /* adjoint of P1, P2, Hermite                 */
    TT_KERNEL_BATCH_BACKWARD = 7, /* (unused: batched calls report per stage) */
    TT_KERNEL_NOS = 8,        /* nearest orthonormal sinogram after the
                                 medians: align, polar iteration, to float  */
    TT_KERNEL_P3_DFT = 9,     /* P3's DFT GEMMs (forward and adjoint) and the
                                 adjoint's spectrum scaling                 */
    TT_KERNEL_RECORDS = 10,   /* adjoint: per-line gradient records         */
    TT_KERNEL_KINDS = 11
};
TT_API tt_status tt_profile_enable(tt_context *ctx, int enable);
TT_API tt_status tt_profile_read(tt_context *ctx, int kind, double *total_ms,
                                 int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
