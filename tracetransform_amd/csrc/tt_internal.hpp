// tt_internal.hpp -- launch descriptors shared by the HIP kernels
// (tt_kernels.hip) and the C-ABI host code (tt_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ttk {

constexpr int kWave = 64;        // CDNA wavefront width
constexpr int kBlock = 256;      // 4 waves per workgroup
constexpr int kMaxP = 16;        // P-functionals per tt_circus call
constexpr int kNumXcd = 8;       // MI355X: 8 XCDs, workgroups round-robin

// Output slot of each T code (T0..T5), -1 when not requested.
struct Slots {
    int s[8];  // output slot of T-functional k (0..7), or -1
};

// Forward T-functional sweep for one padded image (T0..T5).  One wave owns
// 64 adjacent columns of one angle; each lane streams its column's rows.
struct SinoParams {
    const float *img;     // n x n column-major
    const float *imgT;    // n x n row-major copy (transpose of img)
    const float2 *trig;   // [A] (cos(-theta), sin(-theta)) as float
    const float2 *tab345; // [3][n] (re, im) tables of T3, T4, T5
    float *sino;          // [nt][A][n]
    int32_t *med_x;       // [A][n] or null
    int32_t *med_s;       // [A][n] or null
    float2 *z345;         // [3][A][n] complex T3..T5 integrals, or null
    const int *bbox;      // non-zero pixels' bounds {x0, x1, y0, y1} + flag, or null
    float2 *ckpt;         // [sino_ckpt_blocks(n)][A][n] walk prefixes (workspace)
    int n;
    int A;
    int nb;               // column blocks per angle: ceil(n / 64)
    int items;            // A * nb
    Slots slot;
};

// Adjoint.  Per-line record of the rotated-domain gradient of one angle's
// column: G(row) = g0 + [row >= m] (g1 (row-m) + g2 (row-m)^2)
//                + [row > sm] sum_t (alpha_t W_t.re[row-sm] + beta_t W_t.im[row-sm]).
// The first 32 bytes carry what T0 and T3..T5 need.
struct LineRecord {
    float g0;
    int sm;
    float ab[6];          // alpha3, beta3, alpha4, beta4, alpha5, beta5
    float g1, g2;
    int m;
    int pad;
};
// T6 = q x[sm + q] and T7 = x[src7] with the sort held fixed add
// [row == r6] c6 + [row == r7] c7 (c6 = dT6 * q at row r6 = sm + q, c7 = dT7
// at row r7 = src7; rows -1 when absent).
struct LineRecord67 {
    int r6, r7;
    float c6, c7;
};

// One image of a batch (device array, ordered by first item).
struct BatchImage {
    const float *img;     // n x n column-major
    // (the batched forward writes neither medians nor T3..T5 integrals)
    const float *imgT;    // row-major copy (written by the batched transpose)
    float *sino;          // [nt][A][n]
    int32_t *med_x;
    int32_t *med_s;
    const int *bbox;      // non-zero pixels' bounds {x0, x1, y0, y1} + flag, or null
    float2 *ckpt;         // [sino_ckpt_blocks(n)][A][n] walk prefixes (workspace)
    int n;
    int nb;               // column tiles per angle tile (batched line kernel)
    int item0;            // first wave item of this image
    int col0;             // first sinogram column (over all T) of this image
    float *circus;        // [nt][np][A]
    float *spec;          // [nt][A][2h] P3 half spectra, or null
    int32_t *p2src;       // [nt][A] P2 source rows (the adjoint's one-hot), or null
    // adjoint (tt_backward_batched); null in the forward-only batch
    float2 *z345;         // [3][A][n] T3..T5 integrals
    const float *dcircus; // [nt][np][A]
    float *dsino;         // [nt][A][n]
    LineRecord *rec;      // [A][n]
    float *dimg;          // n x n
    int line0;            // first line (angle*n + col) over the batch
    int pix0;             // first pixel over the batch
    int tile0;            // first gather tile over the batch (tiles of the launch's side)
};

struct BatchSinoParams {
    const BatchImage *images;
    int nimg;
    int items;            // total wave items
    int tile_a;           // angles of one wave item (a multiple of 64 / tile_c)
    int tile_c;           // columns of one wave item: 64, 32 or 16
    const float2 *trig;   // [A]
    const float2 *tab345; // [3][tabn]
    int tabn;             // largest n of the batch
    int A;
    unsigned int *counter;
    Slots slot;
};

// P-functionals over every sinogram column of a batch: one workgroup per
// column; P3 twiddles are computed per image in LDS.
struct BatchCircusParams {
    const BatchImage *images;
    int nimg;
    int cols;             // total columns: sum over images of nt * A
    int nt;
    int A;
    int nmax;
    int np;
    int code[kMaxP];
    int p2_wave;          // P2 entries written by circus_p2_batched_kernel (launcher's)
    int with_circus;      // circus_p2_batched_kernel writes circus values
};

// Sort-based T-functionals T6/T7 for one padded image: one workgroup owns one
// (angle, column) line at a time.
struct SortTParams {
    const float *img;     // n x n column-major
    const float *imgT;    // row-major copy
    const float2 *trig;
    float *sino;
    int32_t *med_x;
    int32_t *med_s;
    int32_t *t6_rank;
    int32_t *t6_src;
    int32_t *t7_rank;
    int32_t *t7_src;      // T7 source row (stable order of the tail), or null
    const int *bbox;      // non-zero pixels' bounds {x0, x1, y0, y1} + flag, or null
    // (line, which) queues: a pass reads defer_in (all lines when NULL) and
    // queues the line items (line << 2 | T6/T7 mask) it cannot hold in
    // defer_out (room for A * n items each)
    const int32_t *defer_in;
    const unsigned int *defer_in_count;
    int32_t *defer_out;
    unsigned int *defer_out_count;
    unsigned int *work;   // work-item counter of the pass (set by launch_sino_sort)
    int chunk;            // items taken per atomic on work
    int n;
    int A;
    int slot6;
    int slot7;
};

// P-functionals over sinogram columns: one workgroup per column.
struct CircusParams {
    const float *sino;    // [nsino][ncols][n]
    float *out;           // [nsino][np][ncols]
    int32_t *p2rank;      // [nsino][np][ncols] or null
    const float *p3_spec; // [nsino * ncols][2 * p3_h] half spectra (re, then im), or null
    int p3_h;             // n / 2 + 1
    const double *herm_w; // [nherm][n] Hermite weights
    int n;
    int ncols;
    int nsino;
    int np;
    int code[kMaxP];
    int herm[kMaxP];      // row of herm_w for Hermite entries
    int p2_wave;          // P2 entries written by circus_p2_kernel (launcher's)
};

struct RecordParams {
    const float *dsino;   // [nt][A][n] sinogram gradient
    const int32_t *med_x; // [A][n]
    const int32_t *med_s; // [A][n]
    const int32_t *t6_src; // [A][n] T6 source offset from sm, or null
    const int32_t *t7_src; // [A][n] T7 source row, or null
    const float2 *z345;   // [3][A][n]
    LineRecord *rec;      // [A][n]
    LineRecord67 *rec67;  // [A][n] (T6/T7 requested) or null
    int n;
    int A;
    Slots slot;
};

// Transposed rotation: one wave per (tile x tile pixel tile, angle group);
// with groups > 1 each group writes its partial image into part and
// gather_reduce sums them in group order.
struct GatherParams {
    const LineRecord *rec;      // [A][n]
    const LineRecord67 *rec67;  // [A][n] or null
    const float2 *trig;         // [A]
    const float2 *tab345;       // [3][n]
    float *dimg;                // n x n column-major
    float *part;                // [groups][n][n] (groups > 1)
    int n;
    int A;
    int do345;
    int do12;                   // T1 or T2 requested
    int tile;                   // 16 or 32
    int groups;                 // angle groups
    int waves;                  // waves per workgroup (1..8)
};
// Dynamic LDS bytes of one gather workgroup.
size_t gather_lds_bytes(int tile, int waves, int tabn, bool do345, bool do12, bool do67);

// Circus vector-Jacobian product: one workgroup per sinogram column.
struct CircusVjpParams {
    const float *sino;    // [nsino][ncols][n]
    const float *dcircus; // [nsino][np][ncols]
    float *dsino;         // [nsino][ncols][n] (holds P3's part when p3_done)
    int p3_done;
    const int32_t *p2src; // [nsino][ncols] P2 source rows, or null (then sorted here)
    const double *herm_w; // [nherm][n]
    int n;
    int ncols;
    int nsino;
    int np;
    int code[kMaxP];
    int herm[kMaxP];
};

// P3's transforms as GEMMs on the matrix cores (tt_dft.hip).  One job per
// image: forward, out (2h x cols, ld_out) = W^T in (n x cols, ld_in), the
// half spectra (real parts, then imaginary parts); adjoint, out (n x cols)
// = W in (2h x cols).  W (n x 2h) = [cos(2 pi jk/n) | -sin(2 pi jk/n)],
// h = n / 2 + 1.
struct DftJob {
    const float *in;
    float *out;
    int n;
    int cols;
    int ld_in;
    int ld_out;
    int tile0;            // first workgroup of this job (dft_job_tiles each)
    // p3_vjp_scale: the columns' upstream circus gradients, P-functional q's
    // of column c at g[(c / g_cols) * g_stride + q * g_cols + c % g_cols]
    const float *g;
    int g_cols;
    int g_stride;
};
struct DftParams {
    const DftJob *jobs;   // device array of njobs, or null: job is the one
    int njobs;
    int tiles;            // total workgroups
    int bm;               // output rows per workgroup: 64 or 128
    int bn;               // output columns per workgroup: 64 or 128 (128 when bm is)
    DftJob job;
};
struct P3ScaleParams {
    const DftJob *jobs;   // device array of njobs, or null: job is the one
    int njobs;            // (in: the spectra, out: the scaled spectra)
    DftJob job;
    int np3;              // P3 entries of the P list, at positions q3
    int q3[kMaxP];
};
// Workgroups of one job with bm x bn output tiles, and the tile shape of a
// launch whose jobs need tiles128 workgroups of 128 x 128 and tiles64 of
// 64 x 128.  The shape only partitions the outputs: every output is the same
// sum in the same order whatever the shape.
int dft_job_tiles(int n, int cols, bool forward, int bm, int bn);
void dft_tile_shape(long long tiles128, long long tiles64, int *bm, int *bn);
size_t dft_lds_bytes(int nmax);
hipError_t launch_dft_gemm(const DftParams &p, bool forward, int nmax, hipStream_t s);
hipError_t launch_p3_vjp_scale(const P3ScaleParams &p, int max_cols, hipStream_t s);

// Test hook: mismatches of the line kernel's short square root against
// sqrtf over float bit patterns [lo, lo + count) (mism and first in device
// memory, first initialised to 0xFFFFFFFF by the caller).
hipError_t launch_short_sqrt_check(uint32_t lo, uint64_t count, unsigned long long *mism,
                                   unsigned int *first, hipStream_t s);

// Steps either side of the path and the orthonormal step.
hipError_t launch_resize(const float *in, int in_rows, int in_cols, float *out,
                         int rows, int cols, hipStream_t s);
hipError_t launch_zscore(const float *in, int n, int count, float *out,
                         hipStream_t s);
hipError_t launch_column_medians(const float *in, int n, int cols, int32_t *med,
                                 hipStream_t s);
// Batched forms over count sinograms (n x cols each, back to back) and their
// column medians med[t * cols + c]: align into matrices mstride doubles
// apart (offset origin: the sinogram's largest median); to_float of matrix
// t's aligned height n + |max - sc| + |min - sc| rows into out + t * ostride.
hipError_t launch_align_batched(const float *in, int n, int cols, const int32_t *med,
                                int count, int rows_out, size_t mstride, double *out,
                                hipStream_t s);
hipError_t launch_to_float_batched(const double *in, int ld, size_t mstride, int cols,
                                   const int32_t *med, int n, int sc, int count, int rows_max,
                                   float *out, size_t ostride, hipStream_t s);
hipError_t launch_scale_eigvec(const double *v, const double *lambda, int k, int count,
                               double *out, hipStream_t s);
hipError_t launch_newton_schulz_factor(double *x, int k, int count, hipStream_t s);
// Polar iteration pieces (nearest orthonormal sinogram): coefficients of
// q(y) = c0 + c1 y + c2 y^2 + c3 y^3 + c4 y^4 for one step.
struct PolarStep {
    double c[5];
};
// Gram matrices from row-chunk partials (sum over chunks, in order).
hipError_t launch_nos_reduce(const double *part, int chunks, int k, int count, double *S,
                             hipStream_t s);
// First-step scales s = ||G||_F^-1, r = ||G||_F^(-1/2) per matrix (part:
// kNosSlices * count doubles of scratch).
constexpr int kNosSlices = 64;
hipError_t launch_nos_scales(const double *G, int k, int count, double *part, double *scale,
                             double *r, hipStream_t s);
// part[m * kNosSlices + j] = max |S_m - I| over slice j (the host takes the
// maximum).
hipError_t launch_nos_deviation(const double *S, int k, int count, double *part,
                                hipStream_t s);
hipError_t launch_nos_poly(const double *S, const double *S2, int k, int count,
                           const PolarStep &c, const double *scale, const double *r, double *T,
                           double *Q, hipStream_t s);
// Batched FP64 products C = alpha op(A) B + beta C (column major, op(A) =
// A^T when trans_a) on the matrix cores (tt_nos.hip).  The batch is a
// device table of pointer triples, or (table == nullptr) base pointers and
// element strides.  sym: m == n and the product is known to be symmetric;
// only tiles on and below the diagonal are computed and each value is
// written to (i, j) and (j, i).  Without trans_a, B must be symmetric (it
// is read transposed) when k is a multiple of 4.
struct NosGemmJob {
    const double *a, *b;
    double *c;
};
struct NosGemmBatch {
    const NosGemmJob *table;
    const double *a, *b;
    double *c;
    long long sa, sb, sc;
};
hipError_t launch_nos_gemm(const NosGemmBatch &bt, int count, bool trans_a, bool sym, int m,
                           int n, int k, int lda, int ldb, int ldc, double alpha, double beta,
                           hipStream_t s);
// S2 = S S for a strided batch of k x k symmetric S (k a multiple of 4) with
// a polynomial step's elementwise terms in the same kernel: with a = s S,
// b = s^2 S2 (s = scale[m], r = r[m] on the first step, else 1),
// T = s^2 r (c3 a + c4 b) and Q = r (c0 I + c1 a + c2 b), so Q + S2 T = q(S).
struct NosPolyOut {
    double *T, *Q;
    long long stride;  // elements between matrices (k * k)
    const double *scale, *r;
    double c[5];
};
hipError_t launch_nos_square_poly(const double *S, double *S2, int k, int count,
                                  const NosPolyOut &po, hipStream_t s);

hipError_t launch_records(const RecordParams &p, hipStream_t s);
hipError_t launch_gather(const GatherParams &p, hipStream_t s);
hipError_t launch_circus_vjp(const CircusVjpParams &p, hipStream_t s);

hipError_t launch_transpose(const float *in, float *out, int n, hipStream_t s);
hipError_t launch_pad(const float *in, int rows, int cols, float *out, int n,
                      int dfx, int dfy, hipStream_t s);
// Bounding box {x0, x1, y0, y1} (columns, rows) of the non-zero pixels of an
// n x n column-major image, for the line kernels' row ranges; an all-zero
// image gives x0 > x1.  Word 4 is 1 when some pixel is negative or in
// (0, 2^-45) (then a bilinear sample can be a negative or a positive value
// below 2^-96), else 0.  kBboxInts words per image.
constexpr int kBboxInts = 8;
hipError_t launch_image_bbox(const float *img, int n, int *bbox, hipStream_t s);
// The same for every image of a batch (bbox of image k at images[k].bbox).
hipError_t launch_image_bbox_batched(const BatchImage *images, int nimg, int nmax,
                                     hipStream_t s);
// Checkpoints of the median walks the T0..T5 line kernel keeps per line:
// the prefixes at the starts of its 64-step blocks after the first, up to
// this many (a wave of one angle and 64 columns spans at most n + 64 steps
// with its row skew; a wave spanning more keeps the first ones).
__host__ __device__ inline int sino_ckpt_blocks(int n)
{
    const int b = (n + 62) / 64;
    return b > 1 ? b : 1;
}

// Line tiles of the batched T0..T5 kernel: tile_c columns of 64, 32 or 16
// and tile_a angles, a multiple of 64 / tile_c up to 16.
inline bool line_tile_ok(int tile_a, int tile_c)
{
    if (tile_c != 64 && tile_c != 32 && tile_c != 16)
        return false;
    return tile_a >= 1 && tile_a <= 16 && tile_a % (kWave / tile_c) == 0;
}
// Wave items of an image of side n at A angles, and the column tiles per
// angle tile.
inline int line_tile_items(int n, int A, int tile_a, int tile_c, int *nb)
{
    *nb = (n + tile_c - 1) / tile_c;
    return *nb * ((A + tile_a - 1) / tile_a);
}

hipError_t launch_sino_t05(const SinoParams &p, bool need_x, bool need_s,
                           bool do345, hipStream_t s);
// counts: kSortCounters counters (always); n > 512 also needs two queues of
// A * n line items.
constexpr int kSortCounters = 5;
hipError_t launch_sino_sort(const SortTParams &p, int32_t *queue_a, int32_t *queue_b,
                            unsigned int *counts, hipStream_t s);
hipError_t launch_circus(const CircusParams &p, hipStream_t s);


// Batched path; `waves_per_block` in {1, 2, 4} and `blocks` (persistent grid)
// are the autotuned launch shape.
hipError_t launch_transpose_batched(const BatchImage *d_images, int nimg,
                                    int nmax, hipStream_t s);
hipError_t launch_reset_counter(unsigned int *counter, hipStream_t s);
hipError_t launch_sino_t05_batched(const BatchSinoParams &p, bool need_x,
                                   bool need_s, bool do345, int waves_per_block,
                                   int blocks, hipStream_t s);
hipError_t launch_circus_batched(const BatchCircusParams &p, hipStream_t s);
// P2's source rows (stable order) of n-row columns, n <= 2048: src[cols].
hipError_t launch_p2_source(const float *sino, int n, int cols, int32_t *src,
                            hipStream_t s);
// The same for every column of a batch into images[].p2src (nmax <= 2048).
hipError_t launch_p2_source_batched(const BatchCircusParams &p, hipStream_t s);

// Batched adjoint stages (same descriptors as the forward batch).
struct BatchAdjointParams {
    const BatchImage *images;
    int nimg;
    int A;
    int nt;
    int np;
    int nmax;
    int lines;            // total lines (sum of A * n)
    int pixels;           // total pixels (sum of n * n)
    int tiles;            // total gather tiles (of the launch's tile side)
    int cols;             // total sinogram columns (sum of nt * A)
    const float2 *trig;
    const float2 *tab345; // [3][tabn]
    int tabn;
    int do345;
    Slots slot;
    int code[kMaxP];
    int p3_done;          // dsino holds P3's part (the DFT GEMMs ran first)
};
hipError_t launch_circus_vjp_batched(const BatchAdjointParams &p, hipStream_t s);
hipError_t launch_records_batched(const BatchAdjointParams &p, hipStream_t s);
// Gather of a batch with tiles of side `tile` (16 or 32; images[].tile0 count
// such tiles) and `waves` waves per workgroup.
hipError_t launch_gather_batched(const BatchAdjointParams &p, int tile, int waves,
                                 hipStream_t s);

}  // namespace ttk
