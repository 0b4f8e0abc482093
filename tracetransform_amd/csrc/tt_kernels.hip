// tt_kernels.hip -- CDNA4 (gfx950) kernels of the trace transform.
//
// Built with -ffp-contract=off: every product and sum that feeds a
// weighted-median index is rounded exactly as in the reference's float code
// (no FMA), so median and sort indices are bit-identical to the reference's
// arithmetic.  Denormals are kept (no flush) and sqrtf is correctly rounded,
// as on the reference's x86-64 SSE2 build.  Sums that only feed values (the
// T1-T5 and P integrals) use explicit fma and, where it buys accuracy against
// the FP64 oracle, double accumulators.
#include "tt_internal.hpp"

#include <math.h>

namespace ttk {

namespace {

// 8-byte load from a 4-byte aligned address (unaligned dwordx2 is legal for
// global memory on CDNA).
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ f2u ld2(const float *p)
{
    return *reinterpret_cast<const f2u *>(p);
}

__device__ __forceinline__ int wave_id()
{
    return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Workgroups are dealt to the 8 XCDs round-robin; give each XCD a contiguous
// range of logical blocks so the lines one XCD works on share its L2.  The
// grid is padded to a multiple of 8 by the launcher.
__device__ __forceinline__ int xcd_block()
{
    const int g8 = gridDim.x / kNumXcd;
    const int b = blockIdx.x;
    return (b % kNumXcd) * g8 + b / kNumXcd;
}

// ---------------------------------------------------------------------------
// Bilinear sample of the rotated image at (row, col), c++/src/auxiliary.cpp
// :223-255 and :182-196.  x = px*c + py*s + o, y = px*(-s) + py*c + o, with
// pxc = px*c and pxns = px*(-s) precomputed per lane (each is one rounded
// product, as in the reference).  Strict bounds 0 <= x,y < n-1; the four
// taps are evaluated left to right with the reference's association.
// TR selects the row-major copy of the image so that lanes stepping along x
// read contiguous memory; the tap values (hence the result) are identical.
// ---------------------------------------------------------------------------
template <bool TR>
__device__ __forceinline__ float rot_sample(const float *__restrict__ src,
                                            int n, float lim, float o,
                                            float pxc, float pxns, float c,
                                            float s, float py)
{
    float x = pxc + py * s;
    float y = pxns + py * c;
    x = x + o;
    y = y + o;
    const bool inb = (x >= 0.0f) && (x < lim) && (y >= 0.0f) && (y < lim);
    const float xi = __builtin_truncf(x);
    const float yi = __builtin_truncf(y);
    const float fx = x - xi;
    const float fy = y - yi;
    const int X = inb ? (int)xi : 0;
    const int Y = inb ? (int)yi : 0;
    float s00, s01, s10, s11;
    if (!TR) {
        const f2u a = ld2(src + (X * n + Y));
        const f2u b = ld2(src + ((X + 1) * n + Y));
        s00 = a.x; s10 = a.y; s01 = b.x; s11 = b.y;
    } else {
        const f2u a = ld2(src + (Y * n + X));
        const f2u b = ld2(src + ((Y + 1) * n + X));
        s00 = a.x; s01 = a.y; s10 = b.x; s11 = b.y;
    }
    const float ofx = 1.0f - fx;
    const float ofy = 1.0f - fy;
    const float v = s00 * ofx * ofy + s01 * fx * ofy + s10 * ofx * fy +
                    s11 * fx * fy;
    return inb ? v : 0.0f;
}

// ---------------------------------------------------------------------------
// T0..T5 line kernel.  Lane = one column of the rotated image of one angle.
// Pass 1 streams the column and forms Eigen's packet-order sum (8 stride
// accumulators, c++/src/functionals.cpp:22 via Eigen Redux) of x and sqrt(x):
// S_x is T0 and the median threshold; S_sqrt the sqrt-median threshold.
// Pass 2 re-samples the column (the rotated image is never stored), walks
// the sequential float prefix to find findWeightedMedian's first index with
// 2*prefix >= S (c++/src/functionals.cpp:21-30), and from there accumulates
// T1/T2 (c++/src/functionals.cpp:104-129) and T3/T4/T5
// (c++/src/functionals.cpp:200-213) over the same rows as the reference.
// ---------------------------------------------------------------------------
template <bool TR, bool NEED_X, bool NEED_S, bool DO345>
__device__ __forceinline__ void sino_line(const SinoParams &p, const float *src,
                                          const float2 *s_tab, int a, int col,
                                          bool active, float c, float s)
{
    const int n = p.n;
    const float o = (float)((n - 1) / 2.0);
    const float lim = (float)(n - 1);
    const float px = (float)col - o;
    const float pxc = px * c;
    const float pxns = px * (-s);

    // ---- pass 1: Eigen-order sums -------------------------------------
    const int a8 = n & ~7;
    const int a4 = n & ~3;
    float rx[8], rs[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        rx[j] = 0.0f;
        rs[j] = 0.0f;
    }
    for (int r0 = 0; r0 < a8; r0 += 8) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float v = rot_sample<TR>(src, n, lim, o, pxc, pxns, c, s,
                                           (float)(r0 + j) - o);
            rx[j] += v;
            if (NEED_S)
                rs[j] += sqrtf(v);
        }
    }
    float Sx, Ss;
    if (a4 == 0) {
        Sx = 0.0f;
        Ss = 0.0f;
        for (int r = 0; r < n; r++) {
            const float v = rot_sample<TR>(src, n, lim, o, pxc, pxns, c, s,
                                           (float)r - o);
            Sx = r == 0 ? v : Sx + v;
            if (NEED_S)
                Ss = r == 0 ? sqrtf(v) : Ss + sqrtf(v);
        }
    } else {
        float ax[4], as[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            ax[j] = rx[j] + rx[j + 4];
            as[j] = rs[j] + rs[j + 4];
        }
        if (a4 > a8) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float v = rot_sample<TR>(src, n, lim, o, pxc, pxns, c, s,
                                               (float)(a8 + j) - o);
                const float sv = NEED_S ? sqrtf(v) : 0.0f;
                if (n >= 8) {
                    ax[j] += v;
                    as[j] += sv;
                } else {
                    ax[j] = v;
                    as[j] = sv;
                }
            }
        }
        Sx = (ax[0] + ax[2]) + (ax[1] + ax[3]);
        Ss = (as[0] + as[2]) + (as[1] + as[3]);
        for (int r = a4; r < n; r++) {
            const float v = rot_sample<TR>(src, n, lim, o, pxc, pxns, c, s,
                                           (float)r - o);
            Sx += v;
            if (NEED_S)
                Ss += sqrtf(v);
        }
    }

    const size_t line = (size_t)a * n + col;
    const size_t tstride = (size_t)p.A * n;
    if (active && p.slot[0] >= 0)
        p.sino[p.slot[0] * tstride + line] = Sx;
    if (!NEED_X && !NEED_S)
        return;

    // ---- pass 2: median walk + integrals -----------------------------
    // The walk is the reference's float prefix; the integrals only feed
    // values, so they accumulate in double (full rate on CDNA4) and land
    // within ~1e-7 of the FP64 oracle instead of inheriting the reference's
    // sequential float error.
    float Px = 0.0f, Ps = 0.0f;
    double t1 = 0.0, t2 = 0.0;
    double r3 = 0.0, i3 = 0.0, r4 = 0.0, i4 = 0.0, r5 = 0.0, i5 = 0.0;
    int m = -1, sm = -1;
    const bool do3 = p.slot[3] >= 0, do4 = p.slot[4] >= 0, do5 = p.slot[5] >= 0;
    for (int r = 0; r < n; r++) {
        const float v = rot_sample<TR>(src, n, lim, o, pxc, pxns, c, s,
                                       (float)r - o);
        const double vd = (double)v;
        if (NEED_X) {
            if (m < 0) {
                Px += v;
                if (2.0f * Px >= Sx)
                    m = r;
            }
            if (m >= 0) {
                const double rr = (double)(r - m);
                const double vr = vd * rr;
                t1 += vr;
                t2 = fma(vr, rr, t2);
            }
        }
        if (NEED_S) {
            if (sm < 0) {
                Ps += sqrtf(v);
                if (2.0f * Ps >= Ss)
                    sm = r;
            } else if (DO345) {
                const int r1 = r - sm;
                if (do3) {
                    const float2 w = s_tab[r1];
                    r3 = fma((double)w.x, vd, r3);
                    i3 = fma((double)w.y, vd, i3);
                }
                if (do4) {
                    const float2 w = s_tab[n + r1];
                    r4 = fma((double)w.x, vd, r4);
                    i4 = fma((double)w.y, vd, i4);
                }
                if (do5) {
                    const float2 w = s_tab[2 * n + r1];
                    r5 = fma((double)w.x, vd, r5);
                    i5 = fma((double)w.y, vd, i5);
                }
            }
        }
    }
    if (m < 0)
        m = n - 1;
    if (sm < 0)
        sm = n - 1;
    if (!active)
        return;
    if (NEED_X && p.med_x)
        p.med_x[line] = m;
    if (NEED_S && p.med_s)
        p.med_s[line] = sm;
    if (p.slot[1] >= 0)
        p.sino[p.slot[1] * tstride + line] = (float)t1;
    if (p.slot[2] >= 0)
        p.sino[p.slot[2] * tstride + line] = (float)t2;
    if (DO345) {
        if (do3)
            p.sino[p.slot[3] * tstride + line] = (float)hypot(r3, i3);
        if (do4)
            p.sino[p.slot[4] * tstride + line] = (float)hypot(r4, i4);
        if (do5)
            p.sino[p.slot[5] * tstride + line] = (float)hypot(r5, i5);
    }
}

template <bool NEED_X, bool NEED_S, bool DO345>
__global__ __launch_bounds__(kBlock) void sino_t05_kernel(SinoParams p)
{
    extern __shared__ float2 s_tab[];
    if (DO345) {
        for (int i = threadIdx.x; i < 3 * p.n; i += kBlock)
            s_tab[i] = p.tab345[i];
        __syncthreads();
    }
    const int item = xcd_block() * (kBlock / kWave) + wave_id();
    if (item >= p.items)
        return;
    const int a = item / p.nb;
    const int cb = item - a * p.nb;
    const int lane = threadIdx.x & (kWave - 1);
    int col = cb * kWave + lane;
    const bool active = col < p.n;
    if (!active)
        col = p.n - 1;
    const float2 cs = p.trig[a];
    // Lanes step (dx, dy) = (c, -s) per column: read the copy whose fast axis
    // is the one the lanes mostly move along.
    if (fabsf(cs.x) > fabsf(cs.y))
        sino_line<true, NEED_X, NEED_S, DO345>(p, p.imgT, s_tab, a, col, active,
                                              cs.x, cs.y);
    else
        sino_line<false, NEED_X, NEED_S, DO345>(p, p.img, s_tab, a, col, active,
                                               cs.x, cs.y);
}

// ---------------------------------------------------------------------------
// Helpers: transpose (for the row-major copy) and pad.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void transpose_kernel(const float *in,
                                                           float *out, int n)
{
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int col = bx + k, row = by + tx;
        if (col < n && row < n)
            tile[k][tx] = in[(size_t)col * n + row];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int row = by + k, col = bx + tx;
        if (col < n && row < n)
            out[(size_t)row * n + col] = tile[tx][k];
    }
}

__global__ __launch_bounds__(kBlock) void pad_kernel(const float *in, int rows,
                                                     int cols, float *out,
                                                     int n, int dfx, int dfy)
{
    const size_t total = (size_t)n * n;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const int R = (int)(i % n), Cc = (int)(i / n);
        const int r = R - dfy, c = Cc - dfx;
        float v = 0.0f;
        if (r >= 0 && r < rows && c >= 0 && c < cols)
            v = in[(size_t)c * rows + r];
        out[i] = v;
    }
}

// ---------------------------------------------------------------------------
// Block-wide helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ float block_sum(float v, float *s_red)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        s_red[w] = v;
    __syncthreads();
    float r = 0.0f;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; i++)
        r += s_red[i];
    return r;
}

__device__ __forceinline__ double block_sum_d(double v, double *s_red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        s_red[w] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; i++)
        r += s_red[i];
    return r;
}

// Eigen packet-order sum of x[0..n) held in LDS, evaluated by one thread.
__device__ float eigen_sum_lds(const float *x, int n)
{
    if (n <= 0)
        return 0.0f;
    const int a4 = n & ~3, a8 = n & ~7;
    if (a4 == 0) {
        float r = x[0];
        for (int i = 1; i < n; i++)
            r += x[i];
        return r;
    }
    float p0[4], p1[4];
    for (int j = 0; j < 4; j++)
        p0[j] = x[j];
    if (a4 > 4) {
        for (int j = 0; j < 4; j++)
            p1[j] = x[4 + j];
        for (int i = 8; i < a8; i += 8)
            for (int j = 0; j < 4; j++) {
                p0[j] += x[i + j];
                p1[j] += x[i + 4 + j];
            }
        for (int j = 0; j < 4; j++)
            p0[j] += p1[j];
        if (a4 > a8)
            for (int j = 0; j < 4; j++)
                p0[j] += x[a8 + j];
    }
    float r = (p0[0] + p0[2]) + (p0[1] + p0[3]);
    for (int i = a4; i < n; i++)
        r += x[i];
    return r;
}

// findWeightedMedian over x[0..n) in LDS (c++/src/functionals.cpp:21-30).
__device__ int weighted_median_lds(const float *x, int n, float S)
{
    float integral = 0.0f;
    for (int i = 0; i < n; i++) {
        integral += x[i];
        if (2.0f * integral >= S)
            return i;
    }
    return n - 1;
}

// Ascending bitonic sort of s[0..m) (m a power of two) by the whole block.
__device__ void bitonic_sort_lds(float *s, int m)
{
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += kBlock) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const bool up = (i & k) == 0;
                    const float a = s[i], b = s[ixj];
                    if ((a > b) == up) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// P-functionals, one workgroup per sinogram column (c++/src/circus.cpp
// :104-177): P1 (functionals.cpp:287-296), P2 (:303-311), P3 (:332-360),
// Hermite (:369-386).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void circus_kernel(CircusParams p)
{
    extern __shared__ float smem[];
    const int n = p.n;
    int m2 = 1;
    while (m2 < n)
        m2 <<= 1;
    float *s_x = smem;             // n
    float *s_sort = s_x + n;       // m2
    float *s_mod = s_sort + m2;    // n
    float *s_red = s_mod + n;      // 8
    __shared__ int s_rank;
    __shared__ double s_red_d[kBlock / kWave];

    const int colg = blockIdx.x;   // column over all sinograms
    const int sg = colg / p.ncols;
    const int col = colg - sg * p.ncols;
    const float *x = p.sino + (size_t)colg * n;
    for (int i = threadIdx.x; i < n; i += kBlock)
        s_x[i] = x[i];
    __syncthreads();

    for (int q = 0; q < p.np; q++) {
        const int code = p.code[q];
        float res = 0.0f;
        if (code == 1) {
            float acc = 0.0f;
            for (int i = 1 + threadIdx.x; i < n; i += kBlock)
                acc += fabsf(s_x[i - 1] - s_x[i]);
            res = block_sum(acc, s_red);
        } else if (code == 0) {
            const double *w = p.herm_w + (size_t)p.herm[q] * n;
            double acc = 0.0;
            for (int i = threadIdx.x; i < n; i += kBlock)
                acc = fma((double)s_x[i], w[i], acc);
            res = (float)block_sum_d(acc, s_red_d);
        } else if (code == 2) {
            for (int i = threadIdx.x; i < m2; i += kBlock)
                s_sort[i] = i < n ? s_x[i] : INFINITY;
            __syncthreads();
            bitonic_sort_lds(s_sort, m2);
            if (threadIdx.x == 0) {
                const float S = eigen_sum_lds(s_sort, n);
                s_rank = weighted_median_lds(s_sort, n, S);
            }
            __syncthreads();
            res = s_sort[s_rank];
            if (p.p2rank && threadIdx.x == 0)
                p.p2rank[((size_t)sg * p.np + q) * p.ncols + col] = s_rank;
        } else {
            // P3: half spectrum by direct DFT with a float twiddle table,
            // then pow(hypot(Re/N, Im/N), 4) in double and trapz.
            const int half = n / 2;
            const float fn = (float)n;
            for (int k = threadIdx.x; k <= half; k += kBlock) {
                float re = 0.0f, im = 0.0f;
                int idx = 0;
                for (int j = 0; j < n; j++) {
                    const float2 tw = p.dft_tw[idx];
                    re = fmaf(s_x[j], tw.x, re);
                    im = fmaf(-s_x[j], tw.y, im);
                    idx += k;
                    if (idx >= n)
                        idx -= n;
                }
                const double h = hypot((double)(re / fn), (double)(im / fn));
                const double h2 = h * h;
                s_mod[k] = (float)(h2 * h2);
            }
            __syncthreads();
            for (int k = half + 1 + threadIdx.x; k < n; k += kBlock)
                s_mod[k] = s_mod[n - k];
            __syncthreads();
            float acc = 0.0f;
            for (int i = threadIdx.x; i < n - 1; i += kBlock) {
                const float l0 = (float)(-1 + i * 2.0 / (n - 1));
                const float l1 = (float)(-1 + (i + 1) * 2.0 / (n - 1));
                acc += (l1 - l0) * (s_mod[i + 1] + s_mod[i]);
            }
            res = (float)(block_sum(acc, s_red) * 0.5);
        }
        if (threadIdx.x == 0)
            p.out[((size_t)sg * p.np + q) * p.ncols + col] = res;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Sort-based T-functionals T6 (c++/src/functionals.cpp:226-255) and T7
// (:262-276).  One workgroup per line, persistent over lines.  The line's
// column is sampled into LDS with lanes along rows; the medians m and sm come
// from the T0..T5 kernel, which ran first on the same stream.  Sorting is an
// LDS bitonic sort over 64-bit keys (orderable float bits << 32 | index), a
// total order equal to the stable order glibc's merge-sort qsort_r gives
// compareFloat; T7 sorts values only (equal keys are equal values).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ord_key(float f)
{
    if (f == 0.0f)
        f = 0.0f;  // -0 and +0 compare equal in compareFloat
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float ord_val(uint32_t k)
{
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(b);
}

template <class K>
__device__ void bitonic_lds(K *s, int m)
{
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += kBlock) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const bool up = (i & k) == 0;
                    const K a = s[i], b = s[ixj];
                    if ((a > b) == up) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Eigen packet-order sum of x[0..n) in LDS: threads 0..7 run the eight
// stride-8 chains, thread 0 combines (c++/src/functionals.cpp:22 via Eigen
// Redux).  Result broadcast through *s_out.  Must be called by the block.
__device__ void eigen_sum_block(const float *x, int n, float *s_chain, float *s_out)
{
    const int a4 = n & ~3, a8 = n & ~7;
    if (threadIdx.x < 8 && n >= 8) {
        float r = 0.0f;
        for (int i = threadIdx.x; i < a8; i += 8)
            r += x[i];
        s_chain[threadIdx.x] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float res;
        if (a4 == 0) {
            res = n > 0 ? x[0] : 0.0f;
            for (int i = 1; i < n; i++)
                res += x[i];
        } else {
            float ax[4];
            for (int j = 0; j < 4; j++)
                ax[j] = n >= 8 ? s_chain[j] + s_chain[j + 4] : x[j];
            if (n >= 8 && a4 > a8)
                for (int j = 0; j < 4; j++)
                    ax[j] += x[a8 + j];
            res = (ax[0] + ax[2]) + (ax[1] + ax[3]);
            for (int i = a4; i < n; i++)
                res += x[i];
        }
        *s_out = res;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void sino_sort_kernel(SortTParams p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = p.n;
    int m2 = 1;
    while (m2 < n)
        m2 <<= 1;
    uint64_t *s_key = reinterpret_cast<uint64_t *>(smem_raw);      // m2
    float *s_col = reinterpret_cast<float *>(s_key + m2);           // n
    float *s_sq = s_col + n;                                        // n
    __shared__ float s_chain[8];
    __shared__ float s_sum;
    __shared__ int s_k;

    const float o = (float)((n - 1) / 2.0);
    const float lim = (float)(n - 1);
    const size_t tstride = (size_t)p.A * n;
    const int lines = p.A * n;
    for (int line = blockIdx.x; line < lines; line += gridDim.x) {
        const int a = line / n;
        const int col = line - a * n;
        const float2 cs = p.trig[a];
        const float px = (float)col - o;
        const float pxc = px * cs.x, pxns = px * (-cs.y);
        // lanes step (dx, dy) = (s, c) per row
        if (fabsf(cs.y) > fabsf(cs.x)) {
            for (int r = threadIdx.x; r < n; r += kBlock)
                s_col[r] = rot_sample<true>(p.imgT, n, lim, o, pxc, pxns, cs.x,
                                            cs.y, (float)r - o);
        } else {
            for (int r = threadIdx.x; r < n; r += kBlock)
                s_col[r] = rot_sample<false>(p.img, n, lim, o, pxc, pxns, cs.x,
                                             cs.y, (float)r - o);
        }
        __syncthreads();

        if (p.slot6 >= 0) {
            const int sm = p.med_s[line];
            const int L = n - sm;
            int mL = 1;
            while (mL < L)
                mL <<= 1;
            for (int r = threadIdx.x; r < mL; r += kBlock) {
                uint64_t k = ~0ull;
                if (r < L) {
                    const float w = (float)r * s_col[sm + r];
                    k = ((uint64_t)ord_key(w) << 32) | (uint32_t)r;
                }
                s_key[r] = k;
            }
            __syncthreads();
            bitonic_lds(s_key, mL);
            for (int r = threadIdx.x; r < L; r += kBlock)
                s_sq[r] = sqrtf(s_col[sm + (int)(uint32_t)s_key[r]]);
            __syncthreads();
            eigen_sum_block(s_sq, L, s_chain, &s_sum);
            if (threadIdx.x == 0) {
                const float S = s_sum;
                float integral = 0.0f;
                int k = L - 1;
                for (int i = 0; i < L; i++) {
                    integral += s_sq[i];
                    if (2.0f * integral >= S) {
                        k = i;
                        break;
                    }
                }
                const int src = (int)(uint32_t)s_key[k];
                p.sino[p.slot6 * tstride + line] = (float)src * s_col[sm + src];
                if (p.t6_rank)
                    p.t6_rank[line] = k;
                if (p.t6_src)
                    p.t6_src[line] = src;
            }
            __syncthreads();
        }

        if (p.slot7 >= 0) {
            const int m = p.med_x[line];
            const int L = n - m;
            int mL = 1;
            while (mL < L)
                mL <<= 1;
            uint32_t *s_k32 = reinterpret_cast<uint32_t *>(s_key);
            for (int r = threadIdx.x; r < mL; r += kBlock)
                s_k32[r] = r < L ? ord_key(s_col[m + r]) : 0xFFFFFFFFu;
            __syncthreads();
            bitonic_lds(s_k32, mL);
            for (int r = threadIdx.x; r < L; r += kBlock)
                s_sq[r] = sqrtf(ord_val(s_k32[r]));
            __syncthreads();
            eigen_sum_block(s_sq, L, s_chain, &s_sum);
            if (threadIdx.x == 0) {
                const float S = s_sum;
                float integral = 0.0f;
                int k = L - 1;
                for (int i = 0; i < L; i++) {
                    integral += s_sq[i];
                    if (2.0f * integral >= S) {
                        k = i;
                        break;
                    }
                }
                p.sino[p.slot7 * tstride + line] = ord_val(s_k32[k]);
                if (p.t7_rank)
                    p.t7_rank[line] = k;
            }
            __syncthreads();
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_transpose(const float *in, float *out, int n, hipStream_t s)
{
    dim3 grid((n + 31) / 32, (n + 31) / 32);
    transpose_kernel<<<grid, kBlock, 0, s>>>(in, out, n);
    return hipGetLastError();
}

hipError_t launch_pad(const float *in, int rows, int cols, float *out, int n,
                      int dfx, int dfy, hipStream_t s)
{
    size_t total = (size_t)n * n;
    int blocks = (int)((total + kBlock - 1) / kBlock);
    if (blocks > 4096)
        blocks = 4096;
    pad_kernel<<<blocks, kBlock, 0, s>>>(in, rows, cols, out, n, dfx, dfy);
    return hipGetLastError();
}

hipError_t launch_sino_t05(const SinoParams &p, bool need_x, bool need_s,
                           bool do345, hipStream_t s)
{
    const int wpb = kBlock / kWave;
    int blocks = (p.items + wpb - 1) / wpb;
    blocks = (blocks + kNumXcd - 1) / kNumXcd * kNumXcd;
    const size_t shm = do345 ? sizeof(float2) * 3 * (size_t)p.n : 0;
#define TT_LAUNCH(X, S, D)                                                     \
    sino_t05_kernel<X, S, D><<<blocks, kBlock, shm, s>>>(p)
    if (need_x && need_s && do345)
        TT_LAUNCH(true, true, true);
    else if (need_x && need_s)
        TT_LAUNCH(true, true, false);
    else if (need_x)
        TT_LAUNCH(true, false, false);
    else if (need_s && do345)
        TT_LAUNCH(false, true, true);
    else if (need_s)
        TT_LAUNCH(false, true, false);
    else
        TT_LAUNCH(false, false, false);
#undef TT_LAUNCH
    return hipGetLastError();
}

hipError_t launch_circus(const CircusParams &p, hipStream_t s)
{
    int m2 = 1;
    while (m2 < p.n)
        m2 <<= 1;
    const size_t shm = sizeof(float) * ((size_t)p.n * 2 + m2 + 8);
    circus_kernel<<<p.ncols * p.nsino, kBlock, shm, s>>>(p);
    return hipGetLastError();
}

hipError_t launch_sino_sort(const SortTParams &p, hipStream_t s)
{
    int m2 = 1;
    while (m2 < p.n)
        m2 <<= 1;
    const size_t shm = sizeof(uint64_t) * (size_t)m2 + sizeof(float) * 2 * (size_t)p.n;
    const int lines = p.A * p.n;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int blocks = cus * 8;
    if (blocks > lines)
        blocks = lines;
    sino_sort_kernel<<<blocks, kBlock, shm, s>>>(p);
    return hipGetLastError();
}

}  // namespace ttk
