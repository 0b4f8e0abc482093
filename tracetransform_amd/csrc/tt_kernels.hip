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
#include <type_traits>

namespace ttk {

namespace {

// 8-byte load from a 4-byte aligned address (unaligned dwordx2 is legal for
// global memory on CDNA).
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));

// A pointer every lane of the wave holds, made visibly wave-uniform (its
// two halves read from the first lane), so that loads from it can take the
// scalar-base form.
__device__ __forceinline__ const float *uniform_ptr(const float *p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<const float *>(((uint64_t)hi << 32) | lo);
}

// Load at a 32-bit byte offset from a wave-uniform base in global memory:
// the scalar-base + vector-offset form of global loads (no 64-bit VALU
// math, no flat addressing).
__device__ __forceinline__ f2u ld2b(const float *base, uint32_t byte_off)
{
    typedef const char __attribute__((address_space(1))) *gbyte;
    typedef const float __attribute__((address_space(1))) *gfloat;
    const gfloat q = (gfloat)((gbyte)base + byte_off);
    return f2u{q[0], q[1]};
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
    // 0 <= x < lim as one unsigned compare of the bit patterns: x is never
    // -0 (a sum with o > 0 rounds an exact zero to +0), non-negative floats
    // order as their bits, and negative values and NaN compare above lim
    const bool inb = __float_as_uint(x) < __float_as_uint(lim) &&
                     __float_as_uint(y) < __float_as_uint(lim);
    // for 0 <= x the integer part is the truncating conversion and the
    // fraction x - floor(x) is exact (modf's split)
    const float fx = __builtin_amdgcn_fractf(x);
    const float fy = __builtin_amdgcn_fractf(y);
    // unsigned 32-bit offsets from a wave-uniform base: global loads in
    // scalar-base + vector-offset form, no 64-bit address arithmetic
    const uint32_t X = inb ? (uint32_t)x : 0u;
    const uint32_t Y = inb ? (uint32_t)y : 0u;
    const uint32_t un = (uint32_t)n;
    float s00, s01, s10, s11;
    // X, Y < n <= 2730, so X * n fits the 24-bit multiplier
    if (!TR) {
        const uint32_t o0 = (__umul24(X, un) + Y) * 4u;
        const f2u a = ld2b(src, o0);
        const f2u b = ld2b(src, o0 + 4u * un);
        s00 = a.x; s10 = a.y; s01 = b.x; s11 = b.y;
    } else {
        const uint32_t o0 = (__umul24(Y, un) + X) * 4u;
        const f2u a = ld2b(src, o0);
        const f2u b = ld2b(src, o0 + 4u * un);
        s00 = a.x; s01 = a.y; s10 = b.x; s11 = b.y;
    }
    const float ofx = 1.0f - fx;
    const float ofy = 1.0f - fy;
    const float v = s00 * ofx * ofy + s01 * fx * ofy + s10 * ofx * fy +
                    s11 * fx * fy;
    return inb ? v : 0.0f;
}

// Square root of a line sample, correctly rounded.  FS: v_sqrt_f32 and one
// residual test per neighbour (x - s' s against 0 picks the float nearest
// the true root), without the correctly rounded sequence's scaling of small
// inputs and its zero/infinity select.  Equal to that sequence on every
// float except x in (0, 2^-96) and the negative denormals (checked on the
// GPU over all 2^32 inputs); a bilinear sample falls there only if some
// pixel is negative or in (0, 2^-45) (each non-zero tap weight is a product
// of two factors >= 2^-25), which the image's bounding-box pass flags.
template <bool FS>
__device__ __forceinline__ float line_sqrt(float x)
{
    if constexpr (!FS) {
        return sqrtf(x);
    } else {
        const float s = __builtin_amdgcn_sqrtf(x);
        const float sd = __uint_as_float(__float_as_uint(s) - 1u);
        const float su = __uint_as_float(__float_as_uint(s) + 1u);
        const float r = fmaf(-sd, s, x) <= 0.0f ? sd : s;
        return fmaf(-su, s, x) > 0.0f ? su : r;
    }
}

// Test hook: line_sqrt<true> against the correctly rounded sqrtf on the
// float bit patterns [lo, lo + count), outside its documented exceptions
// ((0, 2^-96) and the negative denormals).  Counts the patterns whose
// results differ (NaN results compare equal) and keeps the smallest one.
__global__ __launch_bounds__(kBlock) void short_sqrt_check_kernel(uint32_t lo, uint64_t count,
                                                                 unsigned long long *mism,
                                                                 unsigned int *first)
{
    unsigned long long bad = 0;
    unsigned int firstbad = 0xFFFFFFFFu;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const uint32_t b = lo + (uint32_t)i;
        const bool tiny = b > 0u && b < 0x0F800000u;           // (0, 2^-96)
        const bool negden = b > 0x80000000u && b < 0x80800000u;  // negative denormals
        if (tiny || negden)
            continue;
        const float x = __uint_as_float(b);
        const float a = line_sqrt<true>(x), r = line_sqrt<false>(x);
        const bool same = __float_as_uint(a) == __float_as_uint(r) || (a != a && r != r);
        if (!same) {
            bad++;
            firstbad = b < firstbad ? b : firstbad;
        }
    }
    if (bad) {
        atomicAdd(mism, bad);
        atomicMin(first, firstbad);
    }
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
// Where one line's results go.  slot[t] is the output slot of T-code t
// (-1 if not requested); tabn is the row stride of the T3..T5 tables in LDS.
__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int m = kWave / 2; m > 0; m >>= 1) {
        const int o = __shfl_xor(v, m, kWave);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int m = kWave / 2; m > 0; m >>= 1) {
        const int o = __shfl_xor(v, m, kWave);
        v = o < v ? o : v;
    }
    return v;
}

// Rows r with base + (r - o) * slope in [lo_v, hi_v] (exact arithmetic),
// intersected into [*lo, *hi].
__device__ __forceinline__ void slab_rows(double base, double slope, double lo_v,
                                          double hi_v, double o, double *lo, double *hi)
{
    if (fabs(slope) < 1e-30) {
        if (base < lo_v || base > hi_v)
            *lo = 1.0, *hi = 0.0;
        return;
    }
    double r1 = (lo_v - base) / slope + o, r2 = (hi_v - base) / slope + o;
    if (r1 > r2) {
        const double t = r1;
        r1 = r2;
        r2 = t;
    }
    *lo = r1 > *lo ? r1 : *lo;
    *hi = r2 < *hi ? r2 : *hi;
}

// Rows of column col (angle with trig c, s) whose sample can be non-zero:
// the sample at row r reads taps (trunc(x), trunc(y)) + {0,1}^2 with
// x = px c + (r - o) s + o, y = -px s + (r - o) c + o.  Outside the rows
// where the exact x and y lie within 2 pixels of the bounding box of the
// non-zero pixels (float rounding of x, y is below 1e-3 here), every tap
// lies outside that box, so the sample is exactly 0.  Returns [lo, hi]
// (lo > hi when empty); no box means every row.
__device__ __forceinline__ int2 line_rows(const int *bbox, int n, int col, float c, float s)
{
    if (!bbox)
        return int2{0, n - 1};
    const int bx0 = bbox[0], bx1 = bbox[1], by0 = bbox[2], by1 = bbox[3];
    if (bx0 > bx1 || by0 > by1)
        return int2{n, -1};
    const double o = (n - 1) / 2.0, px = (double)col - o;
    double lo = 0.0, hi = (double)(n - 1);
    slab_rows(px * (double)c + o, (double)s, bx0 - 2.0, bx1 + 2.0, o, &lo, &hi);
    slab_rows(-px * (double)s + o, (double)c, by0 - 2.0, by1 + 2.0, o, &lo, &hi);
    if (!(lo <= hi))
        return int2{n, -1};
    return int2{(int)ceil(lo), (int)floor(hi)};
}

// Outputs in global memory, typed so (pointers from a batch descriptor are
// generic): their stores and loads are global, not flat, operations.
typedef float __attribute__((address_space(1))) *gf32;
typedef int32_t __attribute__((address_space(1))) *gi32;
typedef float f2v __attribute__((ext_vector_type(2)));
typedef f2v __attribute__((address_space(1))) *gf2;

struct LineDst {
    gf32 sino;
    gi32 med_x;
    gi32 med_s;
    gf2 z345;
    gf2 ckpt;  // walk checkpoints (with CK), see sino_line
    int n;
    int A;
    int tabn;
    Slots slot;

    __device__ LineDst(float *sino_, int32_t *med_x_, int32_t *med_s_, float2 *z345_,
                       float2 *ckpt_, int n_, int A_, int tabn_, Slots slot_)
        : sino((gf32)sino_), med_x((gi32)med_x_), med_s((gi32)med_s_), z345((gf2)z345_),
          ckpt((gf2)ckpt_), n(n_), A(A_), tabn(tabn_), slot(slot_)
    {
    }
};

// Row schedule of a wave.  At step v a lane handles row v - d of its line,
// d = rint(col t) + 64 K, t = -s/c for a line with |c| > |s| and c/s
// otherwise (|t| <= 1, from the line's own angle), K a wave-wide integer
// that makes the smallest d of the wave fall in [0, 64).  Lanes of one
// angle and consecutive columns then take samples along one row of the
// image copy the wave reads (TR: y constant, else x), a few cache lines per
// load, instead of spanning one image row per sample when the lanes all
// take the same row.  Each line still sees its rows in increasing order, so
// every sum, prefix and walk is the unskewed one, and d mod 64 depends on
// the line alone, so the 64-step blocks of pass 2 (v in [64 k, 64 k + 64))
// hold the same rows of a line whatever other lines share its wave: the
// results of a line do not depend on the tile that computes it.  Steps run
// from G, the smallest first row plus skew in the wave, in groups of 8: slot
// j of the stride-8 accumulators takes the rows of steps G + 8k + j, i.e.
// Eigen's accumulator (j + G - d) mod 8, and the slots are rotated back
// after the pass.  Rows outside [0, n) (and outside [0, a8) for the stride-8
// accumulators) are masked; a group all of whose lanes are inside skips the
// masks.
//
// With CK, pass 1 also runs the median walks' sequential prefixes (the
// same float additions in the same row order as the walk) and stores them
// at the start of every 64-step block from G after the first (at most
// sino_ckpt_blocks(n) of them).  When every sample
// a line has before a block is non-negative its prefixes never decrease, so
// a block whose starting prefix fails a lane's test 2 * prefix >= S holds
// none of that lane's earlier passes: pass 2 starts at the last block start
// where no lane has passed yet, from the stored prefixes, and samples only
// the steps from there on.  A negative or NaN sample (S of the square roots
// NaN, or a negative minimum) keeps the full walk.  [lo, hi] is the lane's
// row range that can hold non-zero samples (lo > hi: none).
template <bool TR, bool NEED_X, bool NEED_S, bool DO345, bool CK, bool FS>
__device__ __forceinline__ void sino_line(const LineDst &p, const float *src,
                                          const float2 *s_tab, int a, int col,
                                          bool active, float c, float s, int lo, int hi)
{
    const int n = p.n;
    src = uniform_ptr(src);
    const float o = (float)((n - 1) / 2.0);
    const float lim = (float)(n - 1);
    const float px = (float)col - o;
    const float pxc = px * c;
    const float pxns = px * (-s);
    const size_t line = (size_t)a * n + col;
    const size_t plane = (size_t)p.A * n;
    constexpr int kNone = 1 << 30;

    // ---- schedule --------------------------------------------------------
    int dl = (int)rintf((float)col * (fabsf(c) > fabsf(s) ? -s / c : c / s));
    dl -= wave_min_i(dl) & ~63;
    const int dmin = wave_min_i(dl), dmax = wave_max_i(dl);
    const int ncap = sino_ckpt_blocks(n);
    const int a8 = n & ~7;
    const int a4 = n & ~3;
    const int r0 = lo > 0 ? lo : 0;
    const int r1 = hi + 1 < n ? hi + 1 : n;
    const bool any = active && r1 > r0;  // the lanes past the last column follow it
    const int G = wave_min_i(any ? r0 + dl : kNone);
    const int V1 = wave_max_i(any ? (r1 < a8 ? r1 : a8) + dl : -kNone);
    const int V2 = wave_max_i(any ? r1 + dl : -kNone);

    // ---- pass 1: Eigen-order sums -------------------------------------
    float rx[8], rs[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        rx[j] = 0.0f;
        rs[j] = 0.0f;
    }
    constexpr bool WALK = CK && (NEED_X || NEED_S);
    float cx = 0.0f, cs = 0.0f, vmin = 0.0f;  // walk prefixes, smallest sample
    // one group of 8 steps from step v (MASK: some lane has a row outside
    // [0, a8)); row coordinates py = row - o are integers or half-integers
    // far below 2^23, so stepping them by whole rows is exact
    auto group = [&](auto mask_c, int v) {
        constexpr bool MASK = decltype(mask_c)::value;
        const int rb = v - dl;
        const float py0 = (float)rb - o;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            float x = rot_sample<TR>(src, n, lim, o, pxc, pxns, c, s, py0 + (float)j);
            bool chain = true;
            if constexpr (MASK) {
                x = (unsigned)(rb + j) < (unsigned)n ? x : 0.0f;
                chain = (unsigned)(rb + j) < (unsigned)a8;
            }
            const float sv = NEED_S ? line_sqrt<FS>(x) : 0.0f;
            rx[j] += chain ? x : 0.0f;
            if (NEED_S)
                rs[j] += chain ? sv : 0.0f;
            if (WALK && NEED_X)
                cx += x;
            if (WALK && NEED_S)
                cs += sv;
            if (WALK && !NEED_S)
                vmin = fminf(vmin, x);
        }
    };
    for (int v = G; v < V1; v += 8) {
        if (WALK && v > G && ((v - G) & 63) == 0 && active && (v - G) / 64 <= ncap)
            p.ckpt[(size_t)((v - G) / 64 - 1) * plane + line] = f2v{cx, cs};
        if (v >= dmax && v + 8 - dmin <= a8)
            group(std::false_type{}, v);
        else
            group(std::true_type{}, v);
    }
    // slot j holds accumulator (j + d) mod 8: rotate it back
    const int d = (G - dl) & 7;
#pragma unroll
    for (int b = 1; b < 8; b <<= 1) {
        const bool on = (d & b) != 0;
        float tx[8], ts[8];
#pragma unroll
        for (int m = 0; m < 8; m++) {
            tx[m] = on ? rx[(m - b) & 7] : rx[m];
            ts[m] = on ? rs[(m - b) & 7] : rs[m];
        }
#pragma unroll
        for (int m = 0; m < 8; m++) {
            rx[m] = tx[m];
            rs[m] = ts[m];
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
                Ss = r == 0 ? line_sqrt<FS>(v) : Ss + line_sqrt<FS>(v);
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
                const float sv = NEED_S ? line_sqrt<FS>(v) : 0.0f;
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
                Ss += line_sqrt<FS>(v);
        }
    }

    const size_t tstride = plane;
    if (active && p.slot.s[0] >= 0)
        p.sino[p.slot.s[0] * tstride + line] = Sx;
    if (!NEED_X && !NEED_S)
        return;

    // ---- where pass 2 starts: the last checkpoint no lane passes --------
    int start = G;
    float Px = 0.0f, Ps = 0.0f;
    if constexpr (WALK) {
        // NaN sums flag negative (square root NaN) or NaN samples
        const bool bad = (NEED_S && Ss != Ss) || Sx != Sx || vmin < 0.0f;
        int nck = V1 > G ? (V1 - G - 1) / 64 : 0;
        nck = nck < ncap ? nck : ncap;
        int b0 = 0;
        float2 cb = {0.0f, 0.0f};
        if (nck > 0 && !__ballot(active && bad)) {
            for (int b = 1; b <= nck; b += 4) {
                float2 ck[4];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (b + u <= nck)
                    {
                        const f2v cv = p.ckpt[(size_t)(b + u - 1) * plane + line];
                        ck[u] = float2{cv.x, cv.y};
                    }
                bool stop = false;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (stop || b + u > nck)
                        break;
                    const bool pass = (NEED_X && 2.0f * ck[u].x >= Sx) ||
                                      (NEED_S && 2.0f * ck[u].y >= Ss);
                    if (__ballot(active && pass)) {
                        stop = true;
                    } else {
                        b0 = b + u;
                        cb = ck[u];
                    }
                }
                if (stop)
                    break;
            }
        }
        if (b0 > 0) {
            start = G + 64 * b0;
            Px = cb.x;
            Ps = cb.y;
        }
    }

    // ---- pass 2: median walk + integrals -----------------------------
    // The walk is the reference's float prefix.  The integrals only feed
    // values: they accumulate with float FMAs over blocks of 64 steps, each
    // block flushed into a double, so they land within ~1e-7 of the FP64
    // oracle instead of inheriting the reference's sequential float error.
    // The blocks are the 64-step grid v in [64 k, 64 k + 64) whatever step
    // pass 2 starts at (every integral starts at or after it), so the
    // integrals depend neither on the checkpoints nor on the wave's other
    // lines.
    double t1 = 0.0, t2 = 0.0;
    double r3 = 0.0, i3 = 0.0, r4 = 0.0, i4 = 0.0, r5 = 0.0, i5 = 0.0;
    int m = -1, sm = -1;
    const bool do3 = p.slot.s[3] >= 0, do4 = p.slot.s[4] >= 0, do5 = p.slot.s[5] >= 0;
    if (lo > 0 || lo > hi) {
        // the walk's first row is a zero: it passes there iff S <= 0
        if (NEED_X && 2.0f * 0.0f >= Sx)
            m = 0;
        if (NEED_S && 2.0f * 0.0f >= Ss)
            sm = 0;
    }
    for (int vb = start, ve; vb < V2; vb = ve) {
        ve = (vb & ~63) + 64;
        ve = ve < V2 ? ve : V2;
        // every lane's rows of the block inside [0, n)
        const bool full = vb >= dmax && ve - dmin <= n;
        float f1 = 0.0f, f2 = 0.0f;
        float f3r = 0.0f, f3i = 0.0f, f4r = 0.0f, f4i = 0.0f, f5r = 0.0f, f5i = 0.0f;
        int r = vb - dl;
        float py = (float)r - o;
        for (int v = vb; v < ve; v++, r++, py += 1.0f) {
            if (!full && (unsigned)r >= (unsigned)n)
                continue;
            const float x = rot_sample<TR>(src, n, lim, o, pxc, pxns, c, s, py);
            if (NEED_X) {
                if (m < 0) {
                    Px += x;
                    if (2.0f * Px >= Sx)
                        m = r;
                }
                if (m >= 0) {
                    const float rr = (float)(r - m);
                    const float vr = x * rr;
                    f1 += vr;
                    f2 = fmaf(vr, rr, f2);
                }
            }
            if (NEED_S) {
                if (sm < 0) {
                    Ps += line_sqrt<FS>(x);
                    if (2.0f * Ps >= Ss)
                        sm = r;
                } else if (DO345) {
                    // one 16-byte and one 8-byte LDS read per row for all three
                    // tables (rows interleaved by the kernel prologue)
                    const int q = r - sm;
                    const float4 w34 = reinterpret_cast<const float4 *>(s_tab)[q];
                    const float2 w5 = s_tab[2 * p.tabn + q];
                    if (do3) {
                        f3r = fmaf(w34.x, x, f3r);
                        f3i = fmaf(w34.y, x, f3i);
                    }
                    if (do4) {
                        f4r = fmaf(w34.z, x, f4r);
                        f4i = fmaf(w34.w, x, f4i);
                    }
                    if (do5) {
                        f5r = fmaf(w5.x, x, f5r);
                        f5i = fmaf(w5.y, x, f5i);
                    }
                }
            }
        }
        t1 += f1;
        t2 += f2;
        r3 += f3r;
        i3 += f3i;
        r4 += f4r;
        i4 += f4i;
        r5 += f5r;
        i5 += f5i;
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
    if (p.slot.s[1] >= 0)
        p.sino[p.slot.s[1] * tstride + line] = (float)t1;
    if (p.slot.s[2] >= 0)
        p.sino[p.slot.s[2] * tstride + line] = (float)t2;
    if (DO345) {
        if (do3)
            p.sino[p.slot.s[3] * tstride + line] = (float)hypot(r3, i3);
        if (do4)
            p.sino[p.slot.s[4] * tstride + line] = (float)hypot(r4, i4);
        if (do5)
            p.sino[p.slot.s[5] * tstride + line] = (float)hypot(r5, i5);
        if (p.z345) {
            const size_t plane = (size_t)p.A * n;
            p.z345[line] = f2v{(float)r3, (float)i3};
            p.z345[plane + line] = f2v{(float)r4, (float)i4};
            p.z345[2 * plane + line] = f2v{(float)r5, (float)i5};
        }
    }
}

// One line, with the image copy whose fast axis the lanes step along (TR)
// and the square root the image's bounding-box flag allows (FS).
template <bool NEED_X, bool NEED_S, bool DO345, bool CK>
__device__ __forceinline__ void sino_line_any(const LineDst &dst, const float *img,
                                              const float *imgT, const int *bbox,
                                              const float2 *s_tab, int a, int col,
                                              bool active, float2 cs, int lo, int hi)
{
    // Lanes step (dx, dy) = (c, -s) per column; one copy for the whole wave
    const bool tr = __builtin_amdgcn_readfirstlane(fabsf(cs.x) > fabsf(cs.y));
    const bool fs = bbox && bbox[4] == 0;
    if (tr && fs)
        sino_line<true, NEED_X, NEED_S, DO345, CK, true>(dst, imgT, s_tab, a, col, active,
                                                         cs.x, cs.y, lo, hi);
    else if (tr)
        sino_line<true, NEED_X, NEED_S, DO345, CK, false>(dst, imgT, s_tab, a, col, active,
                                                          cs.x, cs.y, lo, hi);
    else if (fs)
        sino_line<false, NEED_X, NEED_S, DO345, CK, true>(dst, img, s_tab, a, col, active,
                                                          cs.x, cs.y, lo, hi);
    else
        sino_line<false, NEED_X, NEED_S, DO345, CK, false>(dst, img, s_tab, a, col, active,
                                                           cs.x, cs.y, lo, hi);
}

// LDS copy of the T3..T5 tables ([3][tabn] float2 in global memory) with the
// T3 and T4 entries of a row interleaved as one float4 (rows 0..tabn-1) and
// T5 after them (float2 at 2*tabn + r).
__device__ __forceinline__ void stage_tab345(float2 *s_tab, const float2 *g, int tabn,
                                             int nthreads)
{
    float4 *s4 = reinterpret_cast<float4 *>(s_tab);
    for (int r = threadIdx.x; r < tabn; r += nthreads) {
        const float2 a = g[r], b = g[tabn + r];
        s4[r] = float4{a.x, a.y, b.x, b.y};
        s_tab[2 * tabn + r] = g[2 * tabn + r];
    }
}

template <bool NEED_X, bool NEED_S, bool DO345>
__global__ __launch_bounds__(kBlock) void sino_t05_kernel(SinoParams p)
{
    extern __shared__ __align__(16) float2 s_tab[];
    if (DO345) {
        stage_tab345(s_tab, p.tab345, p.n, kBlock);
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
    const LineDst dst = {p.sino, p.med_x, p.med_s, p.z345, p.ckpt, p.n, p.A, p.n, p.slot};
    const int2 rows = line_rows(p.bbox, p.n, col, cs.x, cs.y);
    sino_line_any<NEED_X, NEED_S, DO345, true>(dst, p.img, p.imgT, p.bbox, s_tab, a, col,
                                               active, cs, rows.x, rows.y);
}

// ---------------------------------------------------------------------------
// Batched T0..T5: one persistent launch over every (image, angle tile,
// column tile) wave item of a batch of padded images of uneven side.  A
// tile is tile_a consecutive angles by tile_c consecutive columns (64, 32
// or 16): the wave takes it in tile_a tile_c / 64 passes, in each pass lane
// l on column l % tile_c of angle l / tile_c of the pass.  Waves take items
// from a device counter (dynamic load balance; items are ordered largest
// image first) and find their image by binary search over the descriptors'
// first-item offsets.  The counter is zeroed by reset_counter_kernel ahead
// of each launch.  The image copy a pass reads (row- or column-major) is
// the one that suits the angle of its first lane; the results of a line do
// not depend on the tile.
// ---------------------------------------------------------------------------
template <bool NEED_X, bool NEED_S, bool DO345>
__global__ __launch_bounds__(kBlock) void sino_t05_batched_kernel(BatchSinoParams p)
{
    extern __shared__ __align__(16) float2 s_tab[];
    if (DO345) {
        stage_tab345(s_tab, p.tab345, p.tabn, blockDim.x);
        __syncthreads();
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int sub = kWave / p.tile_c;  // angles per pass
    const int lc = lane & (p.tile_c - 1), la = lane / p.tile_c;
    // one pass per iteration (an item's passes in a row): no loop nested
    // inside the item loop, whose invariants would stay live across the
    // line code
    int item = 0, pass = 0, passes = 0;
    for (;;) {
        if (pass == passes) {
            unsigned int v = 0;
            if (lane == 0)
                v = atomicAdd(p.counter, 1u);
            item = __builtin_amdgcn_readfirstlane(__shfl(v, 0, kWave));
            if (item >= p.items)
                break;
            pass = 0;
        }
        int lo = 0, hi = p.nimg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (p.images[mid].item0 <= item)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchImage im = p.images[lo];
        const int local = item - im.item0;
        const int at = local / im.nb;
        const int cb = local - at * im.nb;
        const int a_first = at * p.tile_a;
        const int a_end = a_first + p.tile_a < p.A ? a_first + p.tile_a : p.A;
        passes = (a_end - a_first + sub - 1) / sub;
        int a = a_first + pass * sub + la;
        pass++;
        int col = cb * p.tile_c + lc;
        const bool active = col < im.n && a < a_end;
        col = col < im.n ? col : im.n - 1;
        a = a < a_end ? a : a_end - 1;
        const float2 cs = p.trig[a];
        const LineDst dst = {im.sino, im.med_x, im.med_s, im.z345, im.ckpt, im.n, p.A, p.tabn,
                             p.slot};
        const int2 rows = line_rows(im.bbox, im.n, col, cs.x, cs.y);
        sino_line_any<NEED_X, NEED_S, DO345, true>(dst, im.img, im.imgT, im.bbox, s_tab, a,
                                                   col, active, cs, rows.x, rows.y);
    }
}

__global__ void reset_counter_kernel(unsigned int *c)
{
    if (threadIdx.x == 0)
        c[0] = 0u;
}

__global__ void reset_counters_kernel(unsigned int *c, int count)
{
    if ((int)threadIdx.x < count)
        c[threadIdx.x] = 0u;
}

// Row-major copies of every image of a batch: grid.y walks the images.
__global__ __launch_bounds__(kBlock) void transpose_batched_kernel(
    const BatchImage *images)
{
    __shared__ float tile[32][33];
    const BatchImage im = images[blockIdx.y];
    const int n = im.n;
    const int tiles = (n + 31) / 32;
    float *out = const_cast<float *>(im.imgT);
    for (int t = blockIdx.x; t < tiles * tiles; t += gridDim.x) {
        const int bx = (t % tiles) * 32, by = (t / tiles) * 32;
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
        for (int k = ty; k < 32; k += 8) {
            const int col = bx + k, row = by + tx;
            if (col < n && row < n)
                tile[k][tx] = im.img[(size_t)col * n + row];
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            const int row = by + k, col = bx + tx;
            if (col < n && row < n)
                out[(size_t)row * n + col] = tile[tx][k];
        }
        __syncthreads();
    }
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

// findWeightedMedian's walk over x[0..n) in LDS (c++/src/functionals.cpp
// :21-30): first i with 2 * (sequential float prefix) >= S.  One thread;
// x must be 16-byte aligned so the rows come in four at a time.
__device__ int weighted_median_walk(const float *x, int n, float S)
{
    float integral = 0.0f;
    int i = 0;
    for (; i + 4 <= n; i += 4) {
        const float4 v = *reinterpret_cast<const float4 *>(x + i);
        integral += v.x;
        if (2.0f * integral >= S)
            return i;
        integral += v.y;
        if (2.0f * integral >= S)
            return i + 1;
        integral += v.z;
        if (2.0f * integral >= S)
            return i + 2;
        integral += v.w;
        if (2.0f * integral >= S)
            return i + 3;
    }
    for (; i < n; i++) {
        integral += x[i];
        if (2.0f * integral >= S)
            return i;
    }
    return n - 1;
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
// LDS layout of one circus workgroup for columns of height up to nmax.
struct CircusLds {
    float *x;      // nmax4
    float *sort;   // pow2 >= nmax
    float *mod;    // nmax4
};

__device__ __forceinline__ CircusLds circus_lds(float *smem, int nmax)
{
    const int n4 = (nmax + 3) & ~3;
    int m2 = 1;
    while (m2 < nmax)
        m2 <<= 1;
    CircusLds l;
    l.x = smem;
    l.sort = l.x + n4;
    l.mod = l.sort + m2;
    return l;
}

// All P-functionals of one sinogram column x[0..n) (already in L.x; P3 reads
// the column's half spectrum spec: real parts, then imaginary parts from
// spec_h on).  Result q goes to out[q * out_stride].
__device__ void circus_column(const CircusLds &L, int n, int np, const int *code,
                              const int *herm, const double *herm_w, float *out,
                              size_t out_stride, int32_t *p2rank,
                              const float *spec, int spec_h, bool skip_p2)
{
    __shared__ float s_red[kBlock / kWave];
    __shared__ double s_red_d[kBlock / kWave];
    __shared__ float s_chain[8];
    __shared__ float s_sum;
    __shared__ int s_rank;
    int m2 = 1;
    while (m2 < n)
        m2 <<= 1;
    for (int q = 0; q < np; q++) {
        const int cd = code[q];
        if (cd == 2 && skip_p2)
            continue;  // written by circus_p2_kernel
        float res = 0.0f;
        if (cd == 1) {
            float acc = 0.0f;
            for (int i = 1 + threadIdx.x; i < n; i += kBlock)
                acc += fabsf(L.x[i - 1] - L.x[i]);
            res = block_sum(acc, s_red);
        } else if (cd == 0) {
            const double *w = herm_w + (size_t)herm[q] * n;
            double acc = 0.0;
            for (int i = threadIdx.x; i < n; i += kBlock)
                acc = fma((double)L.x[i], w[i], acc);
            res = (float)block_sum_d(acc, s_red_d);
        } else if (cd == 2) {
            for (int i = threadIdx.x; i < m2; i += kBlock)
                L.sort[i] = i < n ? L.x[i] : INFINITY;
            __syncthreads();
            bitonic_sort_lds(L.sort, m2);
            eigen_sum_block(L.sort, n, s_chain, &s_sum);
            if (threadIdx.x == 0)
                s_rank = weighted_median_walk(L.sort, n, s_sum);
            __syncthreads();
            res = L.sort[s_rank];
            if (p2rank && threadIdx.x == 0)
                p2rank[q * out_stride] = s_rank;
        } else {
            // P3 from the column's half spectrum (tt_dft.hip):
            // pow(hypot(Re/N, Im/N), 4) in double, mirrored onto k > N/2,
            // and the trapz sum over the reference's float abscissae
            // lin[p] = -1 + 2p/(N-1) (c++/src/functionals.cpp:53-60, 332-360)
            const int half = n / 2;
            for (int k = threadIdx.x; k <= half; k += kBlock) {
                const double hy = hypot((double)spec[k] / n, (double)spec[spec_h + k] / n);
                const double h2 = hy * hy;
                L.mod[k] = (float)(h2 * h2);
            }
            __syncthreads();
            for (int k = half + 1 + threadIdx.x; k < n; k += kBlock)
                L.mod[k] = L.mod[n - k];
            __syncthreads();
            float acc = 0.0f;
            for (int i = threadIdx.x; i < n - 1; i += kBlock) {
                const float l0 = (float)(-1 + i * 2.0 / (n - 1));
                const float l1 = (float)(-1 + (i + 1) * 2.0 / (n - 1));
                acc += (l1 - l0) * (L.mod[i + 1] + L.mod[i]);
            }
            res = (float)(block_sum(acc, s_red) * 0.5);
        }
        if (threadIdx.x == 0)
            out[q * out_stride] = res;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void circus_kernel(CircusParams p)
{
    extern __shared__ __align__(16) float smem[];
    const int n = p.n;
    const CircusLds L = circus_lds(smem, n);
    const int colg = blockIdx.x;   // column over all sinograms
    const int sg = colg / p.ncols;
    const int col = colg - sg * p.ncols;
    const float *x = p.sino + (size_t)colg * n;
    for (int i = threadIdx.x; i < n; i += kBlock)
        L.x[i] = x[i];
    __syncthreads();
    const size_t o = (size_t)sg * p.np * p.ncols + col;
    circus_column(L, n, p.np, p.code, p.herm, p.herm_w, p.out + o, p.ncols,
                  p.p2rank ? p.p2rank + o : nullptr,
                  p.p3_spec ? p.p3_spec + (size_t)colg * 2 * p.p3_h : nullptr, p.p3_h,
                  p.p2_wave != 0);
}

// Every sinogram column of a batch; the image is found by binary search over
// the descriptors' first-column offsets.  P3 reads the image's half spectra
// (im.spec, formed by the DFT GEMM as in the single-image path).
__global__ __launch_bounds__(kBlock) void circus_batched_kernel(BatchCircusParams p)
{
    extern __shared__ __align__(16) float smem[];
    const CircusLds L = circus_lds(smem, p.nmax);
    const int colg = blockIdx.x;
    int lo = 0, hi = p.nimg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.images[mid].col0 <= colg)
            lo = mid;
        else
            hi = mid - 1;
    }
    const BatchImage im = p.images[lo];
    const int n = im.n;
    const int lc = colg - im.col0;
    const int t = lc / p.A;
    const int a = lc - t * p.A;
    const float *x = im.sino + ((size_t)t * p.A + a) * n;
    for (int i = threadIdx.x; i < n; i += kBlock)
        L.x[i] = x[i];
    __syncthreads();
    const int herm_none[kMaxP] = {};
    const int h = n / 2 + 1;
    circus_column(L, n, p.np, p.code, herm_none, nullptr,
                  im.circus + (size_t)t * p.np * p.A + a, p.A, nullptr,
                  im.spec ? im.spec + (size_t)lc * 2 * h : nullptr, h, p.p2_wave != 0);
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

// 64-bit broadcast from one lane.
// Values of one lane given a wave-uniform lane index (v_readlane: no LDS
// round trip, unlike a shuffle).
__device__ __forceinline__ int lane_int(int v, int l)
{
    return __builtin_amdgcn_readlane(v, l);
}

__device__ __forceinline__ float lane_float(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int l)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// 64-bit shuffle (two 32-bit halves).
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m)
{
    const int lo = __shfl_xor((int)(uint32_t)v, m, kWave);
    const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m, kWave);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

__device__ __forceinline__ uint32_t shfl_xor_key(uint32_t v, int m)
{
    return (uint32_t)__shfl_xor((int)v, m, kWave);
}

__device__ __forceinline__ uint64_t shfl_xor_key(uint64_t v, int m)
{
    return shfl_xor64(v, m);
}

// Ascending bitonic sort of 64*KPL keys (uint32_t or uint64_t) held by one
// wave in registers, blocked layout: lane l holds positions l*KPL ..
// l*KPL+KPL-1.  Merge sizes up to KPL stay inside a lane and are unrolled
// with their directions known at compile time; for larger sizes the
// direction depends on the lane only, strides of KPL and more exchange with
// lane l ^ (stride / KPL) through shuffles.  Wave-synchronous: no barriers.
// (DPP forms of the exchanges were measured slower here: they raise the
// VGPR count and cost occupancy.)
template <int KPL, class Key>
__device__ __forceinline__ void wave_bitonic_sort(Key (&k)[KPL], int lane)
{
    constexpr int M = kWave * KPL;
    const bool lane_up = (lane & 1) == 0;  // direction of the size-KPL merge
#pragma unroll
    for (int size = 2; size <= KPL; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int r = 0; r < KPL; r++) {
                const int r2 = r ^ stride;
                if (r2 > r) {
                    const bool up = size == KPL ? lane_up : (r & size) == 0;
                    const Key a = k[r], b = k[r2];
                    const bool sw = (a > b) == up;
                    k[r] = sw ? b : a;
                    k[r2] = sw ? a : b;
                }
            }
        }
    }
    for (int size = 2 * KPL; size <= M; size <<= 1) {
        const bool up = ((lane * KPL) & size) == 0;
        for (int stride = size >> 1; stride >= KPL; stride >>= 1) {
            const int lx = stride / KPL;
            // the lower lane of each pair keeps the minimum when ascending
            const bool keep_min = ((lane & lx) == 0) == up;
#pragma unroll
            for (int r = 0; r < KPL; r++) {
                const Key o = shfl_xor_key(k[r], lx);
                k[r] = ((k[r] < o) == keep_min) ? k[r] : o;
            }
        }
#pragma unroll
        for (int stride = KPL / 2; stride > 0; stride >>= 1) {
#pragma unroll
            for (int r = 0; r < KPL; r++) {
                const int r2 = r ^ stride;
                if (r2 > r) {
                    const Key a = k[r], b = k[r2];
                    const bool sw = (a > b) == up;
                    k[r] = sw ? b : a;
                    k[r2] = sw ? a : b;
                }
            }
        }
    }
}

// Wave-wide reductions.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v)
{
#pragma unroll
    for (int m = kWave / 2; m > 0; m >>= 1) {
        const uint64_t o = shfl_xor64(v, m);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The sorted square roots of a tail of length L as a compact LDS array c of
// its Le non-trailing entries: position j of the full sequence holds c[j]
// for j < ins, 0 for ins <= j < ins + zt, and c[j - zt] beyond.
struct TailView {
    const float *c;
    int ins, zt, L;
    __device__ __forceinline__ float at(int j) const
    {
        return j < ins ? c[j] : (j < ins + zt ? 0.0f : c[j - zt]);
    }
};

// Sequential float sum of c[i0], c[i0 + 8], ... (i < i1), added to acc.
__device__ __forceinline__ float chain_sum(const float *c, int i0, int i1, float acc)
{
    int i = i0;
    for (; i + 24 < i1; i += 32) {
        const float v0 = c[i], v1 = c[i + 8], v2 = c[i + 16], v3 = c[i + 24];
        acc += v0;
        acc += v1;
        acc += v2;
        acc += v3;
    }
    for (; i < i1; i += 8)
        acc += c[i];
    return acc;
}

// Eigen packet-order sum of a TailView's full sequence (as eigen_sum_block),
// for G = 1 or 2 views at once: lanes 8g .. 8g + 7 run the stride-8 chains
// of the view they pass (group g), lane 8g combines them; the sum of group g
// is returned on lane 8g (other lanes: unspecified).  Adding the zero run's +0 leaves a chain of non-negative
// values unchanged, so chain r sums the compact entries before the run
// (positions r, r + 8, ... below ins) and after it (positions congruent to r
// from ins + zt on, entry = position - zt) and skips the run.
template <int G>
__device__ __forceinline__ float eigen_sum_group(const TailView &v, int lane)
{
    const bool active = lane < 8 * G;
    const int n = v.L;
    const int a4 = n & ~3, a8 = n & ~7;
    const int ch_i = lane & 7;
    float r = 0.0f;
    if (active && n >= 8) {
        const int e1 = v.ins < a8 ? v.ins : a8;
        r = chain_sum(v.c, ch_i, e1, r);
        const int q0 = v.ins + v.zt;
        const int p0 = q0 + ((ch_i - q0) & 7);  // first position >= q0 in chain ch_i
        if (p0 < a8)
            r = chain_sum(v.c, p0 - v.zt, a8 - v.zt, r);
    }
    float ch[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        ch[j] = lane_float(r, j);
        if (G == 2 && lane >= 8)
            ch[j] = lane_float(r, 8 + j);
    }
    float res = 0.0f;
    if (active && ch_i == 0) {
        if (a4 == 0) {
            res = n > 0 ? v.at(0) : 0.0f;
            for (int i = 1; i < n; i++)
                res += v.at(i);
        } else {
            float ax[4];
            for (int j = 0; j < 4; j++)
                ax[j] = n >= 8 ? ch[j] + ch[j + 4] : v.at(j);
            if (n >= 8 && a4 > a8)
                for (int j = 0; j < 4; j++)
                    ax[j] += v.at(a8 + j);
            res = (ax[0] + ax[2]) + (ax[1] + ax[3]);
            for (int i = a4; i < n; i++)
                res += v.at(i);
        }
    }
    return res;
}

// The same for one view (lanes 0..7), the sum returned to every lane.
__device__ __forceinline__ float eigen_sum_view(const TailView &v, int lane)
{
    return lane_float(eigen_sum_group<1>(v, lane), 0);
}

// findWeightedMedian's sequential walk over a TailView (one lane).  Adding
// the zero run leaves the prefix unchanged, so the test can first pass there
// only at its first position (when nothing before it passed, that is only
// possible with ins == 0 and S <= 0).
// Walks c[j0..j1); returns the first passing j + off, or -1 with the prefix
// carried in *integral.  A block of 8 is summed first (the same additions in
// the same order as the element walk, so the same prefixes); when none of its
// values is negative or NaN the prefix does not decrease inside it, so a
// block whose last prefix fails the test holds no passing position and is
// skipped.  Any other block is walked element by element.  NONNEG: every c
// is known to be >= 0 (square roots of non-negative samples), which drops
// the per-block sign test.
template <bool NONNEG>
__device__ __forceinline__ int walk_segment(const float *c, int j0, int j1, int off, float S,
                                            float *integral)
{
    float acc = *integral;
    int j = j0;
    for (; j + 8 <= j1; j += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            v[u] = c[j + u];
        float t = acc;
        bool mono = true;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            t += v[u];
            if (!NONNEG)
                mono = mono & (v[u] >= 0.0f);
        }
        if (mono && !(2.0f * t >= S)) {
            acc = t;
            continue;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            acc += v[u];
            if (2.0f * acc >= S)
                return j + u + off;
        }
    }
    for (; j < j1; j++) {
        acc += c[j];
        if (2.0f * acc >= S)
            return j + off;
    }
    *integral = acc;
    return -1;
}

template <bool NONNEG>
__device__ __forceinline__ int weighted_median_walk_view(const TailView &v, int Le, float S)
{
    float integral = 0.0f;
    int k = walk_segment<NONNEG>(v.c, 0, v.ins, 0, S, &integral);
    if (k >= 0)
        return k;
    if (v.zt > 0 && 2.0f * integral >= S)
        return v.ins;
    k = walk_segment<NONNEG>(v.c, v.ins, Le, v.zt, S, &integral);
    return k >= 0 ? k : v.L - 1;
}

// Sorted tail of one line for T6 (T6 = true) or T7, given the tail's first
// Le samples in xs[0..Le), no sample negative or NaN, and every sample at
// i >= Le zero (the line has left the rotated image).  Only the first Le keys
// are sorted (K keys per lane, 64*K >= Le); the L - Le trailing zeros are
// placed without sorting: for T7 (value keys) they are the smallest values
// and lead the order; for T6 (w = float(i) * x keys) they have w = 0 and the
// largest indices, so they follow the zero-w keys of the sorted part (stable
// order: ties by ascending index).  Replaces xs[0..Le) by the square roots
// of the sorted part, takes the Eigen-order sum and the sequential walk over
// the full sequence, and returns the key at the weighted-median position k.
template <bool IDX>
using TailKey = typename std::conditional<IDX, uint64_t, uint32_t>::type;

// Sort key of tail sample i: T6 w = float(i) * x with the index (stable
// order, and the source row is an output); T7 the value x, with the index
// when its source row is asked for (IDX7), else alone (equal values are
// interchangeable).  Counts the zero keys of the index-carrying kinds.
template <bool T6, bool IDX7>
__device__ __forceinline__ TailKey<T6 || IDX7> tail_key(const float *xs, int i, bool *zero)
{
    const float x = xs[i];
    const float kv = T6 ? (float)i * x : x;
    *zero = kv == 0.0f;
    if constexpr (T6 || IDX7)
        return ((uint64_t)ord_key(kv) << 32) | (uint32_t)i;
    else
        return ord_key(kv);
}

// Highest lane index whose flag is set plus base, or `prev` when none is:
// the largest row index of a 64-row step (consecutive lanes on consecutive
// rows) from one ballot, no cross-lane reduction.
__device__ __forceinline__ int last_set(bool flag, int base, int prev)
{
    const unsigned long long b = __ballot(flag);
    return b ? base + 63 - __clzll(b) : prev;
}

// The key at position k of a tail's full sequence (L = Le + zt entries,
// the zt trailing zeros at ins .. ins + zt - 1), the sorted part's keys in
// blocked layout with KO per lane (a power of two); returned to every lane as
// (orderable value << 32) | index (index 0 for value-only keys).
template <int KR, bool IDX>
__device__ __forceinline__ uint64_t tail_key_at(const TailKey<IDX> (&key)[KR], int KO, int k,
                                                int ins, int zt, int Le)
{
    if (k >= ins && k < ins + zt)  // a trailing zero: w (or x) = +0, index Le + ...
        return ((uint64_t)ord_key(0.0f) << 32) | (uint32_t)(Le + (k - ins));
    const int pos = k < ins ? k : k - zt;
    // KO is a power of two (a network width): shift and mask, no division
    const int sh = __builtin_ctz((unsigned)KO);
    const int pr = pos & (KO - 1), pl = pos >> sh;
    TailKey<IDX> kk = 0;
#pragma unroll
    for (int r = 0; r < KR; r++)
        kk = (r == pr) ? key[r] : kk;
    if constexpr (IDX)
        return lane_u64(kk, pl);
    else  // value key in the high word (no index asked for)
        return (uint64_t)(uint32_t)lane_int((int)kk, pl) << 32;
}

// Second half of a sorted tail's median: the sorted part's keys are in
// blocked layout (lane l holds positions l*KO .. l*KO + KO - 1, KO <= KR),
// ins is where the trailing zeros go.  Replaces xs[0..Le) by the square
// roots of the sorted samples, takes the Eigen-order sum and the sequential
// walk over the full sequence, and returns the key at the weighted-median
// position k.
template <int KR, bool T6, bool IDX7>
__device__ __forceinline__ uint64_t tail_finish(const TailKey<T6 || IDX7> (&key)[KR], int KO,
                                                int ins, float *xs, int L, int Le, int lane,
                                                int *k_out, bool fs)
{
    constexpr bool IDX = T6 || IDX7;
    const int zt = L - Le;
    // gather the sorted samples before overwriting the slice with roots
    float sq[KR];
    auto roots = [&](auto fsc) {
        constexpr bool FS = decltype(fsc)::value;
#pragma unroll
        for (int r = 0; r < KR; r++) {
            const int pos = lane * KO + r;
            sq[r] = 0.0f;
            if (r < KO && pos < Le) {
                if constexpr (T6)
                    sq[r] = line_sqrt<FS>(xs[(uint32_t)key[r]]);
                else if constexpr (IDX7)
                    sq[r] = line_sqrt<FS>(ord_val((uint32_t)(key[r] >> 32)));
                else
                    sq[r] = line_sqrt<FS>(ord_val(key[r]));
            }
        }
    };
    if (fs)
        roots(std::true_type{});
    else
        roots(std::false_type{});
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < KR; r++) {
        const int pos = lane * KO + r;
        if (r < KO && pos < Le)
            xs[pos] = sq[r];
    }
    wave_lds_sync();
    const TailView view = {xs, ins, zt, L};
    const float S = eigen_sum_view(view, lane);
    int k = 0;
    if (lane == 0)
        k = weighted_median_walk_view<true>(view, Le, S);
    k = lane_int(k, 0);
    *k_out = k;
    return tail_key_at<KR, IDX>(key, KO, k, ins, zt, Le);
}

// Sorted tail of one line for T6 (T6 = true) or T7, given the tail's first
// Le samples in xs[0..Le), no sample negative or NaN, and every sample at
// i >= Le zero (the line has left the rotated image).  Only the first Le keys
// are sorted (K keys per lane, 64*K >= Le); the L - Le trailing zeros are
// placed without sorting: for T7 (value keys) they are the smallest values
// and lead the order; for T6 (w = float(i) * x keys) they have w = 0 and the
// largest indices, so they follow the zero-w keys of the sorted part (stable
// order: ties by ascending index).
template <int K, bool T6, bool IDX7>
__device__ __forceinline__ uint64_t sorted_tail_median(float *xs, int L, int Le, int lane,
                                                       int *k_out, bool fs)
{
    constexpr bool IDX = T6 || IDX7;
    using Key = TailKey<IDX>;
    Key key[K];
    int zin = 0;  // zero keys of the sorted part (a popcount per slot)
#pragma unroll
    for (int r = 0; r < K; r++) {
        const int i = r * kWave + lane;  // any placement: the sort orders them
        key[r] = (Key)~(Key)0;
        bool z = false;
        if (i < Le)
            key[r] = tail_key<T6, IDX7>(xs, i, &z);
        if constexpr (IDX)
            zin += __popcll(__ballot(z));
    }
    wave_bitonic_sort<K>(key, lane);
    // where the trailing zeros go: after the sorted part's zero keys in the
    // stable (index) order, or first when the order of equal values is free
    const int ins = IDX ? zin : 0;
    return tail_finish<K, T6, IDX7>(key, K, ins, xs, L, Le, lane, k_out, fs);
}

// The sorted tail with the smallest register network that holds its Le
// sorted keys (at least KMIN, at most KMAX per lane; the networks below
// KMIN are not instantiated).
template <int KMAX, int KMIN, bool T6, bool IDX7>
__device__ __forceinline__ uint64_t sorted_tail_dispatch(float *xs, int L, int Le, int lane,
                                                         int *k, bool fs)
{
    const int kneed = (Le + kWave - 1) / kWave;
    if constexpr (KMIN <= 1)
        if (kneed <= 1 || KMAX == 1)
            return sorted_tail_median<1, T6, IDX7>(xs, L, Le, lane, k, fs);
    if constexpr (KMAX >= 2 && KMIN <= 2)
        if (kneed <= 2 || KMAX == 2)
            return sorted_tail_median<2, T6, IDX7>(xs, L, Le, lane, k, fs);
    if constexpr (KMAX >= 4 && KMIN <= 4)
        if (kneed <= 4 || KMAX == 4)
            return sorted_tail_median<4, T6, IDX7>(xs, L, Le, lane, k, fs);
    if constexpr (KMAX >= 8 && KMIN <= 8)
        if (kneed <= 8 || KMAX == 8)
            return sorted_tail_median<8, T6, IDX7>(xs, L, Le, lane, k, fs);
    if constexpr (KMAX >= 16 && KMIN <= 16)
        if (kneed <= 16 || KMAX == 16)
            return sorted_tail_median<16, T6, IDX7>(xs, L, Le, lane, k, fs);
    if constexpr (KMAX >= 32)
        return sorted_tail_median<32, T6, IDX7>(xs, L, Le, lane, k, fs);
    return 0;  // not reached: kneed <= KMAX
}

// Keys per lane of the network for Le keys: KMIN, 2 KMIN, ... up to KMAX.
template <int KMAX, int KMIN = 1>
__device__ __forceinline__ int keys_per_lane(int Le)
{
    int K = KMIN;
    while (K < KMAX && K * kWave < Le)
        K <<= 1;
    return K;
}

// Sorts the first K of KMAX keys per lane (the rest hold the largest key)
// with the K-per-lane network (K a power of two from KMIN up to KMAX).
template <int KMAX, int KMIN = 1, class Key>
__device__ __forceinline__ void sort_prefix(Key (&k)[KMAX], int K, int lane)
{
    if constexpr (KMIN <= 1)
        if (K <= 1 || KMAX == 1) {
            wave_bitonic_sort<1>(reinterpret_cast<Key(&)[1]>(k), lane);
            return;
        }
    if constexpr (KMAX >= 2 && KMIN <= 2)
        if (K <= 2 || KMAX == 2) {
            wave_bitonic_sort<2>(reinterpret_cast<Key(&)[2]>(k), lane);
            return;
        }
    if constexpr (KMAX >= 4 && KMIN <= 4)
        if (K <= 4 || KMAX == 4) {
            wave_bitonic_sort<4>(reinterpret_cast<Key(&)[4]>(k), lane);
            return;
        }
    if constexpr (KMAX >= 8 && KMIN <= 8)
        if (K <= 8 || KMAX == 8) {
            wave_bitonic_sort<8>(reinterpret_cast<Key(&)[8]>(k), lane);
            return;
        }
    if constexpr (KMAX >= 16)
        wave_bitonic_sort<16>(reinterpret_cast<Key(&)[16]>(k), lane);
}

// P2 of one sinogram column per wave, columns of up to kP2Max rows
// (functionals.cpp:303-311, the weighted median of the sorted column with
// the values as weights, :21-30): the column sorted in the wave's registers
// (32-bit order keys: equal values are interchangeable), the sorted values
// into the wave's LDS row, the Eigen-order sum on lanes 0..7 and the
// sequential walk on lane 0.  Entry q of the P list; four columns per
// workgroup.
constexpr int kP2Max = 2048;

// Source row of P2's element (the adjoint's one-hot): the stable order of
// compareFloat (c++/src/functionals.cpp:32-43, ties by ascending index, -0
// equal to +0) puts the r-th occurrence in column order of the value v at
// rank lo + r, lo the number of elements below v.  -1 when v occurs nowhere
// (a NaN median).  Wave-uniform result.
__device__ __forceinline__ int p2_source(const float *x, int n, float v, int k, int lane)
{
    int lo = 0;
    for (int i0 = 0; i0 < n; i0 += kWave) {
        const int i = i0 + lane;
        lo += __popcll(__ballot(i < n && x[i] < v));
    }
    int r = k - lo;
    for (int i0 = 0; i0 < n; i0 += kWave) {
        const int i = i0 + lane;
        uint64_t eq = __ballot(i < n && x[i] == v);
        const int c = __popcll(eq);
        if (r < c) {
            for (int j = 0; j < r; j++)
                eq &= eq - 1;
            return i0 + __ffsll((unsigned long long)eq) - 1;
        }
        r -= c;
    }
    return -1;
}

// P2 of the column x[0..n) (n <= kP2Max) by one wave, sorted values in xs;
// the value and its rank are valid on lane 0, and with src its source row
// (every lane).
__device__ __forceinline__ void p2_wave(const float *x, int n, float *xs, int lane,
                                        float *res, int *rank, int *src = nullptr)
{
    const int K = keys_per_lane<32>(n);
    uint32_t key[32];
#pragma unroll
    for (int r = 0; r < 32; r++) {
        const int i = r * kWave + lane;
        key[r] = r < K && i < n ? ord_key(x[i]) : ~0u;
    }
    if (K <= 16)
        sort_prefix<16>(reinterpret_cast<uint32_t(&)[16]>(key), K, lane);
    else
        wave_bitonic_sort<32>(key, lane);
#pragma unroll
    for (int r = 0; r < 32; r++) {
        const int pos = lane * K + r;
        if (r < K && pos < n)
            xs[pos] = ord_val(key[r]);
    }
    wave_lds_sync();
    const TailView v = {xs, 0, 0, n};
    const float S = eigen_sum_view(v, lane);
    int k = 0;
    if (lane == 0)
        k = weighted_median_walk_view<false>(v, n, S);
    k = lane_int(k, 0);
    *res = xs[k];
    *rank = k;
    if (src)
        *src = p2_source(x, n, xs[k], k, lane);
}

__global__ __launch_bounds__(kBlock) void circus_p2_kernel(CircusParams p, int q)
{
    __shared__ float s_x[kBlock / kWave][kP2Max];
    const int w = wave_id(), lane = threadIdx.x & (kWave - 1);
    const int colg = blockIdx.x * (kBlock / kWave) + w;
    if (colg >= p.ncols * p.nsino)
        return;
    float res;
    int k;
    p2_wave(p.sino + (size_t)colg * p.n, p.n, s_x[w], lane, &res, &k);
    if (lane == 0) {
        const int sg = colg / p.ncols, col = colg - sg * p.ncols;
        const size_t o = ((size_t)sg * p.np + q) * p.ncols + col;
        p.out[o] = res;
        if (p.p2rank)
            p.p2rank[o] = k;
    }
}

// P2's source row of every sinogram column (the adjoint's one-hot), columns
// up to kP2Max rows: src[colg] for the n-row columns of sino.
__global__ __launch_bounds__(kBlock) void p2_source_kernel(const float *sino, int n, int cols,
                                                           int32_t *src)
{
    __shared__ float s_x[kBlock / kWave][kP2Max];
    const int w = wave_id(), lane = threadIdx.x & (kWave - 1);
    const int colg = blockIdx.x * (kBlock / kWave) + w;
    if (colg >= cols)
        return;
    float res;
    int k, sr;
    p2_wave(sino + (size_t)colg * n, n, s_x[w], lane, &res, &k, &sr);
    if (lane == 0)
        src[colg] = sr;
}

// The same over every sinogram column of a batch (image found by binary
// search over the descriptors' first-column offsets; every n <= kP2Max).
// Writes the circus value (with circus) and the source row (into the
// image's p2src [nt][A], when set).
__global__ __launch_bounds__(kBlock) void circus_p2_batched_kernel(BatchCircusParams p, int q)
{
    __shared__ float s_x[kBlock / kWave][kP2Max];
    const int w = wave_id(), lane = threadIdx.x & (kWave - 1);
    const int colg = blockIdx.x * (kBlock / kWave) + w;
    if (colg >= p.cols)
        return;
    int lo = 0, hi = p.nimg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.images[mid].col0 <= colg)
            lo = mid;
        else
            hi = mid - 1;
    }
    const BatchImage &im = p.images[lo];
    const int n = im.n;
    const int lc = colg - im.col0;
    const int t = lc / p.A;
    const int a = lc - t * p.A;
    float res;
    int k, sr = -1;
    p2_wave(im.sino + ((size_t)t * p.A + a) * n, n, s_x[w], lane, &res, &k,
            im.p2src ? &sr : nullptr);
    if (lane == 0) {
        if (p.with_circus)
            im.circus[((size_t)t * p.np + q) * p.A + a] = res;
        if (im.p2src)
            im.p2src[lc] = sr;
    }
}

// findWeightedMedian of one column per wave (columns up to kP2Max rows, in
// their own order, c++/src/circus.cpp:74-78): the column into the wave's LDS
// row, the Eigen-order sum on lanes 0..7 and the walk on lane 0.
__global__ __launch_bounds__(kBlock) void column_median_wave_kernel(const float *in, int n,
                                                                    int cols, int32_t *med)
{
    __shared__ float s_x[kBlock / kWave][kP2Max];
    const int w = wave_id(), lane = threadIdx.x & (kWave - 1);
    const int col = blockIdx.x * (kBlock / kWave) + w;
    if (col >= cols)
        return;
    const float *x = in + (size_t)col * n;
    float *xs = s_x[w];
    for (int i = lane; i < n; i += kWave)
        xs[i] = x[i];
    wave_lds_sync();
    const TailView v = {xs, 0, 0, n};
    const float S = eigen_sum_view(v, lane);
    if (lane == 0)
        med[col] = weighted_median_walk_view<false>(v, n, S);
}

// The largest key of a tail whose first Ls samples are in xs (the rest of
// its L are zero): with a negative or NaN sample, sqrt makes the Eigen sum
// NaN, no prefix passes the test and findWeightedMedian returns the last
// position, i.e. the largest key.
__device__ __forceinline__ uint64_t tail_max_key(const float *xs, int Ls, int L, bool t6,
                                                 int lane)
{
    // the unsampled zeros' largest key: (0, L - 1)
    uint64_t kmax = Ls < L ? (((uint64_t)ord_key(0.0f) << 32) | (uint32_t)(L - 1)) : 0ull;
    for (int i = lane; i < Ls; i += kWave) {
        const float x = xs[i];
        const float kv = t6 ? (float)i * x : x;
        const uint64_t kk = ((uint64_t)ord_key(kv) << 32) | (uint32_t)i;
        kmax = kk > kmax ? kk : kmax;
    }
    return wave_max_u64(kmax);
}

// Lane 0 stores one tail's result: the value, its rank k and source row.
template <bool IDX7>
__device__ __forceinline__ void store_tail_result(const SortTParams &p, int line, int which,
                                                  uint64_t res, int k, int base, int lane)
{
    if (lane != 0)
        return;
    const int slotw = which == 0 ? p.slot6 : p.slot7;
    p.sino[(size_t)slotw * p.A * p.n + line] = ord_val((uint32_t)(res >> 32));
    if (which == 0) {
        if (p.t6_rank)
            p.t6_rank[line] = k;
        if (p.t6_src)
            p.t6_src[line] = (int)(uint32_t)res;
    } else {
        if (p.t7_rank)
            p.t7_rank[line] = k;
        if (IDX7)
            p.t7_src[line] = base + (int)(uint32_t)res;
    }
}

// T6 (which = 0, tail from the sqrt-median sm) or T7 (which = 1, tail from
// the median m) of one line by one wave.  The wave samples the tail into its
// LDS slice, then
//   * if any sample is negative, sqrt makes the Eigen sum NaN, no prefix
//     passes the test and findWeightedMedian returns the last position: the
//     result is the largest key (no sort needed);
//   * otherwise it sorts the part before the trailing zeros with the
//     smallest register network that holds it (K = 1..KMAX keys per lane).
// With DEFER, an entry whose sorted part needs more than KMAX keys per lane
// is left for a wider pass: sort_line then returns true (nothing written).
template <int KMAX, int KMIN, bool DEFER, bool IDX7>
__device__ __forceinline__ bool sort_line(const SortTParams &p, int line, int which,
                                          float *xs, int lane)
{
    constexpr int M = kWave * KMAX;  // capacity of the wave's LDS slice
    const int n = p.n;
    const int slotw = which == 0 ? p.slot6 : p.slot7;
    if (slotw < 0)
        return false;
    const float o = (float)((n - 1) / 2.0);
    const float lim = (float)(n - 1);
    const int a = line / n;
    const int col = line - a * n;
    const float2 cs = p.trig[a];
    const float px = (float)col - o;
    const float pxc = px * cs.x, pxns = px * (-cs.y);
    const bool tr = fabsf(cs.y) > fabsf(cs.x);
    const float *src = tr ? p.imgT : p.img;
    const bool t6 = which == 0;
    const int base = t6 ? p.med_s[line] : p.med_x[line];
    const int L = n - base;
    // rows past the line's last possibly non-zero row sample exactly 0
    const int2 rows = line_rows(p.bbox, n, col, cs.x, cs.y);
    const int Ls = rows.y + 1 - base < L ? rows.y + 1 - base : L;
    // a sampled part longer than the slice cannot sort here: queue it for a
    // wider pass before sampling (its sorted part may turn out shorter;
    // the wider pass then sorts it with the smaller network)
    if (DEFER && Ls > M)
        return true;
    // consecutive lanes take consecutive rows (neighbouring taps)
    int last_nz = -1;
    bool bad = false;
    // (a longer sampled part is queued before sampling, or cannot occur)
    auto sample_tail = [&](auto trc) {
        constexpr bool TRC = decltype(trc)::value;
        float pyo = (float)(base + lane) - o;  // exact steps of 64 rows
        for (int i = lane; i < Ls; i += kWave, pyo += (float)kWave) {
            const float x = rot_sample<TRC>(src, n, lim, o, pxc, pxns, cs.x, cs.y, pyo);
            if (i < M)
                xs[i] = x;
            last_nz = last_set(x != 0.0f, i - lane, last_nz);
            bad |= !(x >= 0.0f);  // negative or NaN
        }
    };
    if (tr)
        sample_tail(std::true_type{});
    else
        sample_tail(std::false_type{});
    int k = L - 1;
    uint64_t res;
    if (__any(bad)) {
        wave_lds_sync();
        res = tail_max_key(xs, Ls < M ? Ls : M, L, t6, lane);
    } else {
        // lane 0 took part in every step of the loop: its copy is complete
        const int Le = __builtin_amdgcn_readfirstlane(last_nz) + 1;
        const int kneed = (Le + kWave - 1) / kWave;
        if (DEFER && kneed > KMAX)
            return true;
        wave_lds_sync();
        const bool fs = p.bbox && p.bbox[4] == 0;
        res = t6 ? sorted_tail_dispatch<KMAX, KMIN, true, false>(xs, L, Le, lane, &k, fs)
                 : sorted_tail_dispatch<KMAX, KMIN, false, IDX7>(xs, L, Le, lane, &k, fs);
    }
    store_tail_result<IDX7>(p, line, which, res, k, base, lane);
    return false;
}

// T6 and T7 of one line by one wave (a first pass, both requested).  The
// tails [sm, n) and [m, n) share their rows from max(sm, m) on, so the wave
// samples the union [min(sm, m), E) once into xs (E: the row after the
// line's last possibly non-zero row), sorts both tails' keys, writes the T6
// roots into xs2 and the T7 roots over xs, and runs the two Eigen sums on
// lanes 0..7 and 8..15 and the two walks on lanes 0 and 1 together.
// Returns the entries left for a wider pass (bit 0: T6, bit 1: T7; nothing
// written for them).  A pair whose union exceeds the slice goes entry by
// entry; so do the tails of a pair with a negative or NaN sample (largest
// key), from the shared samples.
template <int KMAX, int KMIN, bool DEFER, bool IDX7>
__device__ __forceinline__ int sort_pair(const SortTParams &p, int line, float *xs, float *xs2,
                                         int lane)
{
    constexpr int M = kWave * KMAX;
    const int n = p.n;
    const float o = (float)((n - 1) / 2.0);
    const float lim = (float)(n - 1);
    const int a = line / n;
    const int col = line - a * n;
    const float2 cs = p.trig[a];
    const float px = (float)col - o;
    const float pxc = px * cs.x, pxns = px * (-cs.y);
    const bool tr = fabsf(cs.y) > fabsf(cs.x);
    const float *src = tr ? p.imgT : p.img;
    const int sm = p.med_s[line], m = p.med_x[line];
    const bool fs = p.bbox && p.bbox[4] == 0;  // short square roots (line_sqrt)
    const int L6 = n - sm, L7 = n - m;
    const int2 rows = line_rows(p.bbox, n, col, cs.x, cs.y);
    const int E = rows.y + 1 < n ? rows.y + 1 : n;
    const int Ls6 = E - sm, Ls7 = E - m;  // sampled lengths (<= 0: no sample)
    int deferred = 0;
    if (DEFER)
        deferred = (Ls6 > M ? 1 : 0) | (Ls7 > M ? 2 : 0);
    const int b0 = sm < m ? sm : m;
    const int Lu = E - b0;
    if (deferred || Lu > M) {
        for (int which = 0; which < 2; which++) {
            if (!((deferred >> which) & 1) &&
                sort_line<KMAX, KMIN, DEFER, IDX7>(p, line, which, xs, lane))
                deferred |= 1 << which;
            wave_lds_sync();
        }
        return deferred;
    }
    int last_nz = -1, last_bad = -1;
    auto sample_union = [&](auto trc) {
        constexpr bool TRC = decltype(trc)::value;
        float pyo = (float)(b0 + lane) - o;  // exact steps of 64 rows
        for (int i = lane; i < Lu; i += kWave, pyo += (float)kWave) {
            const float x = rot_sample<TRC>(src, n, lim, o, pxc, pxns, cs.x, cs.y, pyo);
            xs[i] = x;
            last_nz = last_set(x != 0.0f, i - lane, last_nz);
            last_bad = last_set(!(x >= 0.0f), i - lane, last_bad);  // negative or NaN
        }
    };
    if (tr)
        sample_union(std::true_type{});
    else
        sample_union(std::false_type{});
    const int o6 = sm - b0, o7 = m - b0;
    // lane 0 took part in every step of the loop: its copies are complete
    const int lnz = __builtin_amdgcn_readfirstlane(last_nz);
    const int lbad = __builtin_amdgcn_readfirstlane(last_bad);
    const int Le6 = lnz + 1 - o6 > 0 ? lnz + 1 - o6 : 0;
    const int Le7 = lnz + 1 - o7 > 0 ? lnz + 1 - o7 : 0;
    wave_lds_sync();
    if (lbad >= o6 || lbad >= o7) {
        // the bad tails read the shared samples first; then at most one
        // good tail sorts (and overwrites them)
        const bool bad6 = lbad >= o6, bad7 = lbad >= o7;
        uint64_t r6 = 0, r7 = 0;
        int k6 = L6 - 1, k7 = L7 - 1;
        if (bad6)
            r6 = tail_max_key(xs + o6, Ls6, L6, true, lane);
        if (bad7)
            r7 = tail_max_key(xs + o7, Ls7, L7, false, lane);
        wave_lds_sync();
        if (!bad6)
            r6 = sorted_tail_dispatch<KMAX, KMIN, true, false>(xs + o6, L6, Le6, lane, &k6, fs);
        if (!bad7)
            r7 = sorted_tail_dispatch<KMAX, KMIN, false, IDX7>(xs + o7, L7, Le7, lane, &k7, fs);
        store_tail_result<IDX7>(p, line, 0, r6, k6, sm, lane);
        store_tail_result<IDX7>(p, line, 1, r7, k7, m, lane);
        return 0;
    }
    using Key7 = TailKey<IDX7>;
    const int K6 = keys_per_lane<KMAX, KMIN>(Le6), K7 = keys_per_lane<KMAX, KMIN>(Le7);
    // T6 first: keys, sort, its samples gathered by index, roots into xs2
    // (the samples in xs stay for T7's keys, so the two key sets are never
    // both being sorted)
    uint64_t key6[KMAX];
    int ins6 = 0;  // zero keys (a popcount per slot)
#pragma unroll
    for (int r = 0; r < KMAX; r++) {
        const int i = r * kWave + lane;  // < 64 * K: within the first K slots
        bool z = false;
        key6[r] = ~0ull;
        if (i < Le6)
            key6[r] = tail_key<true, false>(xs + o6, i, &z);
        ins6 += __popcll(__ballot(z));
    }
    sort_prefix<KMAX, KMIN>(key6, K6, lane);
    auto roots6 = [&](auto fsc) {
        constexpr bool FS = decltype(fsc)::value;
#pragma unroll
        for (int r = 0; r < KMAX; r++) {
            const int p6 = lane * K6 + r;
            if (r < K6 && p6 < Le6)
                xs2[p6] = line_sqrt<FS>(xs[o6 + (uint32_t)key6[r]]);
        }
    };
    if (fs)
        roots6(std::true_type{});
    else
        roots6(std::false_type{});
    // T7: keys, sort, roots from the key values over xs (after every lane
    // has read its samples)
    Key7 key7[KMAX];
    int ins7 = 0;  // zero keys when the order of equal values matters
#pragma unroll
    for (int r = 0; r < KMAX; r++) {
        const int i = r * kWave + lane;
        bool z = false;
        key7[r] = (Key7)~(Key7)0;
        if (i < Le7)
            key7[r] = tail_key<false, IDX7>(xs + o7, i, &z);
        if constexpr (IDX7)
            ins7 += __popcll(__ballot(z));
    }
    sort_prefix<KMAX, KMIN>(key7, K7, lane);
    wave_lds_sync();
    auto roots7 = [&](auto fsc) {
        constexpr bool FS = decltype(fsc)::value;
#pragma unroll
        for (int r = 0; r < KMAX; r++) {
            const int p7 = lane * K7 + r;
            if (r < K7 && p7 < Le7) {
                if constexpr (IDX7)
                    xs[p7] = line_sqrt<FS>(ord_val((uint32_t)(key7[r] >> 32)));
                else
                    xs[p7] = line_sqrt<FS>(ord_val(key7[r]));
            }
        }
    };
    if (fs)
        roots7(std::true_type{});
    else
        roots7(std::false_type{});
    wave_lds_sync();
    const TailView v6 = {xs2, ins6, L6 - Le6, L6}, v7 = {xs, ins7, L7 - Le7, L7};
    const float sg = eigen_sum_group<2>(lane >= 8 ? v7 : v6, lane);
    const float S6 = lane_float(sg, 0), S7 = lane_float(sg, 8);
    int kw = 0;
    if (lane < 2)  // lane 0 walks T6, lane 1 T7
        kw = weighted_median_walk_view<true>(lane ? v7 : v6, lane ? Le7 : Le6, lane ? S7 : S6);
    const int k6 = lane_int(kw, 0), k7 = lane_int(kw, 1);
    const uint64_t r6 = tail_key_at<KMAX, true>(key6, K6, k6, ins6, L6 - Le6, Le6);
    const uint64_t r7 = tail_key_at<KMAX, IDX7>(key7, K7, k7, ins7, L7 - Le7, Le7);
    store_tail_result<IDX7>(p, line, 0, r6, k6, sm, lane);
    store_tail_result<IDX7>(p, line, 1, r7, k7, m, lane);
    return 0;
}

// Persistent waves take work items from the counter p.work (zeroed ahead of
// the launch), p.chunk at a time: a line (T6 and T7) or, with p.defer_in
// set, a queued line item (line << 2 | mask, mask bit 0 for T6 and bit 1
// for T7).  One counter serialises its atomics across the device (tens of
// nanoseconds each), so the cheap items of the first pass go in chunks of
// several lines.  With DEFER, tails whose sorted part needs more than KMAX
// keys per lane are queued in p.defer_out for a wider pass (a wide network's
// VGPR budget would otherwise set the occupancy of every wave); a wave
// collects a chunk's queued items in one register (lane i holds the i-th)
// and queues them with one atomic.  Kernels of at most 16 keys per lane take
// the two tails of a line together when both are requested.  The passes over
// queued items (KMIN = KMAX) build only their widest network: a smaller code
// and fewer spills than one network per width.  The occupancy hints are the
// measured best (more waves beat the spills they cost).
template <int KMAX, int KMIN, bool DEFER, bool IDX7>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(KMAX <= 4 ? 5 : KMAX <= 8 ? 6 : KMAX <= 16 ? 4 : 1)))
void sino_sort_kernel(SortTParams p)
{
    constexpr int M = kWave * KMAX;
    // pairing kernels hold two slices per wave
    constexpr bool PAIR = KMAX <= 16;
    constexpr int W = PAIR ? 2 * M : M;
    __shared__ __align__(16) float s_all[(kBlock / kWave) * W];
    const int lane = threadIdx.x & (kWave - 1);
    float *xs = s_all + wave_id() * W;
    const int total = p.defer_in ? (int)*p.defer_in_count : p.A * p.n;
    const int chunk = p.chunk;  // at most 64: one queued item per line fits a wave
    for (;;) {
        int j0 = 0;
        if (lane == 0)
            j0 = (int)atomicAdd(p.work, (unsigned int)chunk);
        j0 = __builtin_amdgcn_readfirstlane(j0);
        if (j0 >= total)
            break;
        const int j1 = j0 + chunk < total ? j0 + chunk : total;
        int queued = 0, nq = 0;
        for (int j = j0; j < j1; j++) {
            int line = j, mask = 3;
            if (p.defer_in) {
                const int e = p.defer_in[j];
                line = e >> 2;
                mask = e & 3;
            }
            mask &= (p.slot6 >= 0 ? 1 : 0) | (p.slot7 >= 0 ? 2 : 0);
            int d = 0;
            if (PAIR && mask == 3) {
                d = sort_pair<KMAX, KMIN, DEFER, IDX7>(p, line, xs, xs + M, lane);
                wave_lds_sync();
            } else {
                for (int which = 0; which < 2; which++) {
                    if (((mask >> which) & 1) &&
                        sort_line<KMAX, KMIN, DEFER, IDX7>(p, line, which, xs, lane))
                        d |= 1 << which;
                    wave_lds_sync();
                }
            }
            if (d) {
                queued = lane == nq ? (line << 2) | d : queued;
                nq++;
            }
        }
        if (DEFER && nq > 0) {
            unsigned int at = 0;
            if (lane == 0)
                at = atomicAdd(p.defer_out_count, (unsigned int)nq);
            at = __builtin_amdgcn_readfirstlane(at);
            if (lane < nq)
                p.defer_out[at + lane] = queued;
        }
    }
}

// ---------------------------------------------------------------------------
// Adjoint kernels.  The medians, the T3..T5 integrals and the sample
// positions are those of a forward recompute on the same image (held fixed,
// as the medians are piecewise constant in the image).
// ---------------------------------------------------------------------------

// One LineRecord per (angle, column): folds the sinogram gradient of every
// requested T0..T5 into the coefficients of G(row); the T6/T7 one-sample
// terms go to a LineRecord67 when those functionals are requested.
__device__ __forceinline__ void make_record(const float *dsino, const int32_t *med_x,
                                            const int32_t *med_s, const int32_t *t6_src,
                                            const int32_t *t7_src, const float2 *z345,
                                            size_t plane, size_t line, Slots slot,
                                            LineRecord *out, LineRecord67 *out67)
{
    LineRecord r = {};
    r.m = med_x ? med_x[line] : 0;
    r.sm = med_s ? med_s[line] : 0;
    if (slot.s[0] >= 0)
        r.g0 = dsino[slot.s[0] * plane + line];
    if (slot.s[1] >= 0)
        r.g1 = dsino[slot.s[1] * plane + line];
    if (slot.s[2] >= 0)
        r.g2 = dsino[slot.s[2] * plane + line];
    for (int k = 0; k < 3; k++) {
        const int sl = slot.s[3 + k];
        if (sl < 0)
            continue;
        const float2 z = z345[k * plane + line];
        const float mag = hypotf(z.x, z.y);
        if (mag > 0.0f) {
            const float g = dsino[sl * plane + line] / mag;
            r.ab[2 * k] = g * z.x;
            r.ab[2 * k + 1] = g * z.y;
        }
    }
    out[line] = r;
    if (out67) {
        LineRecord67 q = {-1, -1, 0.0f, 0.0f};
        if (slot.s[6] >= 0 && t6_src) {
            const int s6 = t6_src[line];
            q.r6 = r.sm + s6;
            q.c6 = dsino[slot.s[6] * plane + line] * (float)s6;
        }
        if (slot.s[7] >= 0 && t7_src) {
            q.r7 = t7_src[line];
            q.c7 = dsino[slot.s[7] * plane + line];
        }
        out67[line] = q;
    }
}

__global__ __launch_bounds__(kBlock) void records_kernel(RecordParams p)
{
    const size_t lines = (size_t)p.A * p.n;
    for (size_t line = (size_t)blockIdx.x * kBlock + threadIdx.x; line < lines;
         line += (size_t)gridDim.x * kBlock)
        make_record(p.dsino, p.med_x, p.med_s, p.t6_src, p.t7_src, p.z345, lines, line,
                    p.slot, p.rec, p.rec67);
}

template <class F>
__device__ __forceinline__ int find_image(const BatchImage *images, int nimg, int v, F key)
{
    int lo = 0, hi = nimg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(images[mid]) <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void records_batched_kernel(BatchAdjointParams p)
{
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < p.lines; g += gridDim.x * kBlock) {
        const int k = find_image(p.images, p.nimg, g,
                                 [](const BatchImage &b) { return b.line0; });
        const BatchImage &im = p.images[k];
        const size_t plane = (size_t)p.A * im.n;
        const size_t line = (size_t)(g - im.line0);
        make_record(im.dsino, im.med_x, im.med_s, nullptr, nullptr, im.z345, plane, line,
                    p.slot, im.rec, nullptr);
    }
}

// ---------------------------------------------------------------------------
// Transposed bilinear rotation (the adjoint of rotate + interpolate,
// c++/src/auxiliary.cpp:223-255 and :182-196) as a scatter into a pixel tile
// held in LDS, one wave per tile.
//
// For every angle the wave visits the rotated samples (col, row) whose taps
// can touch its T x T pixel tile: those whose exact source position lies in
// the tile grown by one pixel, a square of side T + 1 turned by the angle.
// They are enumerated as a parallelogram: the columns the square spans, and
// in each column the rows of the narrower of the square's two slabs (one
// slab is ignored, so up to twice the square's area at 45 degrees, 1.45x on
// average).  Each sample is evaluated exactly as the forward evaluates it
// (rot_sample's float operations: coordinates, strict bounds, truncated taps
// and modf fractions), its rotated-domain gradient G(row) comes from its
// column's line record, and G times each of the four tap weights is added
// into the tile's accumulators (LDS, with a one-pixel halo ring that is
// discarded: those pixels belong to the neighbouring tiles).
//
// Determinism without atomics: two samples whose lattice positions agree
// modulo 3 in both coordinates are at least 3 apart, more than the 2 sqrt 2
// that two samples with overlapping 2x2 tap footprints can be apart.  The
// samples are visited one residue class (column mod 3, row mod 3) at a
// time, so the 64 lanes of a step read, add to and write back 256 distinct
// accumulators, and a pixel's additions follow program order.  Only the
// owning wave touches a tile's accumulators, so each pixel's sum is formed
// in one fixed order.  Every kGatherFlush angles the float accumulators are
// added into per-lane doubles and cleared.
// ---------------------------------------------------------------------------
constexpr int kGatherFlush = 4;
constexpr int kGatherTile = 32;
constexpr int kGatherTabWords = 6;  // LDS floats per row of the T3..T5 tables

template <int T>
struct GatherShape {
    static constexpr int P = T + 2;                    // accumulator side, with halo
    static constexpr int ACC = ((P * P + 3) / 4) * 4;  // floats, padded to float4
    static constexpr int KMAX = T == 16 ? 32 : 48;     // footprint columns per angle
    static constexpr int NPIX = T * T / kWave;         // interior pixels per lane
    static_assert(T == 16 || T == 32, "tiles of 16 or 32 pixels");
};

// LDS words of one wave: accumulators, then the staged column data
// (per column: (g0, sm, rlo, rhi), (ab0..ab3), (ab4, ab5), and with T1/T2
// (g1, g2, m, -), with T6/T7 (r6, r7, c6, c7)).
__host__ __device__ constexpr int gather_wave_words(int T, bool do12, bool do67)
{
    return (((T + 2) * (T + 2) + 3) / 4) * 4 +
           (T == 16 ? 32 : 48) * (4 + 4 + 2 + (do12 ? 4 : 0) + (do67 ? 4 : 0));
}

struct GatherImg {
    const LineRecord *rec;      // [A][n]
    const LineRecord67 *rec67;  // [A][n] or null
    float *out;                 // n x n (or this angle group's partial)
    int n;
};

template <int T, bool DO345, bool DO12>
__device__ __forceinline__ void gather_tile_wave(const GatherImg &im, const float2 *trig,
                                                 const float2 *s_tab, int tabn, int a0,
                                                 int a1, int X0, int Y0, float *wl)
{
    constexpr bool do12 = DO12;
    using S = GatherShape<T>;
    constexpr int P = S::P;
    const int lane = threadIdx.x & (kWave - 1);
    const int n = im.n;
    const bool do67 = im.rec67 != nullptr;
    float *acc = wl;
    float4 *sA = reinterpret_cast<float4 *>(wl + S::ACC);
    float4 *sB = sA + S::KMAX;
    float2 *sC = reinterpret_cast<float2 *>(sB + S::KMAX);
    float4 *sD = reinterpret_cast<float4 *>(sC + S::KMAX);
    float4 *sE = sD + (do12 ? S::KMAX : 0);
    const float o = (float)((n - 1) / 2.0);
    const float lim = (float)(n - 1);
    // the tile grown by one pixel, plus a margin far above the float
    // rounding of any sample position (< 1e-3 for n <= 2730), relative to o
    const float dm = 0.01f;
    const int X0m = X0 - 1, Y0m = Y0 - 1;  // accumulator cell 0 (the halo)
    const bool interior = X0 >= 2 && X0 + T <= n - 2 && Y0 >= 2 && Y0 + T <= n - 2;
    const float xa = (float)(X0 - 1) - o - dm, xb = (float)(X0 + T) - o + dm;
    const float ya = (float)(Y0 - 1) - o - dm, yb = (float)(Y0 + T) - o + dm;
    const float4 z4 = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = lane; i < S::ACC / 4; i += kWave)
        reinterpret_cast<float4 *>(acc)[i] = z4;
    double accd[S::NPIX];
#pragma unroll
    for (int i = 0; i < S::NPIX; i++)
        accd[i] = 0.0;
    wave_lds_sync();

    // Footprint of angle a: the columns [c_lo, c_lo + K) the grown tile
    // spans, px = c (x - o) - s (y - o) over its corners, and in each column
    // the rows [rlo, rhi] of the slab kp px + kq py in [lo, hi] with
    // |kq| >= 1/sqrt 2 (y - o = -s px + c py, or x - o = c px + s py), at
    // most W of them.
    struct Foot {
        float c, s, kp, lo, hi, inv;
        int c_lo, K, W;
    };
    auto footprint = [&](int a) {
        Foot f;
        const float2 cs = trig[a];
        f.c = cs.x;
        f.s = cs.y;
        const float p1 = f.c * xa - f.s * ya, p2 = f.c * xb - f.s * ya;
        const float p3 = f.c * xa - f.s * yb, p4 = f.c * xb - f.s * yb;
        const float pmin = fminf(fminf(p1, p2), fminf(p3, p4));
        const float pmax = fmaxf(fmaxf(p1, p2), fmaxf(p3, p4));
        int c_lo = (int)ceilf(pmin + o), c_hi = (int)floorf(pmax + o);
        c_lo = c_lo < 0 ? 0 : c_lo;
        c_hi = c_hi > n - 1 ? n - 1 : c_hi;
        // (T + 1.02) (|c| + |s|) bounds the span: at most 25 / 47 columns
        // for T = 16 / 32, so the clamp never drops one (and a class of
        // every third column has at most 16)
        const int K = c_hi - c_lo + 1;
        f.c_lo = c_lo;
        f.K = K > S::KMAX ? S::KMAX : K;
        const bool usey = fabsf(f.c) >= fabsf(f.s);
        const float kq = usey ? f.c : f.s;
        f.kp = usey ? -f.s : f.c;
        f.lo = usey ? ya : xa;
        f.hi = usey ? yb : xb;
        f.inv = 1.0f / kq;
        f.W = (int)((f.hi - f.lo) * fabsf(f.inv)) + 2;
        return f;
    };
    // this lane's column record of angle a (lane < K), loaded ahead (global
    // address space: the loads do not wait on LDS traffic)
    typedef const float __attribute__((address_space(1))) *GF;
    auto ld4 = [](GF q) { return float4{q[0], q[1], q[2], q[3]}; };
    float4 pa4 = z4, pb4 = z4, pd4 = z4, pe4 = z4;
    auto prefetch = [&](const Foot &f, int a) {
        if (lane < f.K) {
            const size_t line = (size_t)a * n + f.c_lo + lane;
            GF r = (GF)(im.rec + line);
            pa4 = ld4(r);
            if (DO345 || do12)
                pb4 = ld4(r + 4);
            if (do12)
                pd4 = ld4(r + 8);
            if (do67)
                pe4 = ld4((GF)(im.rec67 + line));
        }
    };
    Foot fn = footprint(a0);
    prefetch(fn, a0);
    for (int a = a0; a < a1; a++) {
        const Foot f = fn;
        // stage this angle's columns: (g0, sm, rlo, rhi), (ab0..ab3),
        // (ab4, ab5), (g1, g2, m, -), (r6, r7, c6, c7)
        if (lane < f.K) {
            const float px = (float)(f.c_lo + lane) - o;
            const float pa = (f.lo - f.kp * px) * f.inv, pb = (f.hi - f.kp * px) * f.inv;
            int rlo = (int)ceilf(fminf(pa, pb) + o), rhi = (int)floorf(fmaxf(pa, pb) + o);
            rlo = rlo < 0 ? 0 : rlo;
            rhi = rhi > n - 1 ? n - 1 : rhi;
            sA[lane] = float4{pa4.x, pa4.y, __int_as_float(rlo), __int_as_float(rhi)};
            if (DO345) {
                sB[lane] = float4{pa4.z, pa4.w, pb4.x, pb4.y};
                sC[lane] = float2{pb4.z, pb4.w};
            }
            if (do12)
                sD[lane] = pd4;
            if (do67)
                sE[lane] = pe4;
        }
        wave_lds_sync();
        if (a + 1 < a1) {
            fn = footprint(a + 1);
            prefetch(fn, a + 1);
        }
        const float c = f.c, s = f.s;
        const int c_lo = f.c_lo, K = f.K;
        // CHK: test the sample positions against the image bounds.  A tile
        // at least 2 pixels inside the image on every side can skip it: a
        // position outside [0, n - 1) truncates (negative values to 0) to a
        // cell outside the tile's grown range, so such a sample is dropped by
        // the tile test alone, as the bounds test would drop it.
        auto classes = [&](auto chk) {
            constexpr bool CHK = decltype(chk)::value;
            const float pxc0 = (float)c_lo - o;
            const int clo3 = c_lo % 3;
            // one residue class (column, row mod 3) at a time: samples of one
            // class are pairwise at least 3 apart, so their 2x2 tap footprints
            // are disjoint and every read-add-write of the class has its own
            // four targets.  Lane l takes the class's column l % 16 (at most
            // 16: K <= 47) and its rows 4 apart in the class from l / 16 on.
            const int kl = lane & 15, jl = lane >> 4;
            // one column class (col mod 3) at a time: the column's data
            // serves its three row classes
            for (int cc = 0; cc < 3; cc++) {
                const int koff = cc >= clo3 ? cc - clo3 : cc - clo3 + 3;
                const int K3 = (K - koff + 2) / 3;
                const bool colok = kl < K3;
                const int k = colok ? 3 * kl + koff : 0;
                const float4 ea = sA[k];
                const float g0 = ea.x;
                const int sm = __float_as_int(ea.y);
                const int rlo = __float_as_int(ea.z);
                const int rhi = colok ? __float_as_int(ea.w) : -1;
                const int r3 = rlo - 3 * (int)(((uint32_t)rlo * 0xAAABu) >> 17);
                float4 eb = {}, ed = {}, ee = {};
                float2 ec = {};
                if (DO345) {
                    eb = sB[k];
                    ec = sC[k];
                }
                if (do12)
                    ed = sD[k];
                if (do67)
                    ee = sE[k];
                // the forward's sample position (rot_sample), bit for bit
                const float px = pxc0 + (float)k;
                const float pxc = px * c, pxns = px * (-s);
                // One sample: its accumulator cell and the four weighted
                // gradient values of its taps (use: its taps touch the tile).
                struct Hit {
                    int cell;  // byte offset of the sample's first accumulator
                    float w00, w01, w10, w11;
                    bool use;
                };
                // py = row - o, stepped by whole rows by the caller (integers
                // or half-integers far below 2^23: exact)
                auto sample = [&](int row, float py, int q) {
                    Hit h;
                    float x = pxc + py * s;
                    float y = pxns + py * c;
                    x = x + o;
                    y = y + o;
                    int Xb = (int)(uint32_t)x, Yb = (int)(uint32_t)y;
                    if constexpr (CHK) {
                        const bool inb = __float_as_uint(x) < __float_as_uint(lim) &&
                                         __float_as_uint(y) < __float_as_uint(lim);
                        Xb = inb ? Xb : -8;
                        Yb = inb ? Yb : -8;
                    }
                    const float fx = __builtin_amdgcn_fractf(x);
                    const float fy = __builtin_amdgcn_fractf(y);
                    const int xi = Xb - X0m, yi = Yb - Y0m;  // accumulator cell
                    h.use = (unsigned)xi <= (unsigned)T && (unsigned)yi <= (unsigned)T;
                    // G(row): g0 + [row >= m] (g1 d + g2 d^2)
                    //   + [row > sm] sum_t (alpha_t W_t.re + beta_t W_t.im)[row - sm]
                    //   + [row == r6] c6 + [row == r7] c7
                    float G = g0;
                    if (do12) {
                        const int m = __float_as_int(ed.z);
                        const float d = (float)(row - m);
                        const float hh = fmaf(d, fmaf(ed.y, d, ed.x), G);
                        G = row >= m ? hh : G;
                    }
                    // (a wave whose active lanes all sit at or before their
                    // sm skips the term: a uniform branch)
                    if (DO345 && __builtin_amdgcn_ballot_w64(q > 0) != 0) {
                        // q = row - sm, stepped with the row; table entry 0 is
                        // zero, so rows at or before sm add nothing; q <= n - 1
                        // < tabn (row < n <= tabn)
                        const int qc = q > 0 ? q : 0;
                        const float4 w34 = reinterpret_cast<const float4 *>(s_tab)[qc];
                        const float2 w5 = s_tab[2 * tabn + qc];
                        G = fmaf(eb.x, w34.x, G);
                        G = fmaf(eb.y, w34.y, G);
                        G = fmaf(eb.z, w34.z, G);
                        G = fmaf(eb.w, w34.w, G);
                        G = fmaf(ec.x, w5.x, G);
                        G = fmaf(ec.y, w5.y, G);
                    }
                    if (do67) {
                        G += row == __float_as_int(ee.x) ? ee.z : 0.0f;
                        G += row == __float_as_int(ee.y) ? ee.w : 0.0f;
                    }
                    const float ofx = 1.0f - fx, ofy = 1.0f - fy;
                    const float ga = G * ofx, gb = G * fx;
                    // byte offset 4 (xi P + yi) (read only where used, xi <= T):
                    // a 24-bit multiply, which the compiler would otherwise
                    // widen to a quarter-rate 32-bit one
                    int xp;
                    asm volatile("v_mul_u32_u24 %0, %1, %2" : "=v"(xp) : "n"(4 * P), "v"(xi));
                    h.cell = xp + 4 * yi;
                    h.w00 = ga * ofy;
                    h.w01 = ga * fy;
                    h.w10 = gb * ofy;
                    h.w11 = gb * fy;
                    return h;
                };
                for (int rc = 0; rc < 3; rc++) {
                    // one residue class (column, row mod 3): its samples are
                    // pairwise at least 3 apart, so their 2x2 tap footprints
                    // are disjoint and every read-add-write of the class has
                    // its own four targets
                    const int roff = rc >= r3 ? rc - r3 : rc - r3 + 3;
                    const int row0 = rlo + roff + 3 * jl;
                    float py = (float)row0 - o;
                    int q = row0 - sm;
                    for (int row = row0; row <= rhi; row += 12, py += 12.0f, q += 12) {
                        const Hit a = sample(row, py, q);
                        if (a.use) {
                            float *ca = reinterpret_cast<float *>(
                                reinterpret_cast<char *>(acc) + a.cell);
                            const float a00 = ca[0], a01 = ca[1], a10 = ca[P], a11 = ca[P + 1];
                            ca[0] = a00 + a.w00;
                            ca[1] = a01 + a.w01;
                            ca[P] = a10 + a.w10;
                            ca[P + 1] = a11 + a.w11;
                        }
                    }
                }
            }
        };
        if (interior)
            classes(std::false_type{});
        else
            classes(std::true_type{});
        wave_lds_sync();  // staged columns are read; the next angle rewrites them
        if ((a - a0) % kGatherFlush == kGatherFlush - 1 || a == a1 - 1) {
            // flush the float accumulators of these angles into the doubles
#pragma unroll
            for (int i = 0; i < S::NPIX; i++) {
                const int p = lane + kWave * i;
                accd[i] += (double)acc[(p / T + 1) * P + p % T + 1];
            }
            wave_lds_sync();
            for (int i = lane; i < S::ACC / 4; i += kWave)
                reinterpret_cast<float4 *>(acc)[i] = z4;
            wave_lds_sync();
        }
    }
#pragma unroll
    for (int i = 0; i < S::NPIX; i++) {
        const int p = lane + kWave * i;
        const int X = X0 + p / T, Y = Y0 + p % T;
        if (X < n && Y < n)
            im.out[(size_t)X * n + Y] = (float)accd[i];
    }
}

// Stages the T3..T5 tables into the workgroup's LDS (every wave helps; T3
// and T4 of a row as one float4, T5 after them, as stage_tab345), then
// returns this wave's slice of the dynamic LDS.
__device__ __forceinline__ float *gather_lds(float *smem, const float2 *tab345, int tabn,
                                             bool do345, int wave_words)
{
    const int tab_words = do345 ? kGatherTabWords * tabn : 0;
    if (do345)
        stage_tab345(reinterpret_cast<float2 *>(smem), tab345, tabn, blockDim.x);
    __syncthreads();
    return smem + ((tab_words + 3) & ~3) + wave_id() * wave_words;
}

template <bool DO345, bool DO12>
__global__ __launch_bounds__(1024) void gather_kernel(GatherParams p)
{
    constexpr int T = kGatherTile;
    extern __shared__ float4 smem4[];
    float *smem = reinterpret_cast<float *>(smem4);
    const int words = gather_wave_words(T, DO12, p.rec67 != nullptr);
    float *wl = gather_lds(smem, p.tab345, p.n, DO345, words);
    const int waves = blockDim.x / kWave;
    const int item = blockIdx.x * waves + wave_id();
    const int tn = (p.n + T - 1) / T;
    if (item >= tn * tn * p.groups)
        return;
    const int g = item / (tn * tn), tile = item - g * tn * tn;
    const int a0 = (int)((long long)p.A * g / p.groups);
    const int a1 = (int)((long long)p.A * (g + 1) / p.groups);
    GatherImg im;
    im.rec = p.rec;
    im.rec67 = p.rec67;
    im.out = p.groups > 1 ? p.part + (size_t)g * p.n * p.n : p.dimg;
    im.n = p.n;
    gather_tile_wave<T, DO345, DO12>(im, p.trig, reinterpret_cast<const float2 *>(smem), p.n,
                                     a0, a1, tile / tn * T, tile % tn * T, wl);
}

// dimg = sum over angle groups of the partial images, in group order.
__global__ __launch_bounds__(kBlock) void gather_reduce_kernel(const float *part, int groups,
                                                               size_t npix, float *dimg)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < npix;
         i += (size_t)gridDim.x * kBlock) {
        double acc = 0.0;
        for (int g = 0; g < groups; g++)
            acc += (double)part[g * npix + i];
        dimg[i] = (float)acc;
    }
}

template <int T, bool DO345, bool DO12>
__global__ __launch_bounds__(1024) void gather_batched_kernel(BatchAdjointParams p)
{
    extern __shared__ float4 smem4[];
    float *smem = reinterpret_cast<float *>(smem4);
    const int words = gather_wave_words(T, DO12, false);
    float *wl = gather_lds(smem, p.tab345, p.tabn, DO345, words);
    const int waves = blockDim.x / kWave;
    const int item = blockIdx.x * waves + wave_id();
    if (item >= p.tiles)
        return;
    const int k = find_image(p.images, p.nimg, item,
                             [](const BatchImage &b) { return b.tile0; });
    const BatchImage &bi = p.images[k];
    const int n = bi.n;
    const int tn = (n + T - 1) / T;
    const int tile = item - bi.tile0;
    GatherImg im;
    im.rec = bi.rec;
    im.rec67 = nullptr;
    im.out = bi.dimg;
    im.n = n;
    gather_tile_wave<T, DO345, DO12>(im, p.trig, reinterpret_cast<const float2 *>(smem),
                                     p.tabn, 0, p.A, tile / tn * T, tile % tn * T, wl);
}

// dS = J_P^T dC for every sinogram column (P1: sign differences; P2: one-hot
// at the source of the weighted-median rank of the stable sort; P3: the
// derivative of sum_p w_p |X_q(p)|^4 / N^4, an inverse-DFT-shaped sum;
// Hermite: its weights).
// LDS layout of one circus-VJP workgroup for columns of height up to nmax.
struct VjpLds {
    uint64_t *key;  // pow2 >= nmax
    float *x;       // nmax4
    float *sq;      // nmax4
    float *ds;      // nmax4
};

__device__ __forceinline__ VjpLds vjp_lds(unsigned char *smem, int nmax)
{
    const int n4 = (nmax + 3) & ~3;
    int m2 = 1;
    while (m2 < nmax)
        m2 <<= 1;
    VjpLds l;
    l.key = reinterpret_cast<uint64_t *>(smem);
    l.x = reinterpret_cast<float *>(l.key + m2);
    l.sq = l.x + n4;
    l.ds = l.sq + n4;
    return l;
}

// J_P^T dC for the column x[0..n) already in L.x; dc[q * dc_stride] is the
// upstream gradient of P-functional q at this column.  P3's part comes from
// the DFT GEMMs (tt_dft.hip) and is already in out when p3_done; the P1, P2
// and Hermite parts are added to it and the sum written back to out.
__device__ void circus_vjp_column(const VjpLds &L, int n, int np, const int *code,
                                  const int *herm, const double *herm_w,
                                  const float *dc, size_t dc_stride, float *out,
                                  bool p3_done, const int32_t *p2src)
{
    __shared__ float s_chain[8];
    __shared__ float s_sum;
    __shared__ int s_src;
    int m2 = 1;
    while (m2 < n)
        m2 <<= 1;
    for (int i = threadIdx.x; i < n; i += kBlock)
        L.ds[i] = p3_done ? out[i] : 0.0f;
    __syncthreads();
    for (int q = 0; q < np; q++) {
        const float g = dc[q * dc_stride];
        const int cd = code[q];
        if (cd == 1) {
            for (int j = threadIdx.x; j < n; j += kBlock) {
                float d = 0.0f;
                if (j >= 1) {
                    const float u = L.x[j - 1] - L.x[j];
                    d -= (float)((u > 0.0f) - (u < 0.0f));
                }
                if (j <= n - 2) {
                    const float u = L.x[j] - L.x[j + 1];
                    d += (float)((u > 0.0f) - (u < 0.0f));
                }
                L.ds[j] += g * d;
            }
        } else if (cd == 0) {
            const double *w = herm_w + (size_t)herm[q] * n;
            for (int j = threadIdx.x; j < n; j += kBlock)
                L.ds[j] += (float)((double)g * w[j]);
        } else if (cd == 2 && p2src) {
            // the source row found by the P2 wave kernel (stable order)
            const int sr = *p2src;
            if (threadIdx.x == 0 && sr >= 0)
                L.ds[sr] += g;
        } else if (cd == 2) {
            for (int i = threadIdx.x; i < m2; i += kBlock)
                L.key[i] = i < n ? (((uint64_t)ord_key(L.x[i]) << 32) | (uint32_t)i)
                                 : ~0ull;
            __syncthreads();
            bitonic_lds(L.key, m2);
            for (int i = threadIdx.x; i < n; i += kBlock)
                L.sq[i] = L.x[(int)(uint32_t)L.key[i]];
            __syncthreads();
            eigen_sum_block(L.sq, n, s_chain, &s_sum);
            if (threadIdx.x == 0) {
                const int m = weighted_median_walk(L.sq, n, s_sum);
                s_src = (int)(uint32_t)L.key[m];
            }
            __syncthreads();
            if (threadIdx.x == 0)
                L.ds[s_src] += g;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += kBlock)
        out[i] = L.ds[i];
}

// dS = J_P^T dC for every sinogram column: P1 sign differences, P2 one-hot
// at the source of the weighted-median rank of the stable sort, Hermite its
// weights, added to P3's part (the derivative of sum_p w_p |X_q(p)|^4 / N^4,
// from the DFT GEMMs) when p3_done.
__global__ __launch_bounds__(kBlock) void circus_vjp_kernel(CircusVjpParams p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = p.n;
    const VjpLds L = vjp_lds(smem_raw, n);
    const int colg = blockIdx.x;
    const int sg = colg / p.ncols;
    const int col = colg - sg * p.ncols;
    const float *x = p.sino + (size_t)colg * n;
    for (int i = threadIdx.x; i < n; i += kBlock)
        L.x[i] = x[i];
    __syncthreads();
    circus_vjp_column(L, n, p.np, p.code, p.herm, p.herm_w,
                      p.dcircus + (size_t)sg * p.np * p.ncols + col, p.ncols,
                      p.dsino + (size_t)colg * n, p.p3_done != 0,
                      p.p2src ? p.p2src + colg : nullptr);
}

__global__ __launch_bounds__(kBlock) void circus_vjp_batched_kernel(BatchAdjointParams p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const VjpLds L = vjp_lds(smem_raw, p.nmax);
    const int colg = blockIdx.x;
    const int k = find_image(p.images, p.nimg, colg,
                             [](const BatchImage &b) { return b.col0; });
    const BatchImage &im = p.images[k];
    const int n = im.n;
    const int lc = colg - im.col0;       // t * A + a
    const int t = lc / p.A;
    const int a = lc - t * p.A;
    const float *x = im.sino + (size_t)lc * n;
    for (int i = threadIdx.x; i < n; i += kBlock)
        L.x[i] = x[i];
    __syncthreads();
    const int herm_none[kMaxP] = {};
    circus_vjp_column(L, n, p.np, p.code, herm_none, nullptr,
                      im.dcircus + (size_t)t * p.np * p.A + a, p.A,
                      im.dsino + (size_t)lc * n, p.p3_done != 0,
                      im.p2src ? im.p2src + lc : nullptr);
}

// ---------------------------------------------------------------------------
// Steps either side of the path.
// ---------------------------------------------------------------------------

// c++/src/auxiliary.cpp:198-221 (resize): interior pixels sample the input at
// ((col,row) + 0.5) * diag(in_rows/rows, in_cols/cols) - 0.5 with the
// reference's float operations; the one-pixel border stays zero.  Negative
// coordinates (upscaling) keep modf's semantics: integer part -0 -> index 0,
// negative fraction.
__global__ __launch_bounds__(kBlock) void resize_kernel(const float *in, int in_rows,
                                                        int in_cols, float *out,
                                                        int rows, int cols)
{
    const size_t total = (size_t)rows * cols;
    const float t00 = (float)in_rows / (float)rows;
    const float t11 = (float)in_cols / (float)cols;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const int row = (int)(i % rows), col = (int)(i / rows);
        float v = 0.0f;
        if (row >= 1 && row < rows - 1 && col >= 1 && col < cols - 1) {
            const float px = (float)col + 0.5f, py = (float)row + 0.5f;
            float x = px * t00 + py * 0.0f;
            float y = px * 0.0f + py * t11;
            x = x - 0.5f;
            y = y - 0.5f;
            const float xi = __builtin_truncf(x), yi = __builtin_truncf(y);
            const float fx = x - xi, fy = y - yi;
            const int X = (int)xi, Y = (int)yi;
            // The reference scales the column coordinate by the row ratio
            // (and vice versa), so a non-square input reaches past the
            // matrix; those taps read 0 here.
            auto tap = [&](int r, int c) {
                return (r >= 0 && r < in_rows && c >= 0 && c < in_cols)
                           ? in[(size_t)c * in_rows + r]
                           : 0.0f;
            };
            const float s00 = tap(Y, X), s01 = tap(Y, X + 1);
            const float s10 = tap(Y + 1, X), s11 = tap(Y + 1, X + 1);
            v = s00 * (1 - fx) * (1 - fy) + s01 * fx * (1 - fy) +
                s10 * (1 - fx) * fy + s11 * fx * fy;
        }
        out[i] = v;
    }
}

// c++/src/auxiliary.cpp:277-318 (zscore): float mean from a sequential float
// sum, variance in double over float differences, float result.  One
// workgroup per vector.
__global__ __launch_bounds__(kBlock) void zscore_kernel(const float *in, int n,
                                                        float *out)
{
    __shared__ float s_mean;
    __shared__ double s_red[kBlock / kWave];
    const float *x = in + (size_t)blockIdx.x * n;
    float *y = out + (size_t)blockIdx.x * n;
    if (threadIdx.x == 0) {
        float sum = 0.0f;
        for (int i = 0; i < n; i++)
            sum += x[i];
        s_mean = sum / (float)n;
    }
    __syncthreads();
    const float mean = s_mean;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const double d = (double)(x[i] - mean);
        acc = fma(d, d, acc);
    }
    const double var = block_sum_d(acc, s_red);
    const float stdev = (float)sqrt(var / (double)(n - 1));
    for (int i = threadIdx.x; i < n; i += kBlock)
        y[i] = (x[i] - mean) / stdev;
}

// findWeightedMedian of every column of an n x cols matrix (one workgroup per
// column), as nearest_orthonormal_sinogram uses it (c++/src/circus.cpp:74-78).
__global__ __launch_bounds__(kBlock) void column_median_kernel(const float *in, int n,
                                                               int32_t *med)
{
    extern __shared__ __align__(16) float s_col[];
    __shared__ float s_chain[8];
    __shared__ float s_sum;
    const float *x = in + (size_t)blockIdx.x * n;
    for (int i = threadIdx.x; i < n; i += kBlock)
        s_col[i] = x[i];
    __syncthreads();
    eigen_sum_block(s_col, n, s_chain, &s_sum);
    if (threadIdx.x == 0)
        med[blockIdx.x] = weighted_median_walk(s_col, n, s_sum);
}

// Smallest and largest of a sinogram's column medians med[0 .. cols), for
// every thread of the block.
__device__ __forceinline__ int2 block_min_max(const int32_t *med, int cols)
{
    __shared__ int s_mm[2][kBlock / kWave];
    int mn = 1 << 30, mx = -(1 << 30);
    for (int c = threadIdx.x; c < cols; c += kBlock) {
        mn = min(mn, med[c]);
        mx = max(mx, med[c]);
    }
    mn = wave_min_i(mn);
    mx = wave_max_i(mx);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_mm[0][threadIdx.x / kWave] = mn;
        s_mm[1][threadIdx.x / kWave] = mx;
    }
    __syncthreads();
    mn = s_mm[0][0];
    mx = s_mm[1][0];
    for (int w = 1; w < kBlock / kWave; w++) {
        mn = min(mn, s_mm[0][w]);
        mx = max(mx, s_mm[1][w]);
    }
    return int2{mn, mx};
}

// aligned(max + row - med[col], col) = in(row, col), zero elsewhere
// (c++/src/circus.cpp:85-93), in double for the SVD, for a batch: grid y =
// sinogram t (sinograms of n x cols back to back, medians med[t * cols + c],
// aligned matrices mstride doubles apart), max the largest median of
// sinogram t.
__global__ __launch_bounds__(kBlock) void align_batched_kernel(const float *in, int n, int cols,
                                                               const int32_t *med,
                                                               int rows_out, size_t mstride,
                                                               double *out)
{
    const int t = blockIdx.y;
    const int32_t *off = med + (size_t)t * cols;
    const int maxoff = block_min_max(off, cols).y;
    const float *src_t = in + (size_t)t * cols * n;
    double *out_t = out + (size_t)t * mstride;
    const size_t total = (size_t)rows_out * cols;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const int r = (int)(i % rows_out), c = (int)(i / rows_out);
        const int src = r - (maxoff - off[c]);
        out_t[i] = (src >= 0 && src < n) ? (double)src_t[(size_t)c * n + src] : 0.0;
    }
}

// Float copy of the aligned rows of a batch of matrices (grid y = matrix t,
// column major with leading dimension ld, mstride doubles apart in, ostride
// floats apart out, leading dimension the matrix's rows): matrix t keeps its
// aligned height n + |max - sc| + |min - sc| of its medians' offsets from
// sc.
__global__ __launch_bounds__(kBlock) void to_float_batched_kernel(const double *in, int ld,
                                                                  size_t mstride, int cols,
                                                                  const int32_t *med, int n,
                                                                  int sc, float *out,
                                                                  size_t ostride)
{
    const int t = blockIdx.y;
    const int2 mm = block_min_max(med + (size_t)t * cols, cols);
    const int rows = n + abs(mm.y - sc) + abs(mm.x - sc);
    const double *in_t = in + (size_t)t * mstride;
    float *out_t = out + (size_t)t * ostride;
    const size_t total = (size_t)rows * cols;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const int r = (int)(i % rows), c = (int)(i / rows);
        out_t[i] = (float)in_t[(size_t)c * ld + r];
    }
}

// Bounding box of the non-zero pixels of an n x n column-major image
// (pixel (row Y, col X) at X * n + Y): bb = {x0, x1, y0, y1}, set to
// {n, -1, n, -1} beforehand.  Block-local bounds, one atomic each per block.
// Block b takes columns b, b + nblocks, ...; its threads stride down rows.
__device__ __forceinline__ void bbox_accumulate(const float *img, int n, int block,
                                                int nblocks, int *bb)
{
    __shared__ int s_b[5][kBlock / kWave];
    int x0 = n, x1 = -1, y0 = n, y1 = -1;
    bool narrow = false;  // a pixel below 0 or in (0, 2^-45)
    for (int X = block; X < n; X += nblocks) {
        const float *col = img + (size_t)X * n;
#pragma unroll 4
        for (int Y = threadIdx.x; Y < n; Y += kBlock) {
            const float v = col[Y];
            narrow = narrow || v < 0.0f || (v > 0.0f && v < 0x1p-45f);
            if (v != 0.0f) {
                x0 = X < x0 ? X : x0;
                x1 = X > x1 ? X : x1;
                y0 = Y < y0 ? Y : y0;
                y1 = Y > y1 ? Y : y1;
            }
        }
    }
    x0 = wave_min_i(x0);
    x1 = wave_max_i(x1);
    y0 = wave_min_i(y0);
    y1 = wave_max_i(y1);
    const int nw = __ballot(narrow) != 0;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (lane == 0) {
        s_b[0][w] = x0;
        s_b[1][w] = x1;
        s_b[2][w] = y0;
        s_b[3][w] = y1;
        s_b[4][w] = nw;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int nr = s_b[4][0];
        for (int k = 1; k < kBlock / kWave; k++) {
            x0 = s_b[0][k] < x0 ? s_b[0][k] : x0;
            x1 = s_b[1][k] > x1 ? s_b[1][k] : x1;
            y0 = s_b[2][k] < y0 ? s_b[2][k] : y0;
            y1 = s_b[3][k] > y1 ? s_b[3][k] : y1;
            nr |= s_b[4][k];
        }
        if (nr)
            atomicOr(bb + 4, 1);
        if (x1 >= 0) {
            atomicMin(bb + 0, x0);
            atomicMax(bb + 1, x1);
            atomicMin(bb + 2, y0);
            atomicMax(bb + 3, y1);
        }
    }
}

__global__ void bbox_init_kernel(int *bb, int n)
{
    if (threadIdx.x == 0) {
        bb[0] = n;
        bb[1] = -1;
        bb[2] = n;
        bb[3] = -1;
        bb[4] = 0;
    }
}

__global__ __launch_bounds__(kBlock) void image_bbox_kernel(const float *img, int n, int *bb)
{
    bbox_accumulate(img, n, blockIdx.x, gridDim.x, bb);
}

__global__ void bbox_init_batched_kernel(const BatchImage *images, int nimg)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nimg) {
        int *bb = const_cast<int *>(images[k].bbox);
        const int n = images[k].n;
        bb[0] = n;
        bb[1] = -1;
        bb[2] = n;
        bb[3] = -1;
        bb[4] = 0;
    }
}

// grid.y walks the images
__global__ __launch_bounds__(kBlock) void image_bbox_batched_kernel(const BatchImage *images)
{
    const BatchImage &im = images[blockIdx.y];
    bbox_accumulate(im.img, im.n, blockIdx.x, gridDim.x, const_cast<int *>(im.bbox));
}

// out(i, j) = v(i, j) / sqrt(lambda_j) for count column-major k x k matrices
// (eigenvectors v, eigenvalues lambda, k per matrix).
__global__ __launch_bounds__(kBlock) void scale_eigvec_kernel(const double *v,
                                                             const double *lambda, int k,
                                                             int count, double *out)
{
    const size_t kk = (size_t)k * k, total = kk * count;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const size_t m = i / kk;
        const int j = (int)((i - m * kk) / k);
        out[i] = v[i] / sqrt(lambda[m * k + j]);
    }
}

// x <- 1.5 I - 0.5 x for count k x k matrices (one Newton-Schulz factor).
__global__ __launch_bounds__(kBlock) void newton_schulz_factor_kernel(double *x, int k,
                                                                     int count)
{
    const size_t kk = (size_t)k * k, total = kk * count;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const size_t e = i % kk;
        const int r = (int)(e % k), c = (int)(e / k);
        x[i] = (r == c ? 1.5 : 0.0) - 0.5 * x[i];
    }
}

// Polar iteration of the nearest orthonormal sinogram (X <- X q(X^T X)).
// The Gram matrices come from a GEMM split over row chunks: S_m = sum over
// c < chunks of part[m * chunks + c], summed in chunk order (deterministic).
__global__ __launch_bounds__(kBlock) void nos_reduce_kernel(const double *part, int chunks,
                                                           size_t kk, int count, double *S)
{
    const size_t total = kk * count;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const size_t m = i / kk, e = i - m * kk;
        const double *p = part + m * chunks * kk + e;
        double acc = p[0];
        for (int c = 1; c < chunks; c++)
            acc += p[c * kk];
        S[i] = acc;
    }
}

// Per-slice partial reductions of each k x k matrix (grid: kNosSlices x
// count): sum of squares (WHAT = 0) or max |S - I| (WHAT = 1, NaN
// propagates), into part[m * kNosSlices + slice].
template <int WHAT>
__global__ __launch_bounds__(kBlock) void nos_slice_kernel(const double *S, int k, double *part)
{
    __shared__ double s_red[kBlock / kWave];
    const size_t kk = (size_t)k * k;
    const double *M = S + kk * blockIdx.y;
    const size_t per = (kk + gridDim.x - 1) / gridDim.x;
    const size_t i0 = per * blockIdx.x, i1 = i0 + per < kk ? i0 + per : kk;
    double acc = 0.0;
    for (size_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
        if (WHAT == 0) {
            acc += M[i] * M[i];
        } else {
            const double v = fabs(M[i] - ((i % k) == (i / k) ? 1.0 : 0.0));
            acc = (v > acc || v != v) ? v : acc;
        }
    }
    if (WHAT == 0) {
        acc = block_sum_d(acc, s_red);
    } else {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double v = __shfl_xor(acc, o, kWave);
            acc = (v > acc || v != v) ? v : acc;
        }
        __syncthreads();
        if ((threadIdx.x & 63) == 0)
            s_red[threadIdx.x >> 6] = acc;
        __syncthreads();
        acc = s_red[0];
        for (int w = 1; w < kBlock / kWave; w++)
            acc = (s_red[w] > acc || s_red[w] != s_red[w]) ? s_red[w] : acc;
    }
    if (threadIdx.x == 0)
        part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = acc;
}

// Scales of the first step from the Gram matrices' squared Frobenius norms
// f2 (sums of the slices, in order): s = ||G||_F^-1 normalises S_0 = s G
// (every eigenvalue in [0, 1]) and r = ||G||_F^(-1/2) scales X_0 = A r; both
// 0 for a zero matrix, which then never converges.
__global__ void nos_scales_kernel(const double *part, int slices, int count, double *s,
                                  double *r)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < count) {
        double f2 = 0.0;
        for (int j = 0; j < slices; j++)
            f2 += part[(size_t)m * slices + j];
        const double f = sqrt(f2);
        s[m] = f > 0.0 ? 1.0 / f : 0.0;
        r[m] = f > 0.0 ? 1.0 / sqrt(f) : 0.0;
    }
}

// The elementwise part of Q = q(S) = c0 I + c1 S + c2 S^2 + c3 S^3 + c4 S^4
// (Paterson-Stockmeyer: Q = (c0 I + c1 S + c2 S^2) + S^2 (c3 S + c4 S^2),
// the product added by a GEMM).  In the first step S and S2 hold G and G G
// unnormalised, S_0 = s G, and Q is scaled by r (X_1 = A r q(S_0)):
// a = s S, b = s^2 S2, Q = r (c0 I + c1 a + c2 b) and T = s^2 r (c3 a +
// c4 b), so that S2 T = S_0^2 r (c3 S_0 + c4 S_0^2).  Later steps: s = r = 1
// (null pointers).
__global__ __launch_bounds__(kBlock) void nos_poly_kernel(const double *S, const double *S2,
                                                         int k, int count, PolarStep c,
                                                         const double *scale, const double *r,
                                                         double *T, double *Q)
{
    const size_t kk = (size_t)k * k, total = kk * count;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (size_t)gridDim.x * kBlock) {
        const size_t m = i / kk, e = i - m * kk;
        const double sc = scale ? scale[m] : 1.0, rr = r ? r[m] : 1.0;
        const double a = sc * S[i], b = sc * sc * S2[i];
        const double d = (e % k) == (e / k) ? c.c[0] : 0.0;
        T[i] = sc * sc * rr * (c.c[3] * a + c.c[4] * b);
        Q[i] = rr * (d + c.c[1] * a + c.c[2] * b);
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

static size_t circus_shm(int nmax)
{
    int m2 = 1;
    while (m2 < nmax)
        m2 <<= 1;
    const size_t n4 = ((size_t)nmax + 3) & ~(size_t)3;
    return sizeof(float) * (2 * n4 + m2);
}

hipError_t launch_circus(const CircusParams &p, hipStream_t s)
{
    // P2 of columns up to kP2Max rows: one wave per column
    CircusParams c = p;
    c.p2_wave = 0;
    bool other = false;
    const int cols = p.ncols * p.nsino;
    const int wpb = kBlock / kWave;
    for (int q = 0; q < p.np; q++) {
        if (p.code[q] == 2 && p.n <= kP2Max) {
            circus_p2_kernel<<<(cols + wpb - 1) / wpb, kBlock, 0, s>>>(p, q);
            c.p2_wave = 1;
        } else {
            other = true;
        }
    }
    if (other)
        circus_kernel<<<cols, kBlock, circus_shm(p.n), s>>>(c);
    return hipGetLastError();
}


hipError_t launch_p2_source(const float *sino, int n, int cols, int32_t *src, hipStream_t s)
{
    if (n > kP2Max)
        return hipErrorInvalidValue;
    const int wpb = kBlock / kWave;
    p2_source_kernel<<<(cols + wpb - 1) / wpb, kBlock, 0, s>>>(sino, n, cols, src);
    return hipGetLastError();
}

hipError_t launch_p2_source_batched(const BatchCircusParams &p, hipStream_t s)
{
    if (p.nmax > kP2Max)
        return hipErrorInvalidValue;
    BatchCircusParams c = p;
    c.with_circus = 0;
    const int wpb = kBlock / kWave;
    circus_p2_batched_kernel<<<(p.cols + wpb - 1) / wpb, kBlock, 0, s>>>(c, 0);
    return hipGetLastError();
}

hipError_t launch_circus_batched(const BatchCircusParams &p, hipStream_t s)
{
    // P2 of columns up to kP2Max rows: one wave per column
    BatchCircusParams c = p;
    c.p2_wave = 0;
    bool other = false;
    const int wpb = kBlock / kWave;
    for (int q = 0; q < p.np; q++) {
        if (p.code[q] == 2 && p.nmax <= kP2Max) {
            circus_p2_batched_kernel<<<(p.cols + wpb - 1) / wpb, kBlock, 0, s>>>(p, q);
            c.p2_wave = 1;
        } else {
            other = true;
        }
    }
    if (other)
        circus_batched_kernel<<<p.cols, kBlock, circus_shm(p.nmax), s>>>(c);
    return hipGetLastError();
}

hipError_t launch_transpose_batched(const BatchImage *d_images, int nimg, int nmax,
                                    hipStream_t s)
{
    const int tiles = (nmax + 31) / 32;
    int gx = tiles * tiles;
    if (gx > 1024)
        gx = 1024;
    dim3 grid(gx, nimg);
    transpose_batched_kernel<<<grid, kBlock, 0, s>>>(d_images);
    return hipGetLastError();
}

hipError_t launch_reset_counter(unsigned int *counter, hipStream_t s)
{
    reset_counter_kernel<<<1, kWave, 0, s>>>(counter);
    return hipGetLastError();
}

hipError_t launch_sino_t05_batched(const BatchSinoParams &p, bool need_x,
                                   bool need_s, bool do345, int waves_per_block,
                                   int blocks, hipStream_t s)
{
    if (!line_tile_ok(p.tile_a, p.tile_c))
        return hipErrorInvalidValue;
    const int threads = kWave * waves_per_block;
    const size_t shm = do345 ? sizeof(float2) * 3 * (size_t)p.tabn : 0;
#define TT_LAUNCH(X, S, D)                                                     \
    sino_t05_batched_kernel<X, S, D><<<blocks, threads, shm, s>>>(p)
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

// Workgroups of one sort pass that fit on the device at once (the waves are
// persistent, so more would only wait).
template <int KMAX, int KMIN, bool DEFER, bool IDX7>
static int sort_blocks(int cus, int items = 0x7fffffff)
{
    static int per_cu = 0;
    if (per_cu == 0) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &b, sino_sort_kernel<KMAX, KMIN, DEFER, IDX7>, kBlock, 0) != hipSuccess || b < 1)
            b = 2;
        per_cu = b;
    }
    const int need = (int)(((long long)items + kBlock / kWave - 1) / (kBlock / kWave));
    return cus * per_cu < need ? cus * per_cu : need;
}

template <bool IDX7>
static hipError_t sino_sort_passes(const SortTParams &p_in, int32_t *queue_a,
                            int32_t *queue_b, unsigned int *counts, hipStream_t s)
{
    SortTParams p = p_in;
    p.defer_in = nullptr;
    p.defer_in_count = nullptr;
    if (p.n > 2048 || !counts)
        return hipErrorInvalidValue;
    const int lines = p.A * p.n;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // counts: [0], [1] queue lengths, [2..4] the passes' work counters
    reset_counters_kernel<<<1, kWave, 0, s>>>(counts, kSortCounters);
    p.work = counts + 2;
    p.chunk = 8;
    if (p.n <= 512) {
#define TT_SORT(K) sino_sort_kernel<K, 1, false, IDX7><<<sort_blocks<K, 1, false, IDX7>(cus, lines), kBlock, 0, s>>>(p)
        if (p.n <= 64)
            TT_SORT(1);
        else if (p.n <= 128)
            TT_SORT(2);
        else if (p.n <= 256)
            TT_SORT(4);
        else
            TT_SORT(8);
#undef TT_SORT
        return hipGetLastError();
    }
    // sorted parts up to 512 keys in the first pass, up to 1024 in the
    // second, the rest in the third; each pass queues what it cannot hold
    if (!queue_a || !queue_b)
        return hipErrorInvalidValue;
    p.defer_out = queue_a;
    p.defer_out_count = counts;
    sino_sort_kernel<8, 1, true, IDX7><<<sort_blocks<8, 1, true, IDX7>(cus, lines), kBlock, 0, s>>>(p);
    p.defer_in = queue_a;
    p.defer_in_count = counts;
    p.work = counts + 3;
    p.chunk = 2;
    if (p.n <= 1024) {
        sino_sort_kernel<16, 16, false, IDX7><<<sort_blocks<16, 16, false, IDX7>(cus), kBlock, 0, s>>>(p);
        return hipGetLastError();
    }
    p.defer_out = queue_b;
    p.defer_out_count = counts + 1;
    sino_sort_kernel<16, 16, true, IDX7><<<sort_blocks<16, 16, true, IDX7>(cus), kBlock, 0, s>>>(p);
    p.defer_in = queue_b;
    p.defer_in_count = counts + 1;
    p.work = counts + 4;
    p.chunk = 1;
    sino_sort_kernel<32, 32, false, IDX7><<<sort_blocks<32, 32, false, IDX7>(cus), kBlock, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launch_sino_sort(const SortTParams &p, int32_t *queue_a, int32_t *queue_b,
                            unsigned int *counts, hipStream_t s)
{
    return p.t7_src ? sino_sort_passes<true>(p, queue_a, queue_b, counts, s)
                    : sino_sort_passes<false>(p, queue_a, queue_b, counts, s);
}

}  // namespace ttk

namespace ttk {

hipError_t launch_records(const RecordParams &p, hipStream_t s)
{
    const size_t lines = (size_t)p.A * p.n;
    int blocks = (int)((lines + kBlock - 1) / kBlock);
    if (blocks > 8192)
        blocks = 8192;
    records_kernel<<<blocks, kBlock, 0, s>>>(p);
    return hipGetLastError();
}

size_t gather_lds_bytes(int tile, int waves, int tabn, bool do345, bool do12, bool do67)
{
    const size_t tab = do345 ? (size_t)kGatherTabWords * tabn : 0;
    return sizeof(float) * (tab + (size_t)waves * gather_wave_words(tile, do12, do67));
}

hipError_t launch_gather(const GatherParams &p, hipStream_t s)
{
    if (p.tile != kGatherTile || p.waves < 1 || p.waves > 16 || p.groups < 1 ||
        (p.groups > 1 && !p.part))
        return hipErrorInvalidValue;
    const int tn = (p.n + p.tile - 1) / p.tile;
    const int items = tn * tn * p.groups;
    const int blocks = (items + p.waves - 1) / p.waves;
    const size_t shm = gather_lds_bytes(p.tile, p.waves, p.n, p.do345, p.do12,
                                        p.rec67 != nullptr);
    const dim3 block(kWave * p.waves);
    if (p.do345) {
        if (p.do12)
            gather_kernel<true, true><<<blocks, block, shm, s>>>(p);
        else
            gather_kernel<true, false><<<blocks, block, shm, s>>>(p);
    } else {
        if (p.do12)
            gather_kernel<false, true><<<blocks, block, shm, s>>>(p);
        else
            gather_kernel<false, false><<<blocks, block, shm, s>>>(p);
    }
    if (p.groups > 1) {
        const size_t npix = (size_t)p.n * p.n;
        size_t gb = (npix + kBlock - 1) / kBlock;
        gb = gb > 4096 ? 4096 : gb;
        gather_reduce_kernel<<<(int)gb, kBlock, 0, s>>>(p.part, p.groups, npix, p.dimg);
    }
    return hipGetLastError();
}

static size_t vjp_shm(int nmax)
{
    int m2 = 1;
    while (m2 < nmax)
        m2 <<= 1;
    const size_t n4 = ((size_t)nmax + 3) & ~(size_t)3;
    return sizeof(uint64_t) * m2 + sizeof(float) * 3 * n4;
}

hipError_t launch_circus_vjp(const CircusVjpParams &p, hipStream_t s)
{
    circus_vjp_kernel<<<p.ncols * p.nsino, kBlock, vjp_shm(p.n), s>>>(p);
    return hipGetLastError();
}

hipError_t launch_circus_vjp_batched(const BatchAdjointParams &p, hipStream_t s)
{
    circus_vjp_batched_kernel<<<p.cols, kBlock, vjp_shm(p.nmax), s>>>(p);
    return hipGetLastError();
}

hipError_t launch_records_batched(const BatchAdjointParams &p, hipStream_t s)
{
    int blocks = (p.lines + kBlock - 1) / kBlock;
    if (blocks > 8192)
        blocks = 8192;
    records_batched_kernel<<<blocks, kBlock, 0, s>>>(p);
    return hipGetLastError();
}

hipError_t launch_gather_batched(const BatchAdjointParams &p, int tile, int waves,
                                 hipStream_t s)
{
    if ((tile != 32 && tile != 16) || waves < 1 || waves > 16)
        return hipErrorInvalidValue;
    const bool do12 = p.slot.s[1] >= 0 || p.slot.s[2] >= 0;
    const size_t shm = gather_lds_bytes(tile, waves, p.tabn, p.do345, do12, false);
    const int blocks = (p.tiles + waves - 1) / waves;
    const dim3 block(kWave * waves);
#define TT_LAUNCH(T)                                                                   \
    do {                                                                               \
        if (p.do345 && do12)                                                           \
            gather_batched_kernel<T, true, true><<<blocks, block, shm, s>>>(p);        \
        else if (p.do345)                                                              \
            gather_batched_kernel<T, true, false><<<blocks, block, shm, s>>>(p);       \
        else if (do12)                                                                 \
            gather_batched_kernel<T, false, true><<<blocks, block, shm, s>>>(p);       \
        else                                                                           \
            gather_batched_kernel<T, false, false><<<blocks, block, shm, s>>>(p);      \
    } while (0)
    if (tile == 32)
        TT_LAUNCH(32);
    else
        TT_LAUNCH(16);
#undef TT_LAUNCH
    return hipGetLastError();
}

}  // namespace ttk

namespace ttk {

static int grid_for(size_t total)
{
    size_t b = (total + kBlock - 1) / kBlock;
    return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

hipError_t launch_short_sqrt_check(uint32_t lo, uint64_t count, unsigned long long *mism,
                                   unsigned int *first, hipStream_t s)
{
    short_sqrt_check_kernel<<<8192, kBlock, 0, s>>>(lo, count, mism, first);
    return hipGetLastError();
}

hipError_t launch_resize(const float *in, int in_rows, int in_cols, float *out,
                         int rows, int cols, hipStream_t s)
{
    resize_kernel<<<grid_for((size_t)rows * cols), kBlock, 0, s>>>(in, in_rows, in_cols,
                                                                   out, rows, cols);
    return hipGetLastError();
}

hipError_t launch_zscore(const float *in, int n, int count, float *out, hipStream_t s)
{
    zscore_kernel<<<count, kBlock, 0, s>>>(in, n, out);
    return hipGetLastError();
}

hipError_t launch_column_medians(const float *in, int n, int cols, int32_t *med,
                                 hipStream_t s)
{
    if (n <= kP2Max) {
        const int wpb = kBlock / kWave;
        column_median_wave_kernel<<<(cols + wpb - 1) / wpb, kBlock, 0, s>>>(in, n, cols, med);
        return hipGetLastError();
    }
    const size_t shm = sizeof(float) * (((size_t)n + 3) & ~(size_t)3);
    column_median_kernel<<<cols, kBlock, shm, s>>>(in, n, med);
    return hipGetLastError();
}

hipError_t launch_image_bbox(const float *img, int n, int *bbox, hipStream_t s)
{
    bbox_init_kernel<<<1, kWave, 0, s>>>(bbox, n);
    // few blocks: each ends with up to five atomics on the same words, which
    // serialise (256 blocks took 13.8 us on a 363 x 363 image, 64 take 7);
    // large images keep 256 so that each thread's strided loop stays short
    const int nb = n <= 512 ? (n < 64 ? n : 64) : 256;
    image_bbox_kernel<<<nb, kBlock, 0, s>>>(img, n, bbox);
    return hipGetLastError();
}

hipError_t launch_image_bbox_batched(const BatchImage *images, int nimg, int nmax,
                                     hipStream_t s)
{
    bbox_init_batched_kernel<<<(nimg + kBlock - 1) / kBlock, kBlock, 0, s>>>(images, nimg);
    image_bbox_batched_kernel<<<dim3(nmax < 64 ? nmax : 64, nimg), kBlock, 0, s>>>(images);
    return hipGetLastError();
}

hipError_t launch_scale_eigvec(const double *v, const double *lambda, int k, int count,
                               double *out, hipStream_t s)
{
    scale_eigvec_kernel<<<grid_for((size_t)k * k * count), kBlock, 0, s>>>(v, lambda, k,
                                                                          count, out);
    return hipGetLastError();
}

hipError_t launch_newton_schulz_factor(double *x, int k, int count, hipStream_t s)
{
    newton_schulz_factor_kernel<<<grid_for((size_t)k * k * count), kBlock, 0, s>>>(x, k,
                                                                                  count);
    return hipGetLastError();
}

hipError_t launch_nos_reduce(const double *part, int chunks, int k, int count, double *S,
                             hipStream_t s)
{
    const size_t kk = (size_t)k * k;
    nos_reduce_kernel<<<grid_for(kk * count), kBlock, 0, s>>>(part, chunks, kk, count, S);
    return hipGetLastError();
}

hipError_t launch_nos_scales(const double *G, int k, int count, double *part, double *scale,
                             double *r, hipStream_t s)
{
    nos_slice_kernel<0><<<dim3(kNosSlices, count), kBlock, 0, s>>>(G, k, part);
    nos_scales_kernel<<<(count + kWave - 1) / kWave, kWave, 0, s>>>(part, kNosSlices, count,
                                                                    scale, r);
    return hipGetLastError();
}

hipError_t launch_nos_deviation(const double *S, int k, int count, double *part,
                                hipStream_t s)
{
    nos_slice_kernel<1><<<dim3(kNosSlices, count), kBlock, 0, s>>>(S, k, part);
    return hipGetLastError();
}

hipError_t launch_nos_poly(const double *S, const double *S2, int k, int count,
                           const PolarStep &c, const double *scale, const double *r, double *T,
                           double *Q, hipStream_t s)
{
    nos_poly_kernel<<<grid_for((size_t)k * k * count), kBlock, 0, s>>>(S, S2, k, count, c,
                                                                       scale, r, T, Q);
    return hipGetLastError();
}


hipError_t launch_align_batched(const float *in, int n, int cols, const int32_t *med,
                                int count, int rows_out, size_t mstride, double *out,
                                hipStream_t s)
{
    const int bx = (grid_for((size_t)rows_out * cols) + count - 1) / count;
    align_batched_kernel<<<dim3(bx, count), kBlock, 0, s>>>(in, n, cols, med, rows_out,
                                                             mstride, out);
    return hipGetLastError();
}

hipError_t launch_to_float_batched(const double *in, int ld, size_t mstride, int cols,
                                   const int32_t *med, int n, int sc, int count, int rows_max,
                                   float *out, size_t ostride, hipStream_t s)
{
    const int bx = (grid_for((size_t)rows_max * cols) + count - 1) / count;
    to_float_batched_kernel<<<dim3(bx, count), kBlock, 0, s>>>(in, ld, mstride, cols, med, n,
                                                                sc, out, ostride);
    return hipGetLastError();
}


}  // namespace ttk
