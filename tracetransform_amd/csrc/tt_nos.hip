// tt_nos.hip -- FP64 products of the nearest-orthonormal-sinogram polar
// iteration on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// The iteration (c++/src/circus.cpp:94-99 takes the polar factor U V^T of
// the aligned sinogram) runs X <- X q(X^T X).  Every 360 x 360 product in
// it multiplies two polynomials of the same Gram matrix, so its result is
// symmetric: the symmetric mode computes only the 64 x 64 tiles on and
// below the diagonal (21 of 36 for 360 columns) and writes each value to
// both (i, j) and (j, i), which also makes the results exactly symmetric.
// The Gram matrices X^T X are products of that kind too (A transposed).
//
// Two kernels.  nos_gemm_kernel (the Gram matrices, and any product whose
// reduction length is not a multiple of 4): one workgroup of four waves
// owns a 64 x 64 output tile, each wave a 32 x 32 quarter as 2 x 2 MFMA
// blocks, the reduction in chunks of 32 staged through LDS (double
// buffered, one barrier per chunk).  nos_gemm_symb_kernel (products with a
// symmetric right operand): direct operand loads and the reduction split
// over the workgroup's waves, described below.  Matrices are column major,
// as in BLAS.
#include "tt_internal.hpp"

namespace ttk {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
// Pointers into global memory: loads and stores through them are global,
// not flat, instructions, so waiting for the operand prefetch and waiting
// for the LDS reads of the MFMA operands are separate.
typedef const double __attribute__((address_space(1))) *gcdouble;
typedef double __attribute__((address_space(1))) *gdouble;

constexpr int kNgB = 64;      // output tile edge
constexpr int kNgK = 32;      // reduction chunk
constexpr int kNgThreads = 256;
constexpr int kNgPer = kNgB * kNgK / kNgThreads;  // operand entries per thread
// LDS layouts of a 64 x 32 operand chunk.  Operands contiguous along the
// reduction (B, and A when transposed) keep that order: row r (an output
// row or column) at r * kNgRowK, with kNgRowK = 34 doubles (68 dwords, 4
// modulo the 64 banks) so the 16 rows and 2 reduction steps of an MFMA
// operand read land in distinct banks.  A untransposed keeps its column
// order: reduction step kk at kk * kNgRowM, kNgRowM = 80 doubles, so rows
// 16 lanes apart fall in the other half of the banks.
constexpr int kNgRowK = kNgK + 2;
constexpr int kNgRowM = kNgB + 16;
constexpr int kNgSizeA = kNgB * kNgRowK > kNgK * kNgRowM ? kNgB * kNgRowK : kNgK * kNgRowM;
constexpr int kNgSizeB = kNgB * kNgRowK;

template <bool TA, bool SYM>
__global__ __launch_bounds__(kNgThreads) void nos_gemm_kernel(NosGemmBatch bt, int m, int n,
                                                              int k, int lda, int ldb, int ldc,
                                                              double alpha, double beta,
                                                              int tiles_m)
{
    __shared__ double s_a[2][kNgSizeA];
    __shared__ double s_b[2][kNgSizeB];
    int bi, bj;
    if (SYM) {  // lower-triangle tile t -> (bi, bj), bj <= bi
        int t = blockIdx.x;
        bi = 0;
        while (t > bi) {
            t -= bi + 1;
            bi++;
        }
        bj = t;
    } else {
        bi = blockIdx.x % tiles_m;
        bj = blockIdx.x / tiles_m;
    }
    gcdouble A, B;
    gdouble C;
    if (bt.table) {
        typedef const NosGemmJob __attribute__((address_space(1))) *gcjob;
        const gcjob j = (gcjob)bt.table + blockIdx.y;
        A = (gcdouble)j->a;
        B = (gcdouble)j->b;
        C = (gdouble)j->c;
    } else {
        A = (gcdouble)(bt.a + (size_t)blockIdx.y * bt.sa);
        B = (gcdouble)(bt.b + (size_t)blockIdx.y * bt.sb);
        C = (gdouble)(bt.c + (size_t)blockIdx.y * bt.sc);
    }
    const int i0 = bi * kNgB, j0 = bj * kNgB;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w & 1, wn = w >> 1;
    double ra[kNgPer], rb[kNgPer];
    // operand chunk k0..k0+31 into registers (zero outside the matrices):
    // op(A)(i, kk) and B(kk, j), each read along its contiguous direction
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < kNgPer; q++) {
            const int idx = tid + kNgThreads * q;
            if (TA) {
                const int gi = i0 + (idx >> 5), gk = k0 + (idx & 31);
                ra[q] = gi < m && gk < k ? A[gk + (size_t)gi * lda] : 0.0;
            } else {
                const int gi = i0 + (idx & 63), gk = k0 + (idx >> 6);
                ra[q] = gi < m && gk < k ? A[gi + (size_t)gk * lda] : 0.0;
            }
            const int gj = j0 + (idx >> 5), gbk = k0 + (idx & 31);
            rb[q] = gj < n && gbk < k ? B[gbk + (size_t)gj * ldb] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kNgPer; q++) {
            const int idx = tid + kNgThreads * q;
            if (TA)
                s_a[buf][(idx >> 5) * kNgRowK + (idx & 31)] = ra[q];
            else
                s_a[buf][(idx >> 6) * kNgRowM + (idx & 63)] = ra[q];
            s_b[buf][(idx >> 5) * kNgRowK + (idx & 31)] = rb[q];
        }
    };
    d4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
            acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
    const int nk = (k + kNgK - 1) / kNgK;
    load(0);
    store(0);
    __syncthreads();
    // MFMA operands: lane l holds op(A)(l % 16, l / 16) and B(l / 16, l % 16)
    // of a 16 x 4 by 4 x 16 block
    const int r16 = lane & 15, k4 = lane >> 4;
    const int oa = TA ? (wm * 32 + r16) * kNgRowK + k4 : k4 * kNgRowM + wm * 32 + r16;
    const int oa1 = TA ? 16 * kNgRowK : 16;
    const int sa = TA ? 4 : 4 * kNgRowM;
    const int ob = (wn * 32 + r16) * kNgRowK + k4;
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk)
            load((kt + 1) * kNgK);
        const double *pa = s_a[cur] + oa, *pb = s_b[cur] + ob;
#pragma unroll
        for (int s = 0; s < kNgK / 4; s++) {
            const double a0 = pa[s * sa], a1 = pa[s * sa + oa1];
            const double b0 = pb[s * 4], b1 = pb[s * 4 + 16 * kNgRowK];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (kt + 1 < nk)
            store(cur ^ 1);
        __syncthreads();
    }
    // result block (r, c): lane l, entry v holds row 4 v + l / 16, column
    // l % 16
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = i0 + wm * 32 + r * 16 + 4 * v + k4;
                const int col = j0 + wn * 32 + c * 16 + r16;
                if (row >= m || col >= n || (SYM && row < col))
                    continue;
                const size_t o = row + (size_t)col * ldc;
                const double val = beta != 0.0 ? alpha * acc[r][c][v] + beta * C[o]
                                                : alpha * acc[r][c][v];
                C[o] = val;
                if (SYM && row != col)
                    C[col + (size_t)row * ldc] = val;
            }
}

// Products whose right operand B is symmetric (every product but the Gram
// matrices): B(kk, j) = B(j, kk) is read along j, so each wave loads its
// MFMA operands straight from global memory, 16 consecutive doubles per
// quarter wave, keeping kNsDepth reduction steps in flight.  A workgroup
// owns a 32 x 32 output block (2 x 2 MFMA blocks); its four waves take
// consecutive quarters of the reduction and wave 0 adds the other three's
// partial sums from LDS in wave order.  Workgroup g takes matrix g % count:
// workgroups go to the eight XCDs in turn, so a batch of eight keeps each
// matrix's operands in one XCD's L2.
constexpr int kNsDepth = 4;
constexpr int kNsWaves = kNgThreads / 64;

// POLY: the product is S S of a polynomial step (A = B = S, strided batch);
// the epilogue also forms the step's elementwise terms from S and S2 = S S:
// with a = s S, b = s^2 S2 (s, r the first step's scales, else 1),
// T = s^2 r (c3 a + c4 b) and Q = r (c0 I + c1 a + c2 b), so that
// Q + S2 T = q(S).
template <bool SYM, bool POLY>
__global__ __launch_bounds__(kNgThreads) void nos_gemm_symb_kernel(NosGemmBatch bt, int count,
                                                                   int m, int n, int k,
                                                                   int lda, int ldb, int ldc,
                                                                   double alpha, double beta,
                                                                   int bm, NosPolyOut po)
{
    __shared__ double s_part[kNsWaves - 1][16][64];
    const int job = blockIdx.x % count;
    const int blk = blockIdx.x / count;
    int bi, bj;
    if (SYM) {
        int t = blk;
        bi = 0;
        while (t > bi) {
            t -= bi + 1;
            bi++;
        }
        bj = t;
    } else {
        bi = blk % bm;
        bj = blk / bm;
    }
    gcdouble A, B;
    gdouble C;
    if (bt.table) {
        typedef const NosGemmJob __attribute__((address_space(1))) *gcjob;
        const gcjob j = (gcjob)bt.table + job;
        A = (gcdouble)j->a;
        B = (gcdouble)j->b;
        C = (gdouble)j->c;
    } else {
        A = (gcdouble)(bt.a + (size_t)job * bt.sa);
        B = (gcdouble)(bt.b + (size_t)job * bt.sb);
        C = (gdouble)(bt.c + (size_t)job * bt.sc);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, k4 = lane >> 4;
    // this wave's reduction steps (of 4): [s_begin, s_end)
    const int nsteps = k / 4;
    const int s_begin = nsteps * w / kNsWaves, s_end = nsteps * (w + 1) / kNsWaves;
    const int ns = s_end - s_begin;
    // rows of A and columns of B past the matrix only feed outputs past it
    // (never stored): their addresses are clamped to the last row / column
    const int ia0 = min(bi * 32 + r16, m - 1), ia1 = min(bi * 32 + 16 + r16, m - 1);
    const int jb0 = min(bj * 32 + r16, n - 1), jb1 = min(bj * 32 + 16 + r16, n - 1);
    const size_t k0 = (size_t)(4 * s_begin + k4);
    const gcdouble pa0 = A + ia0 + k0 * lda, pa1 = A + ia1 + k0 * lda;
    const gcdouble pb0 = B + jb0 + k0 * ldb, pb1 = B + jb1 + k0 * ldb;
    const size_t sa = (size_t)4 * lda, sb = (size_t)4 * ldb;
    double a0[kNsDepth], a1[kNsDepth], b0[kNsDepth], b1[kNsDepth];
#pragma unroll
    for (int d = 0; d < kNsDepth; d++)
        if (d < ns) {
            a0[d] = pa0[d * sa];
            a1[d] = pa1[d * sa];
            b0[d] = pb0[d * sb];
            b1[d] = pb1[d * sb];
        }
    d4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
            acc[r][c] = d4{0.0, 0.0, 0.0, 0.0};
    for (int s0 = 0; s0 < ns; s0 += kNsDepth) {
#pragma unroll
        for (int d = 0; d < kNsDepth; d++) {
            const int st = s0 + d;
            if (st < ns) {
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[d], b0[d], acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[d], b1[d], acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[d], b0[d], acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[d], b1[d], acc[1][1], 0, 0, 0);
                if (st + kNsDepth < ns) {
                    const size_t u = (size_t)(st + kNsDepth);
                    a0[d] = pa0[u * sa];
                    a1[d] = pa1[u * sa];
                    b0[d] = pb0[u * sb];
                    b1[d] = pb1[u * sb];
                }
            }
        }
    }
    if (w > 0) {
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    s_part[w - 1][(r * 2 + c) * 4 + v][lane] = acc[r][c][v];
    }
    __syncthreads();
    if (w > 0)
        return;
#pragma unroll
    for (int q = 0; q < kNsWaves - 1; q++)
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    acc[r][c][v] += s_part[q][(r * 2 + c) * 4 + v][lane];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = bi * 32 + r * 16 + 4 * v + k4;
                const int col = bj * 32 + c * 16 + r16;
                if (row >= m || col >= n || (SYM && row < col))
                    continue;
                const size_t o = row + (size_t)col * ldc;
                const double val = beta != 0.0 ? alpha * acc[r][c][v] + beta * C[o]
                                                : alpha * acc[r][c][v];
                C[o] = val;
                if (SYM && row != col)
                    C[col + (size_t)row * ldc] = val;
                if (POLY) {
                    const double sc = po.scale ? po.scale[job] : 1.0;
                    const double rr = po.r ? po.r[job] : 1.0;
                    const double a = sc * A[row + (size_t)col * lda], b = sc * sc * val;
                    const double d = row == col ? po.c[0] : 0.0;
                    const double tv = sc * sc * rr * (po.c[3] * a + po.c[4] * b);
                    const double qv = rr * (d + po.c[1] * a + po.c[2] * b);
                    const size_t off = (size_t)job * po.stride;
                    po.T[off + o] = tv;
                    po.Q[off + o] = qv;
                    if (row != col) {
                        const size_t om = col + (size_t)row * ldc;
                        po.T[off + om] = tv;
                        po.Q[off + om] = qv;
                    }
                }
            }
}

}  // namespace

hipError_t launch_nos_gemm(const NosGemmBatch &bt, int count, bool trans_a, bool sym, int m,
                           int n, int k, int lda, int ldb, int ldc, double alpha, double beta,
                           hipStream_t s)
{
    if (count <= 0 || m <= 0 || n <= 0 || k <= 0)
        return hipErrorInvalidValue;
    if (sym && m != n)
        return hipErrorInvalidValue;
    if (!trans_a && k % 4 == 0) {  // B symmetric: the direct-load kernel
        const int bm = (m + 31) / 32, bn = (n + 31) / 32;
        const int nblocks = sym ? bm * (bm + 1) / 2 : bm * bn;
        const dim3 grid((unsigned)(nblocks * count));
        const NosPolyOut none{};
        if (sym)
            nos_gemm_symb_kernel<true, false><<<grid, kNgThreads, 0, s>>>(
                bt, count, m, n, k, lda, ldb, ldc, alpha, beta, bm, none);
        else
            nos_gemm_symb_kernel<false, false><<<grid, kNgThreads, 0, s>>>(
                bt, count, m, n, k, lda, ldb, ldc, alpha, beta, bm, none);
        return hipGetLastError();
    }
    const int tm = (m + kNgB - 1) / kNgB, tn = (n + kNgB - 1) / kNgB;
    constexpr int kMaxGridY = 65535;  // the batch runs along grid y
    for (int c0 = 0; c0 < count; c0 += kMaxGridY) {
        NosGemmBatch part = bt;
        if (bt.table) {
            part.table = bt.table + c0;
        } else {
            part.a = bt.a + (size_t)c0 * bt.sa;
            part.b = bt.b + (size_t)c0 * bt.sb;
            part.c = bt.c + (size_t)c0 * bt.sc;
        }
        const int cnt = count - c0 < kMaxGridY ? count - c0 : kMaxGridY;
        const dim3 grid(sym ? tm * (tm + 1) / 2 : tm * tn, cnt);
        if (trans_a) {
            if (sym)
                nos_gemm_kernel<true, true><<<grid, kNgThreads, 0, s>>>(part, m, n, k, lda, ldb,
                                                                       ldc, alpha, beta, tm);
            else
                nos_gemm_kernel<true, false><<<grid, kNgThreads, 0, s>>>(part, m, n, k, lda, ldb,
                                                                        ldc, alpha, beta, tm);
        } else {
            if (sym)
                nos_gemm_kernel<false, true><<<grid, kNgThreads, 0, s>>>(part, m, n, k, lda, ldb,
                                                                        ldc, alpha, beta, tm);
            else
                nos_gemm_kernel<false, false><<<grid, kNgThreads, 0, s>>>(part, m, n, k, lda,
                                                                         ldb, ldc, alpha, beta,
                                                                         tm);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launch_nos_square_poly(const double *S, double *S2, int k, int count,
                                  const NosPolyOut &po, hipStream_t s)
{
    if (count <= 0 || k <= 0 || k % 4 != 0 || po.stride != (long long)k * k)
        return hipErrorInvalidValue;
    const int bm = (k + 31) / 32, nblocks = bm * (bm + 1) / 2;
    const NosGemmBatch bt{nullptr, S, S, S2, po.stride, po.stride, po.stride};
    nos_gemm_symb_kernel<true, true><<<dim3((unsigned)(nblocks * count)), kNgThreads, 0, s>>>(
        bt, count, k, k, k, k, k, k, 1.0, 0.0, bm, po);
    return hipGetLastError();
}

}  // namespace ttk
