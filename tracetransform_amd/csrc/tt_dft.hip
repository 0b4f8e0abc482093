// tt_dft.hip -- P3's discrete Fourier transforms as GEMMs on the CDNA4
// matrix cores (v_mfma_f32_32x32x2_f32), with the DFT matrix generated tile
// by tile from a twiddle table in LDS instead of being read from memory.
//
// P3 (c++/src/functionals.cpp:332-360) needs the half spectrum
// X_k = sum_j x_j e^{-2 pi i jk/n}, k = 0..n/2, of every sinogram column:
//   spec (2h x cols) = W^T x (n x cols),   h = n/2 + 1,
// with W (n x 2h) = [cos(2 pi jk/n) | -sin(2 pi jk/n)].  Its adjoint applies
// the same matrix the other way: ds (n x cols) = W y (2h x cols), y the
// spectrum scaled per column and frequency.  Both are one kernel over a list
// of jobs (one per image of a batch), so the single-image and the batched
// path form their spectra with the same operations in the same order and
// give the same P3 bit for bit.
#include "tt_internal.hpp"

#include <math.h>

namespace ttk {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Pointers into global memory (the job lists' pointers are generic): loads
// and stores through them are global, not flat, instructions, so their
// waits do not hold the LDS reads of the matrix operands.
typedef const float __attribute__((address_space(1))) *gcfloat;
typedef float __attribute__((address_space(1))) *gfloat;

// output rows x columns per workgroup: 128 x 128, 64 x 128 or 64 x 64
constexpr int kDftBK = 16;   // reduction chunk

// Twiddle index (u v) mod n of W(row, kk) for the first of four consecutive
// kk and the step between them: W(row, kk) is the cosine or minus the sine
// of 2 pi t / n.  FWD: out = W^T in, the output row m is a frequency (m < h:
// real part, else imaginary part of frequency m - h) and kk a position;
// otherwise out = W in, the output row is a position and kk runs over the
// 2h spectrum entries (consecutive kk share u, and v steps by one except
// across kk = h, handled by the caller).
__device__ __forceinline__ int phase_mod(int u, int v, int n, float inv_n)
{
    // u v < 2730^2 < 2^23: exact in float; the quotient is off by at most one
    const int uv = (int)__umul24((uint32_t)u, (uint32_t)v);
    int t = uv - (int)__umul24((uint32_t)(int)((float)uv * inv_n), (uint32_t)n);
    t = t < 0 ? t + n : t;
    return t >= n ? t - n : t;
}

template <bool FWD>
__device__ __forceinline__ float dft_weight(const float2 *tw, int n, int h, float inv_n,
                                            int row, int kk)
{
    int u, v;
    bool sn;
    if (FWD) {
        sn = row >= h;
        u = sn ? row - h : row;
        v = kk;
    } else {
        sn = kk >= h;
        u = row;
        v = sn ? kk - h : kk;
    }
    const float2 w = tw[phase_mod(u, v, n, inv_n)];
    return sn ? -w.y : w.x;
}

__device__ __forceinline__ int find_job(const DftJob *jobs, int njobs, int tile)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One BM x 128 output tile per workgroup of four waves (2 x 2 wave tiles of
// BM/2 x 64, BM/16 MFMA accumulators each; BM = 128 for large launches,
// 64 where 128-row tiles would leave CUs idle), reduction in chunks of 16,
// double buffered: while the matrix cores work on one chunk (16 or 32
// MFMAs per wave, the latter long enough to cover the next chunk's global
// loads) the waves generate the next W chunk from the LDS twiddles (BM/16
// consecutive entries per thread, the phase stepped by u) and load the next
// input chunk into registers.
// Twiddles (cos, sin)(2 pi t / n) are rounded from double sincospi.
template <bool FWD, int BM, int BN>
__global__ __launch_bounds__(kBlock) void dft_gemm_kernel(DftParams p)
{
    constexpr int EA = BM / 16;  // W entries per thread and chunk
    constexpr int MB = BM / 64;  // 32-row blocks per wave
    constexpr int NB = BN / 64;  // 32-column blocks per wave
    __shared__ float sA[2][kDftBK][BM];
    __shared__ float sB[2][kDftBK][BN];
    extern __shared__ float2 s_tw[];
    const DftJob &jb = p.jobs ? p.jobs[find_job(p.jobs, p.njobs, blockIdx.x)] : p.job;
    const int n = jb.n, h = n / 2 + 1;
    const int M = FWD ? 2 * h : n, Kd = FWD ? n : 2 * h;
    const int local = blockIdx.x - jb.tile0;
    const int mtiles = (M + BM - 1) / BM;
    const int m0 = (local % mtiles) * BM, c0 = (local / mtiles) * BN;
    const int tid = threadIdx.x;
    for (int i = tid; i < n; i += kBlock) {
        double sn, cn;
        sincospi(2.0 * (double)i / (double)n, &sn, &cn);
        s_tw[i] = float2{(float)cn, (float)sn};
    }
    const float inv_n = 1.0f / (float)n;
    const int lane = tid & (kWave - 1), wave = tid >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
    // W loader: row ar of the tile, reduction entries ak .. ak + EA - 1
    const int ar = tid % BM, ak = (tid / BM) * EA;
    const int arow = m0 + ar;
    // input loader: columns bc (and bc + 64) of the tile, entries bk .. bk + 3
    const int bc = tid & 63, bk = (tid >> 6) * 4;
    const int col0 = c0 + bc, col1 = c0 + 64 + bc;
    const gcfloat src0 = (gcfloat)(jb.in + (size_t)(col0 < jb.cols ? col0 : 0) * jb.ld_in);
    const gcfloat src1 = (gcfloat)(jb.in + (size_t)(col1 < jb.cols ? col1 : 0) * jb.ld_in);
    float ra[EA], rb0[4], rb1[4];
    auto fetch = [&](int k0) {
        const int kk0 = k0 + ak;
        if (arow < M && kk0 + EA - 1 < Kd && (FWD || kk0 + EA - 1 < h || kk0 >= h)) {
            // EA consecutive kk of one half: one modular product, then steps
            int u, v;
            bool sn;
            if (FWD) {
                sn = arow >= h;
                u = sn ? arow - h : arow;
                v = kk0;
            } else {
                sn = kk0 >= h;
                u = arow;
                v = sn ? kk0 - h : kk0;
            }
            int t = phase_mod(u, v, n, inv_n);
            const int du = u >= n ? u - n : u;
#pragma unroll
            for (int e = 0; e < EA; e++) {
                const float2 w = s_tw[t];
                ra[e] = sn ? -w.y : w.x;
                t += du;
                t = t >= n ? t - n : t;
            }
        } else {
#pragma unroll
            for (int e = 0; e < EA; e++) {
                const int kk = kk0 + e;
                ra[e] = arow < M && kk < Kd ? dft_weight<FWD>(s_tw, n, h, inv_n, arow, kk)
                                            : 0.0f;
            }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int kk = k0 + bk + e;
            rb0[e] = col0 < jb.cols && kk < Kd ? src0[kk] : 0.0f;
            if constexpr (NB == 2)
                rb1[e] = col1 < jb.cols && kk < Kd ? src1[kk] : 0.0f;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < EA; e++)
            sA[buf][ak + e][ar] = ra[e];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            sB[buf][bk + e][bc] = rb0[e];
            if constexpr (NB == 2)
                sB[buf][bk + e][64 + bc] = rb1[e];
        }
    };
    f32x16 acc00 = {}, acc01 = {}, acc10 = {}, acc11 = {};
    __syncthreads();  // twiddles
    fetch(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < Kd; k0 += kDftBK) {
        const bool more = k0 + kDftBK < Kd;
        if (more)
            fetch(k0 + kDftBK);
#pragma unroll
        for (int s = 0; s < kDftBK / 2; s++) {
            const int kr = 2 * s + (lane >> 5);
            const float a0 = sA[buf][kr][wm + (lane & 31)];
            const float b0 = sB[buf][kr][wn + (lane & 31)];
            acc00 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc00, 0, 0, 0);
            float b1 = 0.0f;
            if constexpr (NB == 2) {
                b1 = sB[buf][kr][wn + 32 + (lane & 31)];
                acc01 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc01, 0, 0, 0);
            }
            if constexpr (MB == 2) {
                const float a1 = sA[buf][kr][wm + 32 + (lane & 31)];
                acc10 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc10, 0, 0, 0);
                if constexpr (NB == 2)
                    acc11 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc11, 0, 0, 0);
            }
        }
        if (more)
            store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // accumulator register r of this lane: column wn (+ 32) + lane % 32, row
    // wm (+ 32) + 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
    for (int half = 0; half < NB; half++) {
        const int oc = c0 + wn + 32 * half + (lane & 31);
        if (oc < jb.cols) {
            const gfloat dst = (gfloat)(jb.out + (size_t)oc * jb.ld_out);
#pragma unroll
            for (int mh = 0; mh < MB; mh++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int om =
                        m0 + wm + 32 * mh + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const f32x16 &acc = mh ? (half ? acc11 : acc10) : (half ? acc01 : acc00);
                    if (om < M)
                        dst[om] = acc[r];
                }
        }
    }
}

// Scales a column's half spectrum into the input of P3's adjoint GEMM:
// y_k = g c_k Re X_k, y_{h+k} = g c_k Im X_k, with g the upstream gradient of
// the column's P3 (summed over the P list's P3 entries) and c_k = w_k 4 |X_k|^2 / n^4 (the derivative of
// pow(hypot(Re/n, Im/n), 4) = |X_k|^4 / n^4, times the trapz weight of the
// reference's float abscissae lin[p] = -1 + 2p/(n-1), folded onto k from
// p = k and its mirror p = n - k).  x (the spectra, ld_in) and y (ld_out)
// may be the same array.
__global__ __launch_bounds__(kBlock) void p3_vjp_scale_kernel(P3ScaleParams p)
{
    const DftJob &jb = p.jobs ? p.jobs[blockIdx.y] : p.job;
    const int n = jb.n, h = n / 2 + 1;
    const int cols = jb.cols;
    for (int c = blockIdx.x; c < cols; c += gridDim.x) {
        // every P3 entry of the list adds the same derivative times its own
        // upstream gradient
        float g = 0.0f;
        const size_t gb = (size_t)(c / jb.g_cols) * jb.g_stride + c % jb.g_cols;
        for (int i = 0; i < p.np3; i++)
            g += jb.g[gb + (size_t)p.q3[i] * jb.g_cols];
        const gcfloat x = (gcfloat)(jb.in + (size_t)c * jb.ld_in);
        const gfloat y = (gfloat)(jb.out + (size_t)c * jb.ld_out);
        for (int k = threadIdx.x; k < h; k += kBlock) {
            double w = 0.0;
            for (int side = 0; side < 2; side++) {
                const int q = side == 0 ? k : n - k;
                if (q >= n || (side == 1 && (q == k || k == 0)))
                    continue;
                const double l = (float)(-1 + q * 2.0 / (n - 1));
                if (q >= 1)
                    w += 0.5 * (l - (double)(float)(-1 + (q - 1) * 2.0 / (n - 1)));
                if (q <= n - 2)
                    w += 0.5 * ((double)(float)(-1 + (q + 1) * 2.0 / (n - 1)) - l);
            }
            const double re = x[k], im = x[h + k];
            const double n4 = (double)n * n * (double)n * n;
            const double ck = (double)g * w * 4.0 * (re * re + im * im) / n4;
            y[k] = (float)(ck * re);
            y[h + k] = (float)(ck * im);
        }
    }
}

}  // namespace

void dft_tile_shape(long long tiles128, long long tiles64, int *bm, int *bn)
{
    // 128 x 128 tiles once they give every CU several workgroups, 64 x 128
    // while those fill the chip, else 64 x 64 (twice the workgroups)
    *bm = tiles128 >= 1024 ? 128 : 64;
    *bn = *bm == 128 || tiles64 >= 256 ? 128 : 64;
}

size_t dft_lds_bytes(int nmax)
{
    return sizeof(float2) * (size_t)nmax;
}

hipError_t launch_dft_gemm(const DftParams &p, bool forward, int nmax, hipStream_t s)
{
    if (p.tiles <= 0)
        return hipSuccess;
    const int shape = p.bm == 128 && p.bn == 128 ? 2 : p.bm == 64 && p.bn == 128 ? 1
                                                     : p.bm == 64 && p.bn == 64   ? 0
                                                                                  : -1;
    if (shape < 0)
        return hipErrorInvalidValue;
    const size_t shm = dft_lds_bytes(nmax);
    if (forward) {
        if (shape == 2)
            dft_gemm_kernel<true, 128, 128><<<p.tiles, kBlock, shm, s>>>(p);
        else if (shape == 1)
            dft_gemm_kernel<true, 64, 128><<<p.tiles, kBlock, shm, s>>>(p);
        else
            dft_gemm_kernel<true, 64, 64><<<p.tiles, kBlock, shm, s>>>(p);
    } else {
        if (shape == 2)
            dft_gemm_kernel<false, 128, 128><<<p.tiles, kBlock, shm, s>>>(p);
        else if (shape == 1)
            dft_gemm_kernel<false, 64, 128><<<p.tiles, kBlock, shm, s>>>(p);
        else
            dft_gemm_kernel<false, 64, 64><<<p.tiles, kBlock, shm, s>>>(p);
    }
    return hipGetLastError();
}

hipError_t launch_p3_vjp_scale(const P3ScaleParams &p, int max_cols, hipStream_t s)
{
    if (p.njobs <= 0)
        return hipSuccess;
    int gx = max_cols < 2048 ? max_cols : 2048;
    gx = gx < 1 ? 1 : gx;
    p3_vjp_scale_kernel<<<dim3(gx, p.njobs), kBlock, 0, s>>>(p);
    return hipGetLastError();
}

int dft_job_tiles(int n, int cols, bool forward, int bm, int bn)
{
    const int h = n / 2 + 1;
    const int M = forward ? 2 * h : n;
    return ((M + bm - 1) / bm) * ((cols + bn - 1) / bn);
}

}  // namespace ttk
