/*
 * tt_oracle.c -- CPU restatement of the reference trace transform.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check of
 * __graft_entry__ and bench.py's cpu_baseline leg may load this library;
 * the product never does.
 *
 * Two oracles live here:
 *   FP32-faithful: the reference's float arithmetic, operation for operation
 *     (the c++/src sources, GCC branches of the #if __clang__ switches), compiled
 *     with -ffp-contract=off so no product/sum is fused.  Its weighted-median
 *     indices and sort positions are the parity target for the GPU.
 *   FP64: the same algorithm in double, taking the medians and sort
 *     positions from the FP32 run; GPU values are checked against it within
 *     a stated tolerance.  geometry 0 computes every sample position in
 *     double too; geometry 1 keeps the reference's float sample positions
 *     (rotation coordinates and strict bounds, Hermite z grid, P3 trapz
 *     abscissae) and computes the values at them in double, isolating value
 *     arithmetic from the positions that bit-exact medians pin.
 *
 * Pinning status of each function (how a plausible slip would be caught):
 *   tto_pad_geometry/tto_pad  pinned: closed-form size table (256->363,
 *       512->725, 1024->1449, 64->91, 255->361) and offset df=54 at 256.
 *   tto_rotate/tto_interpolate  pinned: angle 0 identity except the last
 *       row/column, constant image, exact affine-ramp reproduction at
 *       dyadic angles, FP64 brute force on tiny images, 90-degree rotation
 *       as an index permutation.
 *   tto_eigen_sum  order restated from Eigen 3.x Redux.h (SSE2 Packet4f,
 *       two packet accumulators, predux (a0+a2)+(a1+a3)).  Eigen is not in
 *       this image, so the ORDER ITSELF IS PARITY UNPINNED; tests pin the
 *       restatement's value against exact sums and against hand-derived
 *       cancellation vectors where sequential and packet order differ.
 *   tto_weighted_median  pinned: all-zero -> 0, spike -> its index,
 *       all-ones length n -> ceil(n/2)-1, no crossing -> n-1.
 *   tto_stable_argsort  pinned against glibc 2.35 qsort_r (stable merge sort)
 *       on tie-heavy keys.
 *   T0..T7  pinned: spike/box closed forms, T3-T5 spike gives r*v, v,
 *       sqrt(r)*v; T6/T7 against exhaustive brute force on short columns.
 *   P1/P2  pinned: monotone column, brute force.
 *   P3  pinned: constant column c -> c^4/(n-1); spike -> closed form;
 *       tto_dft_f64 against numpy.fft.  FFTW's own FP32 rounding is
 *       PARITY UNPINNED (FFTW is absent); this oracle rounds an FP64 DFT.
 *   Hermite  pinned: orthonormality of the basis on a fine grid, H_n(0).
 */
#include "tt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ======================================================================= */
/* Geometry                                                                 */
/* ======================================================================= */

/* c++/src/auxiliary.cpp:257-275 (pad): origin = floor((cols+1)/2)-1,
 * rLast = ceil(hypot(cols-1-ox-1, rows-1-oy-1)) + 1, side = 2*rLast+1,
 * image placed at (df.y, df.x) = origin_padded - origin. */
int tto_pad_geometry(int rows, int cols, int *n, int *dfx, int *dfy)
{
    if (rows <= 0 || cols <= 0)
        return -1;
    float ox = (float)(floor((cols + 1) / 2.0) - 1);
    float oy = (float)(floor((rows + 1) / 2.0) - 1);
    float hx = (float)cols - 1 - ox - 1;
    float hy = (float)rows - 1 - oy - 1;
    int rlast = (int)ceilf(hypotf(hx, hy)) + 1;
    int nb = 2 * rlast + 1;
    float opx = (float)(floor((nb + 1) / 2.0) - 1);
    float opy = (float)(floor((nb + 1) / 2.0) - 1);
    *n = nb;
    *dfx = (int)(opx - ox);
    *dfy = (int)(opy - oy);
    return 0;
}

void tto_pad(const float *img, int rows, int cols, float *out)
{
    int n, dfx, dfy;
    tto_pad_geometry(rows, cols, &n, &dfx, &dfy);
    memset(out, 0, sizeof(float) * (size_t)n * n);
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++)
            out[(size_t)(r + dfy) + (size_t)(c + dfx) * n] =
                img[(size_t)r + (size_t)c * rows];
}

/* c++/src/auxiliary.cpp:182-196 (interpolate) on a column-major source with
 * leading dimension ld: modf split, then the four bilinear terms evaluated
 * left to right exactly as written in the reference. */
static float interp_ld(const float *src, int ld, float x, float y)
{
    float x_int, y_int;
    float x_fract = modff(x, &x_int);
    float y_fract = modff(y, &y_int);
    long xi = (long)x_int, yi = (long)y_int;
    float s00 = src[yi + xi * ld];
    float s01 = src[yi + (xi + 1) * ld];
    float s10 = src[(yi + 1) + xi * ld];
    float s11 = src[(yi + 1) + (xi + 1) * ld];
    return s00 * (1 - x_fract) * (1 - y_fract) + s01 * x_fract * (1 - y_fract) +
           s10 * (1 - x_fract) * y_fract + s11 * x_fract * y_fract;
}

float tto_interpolate(const float *src, int n, float x, float y)
{
    return interp_ld(src, n, x, y);
}

/* Bilinear sample whose taps outside the in_rows x in_cols matrix read 0. */
static float interp_zero_ext(const float *src, int in_rows, int in_cols, float x,
                             float y)
{
    float x_int, y_int;
    float x_fract = modff(x, &x_int);
    float y_fract = modff(y, &y_int);
    long xi = (long)x_int, yi = (long)y_int;
#define TAP(r, c) (((r) >= 0 && (r) < in_rows && (c) >= 0 && (c) < in_cols) \
                       ? src[(r) + (c) * (long)in_rows] : 0.0f)
    float s00 = TAP(yi, xi), s01 = TAP(yi, xi + 1);
    float s10 = TAP(yi + 1, xi), s11 = TAP(yi + 1, xi + 1);
#undef TAP
    return s00 * (1 - x_fract) * (1 - y_fract) + s01 * x_fract * (1 - y_fract) +
           s10 * (1 - x_fract) * y_fract + s11 * x_fract * y_fract;
}

/* c++/src/auxiliary.cpp:198-221 (resize): borders stay zero; interior
 * p = ((col,row) + 0.5) * diag(in_rows/rows, in_cols/cols) - 0.5.  The
 * reference scales the column coordinate by the row ratio and vice versa,
 * so a non-square input reads outside the matrix (undefined there); here
 * such taps read 0.  Square inputs never reach them. */
void tto_resize(const float *in, int in_rows, int in_cols, float *out,
                int rows, int cols)
{
    float t00 = ((float)in_rows) / (float)rows;
    float t11 = ((float)in_cols) / (float)cols;
    memset(out, 0, sizeof(float) * (size_t)rows * cols);
    for (int col = 1; col < cols - 1; col++)
        for (int row = 1; row < rows - 1; row++) {
            float px = (float)col + 0.5f, py = (float)row + 0.5f;
            float x = px * t00 + py * 0.0f;
            float y = px * 0.0f + py * t11;
            x = x - 0.5f;
            y = y - 0.5f;
            out[(size_t)row + (size_t)col * rows] =
                interp_zero_ext(in, in_rows, in_cols, x, y);
        }
}

/* c++/src/auxiliary.cpp:180: (degrees * M_PI / 180) in double, returned as
 * float. */
float tto_deg2rad(float degrees) { return (float)(degrees * M_PI / 180); }

/* c++/src/sinogram.cpp:112: float a = a_step * angle_stepsize.  For integer
 * steps (the reference's unsigned steps) the product is exact; fractional
 * steps (0.5 degree) follow the same rule in double then float. */
float tto_angle_deg(int a_step, float step_deg)
{
    return (float)((double)a_step * (double)step_deg);
}

/* c++/src/auxiliary.cpp:235-236 (GCC branch): std::cos(-angle),
 * std::sin(-angle) on a float argument, i.e. glibc cosf/sinf. */
void tto_rotation_trig(float angle_rad, float *c, float *s)
{
    *c = cosf(-angle_rad);
    *s = sinf(-angle_rad);
}

/* c++/src/auxiliary.cpp:223-255 (rotate) with origin ((n-1)/2, (n-1)/2)
 * from c++/src/sinogram.cpp:73-74.  Eigen row-vector product p * T with
 * T = [[c, -s], [s, c]] (comma initializer fills row-wise):
 *   x = px*c + py*s,  y = px*(-s) + py*c,  then += origin.
 * Strict bounds 0 <= x < n-1, 0 <= y < n-1, else 0. */
void tto_rotate(const float *in, int n, float c, float s, float *out)
{
    const float o = (float)((n - 1) / 2.0);
    const float ns = -s;
    const float lim = (float)(n - 1);
    for (int col = 0; col < n; col++)
        for (int row = 0; row < n; row++) {
            float px = (float)col - o;
            float py = (float)row - o;
            float x = px * c + py * s;
            float y = px * ns + py * c;
            x = x + o;
            y = y + o;
            float v = 0.0f;
            if (x >= 0 && x < lim && y >= 0 && y < lim)
                v = interp_ld(in, n, x, y);
            out[(size_t)row + (size_t)col * n] = v;
        }
}

/* ======================================================================= */
/* Reductions and the weighted median                                       */
/* ======================================================================= */

/* Eigen's VectorXf::sum() with SSE2 packets (Redux.h, LinearVectorized
 * traversal, aligned start 0 because every summed vector in the reference is
 * a freshly allocated VectorXf):
 *   n < 4: sequential.
 *   acc0 = x[0..3]; if A4 > 4 { acc1 = x[4..7]; for i=8; i<A8; i+=8
 *   { acc0 += x[i..i+3]; acc1 += x[i+4..i+7] } acc0 += acc1;
 *   if A4 > A8: acc0 += x[A8..A8+3] }
 *   res = (acc0[0]+acc0[2]) + (acc0[1]+acc0[3]); res += x[A4..n-1] in order. */
float tto_eigen_sum(const float *x, int n)
{
    if (n <= 0)
        return 0.0f;
    const int a4 = (n / 4) * 4, a8 = (n / 8) * 8;
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

float tto_seq_sum(const float *x, int n)
{
    float r = 0;
    for (int i = 0; i < n; i++)
        r += x[i];
    return r;
}

/* c++/src/functionals.cpp:21-30 */
int tto_weighted_median(const float *x, int n)
{
    float sum = tto_eigen_sum(x, n);
    float integral = 0;
    for (int i = 0; i < n; i++) {
        integral += x[i];
        if (2 * integral >= sum)
            return i;
    }
    return n - 1;
}

/* Stable ascending argsort by float key under compareFloat's order
 * (c++/src/functionals.cpp:32-51): ties keep ascending original index, which
 * is what glibc's merge-sort qsort/qsort_r produces.  Bottom-up merge sort. */
void tto_stable_argsort(const float *key, int n, int *idx, int *tmp)
{
    for (int i = 0; i < n; i++)
        idx[i] = i;
    int *src = idx, *dst = tmp;
    for (int w = 1; w < n; w *= 2) {
        for (int lo = 0; lo < n; lo += 2 * w) {
            int mid = lo + w < n ? lo + w : n;
            int hi = lo + 2 * w < n ? lo + 2 * w : n;
            int i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (key[src[j]] < key[src[i]])
                    dst[k++] = src[j++];
                else
                    dst[k++] = src[i++];
            }
            while (i < mid)
                dst[k++] = src[i++];
            while (j < hi)
                dst[k++] = src[j++];
        }
        int *t = src;
        src = dst;
        dst = t;
    }
    if (src != idx)
        memcpy(idx, src, sizeof(int) * (size_t)n);
}

/* ======================================================================= */
/* T-functionals                                                            */
/* ======================================================================= */

/* c++/src/functionals.cpp:136-198, GCC branch: double math, stored as float.
 * Index 0 is never read by the reference; it is set to 0 here. */
void tto_t345_table(int which, int n, float *re, float *im)
{
    if (n <= 0)
        return;
    re[0] = 0.0f;
    im[0] = 0.0f;
    for (int r = 1; r < n; r++) {
        double lr = log((double)r);
        switch (which) {
        case 3:
            re[r] = (float)(r * cos(5.0 * lr));
            im[r] = (float)(r * sin(5.0 * lr));
            break;
        case 4:
            re[r] = (float)cos(3.0 * lr);
            im[r] = (float)sin(3.0 * lr);
            break;
        default:
            re[r] = (float)(sqrt((double)r) * cos(4.0 * lr));
            im[r] = (float)(sqrt((double)r) * sin(4.0 * lr));
            break;
        }
    }
}

/* c++/src/functionals.cpp:95-97 */
float tto_T0(const float *x, int n) { return tto_eigen_sum(x, n); }

/* c++/src/functionals.cpp:104-113 */
float tto_T1(const float *x, int n, int *m_out)
{
    int median = tto_weighted_median(x, n);
    if (m_out)
        *m_out = median;
    float integral = 0;
    for (int r = 0; r < n - median; r++)
        integral += x[r + median] * r;
    return integral;
}

/* c++/src/functionals.cpp:120-129 */
float tto_T2(const float *x, int n, int *m_out)
{
    int median = tto_weighted_median(x, n);
    if (m_out)
        *m_out = median;
    float integral = 0;
    for (int r = 0; r < n - median; r++)
        integral += x[r + median] * r * r;
    return integral;
}

static int median_of_sqrt(const float *x, int n, float *buf)
{
    for (int i = 0; i < n; i++)
        buf[i] = sqrtf(x[i]);
    return tto_weighted_median(buf, n);
}

/* c++/src/functionals.cpp:200-213 */
float tto_T345(const float *x, int n, const float *re, const float *im,
               int *sm_out)
{
    float *buf = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    int sm = median_of_sqrt(x, n, buf);
    free(buf);
    if (sm_out)
        *sm_out = sm;
    float ir = 0, ii = 0;
    for (int r1 = 1; r1 < n - sm; r1++) {
        ir += re[r1] * x[r1 + sm];
        ii += im[r1] * x[r1 + sm];
    }
    return (float)hypot(ir, ii);
}

/* c++/src/functionals.cpp:226-255 */
float tto_T6(const float *x, int n, int *sm_out, int *rank_out, int *src_out)
{
    size_t sz = sizeof(float) * (size_t)(n > 0 ? n : 1);
    float *buf = (float *)malloc(sz);
    float *w = (float *)malloc(sz);
    float *ds = (float *)malloc(sz);
    int *idx = (int *)malloc(sz);
    int *tmp = (int *)malloc(sz);
    int sm = median_of_sqrt(x, n, buf);
    int len = n - sm;
    for (int r1 = 0; r1 < len; r1++)
        w[r1] = (float)r1 * x[r1 + sm];
    tto_stable_argsort(w, len, idx, tmp);
    for (int r1 = 0; r1 < len; r1++)
        ds[r1] = sqrtf(x[sm + idx[r1]]);
    int k = tto_weighted_median(ds, len);
    float res = w[idx[k]];
    if (sm_out)
        *sm_out = sm;
    if (rank_out)
        *rank_out = k;
    if (src_out)
        *src_out = idx[k];
    free(buf);
    free(w);
    free(ds);
    free(idx);
    free(tmp);
    return res;
}

/* c++/src/functionals.cpp:262-276 */
float tto_T7(const float *x, int n, int *m_out, int *rank_out, int *src_out)
{
    size_t sz = sizeof(float) * (size_t)(n > 0 ? n : 1);
    float *tail = (float *)malloc(sz);
    float *srt = (float *)malloc(sz);
    float *sq = (float *)malloc(sz);
    int *idx = (int *)malloc(sz);
    int *tmp = (int *)malloc(sz);
    int m = tto_weighted_median(x, n);
    int len = n - m;
    for (int i = 0; i < len; i++)
        tail[i] = x[m + i];
    tto_stable_argsort(tail, len, idx, tmp);
    for (int i = 0; i < len; i++) {
        srt[i] = tail[idx[i]];
        sq[i] = sqrtf(srt[i]);
    }
    int k = tto_weighted_median(sq, len);
    float res = srt[k];
    if (m_out)
        *m_out = m;
    if (rank_out)
        *rank_out = k;
    if (src_out)
        *src_out = m + idx[k];
    free(tail);
    free(srt);
    free(sq);
    free(idx);
    free(tmp);
    return res;
}

/* ======================================================================= */
/* P-functionals                                                            */
/* ======================================================================= */

/* c++/src/functionals.cpp:287-296 (fabs through double; the double sum of
 * two floats is exact, so the float-rounded result equals a float add). */
float tto_P1(const float *x, int n)
{
    float sum = 0;
    float previous = x[0];
    for (int p = 1; p < n; p++) {
        float current = x[p];
        sum += (float)fabs((double)(previous - current));
        previous = current;
    }
    return sum;
}

/* c++/src/functionals.cpp:303-311 */
float tto_P2(const float *x, int n, int *rank_out, int *src_out)
{
    size_t sz = sizeof(float) * (size_t)(n > 0 ? n : 1);
    float *srt = (float *)malloc(sz);
    int *idx = (int *)malloc(sz);
    int *tmp = (int *)malloc(sz);
    tto_stable_argsort(x, n, idx, tmp);
    for (int i = 0; i < n; i++)
        srt[i] = x[idx[i]];
    int m = tto_weighted_median(srt, n);
    float res = srt[m];
    if (rank_out)
        *rank_out = m;
    if (src_out)
        *src_out = idx[m];
    free(srt);
    free(idx);
    free(tmp);
    return res;
}

/* ---- FFT (double) used by P3 ------------------------------------------- */

static void fft_pow2(double *re, double *im, int m, int inverse)
{
    for (int i = 1, j = 0; i < m; i++) {
        int bit = m >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i];
            re[i] = re[j];
            re[j] = t;
            t = im[i];
            im[i] = im[j];
            im[j] = t;
        }
    }
    for (int len = 2; len <= m; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            double ang = (inverse ? 2.0 : -2.0) * M_PI * k / len;
            double wr = cos(ang), wi = sin(ang);
            for (int i = 0; i < m; i += len) {
                double ur = re[i + k], ui = im[i + k];
                double vr = re[i + k + half] * wr - im[i + k + half] * wi;
                double vi = re[i + k + half] * wi + im[i + k + half] * wr;
                re[i + k] = ur + vr;
                im[i + k] = ui + vi;
                re[i + k + half] = ur - vr;
                im[i + k + half] = ui - vi;
            }
        }
    }
}

/* Full DFT X_k = sum_j x_j exp(-2 pi i j k / n), k = 0..n-1, in double.
 * Small n: direct sum with exact angle reduction (j*k mod n).  Otherwise
 * Bluestein's chirp-z on a power-of-two FFT (chirp angle k^2 mod 2n). */
void tto_dft_f64(const double *x, int n, double *re, double *im)
{
    if (n <= 0)
        return;
    if (n <= 64) {
        for (int k = 0; k < n; k++) {
            double sr = 0, si = 0;
            for (int j = 0; j < n; j++) {
                long jk = ((long)j * k) % n;
                double ang = -2.0 * M_PI * (double)jk / n;
                sr += x[j] * cos(ang);
                si += x[j] * sin(ang);
            }
            re[k] = sr;
            im[k] = si;
        }
        return;
    }
    int m = 1;
    while (m < 2 * n - 1)
        m <<= 1;
    double *ar = (double *)calloc((size_t)m, sizeof(double));
    double *ai = (double *)calloc((size_t)m, sizeof(double));
    double *br = (double *)calloc((size_t)m, sizeof(double));
    double *bi = (double *)calloc((size_t)m, sizeof(double));
    double *wr = (double *)malloc(sizeof(double) * (size_t)n);
    double *wi = (double *)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; k++) {
        long kk = ((long)k * k) % (2L * n);
        double ang = -M_PI * (double)kk / n;
        wr[k] = cos(ang);
        wi[k] = sin(ang);
    }
    for (int j = 0; j < n; j++) {
        ar[j] = x[j] * wr[j];
        ai[j] = x[j] * wi[j];
        br[j] = wr[j];
        bi[j] = -wi[j];
        if (j > 0) {
            br[m - j] = wr[j];
            bi[m - j] = -wi[j];
        }
    }
    fft_pow2(ar, ai, m, 0);
    fft_pow2(br, bi, m, 0);
    for (int i = 0; i < m; i++) {
        double r = ar[i] * br[i] - ai[i] * bi[i];
        double q = ar[i] * bi[i] + ai[i] * br[i];
        ar[i] = r;
        ai[i] = q;
    }
    fft_pow2(ar, ai, m, 1);
    for (int k = 0; k < n; k++) {
        double cr = ar[k] / m, ci = ai[k] / m;
        re[k] = cr * wr[k] - ci * wi[k];
        im[k] = cr * wi[k] + ci * wr[k];
    }
    free(ar);
    free(ai);
    free(br);
    free(bi);
    free(wr);
    free(wi);
}

/* c++/src/functionals.cpp:53-60 (trapz, float accumulation) */
static float trapz_f(const float *x, const float *y, int n)
{
    float sum = 0;
    for (int i = 0; i < n - 1; i++)
        sum += (x[i + 1] - x[i]) * (y[i + 1] + y[i]);
    return (float)(sum * 0.5);
}

/* c++/src/functionals.cpp:332-360.  The r2c transform is an FP64 DFT rounded
 * to float (FFTW's FP32 rounding is not reproduced).  modifier[p] =
 * pow(hypot(Re/N, Im/N), 4) in double, mirrored X[N-p] above N/2;
 * linspace[p] = -1 + p*2.0/(N-1). */
float tto_P3(const float *x, int n)
{
    if (n <= 1)
        return 0.0f;
    double *xd = (double *)malloc(sizeof(double) * (size_t)n);
    double *re = (double *)malloc(sizeof(double) * (size_t)n);
    double *im = (double *)malloc(sizeof(double) * (size_t)n);
    float *lin = (float *)malloc(sizeof(float) * (size_t)n);
    float *mod = (float *)malloc(sizeof(float) * (size_t)n);
    for (int i = 0; i < n; i++)
        xd[i] = x[i];
    tto_dft_f64(xd, n, re, im);
    const float fn = (float)n;
    for (int p = 0; p < n; p++) {
        int q = p <= n / 2 ? p : n - p;
        float fr = (float)re[q], fi = (float)im[q];
        mod[p] = (float)pow(hypot((double)(fr / fn), (double)(fi / fn)), 4);
        lin[p] = (float)(-1 + p * 2.0 / (n - 1));
    }
    float s = trapz_f(lin, mod, n);
    free(xd);
    free(re);
    free(im);
    free(lin);
    free(mod);
    return s;
}

/* c++/src/functionals.cpp:62-84: recursion in double, each level returned
 * as float; normalisation sqrt(2^n n! sqrt(pi)) * exp(z*z/2) in double with
 * z*z/2 evaluated in float. */
static float herm_poly(unsigned order, float x)
{
    switch (order) {
    case 0:
        return 1.0f;
    case 1:
        return (float)(2.0 * x);
    default:
        return (float)(2.0 * x * herm_poly(order - 1, x) -
                       2.0 * (order - 1) * herm_poly(order - 2, x));
    }
}

static unsigned factorial_u(unsigned n)
{
    return (n == 1 || n == 0) ? 1 : factorial_u(n - 1) * n;
}

float tto_hermite_function(unsigned order, float z)
{
    float zz = z * z / 2;
    return (float)(herm_poly(order, z) /
                   (sqrt(pow(2, order) * factorial_u(order) * sqrt(M_PI)) *
                    exp((double)zz)));
}

/* c++/src/functionals.cpp:369-386 */
float tto_Phermite(const float *x, int n, unsigned order, int center)
{
    float stepsize_lower = (float)(10.0 / center);
    float stepsize_upper = (float)(10.0 / (double)(n - 1 - center));
    float integral = 0;
    float z;
    for (int p = 0; p < n; p++) {
        if (p < center)
            z = p * stepsize_lower - 10;
        else
            z = (p - center) * stepsize_upper;
        integral += x[p] * tto_hermite_function(order, z);
    }
    return integral;
}

/* ======================================================================= */
/* Drivers                                                                  */
/* ======================================================================= */

static int valid_tcodes(const int *tcodes, int nt)
{
    for (int t = 0; t < nt; t++)
        if (tcodes[t] < 0 || tcodes[t] > 7)
            return 0;
    return 1;
}

/* c++/src/sinogram.cpp:67-178; the angle loop is the OpenMP one of
 * openmp/src/sinogram.cpp:110. */
int tto_sinograms(const float *img, int n, int n_angles, float step_deg,
                  const int *tcodes, int nt, float *sino, int32_t *med_x,
                  int32_t *med_s, int32_t *t6_rank, int32_t *t6_src,
                  int32_t *t7_rank, int32_t *t7_src, int nthreads)
{
    if (n <= 0 || n_angles <= 0 || nt < 0 || !valid_tcodes(tcodes, nt))
        return -1;
    float *tre[8] = {0}, *tim[8] = {0};
    for (int t = 0; t < nt; t++) {
        int code = tcodes[t];
        if (code >= 3 && code <= 5 && !tre[code]) {
            tre[code] = (float *)malloc(sizeof(float) * (size_t)n);
            tim[code] = (float *)malloc(sizeof(float) * (size_t)n);
            tto_t345_table(code, n, tre[code], tim[code]);
        }
    }
    if (nthreads <= 0)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        float *rot = (float *)malloc(sizeof(float) * (size_t)n * n);
        float *col = (float *)malloc(sizeof(float) * (size_t)n);
#pragma omp for schedule(static)
        for (int a = 0; a < n_angles; a++) {
            float c, s;
            tto_rotation_trig(tto_deg2rad(tto_angle_deg(a, step_deg)), &c, &s);
            tto_rotate(img, n, c, s, rot);
            for (int cc = 0; cc < n; cc++) {
                memcpy(col, rot + (size_t)cc * n, sizeof(float) * (size_t)n);
                size_t aux = (size_t)a * n + cc;
                for (int t = 0; t < nt; t++) {
                    int m = -1, sm = -1, rk = -1, src = -1;
                    float r = 0;
                    switch (tcodes[t]) {
                    case 0:
                        r = tto_T0(col, n);
                        break;
                    case 1:
                        r = tto_T1(col, n, &m);
                        break;
                    case 2:
                        r = tto_T2(col, n, &m);
                        break;
                    case 3:
                    case 4:
                    case 5:
                        r = tto_T345(col, n, tre[tcodes[t]], tim[tcodes[t]], &sm);
                        break;
                    case 6:
                        r = tto_T6(col, n, &sm, &rk, &src);
                        if (t6_rank)
                            t6_rank[aux] = rk;
                        if (t6_src)
                            t6_src[aux] = src;
                        break;
                    case 7:
                        r = tto_T7(col, n, &m, &rk, &src);
                        if (t7_rank)
                            t7_rank[aux] = rk;
                        if (t7_src)
                            t7_src[aux] = src;
                        break;
                    }
                    if (m >= 0 && med_x)
                        med_x[aux] = m;
                    if (sm >= 0 && med_s)
                        med_s[aux] = sm;
                    sino[(size_t)t * n * n_angles + aux] = r;
                }
            }
        }
        free(rot);
        free(col);
    }
    for (int k = 0; k < 8; k++) {
        free(tre[k]);
        free(tim[k]);
    }
    return 0;
}

/* c++/src/circus.cpp:104-177; column loop parallel as in
 * openmp/src/circus.cpp:129. */
int tto_circus(const float *sino, int n, int n_cols, const int *pcodes,
               const int *orders, const int *centers, int np, float *out,
               int32_t *p2_rank, int32_t *p2_src, int nthreads)
{
    if (n <= 0 || n_cols <= 0 || np < 0)
        return -1;
    for (int p = 0; p < np; p++) {
        if (pcodes[p] < 0 || pcodes[p] > 3)
            return -1;
        if (pcodes[p] == 0 && (!orders || !centers))
            return -1;
    }
    if (nthreads <= 0)
        nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int a = 0; a < n_cols; a++) {
        const float *x = sino + (size_t)a * n;
        for (int p = 0; p < np; p++) {
            float r = 0;
            int rk = -1, src = -1;
            switch (pcodes[p]) {
            case 1:
                r = tto_P1(x, n);
                break;
            case 2:
                r = tto_P2(x, n, &rk, &src);
                if (p2_rank)
                    p2_rank[(size_t)p * n_cols + a] = rk;
                if (p2_src)
                    p2_src[(size_t)p * n_cols + a] = src;
                break;
            case 3:
                r = tto_P3(x, n);
                break;
            case 0:
                r = tto_Phermite(x, n, (unsigned)orders[p], centers[p]);
                break;
            }
            out[(size_t)p * n_cols + a] = r;
        }
    }
    return 0;
}

/* c++/src/auxiliary.cpp:277-318: float mean, double variance (n-1). */
void tto_zscore(const float *in, int n, float *out)
{
    if (n <= 0)
        return;
    float sum = 0;
    for (int i = 0; i < n; i++)
        sum += in[i];
    float mean = sum / n;
    double acc = 0;
    for (int i = 0; i < n; i++) {
        double diff = in[i] - mean;
        acc += diff * diff;
    }
    float stdev = (float)sqrt(acc / (n - 1));
    for (int i = 0; i < n; i++)
        out[i] = (in[i] - mean) / stdev;
}

/* ======================================================================= */
/* FP64 oracle                                                              */
/* ======================================================================= */

/* geometry 0: coordinates and bounds in double.  geometry 1: the
 * reference's float sample positions (tto_rotate's x, y, bounds and modf
 * fractions, all exact restatements), with the bilinear values in double. */
static void rotate_f64(const float *in, int n, float c, float s, int geometry,
                       double *out)
{
    if (geometry) {
        const float o = (float)((n - 1) / 2.0), ns = -s, lim = (float)(n - 1);
        for (int col = 0; col < n; col++)
            for (int row = 0; row < n; row++) {
                float px = (float)col - o, py = (float)row - o;
                float x = px * c + py * s;
                float y = px * ns + py * c;
                x = x + o;
                y = y + o;
                double v = 0.0;
                if (x >= 0 && x < lim && y >= 0 && y < lim) {
                    float xi, yi;
                    double fx = modff(x, &xi), fy = modff(y, &yi);
                    long X = (long)xi, Y = (long)yi;
                    v = (double)in[Y + X * n] * (1 - fx) * (1 - fy) +
                        (double)in[Y + (X + 1) * n] * fx * (1 - fy) +
                        (double)in[(Y + 1) + X * n] * (1 - fx) * fy +
                        (double)in[(Y + 1) + (X + 1) * n] * fx * fy;
                }
                out[(size_t)row + (size_t)col * n] = v;
            }
        return;
    }
    const double o = (n - 1) / 2.0, cd = c, sd = s, lim = n - 1;
    for (int col = 0; col < n; col++)
        for (int row = 0; row < n; row++) {
            double px = col - o, py = row - o;
            double x = px * cd + py * sd + o;
            double y = -px * sd + py * cd + o;
            double v = 0.0;
            if (x >= 0 && x < lim && y >= 0 && y < lim) {
                double xi = floor(x), yi = floor(y);
                double fx = x - xi, fy = y - yi;
                long X = (long)xi, Y = (long)yi;
                v = in[Y + X * n] * (1 - fx) * (1 - fy) +
                    in[Y + (X + 1) * n] * fx * (1 - fy) +
                    in[(Y + 1) + X * n] * (1 - fx) * fy +
                    in[(Y + 1) + (X + 1) * n] * fx * fy;
            }
            out[(size_t)row + (size_t)col * n] = v;
        }
}

static void t345_table_f64(int which, int n, double *re, double *im)
{
    re[0] = im[0] = 0.0;
    for (int r = 1; r < n; r++) {
        double lr = log((double)r);
        double k = which == 3 ? 5.0 : which == 4 ? 3.0 : 4.0;
        double amp = which == 3 ? (double)r : which == 4 ? 1.0 : sqrt((double)r);
        re[r] = amp * cos(k * lr);
        im[r] = amp * sin(k * lr);
    }
}

int tto_sinograms_f64(const float *img, int n, int n_angles, float step_deg,
                      const int *tcodes, int nt, const int32_t *med_x,
                      const int32_t *med_s, const int32_t *t6_src,
                      const int32_t *t7_src, int geometry, double *sino,
                      int nthreads)
{
    if (n <= 0 || n_angles <= 0 || nt < 0 || !valid_tcodes(tcodes, nt))
        return -1;
    double *tre[8] = {0}, *tim[8] = {0};
    for (int t = 0; t < nt; t++) {
        int code = tcodes[t];
        if (code >= 3 && code <= 5 && !tre[code]) {
            tre[code] = (double *)malloc(sizeof(double) * (size_t)n);
            tim[code] = (double *)malloc(sizeof(double) * (size_t)n);
            t345_table_f64(code, n, tre[code], tim[code]);
        }
    }
    if (nthreads <= 0)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double *rot = (double *)malloc(sizeof(double) * (size_t)n * n);
#pragma omp for schedule(static)
        for (int a = 0; a < n_angles; a++) {
            float c, s;
            tto_rotation_trig(tto_deg2rad(tto_angle_deg(a, step_deg)), &c, &s);
            rotate_f64(img, n, c, s, geometry, rot);
            for (int cc = 0; cc < n; cc++) {
                const double *x = rot + (size_t)cc * n;
                size_t aux = (size_t)a * n + cc;
                for (int t = 0; t < nt; t++) {
                    double r = 0;
                    int code = tcodes[t];
                    if (code == 0) {
                        for (int i = 0; i < n; i++)
                            r += x[i];
                    } else if (code == 1 || code == 2) {
                        int m = med_x[aux];
                        for (int q = 0; q < n - m; q++)
                            r += x[m + q] * (code == 1 ? (double)q : (double)q * q);
                    } else if (code >= 3 && code <= 5) {
                        int sm = med_s[aux];
                        double ir = 0, ii = 0;
                        for (int q = 1; q < n - sm; q++) {
                            ir += tre[code][q] * x[q + sm];
                            ii += tim[code][q] * x[q + sm];
                        }
                        r = hypot(ir, ii);
                    } else if (code == 6) {
                        int sm = med_s[aux], q = t6_src[aux];
                        r = (double)q * x[sm + q];
                    } else {
                        r = x[t7_src[aux]];
                    }
                    sino[(size_t)t * n * n_angles + aux] = r;
                }
            }
        }
        free(rot);
    }
    for (int k = 0; k < 8; k++) {
        free(tre[k]);
        free(tim[k]);
    }
    return 0;
}

static double hermite_f64(unsigned order, double z)
{
    double h0 = 1.0, h1 = 2.0 * z;
    if (order == 0)
        h1 = 1.0;
    for (unsigned k = 2; k <= order; k++) {
        double h2 = 2.0 * z * h1 - 2.0 * (k - 1) * h0;
        h0 = h1;
        h1 = h2;
    }
    double fact = 1.0;
    for (unsigned k = 2; k <= order; k++)
        fact *= k;
    return h1 / (sqrt(pow(2.0, order) * fact * sqrt(M_PI)) * exp(z * z / 2));
}

int tto_circus_f64(const double *sino, int n, int n_cols, const int *pcodes,
                   const int *orders, const int *centers, int np,
                   const int32_t *p2_src, int geometry, double *out,
                   int nthreads)
{
    if (n <= 0 || n_cols <= 0 || np < 0)
        return -1;
    if (nthreads <= 0)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double *re = (double *)malloc(sizeof(double) * (size_t)n);
        double *im = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
        for (int a = 0; a < n_cols; a++) {
            const double *x = sino + (size_t)a * n;
            for (int p = 0; p < np; p++) {
                double r = 0;
                switch (pcodes[p]) {
                case 1:
                    for (int i = 1; i < n; i++)
                        r += fabs(x[i - 1] - x[i]);
                    break;
                case 2:
                    r = x[p2_src[(size_t)p * n_cols + a]];
                    break;
                case 3: {
                    tto_dft_f64(x, n, re, im);
                    double prev_l = 0, prev_m = 0;
                    for (int q = 0; q < n; q++) {
                        int k = q <= n / 2 ? q : n - q;
                        double mr = re[k] / n, mi = im[k] / n;
                        double mm = mr * mr + mi * mi;
                        double md = mm * mm;
                        double l = -1 + q * 2.0 / (n - 1);
                        if (geometry)
                            l = (float)l;
                        if (q > 0)
                            r += (l - prev_l) * (md + prev_m);
                        prev_l = l;
                        prev_m = md;
                    }
                    r *= 0.5;
                    break;
                }
                case 0: {
                    int c = centers[p];
                    double lo = 10.0 / c, hi = 10.0 / (double)(n - 1 - c);
                    float lof = (float)lo, hif = (float)hi;
                    for (int q = 0; q < n; q++) {
                        double z = q < c ? q * lo - 10 : (q - c) * hi;
                        if (geometry)
                            z = q < c ? (float)(q * lof - 10) : (float)((q - c) * hif);
                        r += x[q] * hermite_f64((unsigned)orders[p], z);
                    }
                    break;
                }
                }
                out[(size_t)p * n_cols + a] = r;
            }
        }
        free(re);
        free(im);
    }
    return 0;
}

/* c++/src/transform.cpp:49-112 without file output: sinograms, then per
 * T-functional the circus functions, each z-scored into
 * signatures(:, t*np + p) (A rows, column-major). */
int tto_transform(const float *img, int n, int n_angles, float step_deg,
                  const int *tcodes, int nt, const int *pcodes,
                  const int *orders, const int *centers, int np,
                  float *signatures, int nthreads)
{
    float *sino = (float *)malloc(sizeof(float) * (size_t)nt * n * n_angles);
    float *circ = (float *)malloc(sizeof(float) * (size_t)(np > 0 ? np : 1) * n_angles);
    int rc = tto_sinograms(img, n, n_angles, step_deg, tcodes, nt, sino, NULL,
                           NULL, NULL, NULL, NULL, NULL, nthreads);
    for (int t = 0; rc == 0 && t < nt; t++) {
        rc = tto_circus(sino + (size_t)t * n * n_angles, n, n_angles, pcodes,
                        orders, centers, np, circ, NULL, NULL, nthreads);
        for (int p = 0; rc == 0 && p < np; p++)
            tto_zscore(circ + (size_t)p * n_angles, n_angles,
                       signatures + (size_t)(t * np + p) * n_angles);
    }
    free(sino);
    free(circ);
    return rc;
}

/* ======================================================================= */
/* FP64 adjoint (vector-Jacobian product), medians and P2 ranks held fixed  */
/* ======================================================================= */

/* dS = J_P^T dC for one sinogram column x[0..n) (reference geometry: float
 * P3 abscissae, float Hermite z grid).  dc[p] is the upstream gradient of
 * circus function p at this column; src2[p] the P2 source index. */
static void circus_vjp_f64(const double *x, int n, const int *pcodes,
                           const int *orders, const int *centers, int np,
                           const double *dc, const int32_t *src2, double *ds,
                           double *re, double *im)
{
    for (int j = 0; j < n; j++)
        ds[j] = 0.0;
    for (int p = 0; p < np; p++) {
        const double g = dc[p];
        if (g == 0.0)
            continue;
        switch (pcodes[p]) {
        case 1: /* P1 = sum_{j>=1} |x[j-1] - x[j]| */
            for (int j = 0; j < n; j++) {
                double d = 0.0;
                if (j >= 1) {
                    double u = x[j - 1] - x[j];
                    d -= (u > 0) - (u < 0);
                }
                if (j <= n - 2) {
                    double u = x[j] - x[j + 1];
                    d += (u > 0) - (u < 0);
                }
                ds[j] += g * d;
            }
            break;
        case 2:
            ds[src2[p]] += g;
            break;
        case 3: { /* sum_p w_p |X_q(p)|^4 / N^4 */
            tto_dft_f64(x, n, re, im);
            const int half = n / 2;
            double *ck = (double *)calloc((size_t)half + 1, sizeof(double));
            for (int q = 0; q < n; q++) {
                double lq = (float)(-1 + q * 2.0 / (n - 1));
                double w = 0.0;
                if (q >= 1)
                    w += 0.5 * (lq - (double)(float)(-1 + (q - 1) * 2.0 / (n - 1)));
                if (q <= n - 2)
                    w += 0.5 * ((double)(float)(-1 + (q + 1) * 2.0 / (n - 1)) - lq);
                ck[q <= half ? q : n - q] += w;
            }
            const double n4 = (double)n * n * (double)n * n;
            for (int j = 0; j < n; j++) {
                double acc = 0.0;
                for (int k = 0; k <= half; k++) {
                    const double m2 = re[k] * re[k] + im[k] * im[k];
                    const long jk = ((long)j * k) % n;
                    const double th = 2.0 * M_PI * (double)jk / n;
                    acc += ck[k] * 4.0 * m2 / n4 * (re[k] * cos(th) - im[k] * sin(th));
                }
                ds[j] += g * acc;
            }
            free(ck);
            break;
        }
        case 0: {
            const int c = centers[p];
            const float lof = (float)(10.0 / c), hif = (float)(10.0 / (double)(n - 1 - c));
            for (int q = 0; q < n; q++) {
                double z = q < c ? (float)(q * lof - 10) : (float)((q - c) * hif);
                ds[q] += g * hermite_f64((unsigned)orders[p], z);
            }
            break;
        }
        }
    }
}

/* J^T dC of the whole pipeline (sinograms of T codes 0..7, then circus),
 * in FP64 with reference geometry, medians (med_x, med_s), T6/T7 sources
 * (t6_src relative to sm, t7_src absolute, [A][n]) and P2 sources (p2_src
 * [nt][np][A]) fixed at the FP32 run's values.  dcircus is [nt][np][A];
 * dimg is n*n (column-major).  T3..T5 use their double tables. */
int tto_backward_f64(const float *img, int n, int n_angles, float step_deg,
                     const int *tcodes, int nt, const int *pcodes,
                     const int *orders, const int *centers, int np,
                     const int32_t *med_x, const int32_t *med_s,
                     const int32_t *t6_src, const int32_t *t7_src,
                     const int32_t *p2_src, const double *dcircus, double *dimg,
                     int nthreads)
{
    if (n <= 0 || n_angles <= 0 || nt <= 0 || !valid_tcodes(tcodes, nt))
        return -1;
    for (int t = 0; t < nt; t++)
        if ((tcodes[t] == 6 && !t6_src) || (tcodes[t] == 7 && !t7_src))
            return -1;
    const int A = n_angles;
    double *sino = (double *)malloc(sizeof(double) * (size_t)nt * n * A);
    tto_sinograms_f64(img, n, A, step_deg, tcodes, nt, med_x, med_s, t6_src, t7_src, 1,
                      sino, nthreads);
    double *dsino = (double *)calloc((size_t)nt * n * A, sizeof(double));
    if (nthreads <= 0)
        nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double *re = (double *)malloc(sizeof(double) * (size_t)n);
        double *im = (double *)malloc(sizeof(double) * (size_t)n);
        double *dc = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
        int32_t *src = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
#pragma omp for schedule(static)
        for (int ta = 0; ta < nt * A; ta++) {
            const int t = ta / A, a = ta % A;
            for (int p = 0; p < np; p++) {
                dc[p] = dcircus[((size_t)t * np + p) * A + a];
                src[p] = p2_src ? p2_src[((size_t)t * np + p) * A + a] : 0;
            }
            circus_vjp_f64(sino + (size_t)ta * n, n, pcodes, orders, centers, np, dc,
                           src, dsino + (size_t)ta * n, re, im);
        }
        free(re);
        free(im);
        free(dc);
        free(src);
    }
    int rc = tto_sinograms_vjp_f64(img, n, A, step_deg, tcodes, nt, med_x, med_s,
                                   t6_src, t7_src, dsino, dimg, nthreads);
    free(sino);
    free(dsino);
    return rc;
}

/* J_sinograms^T dsino ([nt][A][n]) in FP64, reference geometry, medians
 * and T6/T7 sources fixed; dimg is n*n column-major.  With the sort fixed,
 * T6 = q * x[sm + q] (q = t6_src) and T7 = x[t7_src] are linear in one
 * sample each (c++/src/functionals.cpp:226-276). */
int tto_sinograms_vjp_f64(const float *img, int n, int n_angles, float step_deg,
                          const int *tcodes, int nt, const int32_t *med_x,
                          const int32_t *med_s, const int32_t *t6_src,
                          const int32_t *t7_src, const double *dsino, double *dimg,
                          int nthreads)
{
    if (n <= 0 || n_angles <= 0 || nt <= 0 || !valid_tcodes(tcodes, nt))
        return -1;
    for (int t = 0; t < nt; t++)
        if ((tcodes[t] == 6 && !t6_src) || (tcodes[t] == 7 && !t7_src))
            return -1;
    if (nthreads <= 0)
        nthreads = 1;
    const int A = n_angles;
    double *tre[6] = {0}, *tim[6] = {0};
    for (int t = 0; t < nt; t++) {
        int code = tcodes[t];
        if (code >= 3 && code <= 5 && !tre[code]) {
            tre[code] = (double *)malloc(sizeof(double) * (size_t)n);
            tim[code] = (double *)malloc(sizeof(double) * (size_t)n);
            t345_table_f64(code, n, tre[code], tim[code]);
        }
    }
    memset(dimg, 0, sizeof(double) * (size_t)n * n);
#pragma omp parallel num_threads(nthreads)
    {
        double *acc = (double *)calloc((size_t)n * n, sizeof(double));
        double *rot = (double *)malloc(sizeof(double) * (size_t)n * n);
        double *G = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for schedule(static)
        for (int a = 0; a < A; a++) {
            float c, s;
            tto_rotation_trig(tto_deg2rad(tto_angle_deg(a, step_deg)), &c, &s);
            rotate_f64(img, n, c, s, 1, rot);
            const float o = (float)((n - 1) / 2.0), ns = -s, lim = (float)(n - 1);
            for (int col = 0; col < n; col++) {
                const double *x = rot + (size_t)col * n;
                const size_t line = (size_t)a * n + col;
                for (int r = 0; r < n; r++)
                    G[r] = 0.0;
                for (int t = 0; t < nt; t++) {
                    const double g = dsino[(size_t)t * n * A + line];
                    const int code = tcodes[t];
                    if (g == 0.0)
                        continue;
                    if (code == 0) {
                        for (int r = 0; r < n; r++)
                            G[r] += g;
                    } else if (code == 1 || code == 2) {
                        const int m = med_x[line];
                        for (int r = m; r < n; r++)
                            G[r] += g * (code == 1 ? (double)(r - m)
                                                   : (double)(r - m) * (r - m));
                    } else if (code == 6) {
                        const int q = t6_src[line];
                        G[med_s[line] + q] += g * (double)q;
                    } else if (code == 7) {
                        G[t7_src[line]] += g;
                    } else {
                        const int sm = med_s[line];
                        double zr = 0.0, zi = 0.0;
                        for (int q = 1; q < n - sm; q++) {
                            zr += tre[code][q] * x[sm + q];
                            zi += tim[code][q] * x[sm + q];
                        }
                        const double z = hypot(zr, zi);
                        if (z > 0.0)
                            for (int q = 1; q < n - sm; q++)
                                G[sm + q] += g * (zr * tre[code][q] + zi * tim[code][q]) / z;
                    }
                }
                /* transposed bilinear sampling at the reference's float
                 * coordinates */
                for (int r = 0; r < n; r++) {
                    if (G[r] == 0.0)
                        continue;
                    float px = (float)col - o, py = (float)r - o;
                    float xx = px * c + py * s;
                    float yy = px * ns + py * c;
                    xx = xx + o;
                    yy = yy + o;
                    if (!(xx >= 0 && xx < lim && yy >= 0 && yy < lim))
                        continue;
                    float xi, yi;
                    double fx = modff(xx, &xi), fy = modff(yy, &yi);
                    long X = (long)xi, Y = (long)yi;
                    acc[Y + X * n] += G[r] * (1 - fx) * (1 - fy);
                    acc[Y + (X + 1) * n] += G[r] * fx * (1 - fy);
                    acc[(Y + 1) + X * n] += G[r] * (1 - fx) * fy;
                    acc[(Y + 1) + (X + 1) * n] += G[r] * fx * fy;
                }
            }
        }
#pragma omp critical(tto_backward_reduce)
        for (size_t i = 0; i < (size_t)n * n; i++)
            dimg[i] += acc[i];
        free(acc);
        free(rot);
        free(G);
    }
    for (int k = 0; k < 6; k++) {
        free(tre[k]);
        free(tim[k]);
    }
    return 0;
}
