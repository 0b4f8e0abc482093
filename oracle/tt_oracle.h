/*
 * tt_oracle.h -- CPU restatement of the reference trace transform.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tracetransform_amd/,
 * include/) links, loads or calls this code.  It is used by tests/ as the
 * checker, by __graft_entry__.smoke() as the checker, and by bench.py's
 * cpu_baseline leg as the timed CPU port of the reference's OpenMP variant.
 *
 * Layout conventions follow the reference (Eigen column-major MatrixXf):
 *   image / rotated image   : n x n, element (row, col) at row + col*n
 *   sinogram of T-functional: n rows (column index of the rotated image)
 *                             x A columns (angles), element (c, a) at c + a*n
 *   circus function         : length A
 *
 * Functional codes follow the reference enums:
 *   TFunctional Radon..T7 = 0..7   (c++/src/sinogram.hpp:22-31)
 *   PFunctional Hermite=0, P1=1, P2=2, P3=3   (c++/src/circus.hpp:27-32)
 */
#ifndef TT_ORACLE_H
#define TT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry (c++/src/auxiliary.cpp) ---------------------------------- */
int   tto_pad_geometry(int rows, int cols, int *n, int *dfx, int *dfy);
void  tto_pad(const float *img, int rows, int cols, float *out);
void  tto_resize(const float *in, int in_rows, int in_cols, float *out,
                 int rows, int cols);
float tto_deg2rad(float degrees);
float tto_angle_deg(int a_step, float step_deg);
void  tto_rotation_trig(float angle_rad, float *c, float *s);
float tto_interpolate(const float *src, int n, float x, float y);
void  tto_rotate(const float *in, int n, float c, float s, float *out);

/* ---- reductions and medians (c++/src/functionals.cpp:21-60) ------------ */
float tto_eigen_sum(const float *x, int n);
float tto_seq_sum(const float *x, int n);
int   tto_weighted_median(const float *x, int n);
void  tto_stable_argsort(const float *key, int n, int *idx, int *tmp);

/* ---- T-functionals on one column (c++/src/functionals.cpp:95-276) ------ */
void  tto_t345_table(int which, int n, float *re, float *im);
float tto_T0(const float *x, int n);
float tto_T1(const float *x, int n, int *m_out);
float tto_T2(const float *x, int n, int *m_out);
float tto_T345(const float *x, int n, const float *re, const float *im,
               int *sm_out);
float tto_T6(const float *x, int n, int *sm_out, int *rank_out, int *src_out);
float tto_T7(const float *x, int n, int *m_out, int *rank_out, int *src_out);

/* ---- P-functionals on one column (c++/src/functionals.cpp:287-386) ----- */
float tto_P1(const float *x, int n);
float tto_P2(const float *x, int n, int *rank_out, int *src_out);
float tto_P3(const float *x, int n);
float tto_hermite_function(unsigned order, float z);
float tto_Phermite(const float *x, int n, unsigned order, int center);

/* ---- drivers (c++/src/sinogram.cpp:67-178, c++/src/circus.cpp:104-177) - */
/* aux arrays are [A*n] (index a*n + col) and may be NULL.                  */
int tto_sinograms(const float *img, int n, int n_angles, float step_deg,
                  const int *tcodes, int nt, float *sino,
                  int32_t *med_x, int32_t *med_s,
                  int32_t *t6_rank, int32_t *t6_src,
                  int32_t *t7_rank, int32_t *t7_src, int nthreads);
/* p2 arrays are [np*A] (index p*A + a) and may be NULL. */
int tto_circus(const float *sino, int n, int n_cols, const int *pcodes,
               const int *orders, const int *centers, int np, float *out,
               int32_t *p2_rank, int32_t *p2_src, int nthreads);
void tto_zscore(const float *in, int n, float *out);

/* ---- FP64 oracle: same algorithm in double, indices from the FP32 run --- */
/* geometry 0: sample positions in double; 1: the reference's float sample
 * positions with values in double. */
int tto_sinograms_f64(const float *img, int n, int n_angles, float step_deg,
                      const int *tcodes, int nt, const int32_t *med_x,
                      const int32_t *med_s, const int32_t *t6_src,
                      const int32_t *t7_src, int geometry, double *sino,
                      int nthreads);
int tto_circus_f64(const double *sino, int n, int n_cols, const int *pcodes,
                   const int *orders, const int *centers, int np,
                   const int32_t *p2_src, int geometry, double *out,
                   int nthreads);
void tto_dft_f64(const double *x, int n, double *re, double *im);

/* ---- FP64 adjoint: J^T dC with medians and P2 sources fixed ------------ */
/* dcircus and p2_src are [nt][np][A]; dimg is n*n column-major.  T codes
 * 0..7 (T6 and T7 with their sources t6_src, relative to sm, and t7_src,
 * absolute, fixed as well), P codes 0..3, reference geometry. */
int tto_backward_f64(const float *img, int n, int n_angles, float step_deg,
                     const int *tcodes, int nt, const int *pcodes,
                     const int *orders, const int *centers, int np,
                     const int32_t *med_x, const int32_t *med_s,
                     const int32_t *t6_src, const int32_t *t7_src,
                     const int32_t *p2_src, const double *dcircus, double *dimg,
                     int nthreads);

int tto_sinograms_vjp_f64(const float *img, int n, int n_angles, float step_deg,
                          const int *tcodes, int nt, const int32_t *med_x,
                          const int32_t *med_s, const int32_t *t6_src,
                          const int32_t *t7_src, const double *dsino, double *dimg,
                          int nthreads);

/* ---- full reference step: sinograms -> circus -> zscore ---------------- */
int tto_transform(const float *img, int n, int n_angles, float step_deg,
                  const int *tcodes, int nt, const int *pcodes,
                  const int *orders, const int *centers, int np,
                  float *signatures, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
