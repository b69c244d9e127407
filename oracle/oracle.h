/*
 * oracle.h -- CPU oracle for the cvortex direct-sum hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() (as the checker) and by bench.py (cpu_baseline
 * leg, timed as the CPU baseline).
 *
 * Two halves:
 *
 *   refcpu.c  A restatement of the reference library's OpenMP CPU path:
 *             FP32 pair arithmetic with the reference's own formula order
 *             (powf, expf, sqrtf), regularisations called through function
 *             pointers per pair, FP64 per-target accumulation, the same
 *             OpenMP loop nests.  Functions are prefixed orc_.
 *             This is the "reference" side of every parity test and the
 *             CPU baseline ("port") of bench.py.
 *
 *   fp64.c    The same physics evaluated entirely in double precision, with
 *             a choice of Gaussian regularisation: the reference's
 *             Abramowitz & Stegun 7.1.26 erf evaluated in FP64 (mode 0) or
 *             exact erf with a small-argument series (mode 1).  Functions are
 *             prefixed orc64_.  Used for max/RMS relative error reports.
 *
 * Regularisation ids: 0 singular, 1 winckelmans, 2 planetary, 3 gaussian.
 */
#ifndef CVTX_ORACLE_H
#define CVTX_ORACLE_H

#include <cvortex/libcvtx.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SINGULAR = 0, ORC_WINCKELMANS = 1, ORC_PLANETARY = 2, ORC_GAUSSIAN = 3 };
enum { ORC_FN_G3 = 0, ORC_FN_G2 = 1, ORC_FN_ZETA3 = 2, ORC_FN_ETA3 = 3,
       ORC_FN_ETA2 = 4, ORC_FN_COMB_G = 5, ORC_FN_COMB_ZETA = 6 };

/* ---- refcpu.c: FP32 restatement of the reference CPU path ---- */
float orc_vf_eval(int reg, int fn, float rho);
int orc_num_threads(void);

bsv_V3f orc_P3D_S2S_vel(const cvtx_P3D* self, bsv_V3f mes, int reg, float sigma);
bsv_V3f orc_P3D_S2S_dvort(const cvtx_P3D* self, const cvtx_P3D* induced, int reg, float sigma);
bsv_V3f orc_P3D_S2S_visc_dvort(const cvtx_P3D* self, const cvtx_P3D* induced, int reg,
                               float sigma, float nu);
void orc_P3D_M2M_vel(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nm,
                     bsv_V3f* out, int reg, float sigma);
void orc_P3D_M2M_dvort(const cvtx_P3D** src, int ns, const cvtx_P3D** ind, int ni,
                       bsv_V3f* out, int reg, float sigma);
void orc_P3D_M2M_visc_dvort(const cvtx_P3D** src, int ns, const cvtx_P3D** ind, int ni,
                            bsv_V3f* out, int reg, float sigma, float nu);

bsv_V3f orc_P3D_S2S_vort(const cvtx_P3D* self, bsv_V3f mes, int reg, float sigma);
bsv_V3f orc_P3D_M2S_vort(const cvtx_P3D** src, int ns, bsv_V3f mes, int reg, float sigma);
void orc_P3D_M2M_vort(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nm,
                      bsv_V3f* out, int reg, float sigma);
void orc_P3D_pedrizzetti_relaxation(cvtx_P3D** p, int n, float fdt, int reg, float sigma);

bsv_V3f orc_F3D_S2S_vel(const cvtx_F3D* self, bsv_V3f mes);
bsv_V3f orc_F3D_S2S_dvort(const cvtx_F3D* self, const cvtx_P3D* induced);
void orc_F3D_M2M_vel(const cvtx_F3D** src, int nf, const bsv_V3f* mes, int nm, bsv_V3f* out);
void orc_F3D_M2M_dvort(const cvtx_F3D** src, int nf, const cvtx_P3D** ind, int ni, bsv_V3f* out);
void orc_F3D_inf_mtrx(const cvtx_F3D** src, int nf, const bsv_V3f* mes, const bsv_V3f* dir,
                      int nm, float* out);

/* redistribution: rid 0..4 = lambda0, lambda1, lambda2, lambda3, m4p */
float orc_redist_eval(int rid, float U);
float orc_redist_radius(int rid);
int orc_P3D_redistribute_on_grid(const cvtx_P3D** in, int n, cvtx_P3D* out, int max_out,
                                 int rid, float h, float negligible);
int orc_P2D_redistribute_on_grid(const cvtx_P2D** in, int n, cvtx_P2D* out, int max_out,
                                 int rid, float h, float negligible);

bsv_V2f orc_P2D_S2S_vel(const cvtx_P2D* self, bsv_V2f mes, int reg, float sigma);
float orc_P2D_S2S_visc_dvort(const cvtx_P2D* self, const cvtx_P2D* induced, int reg,
                             float sigma, float nu);
void orc_P2D_M2M_vel(const cvtx_P2D** src, int ns, const bsv_V2f* mes, int nm,
                     bsv_V2f* out, int reg, float sigma);
void orc_P2D_M2M_visc_dvort(const cvtx_P2D** src, int ns, const cvtx_P2D** ind, int ni,
                            float* out, int reg, float sigma, float nu);

/* ---- fp64.c: double-precision oracle ----
 * Inputs are the same caller records (float), promoted to double.  Outputs
 * are double arrays, 3 (or 2, or 1) per target, row-major.  Targets are
 * selected by an index list so that large configurations can be sampled:
 * tgt_idx[k] indexes the target arrays; out has n_idx rows. */
double orc64_vf_eval(int reg, int fn, double rho, int gauss_mode);
void orc64_P3D_M2M(const cvtx_P3D** src, int ns, const cvtx_P3D** ind,
                   const int* tgt_idx, int n_idx, int reg, double sigma, double nu,
                   int gauss_mode, double* vel, double* dvort, double* visc);
void orc64_P3D_M2M_vel_pts(const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                           const int* tgt_idx, int n_idx, int reg, double sigma,
                           int gauss_mode, double* vel);
/* Vorticity field with the reference's 5 sigma box cut-off (the box test
 * is made in FP32 exactly as the reference makes it, the rest in double),
 * and the Pedrizzetti relaxation of n particles: alpha_out[3 n] (double). */
void orc64_P3D_M2M_vort(const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                        const int* tgt_idx, int n_idx, int reg, double sigma,
                        int gauss_mode, double* out);
void orc64_P3D_pedrizzetti(const cvtx_P3D** p, int n, double fdt, int reg, double sigma,
                           int gauss_mode, double* alpha_out);
void orc64_F3D_M2M(const cvtx_F3D** src, int nf, const cvtx_P3D** ind,
                   const int* tgt_idx, int n_idx, double* vel, double* dvort);
/* influence matrix in double: out[k * nf + j] for point tgt_idx[k] */
void orc64_F3D_inf_mtrx(const cvtx_F3D** src, int nf, const bsv_V3f* mes, const bsv_V3f* dir,
                        const int* tgt_idx, int n_idx, double* out);
void orc64_F3D_M2M_vel_pts(const cvtx_F3D** src, int nf, const bsv_V3f* mes,
                           const int* tgt_idx, int n_idx, double* vel);
void orc64_P2D_M2M(const cvtx_P2D** src, int ns, const cvtx_P2D** ind,
                   const int* tgt_idx, int n_idx, int reg, double sigma, double nu,
                   int gauss_mode, double* vel, double* visc);
void orc64_P2D_M2M_vel_pts(const cvtx_P2D** src, int ns, const bsv_V2f* mes,
                           const int* tgt_idx, int n_idx, int reg, double sigma,
                           int gauss_mode, double* vel);

#ifdef __cplusplus
}
#endif

#endif /* CVTX_ORACLE_H */
