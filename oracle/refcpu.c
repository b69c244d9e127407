/*
 * refcpu.c -- FP32 restatement of the reference cvortex OpenMP CPU path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Never linked into the library.
 *
 * Each function states which reference lines it follows.  The point of the
 * restatement is to reproduce the reference's FP32 pair arithmetic in the
 * same operation order (so parity tests compare against what the reference
 * would compute), including its structure: regularisation functions are
 * called through function pointers per pair, sums over sources are FP64,
 * loops over targets are `omp parallel for schedule(static)`.
 *
 * Vector helpers stand in for the un-vendored "bsv" package (reference
 * CMakeLists.txt:69-71).  Their arithmetic is textbook vector algebra; the
 * rounding order chosen here (left-to-right dot products, sqrtf of the dot
 * for the norm, component-wise division for scalar division) is an
 * assumption, so results can differ from a real reference build by a few
 * ULP per pair -- far below the 1e-5 parity tolerance.
 *
 * Planted mistakes: building with -DORACLE_MUTANT=k changes one formula
 * (see the ORC_MUT(k) sites).  tests/test_oracle_mutants.py builds these
 * variants and checks that the pinning tests reject every one of them.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#define ORC_MUT(k) (ORACLE_MUTANT == (k))

/* pi as spelled in reference src/P3D.cpp:47 and src/F3D.cpp:32. */
#define REF_PI_F 3.14159265359f
/* reference src/VortFunc.cpp:49-50 */
#define REF_SQRT_2_OVER_PI 0.7978845608028654f
#define REF_RECIP_SQRT_2 0.7071067811865475f

int orc_num_threads(void) {
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* vector helpers (stand-ins for bsv)                                  */
/* ------------------------------------------------------------------ */
static inline bsv_V3f v3(float a, float b, float c) { bsv_V3f r = {{a, b, c}}; return r; }
static inline bsv_V3f v3_sub(bsv_V3f a, bsv_V3f b) { return v3(a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]); }
static inline bsv_V3f v3_add(bsv_V3f a, bsv_V3f b) { return v3(a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]); }
static inline bsv_V3f v3_scale(bsv_V3f a, float s) { return v3(a.x[0] * s, a.x[1] * s, a.x[2] * s); }
static inline bsv_V3f v3_divs(bsv_V3f a, float s) { return v3(a.x[0] / s, a.x[1] / s, a.x[2] / s); }
static inline float v3_dot(bsv_V3f a, bsv_V3f b) { return a.x[0] * b.x[0] + a.x[1] * b.x[1] + a.x[2] * b.x[2]; }
static inline float v3_norm(bsv_V3f a) { return sqrtf(v3_dot(a, a)); }
static inline bsv_V3f v3_cross(bsv_V3f a, bsv_V3f b) {
	return v3(a.x[1] * b.x[2] - a.x[2] * b.x[1],
	          a.x[2] * b.x[0] - a.x[0] * b.x[2],
	          a.x[0] * b.x[1] - a.x[1] * b.x[0]);
}
static inline int v3_eq(bsv_V3f a, bsv_V3f b) {
	return a.x[0] == b.x[0] && a.x[1] == b.x[1] && a.x[2] == b.x[2];
}
static inline bsv_V2f v2(float a, float b) { bsv_V2f r = {{a, b}}; return r; }
static inline bsv_V2f v2_sub(bsv_V2f a, bsv_V2f b) { return v2(a.x[0] - b.x[0], a.x[1] - b.x[1]); }
static inline float v2_norm(bsv_V2f a) { return sqrtf(a.x[0] * a.x[0] + a.x[1] * a.x[1]); }
static inline int v2_eq(bsv_V2f a, bsv_V2f b) { return a.x[0] == b.x[0] && a.x[1] == b.x[1]; }

/* ------------------------------------------------------------------ */
/* Regularisations: reference src/VortFunc.cpp:52-199                  */
/* ------------------------------------------------------------------ */
typedef struct {
	float (*g3)(float);
	float (*g2)(float);
	float (*zeta3)(float);
	void (*combined3)(float, float*, float*);
	float (*eta3)(float);
	float (*eta2)(float);
} orc_vortfunc;

/* VortFunc.cpp:52-62 (release build: one warning, returns zero) */
static float eta_unsupported(float rho) {
	static int warned = 0;
	(void)rho;
	if (!warned) {
		fprintf(stderr, "oracle: viscous term requested for a regularisation "
		                "without eta; modelled as inviscid (zero).\n");
		warned = 1;
	}
	return 0.f;
}

/* singular, VortFunc.cpp:64-80 */
static float sing_g3(float rho) { (void)rho; return 1.f; }
static float sing_zeta3(float rho) { (void)rho; return 0.f; }
static void sing_comb3(float rho, float* g, float* z) { (void)rho; *g = 1.f; *z = 0.f; }
static float sing_g2(float rho) { (void)rho; return 1.f; }

/* Winckelmans, VortFunc.cpp:82-131 */
static float winck_g3(float rho) {
	float rr = rho * rho;
	float num = (rr + 2.5f) * rho * rr;
	float base = rr + 1.f;
	return num * powf(base, ORC_MUT(1) ? -1.5f : -2.5f);
}
static float winck_zeta3(float rho) {
	return 7.5f * powf(rho * rho + 1.f, -3.5f);
}
static float winck_eta3(float rho) {
	float c = ORC_MUT(9) ? 50.f : 52.5f;
	return c * powf(rho * rho + 1.f, -4.5f);
}
static void winck_comb3(float rho, float* g, float* z) { *g = winck_g3(rho); *z = winck_zeta3(rho); }
static float winck_g2(float rho) {
	float rr = rho * rho;
	float top = rr * rr + rr * (ORC_MUT(12) ? 1.f : 2.f);
	float bot = rr * rr + 2.f * rr + 1.f;
	return top / bot;
}
static float winck_eta2(float rho) {
	/* reference formula: 24 exp(4 a / a^4) / a^4, a = rho^2 + 1 */
	float a = rho * rho + 1.f;
	float ia2 = 1.f / (a * a);
	float e = expf(4.f * a * (ia2 * ia2));
	return 24.f * e * (ia2 * ia2);
}

/* planetary, VortFunc.cpp:133-153 */
static float plan_g3(float rho) { return rho < 1.f ? rho * rho * rho : 1.f; }
static float plan_zeta3(float rho) { return rho < 1.f ? 3.f : 0.f; }
static void plan_comb3(float rho, float* g, float* z) { *g = plan_g3(rho); *z = plan_zeta3(rho); }
static float plan_g2(float rho) { return rho < 1.f ? rho * rho : 1.f; }

/* Gaussian, VortFunc.cpp:155-199 (Abramowitz & Stegun 7.1.26 erf) */
static float gauss_g3(float rho) {
	if (rho > 6.f) return 1.f;
	const float c1 = 0.254829592f, c2 = -0.284496736f, c3 = 1.421413741f;
	const float c4 = -1.453152027f, c5 = 1.061405429f, pp = 0.3275911f;
	float y = rho * REF_RECIP_SQRT_2;
	float t = 1.f / (1.f + pp * y);
	float t2 = t * t, t3 = t2 * t, t4 = t2 * t2, t5 = t3 * t2;
	float ey = expf(-y * y);
	float erfv = 1.f - (c1 * t + c2 * t2 + c3 * t3 + c4 * t4 + c5 * t5) * ey;
	float tail = rho * REF_SQRT_2_OVER_PI * expf(-y * y);
	return erfv - tail;
}
static float gauss_zeta3(float rho) {
	float c = ORC_MUT(2) ? 0.3989422804f : REF_SQRT_2_OVER_PI;
	return c * expf(-rho * rho * 0.5f);
}
static void gauss_comb3(float rho, float* g, float* z) { *g = gauss_g3(rho); *z = gauss_zeta3(rho); }
static float gauss_g2(float rho) { return 1.f - expf(-rho * rho * 0.5f); }
static float gauss_eta2(float rho) { return expf(-rho * rho * (ORC_MUT(11) ? 1.f : 0.5f)); }

/* factories, VortFunc.cpp:201-251 */
static orc_vortfunc orc_vf(int reg) {
	orc_vortfunc f;
	switch (reg) {
	case ORC_SINGULAR:
		f.g3 = sing_g3; f.g2 = sing_g2; f.zeta3 = sing_zeta3; f.combined3 = sing_comb3;
		f.eta3 = eta_unsupported; f.eta2 = eta_unsupported; break;
	case ORC_WINCKELMANS:
		f.g3 = winck_g3; f.g2 = winck_g2; f.zeta3 = winck_zeta3; f.combined3 = winck_comb3;
		f.eta3 = winck_eta3; f.eta2 = winck_eta2; break;
	case ORC_PLANETARY:
		f.g3 = plan_g3; f.g2 = plan_g2; f.zeta3 = plan_zeta3; f.combined3 = plan_comb3;
		f.eta3 = eta_unsupported; f.eta2 = eta_unsupported; break;
	default:
		f.g3 = gauss_g3; f.g2 = gauss_g2; f.zeta3 = gauss_zeta3; f.combined3 = gauss_comb3;
		f.eta3 = gauss_zeta3; f.eta2 = gauss_eta2; break;
	}
	return f;
}

float orc_vf_eval(int reg, int fn, float rho) {
	orc_vortfunc f = orc_vf(reg);
	float g, z;
	switch (fn) {
	case ORC_FN_G3: return f.g3(rho);
	case ORC_FN_G2: return f.g2(rho);
	case ORC_FN_ZETA3: return f.zeta3(rho);
	case ORC_FN_ETA3: return f.eta3(rho);
	case ORC_FN_ETA2: return f.eta2(rho);
	case ORC_FN_COMB_G: f.combined3(rho, &g, &z); return g;
	case ORC_FN_COMB_ZETA: f.combined3(rho, &g, &z); return z;
	}
	return NAN;
}

/* ------------------------------------------------------------------ */
/* P3D pair terms                                                      */
/* ------------------------------------------------------------------ */

/* velocity without the 1/(4 pi) factor; reference src/P3D.cpp:51-72 */
static inline bsv_V3f p3d_vel_pair(const cvtx_P3D* self, bsv_V3f mes,
                                   const orc_vortfunc* f, float recip_sigma) {
	if (v3_eq(self->coord, mes)) return v3(0.f, 0.f, 0.f);
	bsv_V3f d = v3_sub(mes, self->coord);
	float r = v3_norm(d);
	float rho = r * recip_sigma;
	float neg_g = -f->g3(rho);
	float inv_r3 = powf(r, -3.f);
	bsv_V3f c = ORC_MUT(3) ? v3_cross(self->vorticity, d) : v3_cross(d, self->vorticity);
	return v3_scale(c, neg_g * inv_r3);
}

/* reference src/P3D.cpp:74-84 */
bsv_V3f orc_P3D_S2S_vel(const cvtx_P3D* self, bsv_V3f mes, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	bsv_V3f v = p3d_vel_pair(self, mes, &f, 1.f / fabsf(sigma));
	return v3_scale(v, 1.f / (4.f * REF_PI_F));
}

/* reference src/P3D.cpp:86-114 (transpose-scheme stretching) */
static inline bsv_V3f p3d_dvort_pair(const cvtx_P3D* self, const cvtx_P3D* ind,
                                     const orc_vortfunc* f, float sigma) {
	if (v3_eq(self->coord, ind->coord)) return v3(0.f, 0.f, 0.f);
	bsv_V3f d = v3_sub(ind->coord, self->coord);
	float r = v3_norm(d);
	float rho = fabsf(r / sigma);
	float g, z;
	f->combined3(rho, &g, &z);
	bsv_V3f w = v3_cross(ind->vorticity, self->vorticity);
	float pre = 1.f / (4.f * REF_PI_F * powf(sigma, 3.f));
	float rho3 = rho * rho * rho;
	bsv_V3f term1 = v3_divs(v3_scale(w, g), rho3);
	float a = -1.f / (r * r);
	float b = (3.f * g) / rho3 - z;
	float dw = v3_dot(d, w);
	bsv_V3f term2 = v3_scale(d, a * b * dw);
	bsv_V3f sum = ORC_MUT(4) ? term1 : v3_add(term1, term2);
	return v3_scale(sum, pre);
}

bsv_V3f orc_P3D_S2S_dvort(const cvtx_P3D* self, const cvtx_P3D* ind, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	return p3d_dvort_pair(self, ind, &f, sigma);
}

/* reference src/P3D.cpp:116-144 (PSE, one direction, no 1/(4 pi sigma^3)) */
static inline bsv_V3f p3d_visc_pair(const cvtx_P3D* self, const cvtx_P3D* ind,
                                    const orc_vortfunc* f, float sigma, float nu) {
	if (v3_eq(self->coord, ind->coord)) return v3(0.f, 0.f, 0.f);
	bsv_V3f d = v3_sub(self->coord, ind->coord);
	float r = v3_norm(d);
	float rho = fabsf(r / sigma);
	float pre = (ORC_MUT(20) ? 1.f : 2.f) * nu / powf(sigma, 2.f);
	bsv_V3f t1 = v3_scale(self->vorticity, ORC_MUT(21) ? self->volume : ind->volume);
	bsv_V3f t2 = v3_scale(ind->vorticity, (ORC_MUT(5) ? 1.f : -1.f) *
	                                      (ORC_MUT(21) ? ind->volume : self->volume));
	bsv_V3f t = v3_add(t1, t2);
	float eta = f->eta3(rho);
	return v3_scale(v3_scale(t, eta), pre);
}

bsv_V3f orc_P3D_S2S_visc_dvort(const cvtx_P3D* self, const cvtx_P3D* ind, int reg,
                               float sigma, float nu) {
	orc_vortfunc f = orc_vf(reg);
	return p3d_visc_pair(self, ind, &f, sigma, nu);
}

/* ------------------------------------------------------------------ */
/* P3D M2S / M2M: reference src/P3D.cpp:230-296, 324-456               */
/* The inner reduction of M2S_vel is an OpenMP region nested inside    */
/* the outer parallel loop; with default (disabled) nesting it runs    */
/* serially, which is what is restated here.                           */
/* ------------------------------------------------------------------ */
static bsv_V3f p3d_m2s_vel(const cvtx_P3D** src, int ns, bsv_V3f mes,
                           const orc_vortfunc* f, float sigma) {
	double sx = 0, sy = 0, sz = 0;
	float rs = 1.f / fabsf(sigma);
	long n = ORC_MUT(10) ? (long)ns - 1 : (long)ns;
	for (long i = 0; i < n; ++i) {
		bsv_V3f v = p3d_vel_pair(src[i], mes, f, rs);
		sx += v.x[0]; sy += v.x[1]; sz += v.x[2];
	}
	bsv_V3f r = v3((float)sx, (float)sy, (float)sz);
	return v3_scale(r, 1.f / (4.f * REF_PI_F));
}

void orc_P3D_M2M_vel(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nm,
                     bsv_V3f* out, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	long i;
#pragma omp parallel for schedule(static)
	for (i = 0; i < nm; ++i)
		out[i] = p3d_m2s_vel(src, ns, mes[i], &f, sigma);
}

void orc_P3D_M2M_dvort(const cvtx_P3D** src, int ns, const cvtx_P3D** ind, int ni,
                       bsv_V3f* out, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	long i;
#pragma omp parallel for schedule(static)
	for (i = 0; i < ni; ++i) {
		double sx = 0, sy = 0, sz = 0;
		for (long j = 0; j < ns; ++j) {
			bsv_V3f v = p3d_dvort_pair(src[j], ind[i], &f, sigma);
			sx += v.x[0]; sy += v.x[1]; sz += v.x[2];
		}
		out[i] = v3((float)sx, (float)sy, (float)sz);
	}
}

void orc_P3D_M2M_visc_dvort(const cvtx_P3D** src, int ns, const cvtx_P3D** ind, int ni,
                            bsv_V3f* out, int reg, float sigma, float nu) {
	orc_vortfunc f = orc_vf(reg);
	long i;
#pragma omp parallel for schedule(static)
	for (i = 0; i < ni; ++i) {
		double sx = 0, sy = 0, sz = 0;
		for (long j = 0; j < ns; ++j) {
			bsv_V3f v = p3d_visc_pair(src[j], ind[i], &f, sigma, nu);
			sx += v.x[0]; sy += v.x[1]; sz += v.x[2];
		}
		out[i] = v3((float)sx, (float)sy, (float)sz);
	}
}

/* ------------------------------------------------------------------ */
/* P3D vorticity field and Pedrizzetti relaxation: reference           */
/* src/P3D.cpp:146-161, 213-228, 298-322, 458-498, 669-707             */
/* ------------------------------------------------------------------ */

/* reference src/P3D.cpp:146-161: no cut-off, rho = |r| / sigma (signed) */
bsv_V3f orc_P3D_S2S_vort(const cvtx_P3D* self, bsv_V3f mes, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	bsv_V3f d = v3_sub(self->coord, mes);
	float r = v3_norm(d);
	float coeff = f.zeta3(r / sigma);
	float divisor = 4.f * REF_PI_F * sigma * sigma * sigma;
	coeff = coeff / divisor;
	return v3_scale(self->vorticity, coeff);
}

/* reference src/P3D.cpp:298-322: sources outside the box |dx|,|dy|,|dz| <
 * 5 sigma are skipped; the sum is a running FP32 vector sum; the
 * self-pair is included (zeta(0) is finite). */
static bsv_V3f p3d_m2s_vort(const cvtx_P3D** src, int ns, bsv_V3f mes,
                            const orc_vortfunc* f, float sigma) {
	float cutoff = (ORC_MUT(13) ? 4.f : 5.f) * sigma;
	float rsigma = 1 / sigma;
	bsv_V3f sum = v3(0.f, 0.f, 0.f);
	for (long i = 0; i < ns; ++i) {
		bsv_V3f d = v3_sub(src[i]->coord, mes);
		int in = ORC_MUT(14) ? (v3_norm(d) < cutoff)
		                     : (fabsf(d.x[0]) < cutoff && fabsf(d.x[1]) < cutoff &&
		                        fabsf(d.x[2]) < cutoff);
		if (in) {
			float r = v3_norm(d);
			float coeff = f->zeta3(r * rsigma);
			sum = v3_add(v3_scale(src[i]->vorticity, coeff), sum);
		}
	}
	return v3_divs(sum, 4.f * REF_PI_F * sigma * sigma * sigma);
}

bsv_V3f orc_P3D_M2S_vort(const cvtx_P3D** src, int ns, bsv_V3f mes, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	return p3d_m2s_vort(src, ns, mes, &f, sigma);
}

/* reference src/P3D.cpp:458-474 */
void orc_P3D_M2M_vort(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nm,
                      bsv_V3f* out, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	long i;
#pragma omp parallel for schedule(guided)
	for (i = 0; i < nm; ++i)
		out[i] = p3d_m2s_vort(src, ns, mes[i], &f, sigma);
}

/* reference src/P3D.cpp:669-707: omega = M2M_vort at the particles, then
 * alpha <- (1 - fdt) alpha + fdt omega |alpha| / |omega|, and alpha <- 0
 * where |omega| == 0.  Updates the records in place. */
void orc_P3D_pedrizzetti_relaxation(cvtx_P3D** p, int n, float fdt, int reg, float sigma) {
	bsv_V3f* mes = (bsv_V3f*)malloc(sizeof(bsv_V3f) * (size_t)(n > 0 ? n : 1));
	bsv_V3f* om = (bsv_V3f*)malloc(sizeof(bsv_V3f) * (size_t)(n > 0 ? n : 1));
	long i;
	for (i = 0; i < n; ++i) mes[i] = p[i]->coord;
	orc_P3D_M2M_vort((const cvtx_P3D**)p, n, mes, n, om, reg, sigma);
	float tmp = 1.f - fdt;
#pragma omp parallel for
	for (i = 0; i < n; ++i) {
		bsv_V3f ov = p[i]->vorticity;
		float absomega = v3_norm(om[i]);
		float coeff = v3_norm(ov) / absomega;
		bsv_V3f nv = v3_scale(ov, ORC_MUT(15) ? fdt : tmp);
		nv = v3_add(nv, v3_scale(om[i], coeff * fdt));
		p[i]->vorticity = absomega != 0.f ? nv : v3(0.f, 0.f, 0.f);
	}
	free(mes);
	free(om);
}

/* ------------------------------------------------------------------ */
/* F3D: reference src/F3D.cpp:34-202                                   */
/* ------------------------------------------------------------------ */

/* reference src/F3D.cpp:34-54 */
bsv_V3f orc_F3D_S2S_vel(const cvtx_F3D* self, bsv_V3f mes) {
	const float big = 3.40282346e38f;
	bsv_V3f r1 = v3_sub(mes, self->start);
	bsv_V3f r2 = v3_sub(mes, self->end);
	bsv_V3f r0 = v3_sub(r1, r2);
	bsv_V3f c = v3_cross(r1, r2);
	float t1 = self->strength / (4.f * REF_PI_F * powf(v3_norm(c), 2.f));
	float ta = v3_dot(r1, r0) / v3_norm(r1);
	float tb = v3_dot(r2, r0) / v3_norm(r2);
	float t2 = ORC_MUT(6) ? ta + tb : ta - tb;
	if (fabsf(t1) <= big && fabsf(t2) <= big) return v3_scale(c, t1 * t2);
	return v3(0.f, 0.f, 0.f);
}

/* reference src/F3D.cpp:56-85 */
bsv_V3f orc_F3D_S2S_dvort(const cvtx_F3D* self, const cvtx_P3D* ind) {
	bsv_V3f r1 = v3_sub(ind->coord, self->start);
	bsv_V3f r2 = v3_sub(ind->coord, self->end);
	bsv_V3f r0 = v3_sub(r1, r2);
	float t1 = self->strength / (4.f * REF_PI_F);
	bsv_V3f t211 = v3_divs(r0, -powf(v3_norm(v3_cross(r1, r0)), 2.f));
	float t2121 = v3_dot(r0, r1) / v3_norm(r1);
	float t2122 = -v3_dot(r0, r2) / v3_norm(r2);
	float t221 = (ORC_MUT(19) ? 2.0f : 3.0f) / v3_norm(r0);
	float t2221 = v3_norm(v3_cross(r0, r1)) / v3_norm(r1);
	float t2222 = -v3_norm(v3_cross(r0, r1)) / v3_norm(r2);
	float t222 = t2221 + t2222;
	float t212 = t2121 + t2122;
	bsv_V3f A = v3_scale(t211, t1 * t212);
	float B = t221 * t1 * t222;
	bsv_V3f ret = ORC_MUT(7) ? v3_add(v3_scale(ind->vorticity, B), v3_cross(ind->vorticity, A))
	                         : v3_add(v3_scale(ind->vorticity, B), v3_cross(A, ind->vorticity));
	if (!v3_eq(ret, ret) || t222 != t222 || t212 != t212) ret = v3(0.f, 0.f, 0.f);
	return ret;
}

void orc_F3D_M2M_vel(const cvtx_F3D** src, int nf, const bsv_V3f* mes, int nm, bsv_V3f* out) {
	long i;
#pragma omp parallel for schedule(static)
	for (i = 0; i < nm; ++i) {
		double sx = 0, sy = 0, sz = 0;
		for (long j = 0; j < nf; ++j) {
			bsv_V3f v = orc_F3D_S2S_vel(src[j], mes[i]);
			sx += v.x[0]; sy += v.x[1]; sz += v.x[2];
		}
		out[i] = v3((float)sx, (float)sy, (float)sz);
	}
}

void orc_F3D_M2M_dvort(const cvtx_F3D** src, int nf, const cvtx_P3D** ind, int ni, bsv_V3f* out) {
	long i;
#pragma omp parallel for schedule(static)
	for (i = 0; i < ni; ++i) {
		double sx = 0, sy = 0, sz = 0;
		for (long j = 0; j < nf; ++j) {
			bsv_V3f v = orc_F3D_S2S_dvort(src[j], ind[i]);
			sx += v.x[0]; sy += v.x[1]; sz += v.x[2];
		}
		out[i] = v3((float)sx, (float)sy, (float)sz);
	}
}

/* reference src/F3D.cpp:204-227: row-major, row i = measurement point i,
 * result[i * nf + j] = S2S_vel(filament j, mes_i) . dir_i */
void orc_F3D_inf_mtrx(const cvtx_F3D** src, int nf, const bsv_V3f* mes, const bsv_V3f* dir,
                      int nm, float* out) {
	long i;
#pragma omp parallel for schedule(static)
	for (i = 0; i < nm; ++i) {
		for (long j = 0; j < nf; ++j) {
			bsv_V3f v = orc_F3D_S2S_vel(src[j], mes[i]);
			long at = ORC_MUT(16) ? j * (long)nm + i : i * (long)nf + j;
			out[at] = v3_dot(v, dir[i]);
		}
	}
}

/* ------------------------------------------------------------------ */
/* Redistribution on a grid (remeshing): reference src/RedistFunc.cpp:  */
/* 36-96, src/P3D.cpp:509-665, src/P2D.cpp:283-436,                     */
/* src/redistribution_helper_funcs.cpp:33-80, src/array_methods.cpp.    */
/* The reference spreads particles over per-thread octrees and merges   */
/* them; restated here with one thread: contributions are summed per    */
/* grid node in particle order, and nodes are emitted in the tree's     */
/* depth-first order, which is ascending Morton order of the node key   */
/* with the octant index x + 2y + 4z (quadrant x + 2y in 2D).           */
/* ------------------------------------------------------------------ */

static float rd_lambda0(float U) { return U < 0.5f ? 1.f : 0.f; }
static float rd_lambda1(float U) { return U <= 1.f ? 1.f - U : 0.f; }
static float rd_lambda2(float U) {
	return U < 0.5f ? 1.f - U * U : (U < 1.5f ? 0.5f * (1.f - U) * (2.f - U) : 0.f);
}
static float rd_lambda3(float U) {
	return U < 1.f ? 0.5f * (1.f - U * U) * (2.f - U)
	               : (U < 2.f ? (1.f / 6.f) * (1.f - U) * (2.f - U) * (3.f - U) : 0.f);
}
static float rd_m4p(float U) {
	return U < 1.f ? 1.f - (ORC_MUT(17) ? 2.0f : 2.5f) * U * U + 1.5f * U * U * U
	               : (U < 2.f ? 0.5f * (1.f - U) * (2.f - U) * (2.f - U) : 0.f);
}

static float (*const RD_FUNC[5])(float) = {rd_lambda0, rd_lambda1, rd_lambda2, rd_lambda3, rd_m4p};
static const float RD_RADIUS[5] = {0.5f, 1.0f, 1.5f, 2.0f, 2.0f};

float orc_redist_eval(int rid, float U) { return RD_FUNC[rid](U); }
float orc_redist_radius(int rid) { return RD_RADIUS[rid]; }

typedef struct {
	unsigned long long key; /* Morton code */
	long seq;               /* particle order, for a stable sort */
	float s[3];
} rd_entry;

static int rd_cmp(const void* a, const void* b) {
	const rd_entry* x = (const rd_entry*)a;
	const rd_entry* y = (const rd_entry*)b;
	if (x->key != y->key) return x->key < y->key ? -1 : 1;
	return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

static unsigned long long rd_spread3(unsigned v) {
	unsigned long long r = 0;
	for (int b = 0; b < 21; ++b) r |= (unsigned long long)((v >> b) & 1u) << (3 * b);
	return r;
}
static unsigned long long rd_spread2(unsigned v) {
	unsigned long long r = 0;
	for (int b = 0; b < 32; ++b) r |= (unsigned long long)((v >> b) & 1u) << (2 * b);
	return r;
}
static unsigned rd_compact3(unsigned long long m) {
	unsigned v = 0;
	for (int b = 0; b < 21; ++b) v |= (unsigned)((m >> (3 * b)) & 1ull) << b;
	return v;
}
static unsigned rd_compact2(unsigned long long m) {
	unsigned v = 0;
	for (int b = 0; b < 32; ++b) v |= (unsigned)((m >> (2 * b)) & 1ull) << b;
	return v;
}

/* reference src/redistribution_helper_funcs.cpp:33-80 (n > desired) */
static float rd_strength_threshold(const float* strs, int n, int n_desired) {
	double minv, maxv, range;
	float fminv, fmaxv, tol = 0.01f, guesses[1024];
	int g_counts[1024], i, k = 0, interp;
	const int n_guesses = 1024;
	fminv = fmaxv = n > 0 ? strs[0] : 0.f;
	for (i = 0; i < n; ++i) {
		fminv = fminv < strs[i] ? fminv : strs[i];
		fmaxv = fmaxv > strs[i] ? fmaxv : strs[i];
	}
	minv = fminv; maxv = fmaxv;
	if (n < n_desired) return (float)(maxv * 1.05);
	while (1) {
		range = (maxv - minv) * 1.05;
		for (i = 0; i < n_guesses; ++i) {
			g_counts[i] = 0;
			guesses[i] = (float)(minv + i * range / (float)(n_guesses - 1));
		}
		for (i = 0; i < n; ++i) {
			interp = (int)floor((double)(n_guesses - 1) * (strs[i] - minv) / range);
			if (interp < 0) g_counts[0]++;
			else if (interp >= n_guesses) g_counts[n_guesses - 1]++;
			else g_counts[interp]++;
		}
		interp = g_counts[n_guesses - 1];
		for (i = n_guesses - 2; i >= 0; --i) {
			maxv = guesses[i + 1];
			minv = guesses[i];
			interp += g_counts[i];
			g_counts[i] = interp;
			if (interp > n_desired) { k = i + 1; break; }
		}
		if (minv == maxv || g_counts[k] == g_counts[k - 1] ||
		    fabs((float)(n_desired - g_counts[k]) / ((float)n_desired)) < tol * 0.6)
			break;
	}
	return guesses[k];
}

/* reference src/P3D.cpp:638-665 and src/P2D.cpp:410-436: keep particles
 * stronger than the threshold (and of index < max_keep); the removed
 * strength is shared equally by the kept ones.  dim = 3 or 1 components
 * of strength at offset `so` floats into records of `rec` floats. */
static int rd_remove(float* recs, int rec, int so, int dim, const float* strs, int n,
                     float thr, int max_keep) {
	float def[3] = {0.f, 0.f, 0.f};
	int j = 0;
	for (int i = 0; i < n; ++i) {
		float* r = recs + (size_t)i * rec;
		if (strs[i] > thr && i < max_keep) {
			if (j != i) memcpy(recs + (size_t)j * rec, r, sizeof(float) * rec);
			++j;
		} else if (!ORC_MUT(18)) {
			for (int c = 0; c < dim; ++c) def[c] = r[so + c] + def[c];
		}
	}
	for (int c = 0; c < dim; ++c) def[c] = def[c] / (float)j;
	for (int i = 0; i < j; ++i)
		for (int c = 0; c < dim; ++c) recs[(size_t)i * rec + so + c] = recs[(size_t)i * rec + so + c] + def[c];
	return j;
}

static float rd_mean(const float* strs, int n) {
	float ave = 0.f;
	for (int i = 0; i < n; ++i) ave += strs[i];
	return ave / n;
}

int orc_P3D_redistribute_on_grid(const cvtx_P3D** in, int n, cvtx_P3D* out, int max_out,
                                 int rid, float h, float negligible) {
	float (*f)(float) = RD_FUNC[rid];
	const float rh = 1.f / h;
	float mn[3] = {0.f, 0.f, 0.f};
	double sm[3] = {0.0, 0.0, 0.0};
	if (n > 0) for (int c = 0; c < 3; ++c) mn[c] = in[0]->coord.x[c];
	for (int i = 0; i < n; ++i)
		for (int c = 0; c < 3; ++c) {
			const float x = in[i]->coord.x[c];
			mn[c] = mn[c] > x ? x : mn[c];
			sm[c] += (double)x;
		}
	float mean[3];
	for (int c = 0; c < 3; ++c) mean[c] = (float)(sm[c] / (double)n);
	const int R = (int)roundf(RD_RADIUS[rid]);
	for (int c = 0; c < 3; ++c) {
		mn[c] = mn[c] - 1.f * (R * h);
		float dc = (mean[c] - mn[c]) / h;
		dc = roundf(dc) + 5;
		mn[c] = mean[c] - dc * h;
	}
	const int w = 2 * R + 1;
	rd_entry* e = (rd_entry*)malloc(sizeof(rd_entry) * (size_t)(n > 0 ? n : 1) * w * w * w);
	size_t ne = 0;
	for (long i = 0; i < n; ++i) {
		const bsv_V3f p = in[i]->coord;
		const bsv_V3f a = in[i]->vorticity;
		unsigned kx = (unsigned)roundf((p.x[0] - mn[0]) * rh);
		unsigned ky = (unsigned)roundf((p.x[1] - mn[1]) * rh);
		unsigned kz = (unsigned)roundf((p.x[2] - mn[2]) * rh);
		for (int ii = -R; ii <= R; ++ii)
			for (int jj = -R; jj <= R; ++jj)
				for (int kk = -R; kk <= R; ++kk) {
					const unsigned qx = kx + ii, qy = ky + jj, qz = kz + kk;
					const float nx = mn[0] + h * (float)qx, ny = mn[1] + h * (float)qy,
					            nz = mn[2] + h * (float)qz;
					const float U = fabsf((p.x[0] - nx) * rh);
					const float W = fabsf((p.x[1] - ny) * rh);
					const float V = fabsf((p.x[2] - nz) * rh);
					const float frac = f(U) * f(W) * f(V);
					const float s0 = a.x[0] * frac, s1 = a.x[1] * frac, s2 = a.x[2] * frac;
					if (s0 == 0.f && s1 == 0.f && s2 == 0.f) continue;
					rd_entry* t = &e[ne++];
					t->key = rd_spread3(qx) | (rd_spread3(qy) << 1) | (rd_spread3(qz) << 2);
					t->seq = (long)ne;
					t->s[0] = s0; t->s[1] = s1; t->s[2] = s2;
				}
	}
	qsort(e, ne, sizeof(rd_entry), rd_cmp);
	/* nodes: {x, y, z, ax, ay, az, vol} as float records */
	float* nodes = (float*)malloc(sizeof(float) * 7 * (ne > 0 ? ne : 1));
	int nn = 0;
	for (size_t i = 0; i < ne; ++i) {
		if (i == 0 || e[i].key != e[i - 1].key) {
			const unsigned qx = rd_compact3(e[i].key), qy = rd_compact3(e[i].key >> 1),
			               qz = rd_compact3(e[i].key >> 2);
			float* r = nodes + 7 * (size_t)nn++;
			r[0] = mn[0] + h * (float)qx;
			r[1] = mn[1] + h * (float)qy;
			r[2] = mn[2] + h * (float)qz;
			r[3] = r[4] = r[5] = 0.f;
			r[6] = h * h * h;
		}
		float* r = nodes + 7 * (size_t)(nn - 1);
		for (int c = 0; c < 3; ++c) r[3 + c] = r[3 + c] + e[i].s[c];
	}
	free(e);
	float* strs = (float*)malloc(sizeof(float) * (nn > 0 ? nn : 1));
	for (int i = 0; i < nn; ++i) {
		const float* r = nodes + 7 * (size_t)i;
		strs[i] = sqrtf(r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
	}
	float thr = rd_mean(strs, nn) * negligible;
	nn = rd_remove(nodes, 7, 3, 3, strs, nn, thr, nn);
	for (int i = 0; i < nn; ++i) {
		const float* r = nodes + 7 * (size_t)i;
		strs[i] = sqrtf(r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
	}
	if (out != NULL) {
		if (nn > max_out) {
			thr = rd_strength_threshold(strs, nn, max_out);
			nn = rd_remove(nodes, 7, 3, 3, strs, nn, thr, max_out);
		}
		for (int i = 0; i < nn; ++i) {
			const float* r = nodes + 7 * (size_t)i;
			for (int c = 0; c < 3; ++c) {
				out[i].coord.x[c] = r[c];
				out[i].vorticity.x[c] = r[3 + c];
			}
			out[i].volume = r[6];
		}
	}
	free(strs);
	free(nodes);
	return nn;
}

int orc_P2D_redistribute_on_grid(const cvtx_P2D** in, int n, cvtx_P2D* out, int max_out,
                                 int rid, float h, float negligible) {
	float (*f)(float) = RD_FUNC[rid];
	const float rh = 1.f / h;
	float mn[2] = {0.f, 0.f};
	double sm[2] = {0.0, 0.0};
	if (n > 0) for (int c = 0; c < 2; ++c) mn[c] = in[0]->coord.x[c];
	for (int i = 0; i < n; ++i)
		for (int c = 0; c < 2; ++c) {
			const float x = in[i]->coord.x[c];
			mn[c] = mn[c] > x ? x : mn[c];
			sm[c] += (double)x;
		}
	float mean[2];
	for (int c = 0; c < 2; ++c) mean[c] = (float)(sm[c] / (double)n);
	const int R = (int)roundf(RD_RADIUS[rid]);
	for (int c = 0; c < 2; ++c) {
		mn[c] = mn[c] - 1.f * (R * h);
		float dc = (mean[c] - mn[c]) / h;
		dc = roundf(dc) + 5;
		mn[c] = mean[c] - dc * h;
	}
	const int w = 2 * R + 1;
	rd_entry* e = (rd_entry*)malloc(sizeof(rd_entry) * (size_t)(n > 0 ? n : 1) * w * w);
	size_t ne = 0;
	for (long i = 0; i < n; ++i) {
		const bsv_V2f p = in[i]->coord;
		const float a = in[i]->vorticity;
		/* reference UIntKey64.cpp:79-96: the offset is formed in double */
		unsigned kx = (unsigned)roundf((float)(((double)p.x[0] - mn[0]) * rh));
		unsigned ky = (unsigned)roundf((float)(((double)p.x[1] - mn[1]) * rh));
		for (int ii = -R; ii <= R; ++ii)
			for (int jj = -R; jj <= R; ++jj) {
				const unsigned qx = kx + ii, qy = ky + jj;
				const float nx = mn[0] + h * (float)qx, ny = mn[1] + h * (float)qy;
				const float U = fabsf((p.x[0] - nx) * rh);
				const float W = fabsf((p.x[1] - ny) * rh);
				const float s0 = a * (f(U) * f(W));
				if (s0 == 0.f) continue;
				rd_entry* t = &e[ne++];
				t->key = rd_spread2(qx) | (rd_spread2(qy) << 1);
				t->seq = (long)ne;
				t->s[0] = s0;
			}
	}
	qsort(e, ne, sizeof(rd_entry), rd_cmp);
	/* nodes: {x, y, a, area} */
	float* nodes = (float*)malloc(sizeof(float) * 4 * (ne > 0 ? ne : 1));
	int nn = 0;
	for (size_t i = 0; i < ne; ++i) {
		if (i == 0 || e[i].key != e[i - 1].key) {
			float* r = nodes + 4 * (size_t)nn++;
			r[0] = mn[0] + h * (float)rd_compact2(e[i].key);
			r[1] = mn[1] + h * (float)rd_compact2(e[i].key >> 1);
			r[2] = 0.f;
			r[3] = h * h;
		}
		float* r = nodes + 4 * (size_t)(nn - 1);
		r[2] = r[2] + e[i].s[0];
	}
	free(e);
	float* strs = (float*)malloc(sizeof(float) * (nn > 0 ? nn : 1));
	for (int i = 0; i < nn; ++i) strs[i] = fabsf(nodes[4 * (size_t)i + 2]);
	float thr = rd_mean(strs, nn) * negligible;
	nn = rd_remove(nodes, 4, 2, 1, strs, nn, thr, nn);
	for (int i = 0; i < nn; ++i) strs[i] = fabsf(nodes[4 * (size_t)i + 2]);
	if (out != NULL) {
		if (nn > max_out) {
			thr = rd_strength_threshold(strs, nn, max_out);
			nn = rd_remove(nodes, 4, 2, 1, strs, nn, thr, max_out);
		}
		for (int i = 0; i < nn; ++i) {
			const float* r = nodes + 4 * (size_t)i;
			out[i].coord.x[0] = r[0];
			out[i].coord.x[1] = r[1];
			out[i].vorticity = r[2];
			out[i].area = r[3];
		}
	}
	free(strs);
	free(nodes);
	return nn;
}

/* ------------------------------------------------------------------ */
/* P2D: reference src/P2D.cpp:49-276                                   */
/* ------------------------------------------------------------------ */

/* reference src/P2D.cpp:49-69 (without the 1/(2 pi) factor) */
static inline bsv_V2f p2d_vel_pair(const cvtx_P2D* self, bsv_V2f mes,
                                   const orc_vortfunc* f, float recip_sigma) {
	if (v2_eq(self->coord, mes)) return v2(0.f, 0.f);
	bsv_V2f d = v2_sub(mes, self->coord);
	float r = v2_norm(d);
	float rho = r * recip_sigma;
	float g = f->g2(rho);
	bsv_V2f out;
	if (ORC_MUT(8)) {
		out.x[0] = d.x[0] * self->vorticity * g / (r * r);
		out.x[1] = -d.x[1] * self->vorticity * g / (r * r);
	} else {
		out.x[0] = d.x[1] * self->vorticity * g / (r * r);
		out.x[1] = -d.x[0] * self->vorticity * g / (r * r);
	}
	return out;
}

/* reference src/P2D.cpp:71-81 */
bsv_V2f orc_P2D_S2S_vel(const cvtx_P2D* self, bsv_V2f mes, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	bsv_V2f v = p2d_vel_pair(self, mes, &f, 1.f / fabsf(sigma));
	float k = 1.f / (2.f * acosf(-1.f));
	return v2(v.x[0] * k, v.x[1] * k);
}

/* reference src/P2D.cpp:167-194 */
static inline float p2d_visc_pair(const cvtx_P2D* self, const cvtx_P2D* ind,
                                  const orc_vortfunc* f, float sigma, float nu) {
	if (v2_eq(self->coord, ind->coord)) return 0.f;
	bsv_V2f d = v2_sub(self->coord, ind->coord);
	float r = v2_norm(d);
	float rho = fabsf(r / sigma);
	float pre = (ORC_MUT(22) ? 1.f : 2.f) * nu / powf(sigma, 2.f);
	float a = self->vorticity * (ORC_MUT(23) ? self->area : ind->area);
	float b = -ind->vorticity * (ORC_MUT(23) ? ind->area : self->area);
	float t = a + b;
	float eta = f->eta2(rho);
	return t * eta * pre;
}

float orc_P2D_S2S_visc_dvort(const cvtx_P2D* self, const cvtx_P2D* ind, int reg,
                             float sigma, float nu) {
	orc_vortfunc f = orc_vf(reg);
	return p2d_visc_pair(self, ind, &f, sigma, nu);
}

/* reference src/P2D.cpp:99-162 */
void orc_P2D_M2M_vel(const cvtx_P2D** src, int ns, const bsv_V2f* mes, int nm,
                     bsv_V2f* out, int reg, float sigma) {
	orc_vortfunc f = orc_vf(reg);
	float rs = 1.f / fabsf(sigma);
	float k = 1.f / (2.f * acosf(-1.f));
	long i;
#pragma omp parallel for schedule(static)
	for (i = 0; i < nm; ++i) {
		double sx = 0, sy = 0;
		for (long j = 0; j < ns; ++j) {
			bsv_V2f v = p2d_vel_pair(src[j], mes[i], &f, rs);
			sx += v.x[0]; sy += v.x[1];
		}
		out[i] = v2((float)sx * k, (float)sy * k);
	}
}

/* reference src/P2D.cpp:214-250: serial loop over targets, parallel
 * reduction over sources. */
void orc_P2D_M2M_visc_dvort(const cvtx_P2D** src, int ns, const cvtx_P2D** ind, int ni,
                            float* out, int reg, float sigma, float nu) {
	orc_vortfunc f = orc_vf(reg);
	for (long i = 0; i < ni; ++i) {
		double acc = 0.;
		long j;
#pragma omp parallel for reduction(+:acc)
		for (j = 0; j < ns; ++j)
			acc += (double)p2d_visc_pair(src[j], ind[i], &f, sigma, nu);
		out[i] = (float)acc;
	}
}
