/*
 * fp64.c -- double-precision oracle for the cvortex direct-sum hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Never linked into the library.
 *
 * The physics is the reference's (formulas: reference src/P3D.cpp:51-144,
 * src/F3D.cpp:34-85, src/P2D.cpp:49-194, regularisations
 * src/VortFunc.cpp:64-199) evaluated in double precision throughout.  The
 * 3D dvort pair is written with the reduced kernels
 *     G(rho)  = g(rho) / rho^3
 *     Hn(rho) = (3 G(rho) - zeta(rho)) / rho^2
 * so that   dvort = [ w G - d (d.w) Hn / sigma^2 ] / (4 pi sigma^3),
 * w = alpha_t x alpha_s, d = x_t - x_s; this is algebraically identical to
 * reference src/P3D.cpp:97-111 and avoids the small-rho cancellation of
 * 3 g / rho^3 - zeta.  Closed forms of G and Hn are used per kernel.
 *
 * Gaussian regularisation, gauss_mode:
 *   0  the reference's Abramowitz & Stegun 7.1.26 erf (VortFunc.cpp:155-176),
 *      evaluated in double -- "what the reference formula means";
 *   1  exact: erf from libm, with the convergent series
 *        G  = sqrt(2/pi) sum_k (-rho^2/2)^k / (k! (2k+3))
 *        Hn = sqrt(2/pi) sum_k (-rho^2/2)^k / (k! (2k+5))
 *      for rho < 0.5.
 */
#include <math.h>
#include <stddef.h>

#include "oracle.h"

#define PI_D 3.14159265358979323846
#define SQRT_2_OVER_PI_D 0.79788456080286535588

/* ---------------- regularisation kernels in double ---------------- */

static double gauss_g_as(double rho) {
	/* reference VortFunc.cpp:155-176 in double */
	if (rho > 6.0) return 1.0;
	const double c1 = 0.254829592, c2 = -0.284496736, c3 = 1.421413741;
	const double c4 = -1.453152027, c5 = 1.061405429, pp = 0.3275911;
	double y = rho / sqrt(2.0);
	double t = 1.0 / (1.0 + pp * y);
	double poly = t * (c1 + t * (c2 + t * (c3 + t * (c4 + t * c5))));
	double e = exp(-y * y);
	return (1.0 - poly * e) - rho * SQRT_2_OVER_PI_D * e;
}

static double gauss_series(double rho, int shift) {
	/* sqrt(2/pi) sum_k (-rho^2/2)^k / (k! (2k+shift)) */
	double x = -0.5 * rho * rho, term = 1.0, sum = 0.0;
	for (int k = 0; k < 40; ++k) {
		double add = term / (2.0 * k + shift);
		sum += add;
		if (fabs(add) < 1e-18 * fabs(sum)) break;
		term *= x / (k + 1.0);
	}
	return SQRT_2_OVER_PI_D * sum;
}

/* G = g/rho^3, Hn = (3G - zeta)/rho^2, z = zeta, g = g itself */
static void kernels3(int reg, double rho, int gmode, double* G, double* Hn, double* z, double* g) {
	double rr = rho * rho;
	switch (reg) {
	case ORC_SINGULAR:
		*g = 1.0; *z = 0.0;
		*G = 1.0 / (rr * rho); *Hn = 3.0 / (rr * rr * rho);
		return;
	case ORC_WINCKELMANS: {
		double a = rr + 1.0;
		*G = (rr + 2.5) * pow(a, -2.5);
		*z = 7.5 * pow(a, -3.5);
		*Hn = 3.0 * (rr + 3.5) * pow(a, -3.5);
		*g = *G * rr * rho;
		return;
	}
	case ORC_PLANETARY:
		if (rho < 1.0) { *g = rr * rho; *z = 3.0; *G = 1.0; *Hn = 0.0; }
		else { *g = 1.0; *z = 0.0; *G = 1.0 / (rr * rho); *Hn = 3.0 / (rr * rr * rho); }
		return;
	default: {
		*z = SQRT_2_OVER_PI_D * exp(-0.5 * rr);
		if (gmode == 0) {
			*g = gauss_g_as(rho);
			*G = *g / (rr * rho);
			*Hn = (3.0 * *G - *z) / rr;
		} else if (rho < 0.5) {
			*G = gauss_series(rho, 3);
			*Hn = gauss_series(rho, 5);
			*g = *G * rr * rho;
		} else {
			*g = erf(rho / sqrt(2.0)) - rho * *z;
			*G = *g / (rr * rho);
			*Hn = (3.0 * *G - *z) / rr;
		}
		return;
	}
	}
}

static double eta3(int reg, double rho) {
	if (reg == ORC_WINCKELMANS) return 52.5 * pow(rho * rho + 1.0, -4.5);
	if (reg == ORC_GAUSSIAN) return SQRT_2_OVER_PI_D * exp(-0.5 * rho * rho);
	return 0.0;
}

static double g2(int reg, double rho, int gmode) {
	double rr = rho * rho;
	switch (reg) {
	case ORC_SINGULAR: return 1.0;
	case ORC_WINCKELMANS: return (rr * rr + 2.0 * rr) / (rr * rr + 2.0 * rr + 1.0);
	case ORC_PLANETARY: return rho < 1.0 ? rr : 1.0;
	default: return gmode == 0 ? 1.0 - exp(-0.5 * rr) : -expm1(-0.5 * rr);
	}
}

static double eta2(int reg, double rho) {
	double rr = rho * rho;
	if (reg == ORC_WINCKELMANS) {
		/* reference VortFunc.cpp:124-131: 24 exp(4/a^3) / a^4, a = rho^2 + 1 */
		double a = rr + 1.0;
		return 24.0 * exp(4.0 / (a * a * a)) / (a * a * a * a);
	}
	if (reg == ORC_GAUSSIAN) return exp(-0.5 * rr);
	return 0.0;
}

double orc64_vf_eval(int reg, int fn, double rho, int gauss_mode) {
	double G, Hn, z, g;
	switch (fn) {
	case ORC_FN_G3: case ORC_FN_COMB_G:
		kernels3(reg, rho, gauss_mode, &G, &Hn, &z, &g);
		if (reg == ORC_GAUSSIAN && gauss_mode == 1 && rho < 0.5) return G * rho * rho * rho;
		return g;
	case ORC_FN_ZETA3: case ORC_FN_COMB_ZETA:
		kernels3(reg, rho, gauss_mode, &G, &Hn, &z, &g);
		return z;
	case ORC_FN_G2: return g2(reg, rho, gauss_mode);
	case ORC_FN_ETA3: return eta3(reg, rho);
	case ORC_FN_ETA2: return eta2(reg, rho);
	}
	return NAN;
}

/* ------------------------------- P3D ------------------------------- */

static inline void cross3(const double* a, const double* b, double* o) {
	o[0] = a[1] * b[2] - a[2] * b[1];
	o[1] = a[2] * b[0] - a[0] * b[2];
	o[2] = a[0] * b[1] - a[1] * b[0];
}

static void p3d_one(const cvtx_P3D** src, int ns, const float* xt_f, const float* at_f,
                    float vt_f, int reg, double sigma, double nu, int gmode,
                    double* vel, double* dvort, double* visc) {
	double sa = fabs(sigma);
	double xt[3] = {xt_f[0], xt_f[1], xt_f[2]};
	double at[3] = {at_f ? at_f[0] : 0, at_f ? at_f[1] : 0, at_f ? at_f[2] : 0};
	double vt = vt_f;
	double v[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, vs[3] = {0, 0, 0};
	for (int j = 0; j < ns; ++j) {
		const cvtx_P3D* s = src[j];
		if (s->coord.x[0] == xt_f[0] && s->coord.x[1] == xt_f[1] && s->coord.x[2] == xt_f[2])
			continue;
		double as[3] = {s->vorticity.x[0], s->vorticity.x[1], s->vorticity.x[2]};
		double d[3] = {xt[0] - s->coord.x[0], xt[1] - s->coord.x[1], xt[2] - s->coord.x[2]};
		double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		double r = sqrt(r2);
		double rho = r / sa;
		double G, Hn, z, g;
		kernels3(reg, rho, gmode, &G, &Hn, &z, &g);
		if (vel) {
			double c[3];
			cross3(d, as, c);
			double k = -G / (sa * sa * sa);
			v[0] += k * c[0]; v[1] += k * c[1]; v[2] += k * c[2];
		}
		if (dvort) {
			double w[3];
			cross3(at, as, w);
			double dw = d[0] * w[0] + d[1] * w[1] + d[2] * w[2];
			double kh = Hn / (sa * sa);
			for (int c = 0; c < 3; ++c) dv[c] += w[c] * G - d[c] * dw * kh;
		}
		if (visc) {
			double e = eta3(reg, rho);
			double vs_s = s->volume;
			for (int c = 0; c < 3; ++c) vs[c] += (as[c] * vt - at[c] * vs_s) * e;
		}
	}
	if (vel) for (int c = 0; c < 3; ++c) vel[c] = v[c] / (4.0 * PI_D);
	if (dvort) {
		double pre = 1.0 / (4.0 * PI_D * sigma * sigma * sigma);
		for (int c = 0; c < 3; ++c) dvort[c] = dv[c] * pre;
	}
	if (visc) {
		double pre = 2.0 * nu / (sigma * sigma);
		for (int c = 0; c < 3; ++c) visc[c] = vs[c] * pre;
	}
}

void orc64_P3D_M2M(const cvtx_P3D** src, int ns, const cvtx_P3D** ind,
                   const int* tgt_idx, int n_idx, int reg, double sigma, double nu,
                   int gauss_mode, double* vel, double* dvort, double* visc) {
	long k;
#pragma omp parallel for schedule(dynamic, 4)
	for (k = 0; k < n_idx; ++k) {
		const cvtx_P3D* t = ind[tgt_idx[k]];
		p3d_one(src, ns, t->coord.x, t->vorticity.x, t->volume, reg, sigma, nu, gauss_mode,
		        vel ? vel + 3 * k : NULL, dvort ? dvort + 3 * k : NULL, visc ? visc + 3 * k : NULL);
	}
}

void orc64_P3D_M2M_vel_pts(const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                           const int* tgt_idx, int n_idx, int reg, double sigma,
                           int gauss_mode, double* vel) {
	long k;
#pragma omp parallel for schedule(dynamic, 4)
	for (k = 0; k < n_idx; ++k)
		p3d_one(src, ns, mes[tgt_idx[k]].x, NULL, 0.f, reg, sigma, 0.0, gauss_mode,
		        vel + 3 * k, NULL, NULL);
}

/* ------------------------------- F3D ------------------------------- */

static void f3d_one(const cvtx_F3D** src, int nf, const float* xt_f, const float* at_f,
                    double* vel, double* dvort) {
	double v[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
	double at[3] = {at_f ? at_f[0] : 0, at_f ? at_f[1] : 0, at_f ? at_f[2] : 0};
	for (int j = 0; j < nf; ++j) {
		const cvtx_F3D* f = src[j];
		double r1[3], r2[3], r0[3];
		for (int c = 0; c < 3; ++c) {
			r1[c] = (double)xt_f[c] - f->start.x[c];
			r2[c] = (double)xt_f[c] - f->end.x[c];
			r0[c] = r1[c] - r2[c];
		}
		double n1 = sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
		double n2 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
		double r0r1 = r0[0] * r1[0] + r0[1] * r1[1] + r0[2] * r1[2];
		double r0r2 = r0[0] * r2[0] + r0[1] * r2[1] + r0[2] * r2[2];
		if (vel) {
			/* reference F3D.cpp:34-54 */
			double c[3];
			cross3(r1, r2, c);
			double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
			double t1 = f->strength / (4.0 * PI_D * cc);
			double t2 = r0r1 / n1 - r0r2 / n2;
			if (isfinite(t1) && isfinite(t2))
				for (int q = 0; q < 3; ++q) v[q] += c[q] * t1 * t2;
		}
		if (dvort) {
			/* reference F3D.cpp:56-85 */
			double c[3];
			cross3(r0, r1, c);
			double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
			double n0 = sqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
			double t1 = f->strength / (4.0 * PI_D);
			double t212 = r0r1 / n1 - r0r2 / n2;
			double t222 = sqrt(cc) / n1 - sqrt(cc) / n2;
			double A[3];
			for (int q = 0; q < 3; ++q) A[q] = -r0[q] / cc * t1 * t212;
			double B = 3.0 / n0 * t1 * t222;
			double Ax[3];
			cross3(A, at, Ax);
			double out[3];
			int ok = isfinite(t212) && isfinite(t222);
			for (int q = 0; q < 3; ++q) {
				out[q] = at[q] * B + Ax[q];
				ok = ok && isfinite(out[q]);
			}
			if (ok) for (int q = 0; q < 3; ++q) dv[q] += out[q];
		}
	}
	if (vel) for (int c = 0; c < 3; ++c) vel[c] = v[c];
	if (dvort) for (int c = 0; c < 3; ++c) dvort[c] = dv[c];
}

void orc64_F3D_M2M(const cvtx_F3D** src, int nf, const cvtx_P3D** ind,
                   const int* tgt_idx, int n_idx, double* vel, double* dvort) {
	long k;
#pragma omp parallel for schedule(dynamic, 4)
	for (k = 0; k < n_idx; ++k) {
		const cvtx_P3D* t = ind[tgt_idx[k]];
		f3d_one(src, nf, t->coord.x, t->vorticity.x,
		        vel ? vel + 3 * k : NULL, dvort ? dvort + 3 * k : NULL);
	}
}

void orc64_F3D_M2M_vel_pts(const cvtx_F3D** src, int nf, const bsv_V3f* mes,
                           const int* tgt_idx, int n_idx, double* vel) {
	long k;
#pragma omp parallel for schedule(dynamic, 4)
	for (k = 0; k < n_idx; ++k)
		f3d_one(src, nf, mes[tgt_idx[k]].x, NULL, vel + 3 * k, NULL);
}

/* influence matrix, reference src/F3D.cpp:204-227: each element is one
 * filament's velocity (f3d_one over a single filament) dotted with dir */
void orc64_F3D_inf_mtrx(const cvtx_F3D** src, int nf, const bsv_V3f* mes, const bsv_V3f* dir,
                        const int* tgt_idx, int n_idx, double* out) {
	long k;
#pragma omp parallel for schedule(dynamic, 4)
	for (k = 0; k < n_idx; ++k) {
		const int i = tgt_idx[k];
		for (int j = 0; j < nf; ++j) {
			double v[3];
			f3d_one(src + j, 1, mes[i].x, NULL, v, NULL);
			out[k * (long)nf + j] = v[0] * dir[i].x[0] + v[1] * dir[i].x[1] + v[2] * dir[i].x[2];
		}
	}
}

/* ------------------------------- P2D ------------------------------- */

static void p2d_one(const cvtx_P2D** src, int ns, const float* xt_f, float at, float At,
                    int reg, double sigma, double nu, int gmode, double* vel, double* visc) {
	double sa = fabs(sigma);
	double v0 = 0, v1 = 0, vs = 0;
	for (int j = 0; j < ns; ++j) {
		const cvtx_P2D* s = src[j];
		if (s->coord.x[0] == xt_f[0] && s->coord.x[1] == xt_f[1]) continue;
		double d0 = (double)xt_f[0] - s->coord.x[0];
		double d1 = (double)xt_f[1] - s->coord.x[1];
		double r2 = d0 * d0 + d1 * d1;
		double rho = sqrt(r2) / sa;
		if (vel) {
			double k = s->vorticity * g2(reg, rho, gmode) / r2;
			v0 += d1 * k;
			v1 -= d0 * k;
		}
		if (visc) vs += ((double)s->vorticity * At - (double)at * s->area) * eta2(reg, rho);
	}
	if (vel) { vel[0] = v0 / (2.0 * PI_D); vel[1] = v1 / (2.0 * PI_D); }
	if (visc) visc[0] = vs * 2.0 * nu / (sigma * sigma);
}

void orc64_P2D_M2M(const cvtx_P2D** src, int ns, const cvtx_P2D** ind,
                   const int* tgt_idx, int n_idx, int reg, double sigma, double nu,
                   int gauss_mode, double* vel, double* visc) {
	long k;
#pragma omp parallel for schedule(dynamic, 4)
	for (k = 0; k < n_idx; ++k) {
		const cvtx_P2D* t = ind[tgt_idx[k]];
		p2d_one(src, ns, t->coord.x, t->vorticity, t->area, reg, sigma, nu, gauss_mode,
		        vel ? vel + 2 * k : NULL, visc ? visc + k : NULL);
	}
}

void orc64_P2D_M2M_vel_pts(const cvtx_P2D** src, int ns, const bsv_V2f* mes,
                           const int* tgt_idx, int n_idx, int reg, double sigma,
                           int gauss_mode, double* vel) {
	long k;
#pragma omp parallel for schedule(dynamic, 4)
	for (k = 0; k < n_idx; ++k)
		p2d_one(src, ns, mes[tgt_idx[k]].x, 0.f, 0.f, reg, sigma, 0.0, gauss_mode,
		        vel + 2 * k, NULL);
}

/* ---------------- P3D vorticity field and relaxation ----------------
 * reference src/P3D.cpp:298-322 (M2S_vort) and :669-707 (relaxation).  The
 * 5 sigma box test is made in FP32 exactly as the reference makes it (the
 * set of contributing sources is part of the definition); zeta, the sum and
 * the normalisation are in double. */
static double zeta3(int reg, double rho, int gmode) {
	double G, Hn, z, g;
	if (reg == ORC_SINGULAR) return 0.0;
	if (reg == ORC_PLANETARY) return rho < 1.0 ? 3.0 : 0.0;
	kernels3(reg, rho > 0.0 ? rho : 1e-300, gmode, &G, &Hn, &z, &g);
	return z;
}

static void vort64_one(const cvtx_P3D** src, int ns, bsv_V3f mes, int reg, double sigma,
                       int gmode, double* out) {
	const float cutoff = 5.f * (float)sigma;
	double sx = 0.0, sy = 0.0, sz = 0.0;
	for (int i = 0; i < ns; ++i) {
		const float dx = src[i]->coord.x[0] - mes.x[0];
		const float dy = src[i]->coord.x[1] - mes.x[1];
		const float dz = src[i]->coord.x[2] - mes.x[2];
		if (!(fabsf(dx) < cutoff && fabsf(dy) < cutoff && fabsf(dz) < cutoff)) continue;
		const double ex = (double)src[i]->coord.x[0] - mes.x[0];
		const double ey = (double)src[i]->coord.x[1] - mes.x[1];
		const double ez = (double)src[i]->coord.x[2] - mes.x[2];
		const double z = zeta3(reg, sqrt(ex * ex + ey * ey + ez * ez) / sigma, gmode);
		sx += z * src[i]->vorticity.x[0];
		sy += z * src[i]->vorticity.x[1];
		sz += z * src[i]->vorticity.x[2];
	}
	const double k = 1.0 / (4.0 * PI_D * sigma * sigma * sigma);
	out[0] = sx * k;
	out[1] = sy * k;
	out[2] = sz * k;
}

void orc64_P3D_M2M_vort(const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                        const int* tgt_idx, int n_idx, int reg, double sigma,
                        int gauss_mode, double* out) {
	int k;
#pragma omp parallel for schedule(dynamic, 16)
	for (k = 0; k < n_idx; ++k)
		vort64_one(src, ns, mes[tgt_idx[k]], reg, sigma, gauss_mode, out + 3 * (size_t)k);
}

void orc64_P3D_pedrizzetti(const cvtx_P3D** p, int n, double fdt, int reg, double sigma,
                           int gauss_mode, double* alpha_out) {
	int i;
#pragma omp parallel for schedule(dynamic, 16)
	for (i = 0; i < n; ++i) {
		double w[3];
		vort64_one(p, n, p[i]->coord, reg, sigma, gauss_mode, w);
		const double ax = p[i]->vorticity.x[0], ay = p[i]->vorticity.x[1],
		             az = p[i]->vorticity.x[2];
		const double aw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
		const double aa = sqrt(ax * ax + ay * ay + az * az);
		double* o = alpha_out + 3 * (size_t)i;
		if (aw == 0.0) {
			o[0] = o[1] = o[2] = 0.0;
		} else {
			const double c = aa / aw * fdt;
			o[0] = (1.0 - fdt) * ax + c * w[0];
			o[1] = (1.0 - fdt) * ay + c * w[1];
			o[2] = (1.0 - fdt) * az + c * w[2];
		}
	}
}
