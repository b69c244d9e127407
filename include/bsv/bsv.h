/*
 * bsv.h -- small-vector value types used by the cvortex C ABI.
 *
 * The reference library takes these types from the external "bsv" package
 * (hjabird/bsv, not vendored; reference CMakeLists.txt:69-71).  Only the
 * memory layout crosses the ABI, and the reference code uses it exclusively
 * through the `.x[i]` array member (e.g. reference src/P3D.cpp:245-249,
 * test/testparticle.h:36-41), so the layout is fixed as a plain float array:
 *
 *   bsv_V3f : 12 bytes, 4-byte aligned, { float x[3]; }
 *   bsv_V2f :  8 bytes, 4-byte aligned, { float x[2]; }
 *
 * A handful of inline helpers are provided for callers (tests, examples);
 * the library itself never needs them.
 */
#ifndef CVTX_AMD_BSV_H
#define CVTX_AMD_BSV_H

#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x[3]; } bsv_V3f;
typedef struct { float x[2]; } bsv_V2f;

static inline bsv_V3f bsv_V3f_zero(void) {
	bsv_V3f r = {{0.f, 0.f, 0.f}};
	return r;
}

static inline bsv_V3f bsv_V3f_minus(bsv_V3f a, bsv_V3f b) {
	bsv_V3f r = {{a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]}};
	return r;
}

static inline bsv_V3f bsv_V3f_plus(bsv_V3f a, bsv_V3f b) {
	bsv_V3f r = {{a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]}};
	return r;
}

static inline float bsv_V3f_dot(bsv_V3f a, bsv_V3f b) {
	return a.x[0] * b.x[0] + a.x[1] * b.x[1] + a.x[2] * b.x[2];
}

static inline float bsv_V3f_abs(bsv_V3f a) {
	return sqrtf(bsv_V3f_dot(a, a));
}

static inline int bsv_V3f_isequal(bsv_V3f a, bsv_V3f b) {
	return a.x[0] == b.x[0] && a.x[1] == b.x[1] && a.x[2] == b.x[2];
}

static inline bsv_V2f bsv_V2f_minus(bsv_V2f a, bsv_V2f b) {
	bsv_V2f r = {{a.x[0] - b.x[0], a.x[1] - b.x[1]}};
	return r;
}

static inline float bsv_V2f_abs(bsv_V2f a) {
	return sqrtf(a.x[0] * a.x[0] + a.x[1] * a.x[1]);
}

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CVTX_AMD_BSV_H */
