// tile_cull.hpp -- bounding boxes for far-tile pair loops.
//
// A pair kernel streams sources through LDS in tiles of one workgroup's
// size.  When the box around the workgroup's targets and the box around a
// tile's sources are far apart, no pair of that (block, tile) combination
// can fall in a regularisation's near range, and the workgroup can run a
// cheaper loop for the whole tile: no per-pair range tests or selects, and
// one reciprocal shared by four pairs (see shared_rcp4).  The test is
// uniform across the workgroup, so the choice costs no divergence.  For
// inputs without spatial coherence the boxes overlap and the general loop
// runs as before; the boxes then cost a few operations per tile.
//
// Boxes are reduced with wave shuffles and one LDS slot per wave.  A box
// records whether every point in it is finite: a tile or block with a
// non-finite coordinate never takes the far loop, since its boxes say
// nothing about its pairs.
#pragma once

#include <hip/hip_runtime.h>

namespace cvtx {

template <int D>
struct Box {
  float lo[D], hi[D];
  bool finite;
};

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// LDS partials of a box reduction: one entry per wave, plus the wave's
// largest value of one extra per-thread quantity (e.g. a segment length).
template <int D, int WAVES>
struct BoxPartials {
  float lo[WAVES][D];
  float hi[WAVES][D];
  float ex[WAVES];
  int bad[WAVES];
};

// Step 1, before a barrier: this thread's box [lo, hi] (when `valid`; a
// point has lo = hi) and its `extra` value into its wave's partial, written
// by the wave's first lane.  `fin`: the thread's coordinates are finite.
template <int D, int WAVES>
__device__ __forceinline__ void box_partial(const float* lo_t, const float* hi_t, float extra,
                                            bool fin, bool valid, BoxPartials<D, WAVES>& s) {
  const int w = threadIdx.x >> 6;
  float lo[D], hi[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    lo[c] = wave_min(valid ? lo_t[c] : __builtin_inff());
    hi[c] = wave_max(valid ? hi_t[c] : -__builtin_inff());
  }
  const float ex = wave_max(valid ? extra : 0.f);
  const bool bad = __any(valid && !fin);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      s.lo[w][c] = lo[c];
      s.hi[w][c] = hi[c];
    }
    s.ex[w] = ex;
    s.bad[w] = bad ? 1 : 0;
  }
}

// A point's box.
template <int D, int WAVES>
__device__ __forceinline__ void box_partial(const float* x, bool valid, BoxPartials<D, WAVES>& s) {
  bool fin = true;
#pragma unroll
  for (int c = 0; c < D; ++c) fin = fin && __builtin_isfinite(x[c]);
  box_partial<D, WAVES>(x, x, 0.f, fin, valid, s);
}

// Step 2, after the barrier: the box of the whole workgroup (every thread);
// *extra (if given) receives the largest extra value.
template <int D, int WAVES>
__device__ __forceinline__ Box<D> box_combine(const BoxPartials<D, WAVES>& s,
                                              float* extra = nullptr) {
  Box<D> b;
  b.finite = true;
  float ex = s.ex[0];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    b.lo[c] = s.lo[0][c];
    b.hi[c] = s.hi[0][c];
  }
#pragma unroll
  for (int w = 1; w < WAVES; ++w) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      b.lo[c] = fminf(b.lo[c], s.lo[w][c]);
      b.hi[c] = fmaxf(b.hi[c], s.hi[w][c]);
    }
    ex = fmaxf(ex, s.ex[w]);
  }
#pragma unroll
  for (int w = 0; w < WAVES; ++w) b.finite = b.finite && s.bad[w] == 0;
  if (extra) *extra = ex;
  return b;
}

// The box of a workgroup's targets: each thread passes the box of its own
// targets (and whether they are finite).  Has its own barriers; the caller
// issues none around it that its LDS use could race with.
template <int D, int WAVES>
__device__ __forceinline__ Box<D> block_box(const float* lo_t, const float* hi_t, bool fin,
                                            BoxPartials<D, WAVES>& s) {
  box_partial<D, WAVES>(lo_t, hi_t, 0.f, fin, true, s);
  __syncthreads();
  const Box<D> b = box_combine<D, WAVES>(s);
  __syncthreads();
  return b;
}

// Squared smallest and largest distances between points of boxes a and b
// (FP32; a box with no points has lo > hi and is reported as not far).
template <int D>
__device__ __forceinline__ void box_dist2(const Box<D>& a, const Box<D>& b, float& dmin2,
                                          float& dmax2) {
  dmin2 = 0.f;
  dmax2 = 0.f;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const float gap = fmaxf(0.f, fmaxf(a.lo[c] - b.hi[c], b.lo[c] - a.hi[c]));
    const float span = fmaxf(a.hi[c] - b.lo[c], b.hi[c] - a.lo[c]);
    dmin2 = __builtin_fmaf(gap, gap, dmin2);
    dmax2 = __builtin_fmaf(span, span, dmax2);
  }
}

// Whether every pair between boxes a and b is at squared distance >= near2,
// and four squared distances multiply without leaving FP32's normal range
// (each in [2^-31, 2^31], so their product in [2^-124, 2^124]).
template <int D>
__device__ __forceinline__ bool boxes_far(const Box<D>& a, const Box<D>& b, float near2) {
  if (!a.finite || !b.finite) return false;
  float dmin2, dmax2;
  box_dist2(a, b, dmin2, dmax2);
  return dmin2 >= near2 && dmin2 >= 4.6566e-10f && dmax2 <= 2.1474e9f;
}

// Reciprocals of four values from one reciprocal of their product
// (Montgomery's trick): 9 multiplications and one v_rcp instead of four
// v_rcp.  Each result is within a few ulp of the correctly rounded
// reciprocal when the product stays in the normal range (boxes_far).
__device__ __forceinline__ void shared_rcp4(const float* x, float* r) {
  const float m01 = x[0] * x[1];
  const float m23 = x[2] * x[3];
  const float R = __builtin_amdgcn_rcpf(m01 * m23);
  const float R01 = m23 * R;
  const float R23 = m01 * R;
  r[0] = x[1] * R01;
  r[1] = x[0] * R01;
  r[2] = x[3] * R23;
  r[3] = x[2] * R23;
}

// a / x[0], a / x[1], b / x[2], b / x[3] from one reciprocal: each scale
// multiplies a reciprocal of two of the values once (11 multiplications and
// one v_rcp, against 13 for shared_rcp4 and four scalings).  The scales are
// applied after 1/(x0 x1) and 1/(x2 x3) are formed, so an intermediate
// leaves the FP32 range only where a / x is within 2^31 of doing so.
__device__ __forceinline__ void shared_rcp4_scaled(const float* x, float a, float b, float* r) {
  const float m01 = x[0] * x[1];
  const float m23 = x[2] * x[3];
  const float R = __builtin_amdgcn_rcpf(m01 * m23);
  const float A = (m23 * R) * a;  // a / (x0 x1)
  const float B = (m01 * R) * b;  // b / (x2 x3)
  r[0] = x[1] * A;
  r[1] = x[0] * A;
  r[2] = x[3] * B;
  r[3] = x[2] * B;
}

// Reciprocals of N values from one reciprocal of their product (prefix
// products, then one backward pass: 3 (N - 1) multiplications and one
// v_rcp).  Every prefix product must stay in FP32's normal range.
template <int N>
__device__ __forceinline__ void batch_rcp(const float* x, float* r) {
  float pre[N];
  pre[0] = x[0];
#pragma unroll
  for (int i = 1; i < N; ++i) pre[i] = pre[i - 1] * x[i];
  float inv = __builtin_amdgcn_rcpf(pre[N - 1]);
#pragma unroll
  for (int i = N - 1; i > 0; --i) {
    r[i] = inv * pre[i - 1];
    inv = inv * x[i];
  }
  r[0] = inv;
}

}  // namespace cvtx
