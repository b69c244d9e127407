// device_common.hpp -- shared device-side definitions for the gfx950 kernels.
//
// Pair kernels are written in the reduced form
//     vel    = -(1/4pi)       * sum_s  Kv(r2) (d x a_s)
//     dvort  = sign(sigma)/4pi * [ a_t x sum_s Kv a_s  +  sum_s Kh (a_t . (d x a_s)) d ]
//     visc   = (2 nu / sigma^2) * [ V_t sum_s eta a_s  -  a_t sum_s eta V_s ]
// with d = x_t - x_s, Kv = g(rho)/r^3 and Kh = (3 g(rho)/r^3 - zeta(rho)/sigma^3)/r^2.
// This is algebraically the reference's pair arithmetic (reference
// src/P3D.cpp:51-144) after using d.(a_t x a_s) = -a_t.(d x a_s) and
// sum_s (a_t x a_s) Kv = a_t x sum_s Kv a_s; the rearrangement lets vel,
// dvort and visc share one evaluation of d, r2 and the regularisation per
// pair and moves two cross products out of the pair loop.
//
// Each regularisation computes (Kv, Kh, eta) up to per-call constant factors
// that the epilogue applies in FP64 (Scales below).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvtx {

enum Reg : int { REG_SINGULAR = 0, REG_WINCKELMANS = 1, REG_PLANETARY = 2, REG_GAUSSIAN = 3 };
enum OutMask : int { OUT_VEL = 1, OUT_DVORT = 2, OUT_VISC = 4 };
// ACC_PLAIN: one FP32 running sum per accumulator.  ACC_COMP: FP32 tile
// sums folded by TwoSum, and the P3D stretching (and PSE) sums and the P2D
// viscous exchange per pair in FP64.  ACC_TILES: FP32 tile sums folded by
// TwoSum for every accumulator, no per-pair FP64 (F3D and the vorticity
// field have no per-pair FP64, so there it is ACC_COMP).
enum AccMode : int { ACC_PLAIN = 0, ACC_COMP = 1, ACC_TILES = 2 };
inline bool valid_acc(int a) { return a == ACC_PLAIN || a == ACC_COMP || a == ACC_TILES; }

// Gaussian near-field table: cubic pieces of the reduced kernels
//   G(rho) = g(rho)/rho^3,  H(rho) = (3 G(rho) - zeta(rho)) / rho^2
// on rho in [0, GTAB_RHO_MAX) with GTAB_PER_UNIT pieces per unit of rho.
// Piece i covers rho = (i + t)/GTAB_PER_UNIT, t in [0,1):
//   G = ((g3 t + g2) t + g1) t + g0, same for H; stored as two float4.
constexpr int GTAB_PER_UNIT = 16;
constexpr int GTAB_RHO_MAX = 6;
constexpr int GTAB_N = GTAB_PER_UNIT * GTAB_RHO_MAX;  // 96 pieces

// Per-call constants for the regularisations (all FP32, set by the host).
struct RegConsts {
  float s2;        // sigma^2
  float inv_s2;    // 1/sigma^2
  float c1;        // gaussian: -log2(e)/2 | 2D winckelmans eta: 4 log2(e)
  float c2;        // gaussian: sqrt(2/pi)
  float q_far = 208.f;  // gaussian: rho^2 beyond which exp(-rho^2/2) is 0 in FP32
  float q_tab = (float)(GTAB_RHO_MAX * GTAB_RHO_MAX);  // gaussian: table range in rho^2
};

// A uniform value that a pair loop reads on every pair, held in a VGPR.  On
// gfx950 a VALU instruction with an SGPR operand issues at about half the
// rate of one whose operands are VGPRs, inline constants or a literal
// (measured: 1.8 issue slots against 1), and the compiler keeps uniform
// values in SGPRs.  The empty asm makes the value opaque, so it stays in a
// VGPR for the loop.
__device__ __forceinline__ float in_vgpr(float x) {
  asm volatile("" : "+v"(x));
  return x;
}

__device__ __forceinline__ int in_vgpr(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// The constants of `k` as VGPR values (see in_vgpr), for use inside pair loops.
__device__ __forceinline__ RegConsts loop_consts(const RegConsts& k) {
  RegConsts v;
  v.s2 = in_vgpr(k.s2);
  v.inv_s2 = in_vgpr(k.inv_s2);
  v.c1 = in_vgpr(k.c1);
  v.c2 = in_vgpr(k.c2);
  v.q_far = in_vgpr(k.q_far);
  v.q_tab = in_vgpr(k.q_tab);
  return v;
}

// ---- separately rounded FP32 operations --------------------------------
// For code that must reproduce another program's FP32 operation sequence:
// each product and sum is rounded on its own.  (HIP's __fmul_rn family is
// plain arithmetic, which floating-point contraction may fuse into a
// multiply-add with a single rounding.)
__device__ __forceinline__ float rn_mul(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float rn_add(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float rn_sub(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}

// ---- compensated (TwoSum) accumulation ----------------------------------
// Knuth's TwoSum: hi + lo carries the running sum exactly up to the final
// rounding of lo.  Six FP32 operations per update; applied once per source
// tile per accumulator, so its cost is amortised over the tile.
__device__ __forceinline__ void two_sum_add(float& hi, float& lo, float x) {
  float s = hi + x;
  float bp = s - hi;
  float err = (hi - (s - bp)) + (x - bp);
  hi = s;
  lo += err;
}

// ---- regularisation kernels (reduced form) -------------------------------
// eval<SELF>(r2, K, lds_table, Kv, Kh, eta)
//   Kv: velocity / stretching kernel, zero when r2 == 0 (self and
//       coincident pairs are excluded, reference src/P3D.cpp:58,94,127)
//   Kh: gradient kernel; Kh * 0 must be 0 when r2 == 0
//   eta: PSE kernel, zero when r2 == 0
// SELF = false (velocity only) drops the r2 == 0 test where Kv is finite
// there: the velocity term Kv (d x a_s) is then exactly 0 for d = 0, and the
// test is a compare and a select per pair, which are not dual-issued.
template <int REG>
struct RegKernel;

template <>
struct RegKernel<REG_SINGULAR> {
  // Kv = r^-3, Kh = r^-5 (x3 in the epilogue); infinite at r2 = 0, so the
  // test stays for every SELF
  template <bool SELF = true>
  static __device__ __forceinline__ void eval(float r2, const RegConsts&, const float4*,
                                              float& Kv, float& Kh, float& eta) {
    float y = __builtin_amdgcn_rsqf(r2);
    float y2 = y * y;
    float kv = y * y2;
    float kh = kv * y2;
    bool z = (r2 == 0.f);
    Kv = z ? 0.f : kv;
    Kh = z ? 0.f : kh;
    eta = 0.f;
  }
};

template <>
struct RegKernel<REG_WINCKELMANS> {
  // Reduced units (q = rho^2, s = q + 1, y = s^-1/2); the epilogue applies
  // |sigma|^-3 to Kv, 3 |sigma|^-5 to Kh and 52.5 to eta:
  //   Kv  = G(rho)     = (q + 2.5) y^5
  //   Kh  = H(rho) / 3 = (q + 3.5) y^7
  //   eta = y^9
  template <bool SELF = true>
  static __device__ __forceinline__ void eval(float r2, const RegConsts& k, const float4*,
                                              float& Kv, float& Kh, float& eta) {
    float s = __builtin_fmaf(r2, k.inv_s2, 1.f);
    float y = __builtin_amdgcn_rsqf(s);
    float y2 = y * y;
    float y3 = y2 * y;
    float y5 = y3 * y2;
    float kv = (s + 1.5f) * y5;
    Kh = (kv + y5) * y2;  // (s + 2.5) y^7 from Kv: two operations instead of three
    float e = (y5 * y2) * y2;
    if constexpr (SELF) {
      bool z = (r2 == 0.f);
      Kv = z ? 0.f : kv;
      eta = z ? 0.f : e;
    } else {
      Kv = kv;
      eta = e;
    }
  }
};

template <>
struct RegKernel<REG_PLANETARY> {
  // Reduced units, q = rho^2.  Inside the core (q < 1): g = rho^3, zeta = 3,
  // so G = 1 and H = 0; outside: g = 1, zeta = 0, so G = q^-3/2 and
  // H / 3 = q^-5/2.  The epilogue applies |sigma|^-3 and 3 |sigma|^-5.
  template <bool SELF = true>
  static __device__ __forceinline__ void eval(float r2, const RegConsts& k, const float4*,
                                              float& Kv, float& Kh, float& eta) {
    float q = r2 * k.inv_s2;
    float y = __builtin_amdgcn_rsqf(q);
    float y2 = y * y;
    float kv = y * y2;
    float kh = kv * y2;
    bool in = q < 1.f;
    bool z = SELF && (r2 == 0.f);
    Kv = z ? 0.f : (in ? 1.f : kv);
    Kh = in ? 0.f : kh;
    eta = 0.f;
  }
};

template <>
struct RegKernel<REG_GAUSSIAN> {
  // Works in reduced units: Kv = G(rho), Kh = H(rho)/3 (the epilogue scales
  // by |sigma|^-3 and 3 |sigma|^-5), eta = zeta = sqrt(2/pi) exp(-rho^2/2).
  // Three ranges of rho^2, nested so that a wave runs only the ranges its
  // lanes need (for spatially ordered targets, order.hip, a wave's lanes
  // are close together and most sources fall in the first range for all):
  //   rho^2 >= 208: exp(-rho^2/2) < 2^-149 is zero in FP32, so
  //                 G = rho^-3, H/3 = rho^-5, zeta = 0;
  //   rho^2 >= 36:  erf(rho/sqrt2) = 1 to below 2e-9, so
  //                 G = rho^-3 (1 - rho zeta), H/3 = (G - zeta/3) rho^-2;
  //   rho^2 < 36:   cubic pieces of G and H/3 from the LDS table, built on the host from
  //                 exact FP64 values (series below rho = 0.5), so the
  //                 small-rho behaviour has no cancellation.
  // r2 == 0 (self or coincident pair) falls in the last range, where G and
  // zeta are zeroed; H stays finite and only ever multiplies d = 0.
  //
  // eval_ranges also reports the split between the ranges: near(Kv, eta) is
  // called inside the rho^2 < 208 branch (only there is eta nonzero), and
  // Kfar is Kv for the pairs beyond it and 0 for the others.  A caller can so
  // accumulate the near pairs in a wider type at the cost of the branch only.
  // The evaluation in two steps, for callers that interleave pairs:
  // far() is the rho^2 >= 208 form every pair starts from, near() replaces
  // g and h and sets zeta for a pair with rho^2 < 208 (the caller tests
  // q < k.q_far).  eval_ranges is the two combined.
  struct Far {
    float q, pinv, pinv2, g, h;
  };
  static __device__ __forceinline__ Far far(float r2, const RegConsts& k) {
    Far f;
    f.q = __builtin_fmaf(r2, k.inv_s2, 1e-30f);  // rho^2; the 1e-30 keeps rho finite at r2 = 0
    f.pinv = __builtin_amdgcn_rsqf(f.q);         // 1/rho
    f.pinv2 = f.pinv * f.pinv;
    f.g = f.pinv2 * f.pinv;
    f.h = f.g * f.pinv2;
    return f;
  }
  static __device__ __forceinline__ void near(const Far& f, float r2, const RegConsts& k,
                                              const float4* tab, float& g, float& h,
                                              float& z) {
    z = k.c2 * __builtin_amdgcn_exp2f(f.q * k.c1);
    g = f.pinv2 * (f.pinv - z);
    h = __builtin_fmaf(z, -1.f / 3.f, g) * f.pinv2;
    if (f.q < k.q_tab) {
      float t = f.q * f.pinv * (float)GTAB_PER_UNIT;  // rho * 16
      int i = (int)t;
      i = i < (GTAB_N - 1) ? i : (GTAB_N - 1);
      i = i > 0 ? i : 0;
      float u = t - (float)i;
      float4 cg = tab[2 * i];
      float4 ch = tab[2 * i + 1];
      g = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(cg.w, u, cg.z), u, cg.y), u, cg.x);
      h = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ch.w, u, ch.z), u, ch.y), u, ch.x);
      bool zr = (r2 == 0.f);
      g = zr ? 0.f : g;
      z = zr ? 0.f : z;
    }
  }

  // PIN: keep the near range a branch.  With no work of its own in
  // near_cb the compiler if-converts it, and every pair then pays the exp (a
  // transcendental, about three issue slots) and the selects.  Where near_cb
  // does FP64 sums it stays a branch by itself, and the pin would only stop
  // the loop around it from unrolling.  Measured on the 2^20 ring: the
  // compensated vel-only kernel 593 -> 515 ms with the pin; the plain FP32
  // vel+dvort kernel (two targets per lane) 680 -> 700 ms, so eval() does
  // not pin.
  template <bool PIN = false, class Near>
  static __device__ __forceinline__ void eval_ranges(float r2, const RegConsts& k,
                                                     const float4* tab, float& Kv, float& Kh,
                                                     float& eta, float& Kfar, Near&& near_cb) {
    const Far f = far(r2, k);
    float g = f.g, h = f.h, z = 0.f, kf = f.g;
    if (f.q < k.q_far) {
      if constexpr (PIN) asm volatile("");
      near(f, r2, k, tab, g, h, z);
      near_cb(g, z);
      kf = 0.f;
    }
    Kv = g;
    Kh = h;
    eta = z;
    Kfar = kf;
  }

  // The same with the far range's own work in a callback on an else
  // branch, far_cb(G): the caller needs no copy of G that the near branch
  // would zero (a register copy per pair otherwise).  Ring of 2^20, fused
  // vel + dvort: 694.5 -> 679.1 ms (three A/B pairs on one box).
  template <class Near, class FarCb>
  static __device__ __forceinline__ void eval_split(float r2, const RegConsts& k,
                                                    const float4* tab, float& Kv, float& Kh,
                                                    float& eta, Near&& near_cb, FarCb&& far_cb) {
    const Far f = far(r2, k);
    float g = f.g, h = f.h, z = 0.f;
    if (f.q < k.q_far) {
      near(f, r2, k, tab, g, h, z);
      near_cb(g, z);
    } else {
      far_cb(f.g);
    }
    Kv = g;
    Kh = h;
    eta = z;
  }

  template <bool SELF = true>
  static __device__ __forceinline__ void eval(float r2, const RegConsts& k, const float4* tab,
                                              float& Kv, float& Kh, float& eta) {
    float kf;
    eval_ranges(r2, k, tab, Kv, Kh, eta, kf, [](float, float) {});
  }
};

}  // namespace cvtx
