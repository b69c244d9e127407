// p3d_body.hpp -- the P3D pair loop shared by the regular kernel (p3d.hip)
// and the persistent multi-cloud kernel (batched.hip).  See p3d.hip for the
// scheme.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace cvtx {

// Canonical layout of the per-target sums (FP64 partials and epilogue).
template <int MASK>
struct P3DLayout {
  static constexpr int VEL = 0;
  static constexpr int A = (MASK & OUT_VEL) ? 3 : 0;       // sum Kv a_s
  static constexpr int H = A + 3;                           // sum Kh (a_t.c) d
  static constexpr int E = A + ((MASK & OUT_DVORT) ? 6 : 0); // sum eta a_s
  static constexpr int EV = E + 3;                          // sum eta V_s
  static constexpr int N = E + ((MASK & OUT_VISC) ? 4 : 0);
  static constexpr bool NEED_TA = (MASK & (OUT_DVORT | OUT_VISC)) != 0;
};

// Register layout of the accumulators.  FP32 accumulators hold the sums
// whose terms carry no coherent large component: vel (sum Kv d x a_s) and
// the gradient term (sum Kh (a_t.c) d).  With ACC_COMP the stretching sum
// sum Kv a_s and the PSE sums sum eta a_s, sum eta V_s are accumulated per
// pair in FP64: their near-field terms are large and nearly parallel to a_t,
// and only their small component normal to a_t (or the small difference
// V_t E - a_t e) survives, so FP32 tile sums would lose several digits
// there.  With ACC_PLAIN and ACC_TILES everything is FP32 (ACC_TILES folds
// the tile sums by TwoSum, which suits fields whose near terms are not
// coherent, e.g. i.i.d. strengths).
// SPLIT (Gaussian with ACC_COMP): only pairs with rho^2 < 208 go to FP64,
// inside the regularisation's near branch; the stretching terms of the
// pairs beyond (Kv < 208^-3/2 in reduced units, and eta = 0 there) are
// summed in FP32 tiles combined by TwoSum, in the F_A slots, on the far
// branch (RegKernel<REG_GAUSSIAN>::eval_split).
template <int REG, int MASK, int ACC>
struct P3DRegs {
  static constexpr bool DBL = (ACC == ACC_COMP);
  static constexpr bool SPLIT = DBL && REG == REG_GAUSSIAN;
  static constexpr int F_VEL = 0;
  static constexpr int F_A = (MASK & OUT_VEL) ? 3 : 0;  // plain, or the far part with SPLIT
  static constexpr int F_H = F_A + (((!DBL || SPLIT) && (MASK & OUT_DVORT)) ? 3 : 0);
  static constexpr int F_E = F_H + ((MASK & OUT_DVORT) ? 3 : 0);  // plain only
  static constexpr int NF = F_E + ((!DBL && (MASK & OUT_VISC)) ? 4 : 0);
  static constexpr int D_A = 0;
  static constexpr int D_E = (MASK & OUT_DVORT) ? 3 : 0;
  static constexpr int ND = DBL ? (D_E + ((MASK & OUT_VISC) ? 4 : 0)) : 0;
};

struct P3DKParams {
  const float4* __restrict__ src_pv;
  const float4* __restrict__ src_a;
  const float4* __restrict__ tgt_pv;
  const float4* __restrict__ tgt_a;
  const float4* __restrict__ gtab;
  float* __restrict__ out_vel;
  float* __restrict__ out_dvort;
  float* __restrict__ out_visc;
  double* __restrict__ partial;          // indexed by target slot
  const int* __restrict__ tgt_perm;      // slot -> target (nullptr: identity)
  int n_src, n_tgt, split_len, n_split;
  RegConsts rc;
  Scales sc;
};

template <int REG, int MASK, int ACC>
__device__ __forceinline__ void p3d_pair(const float4 sp, const float4 sa, const double2* pxy,
                                         const double2* pzv, const float3 xt, const float3 at,
                                         const RegConsts& rc, const float4* tab, float* acc,
                                         double* dacc) {
  using R = P3DRegs<REG, MASK, ACC>;
  const float dx = xt.x - sp.x;
  const float dy = xt.y - sp.y;
  const float dz = xt.z - sp.z;
  const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
  float Kv, Kh, eta, Kfar = 0.f;
  if constexpr (R::SPLIT && (MASK & OUT_DVORT) != 0) {
    RegKernel<REG_GAUSSIAN>::eval_split(r2, rc, tab, Kv, Kh, eta,
                                        [&](float kn, float en) {
      const double2 sxy = *pxy, szv = *pzv;
      const double k = (double)kn;
      dacc[R::D_A + 0] = __builtin_fma(k, sxy.x, dacc[R::D_A + 0]);
      dacc[R::D_A + 1] = __builtin_fma(k, sxy.y, dacc[R::D_A + 1]);
      dacc[R::D_A + 2] = __builtin_fma(k, szv.x, dacc[R::D_A + 2]);
      if constexpr ((MASK & OUT_VISC) != 0) {
        const double e = (double)en;
        dacc[R::D_E + 0] = __builtin_fma(e, sxy.x, dacc[R::D_E + 0]);
        dacc[R::D_E + 1] = __builtin_fma(e, sxy.y, dacc[R::D_E + 1]);
        dacc[R::D_E + 2] = __builtin_fma(e, szv.x, dacc[R::D_E + 2]);
        dacc[R::D_E + 3] = __builtin_fma(e, szv.y, dacc[R::D_E + 3]);
      }
    }, [&](float kf) {
      acc[R::F_A + 0] = __builtin_fmaf(kf, sa.x, acc[R::F_A + 0]);
      acc[R::F_A + 1] = __builtin_fmaf(kf, sa.y, acc[R::F_A + 1]);
      acc[R::F_A + 2] = __builtin_fmaf(kf, sa.z, acc[R::F_A + 2]);
    });
  } else if constexpr (R::SPLIT) {
    constexpr bool PIN = (MASK & (OUT_DVORT | OUT_VISC)) == 0;  // no FP64 near sums
    RegKernel<REG_GAUSSIAN>::template eval_ranges<PIN>(r2, rc, tab, Kv, Kh, eta, Kfar,
                                                       [&](float kn, float en) {
      // FP64 source strengths are read from LDS here, on the rare path (a
      // conversion of sa would be speculated out of the branch)
      const double2 sxy = *pxy, szv = *pzv;
      if constexpr ((MASK & OUT_DVORT) != 0) {
        const double k = (double)kn;
        dacc[R::D_A + 0] = __builtin_fma(k, sxy.x, dacc[R::D_A + 0]);
        dacc[R::D_A + 1] = __builtin_fma(k, sxy.y, dacc[R::D_A + 1]);
        dacc[R::D_A + 2] = __builtin_fma(k, szv.x, dacc[R::D_A + 2]);
      }
      if constexpr ((MASK & OUT_VISC) != 0) {
        const double e = (double)en;
        dacc[R::D_E + 0] = __builtin_fma(e, sxy.x, dacc[R::D_E + 0]);
        dacc[R::D_E + 1] = __builtin_fma(e, sxy.y, dacc[R::D_E + 1]);
        dacc[R::D_E + 2] = __builtin_fma(e, szv.x, dacc[R::D_E + 2]);
        dacc[R::D_E + 3] = __builtin_fma(e, szv.y, dacc[R::D_E + 3]);
      }
    });
  } else {
    RegKernel<REG>::template eval<(MASK & (OUT_DVORT | OUT_VISC)) != 0>(r2, rc, tab, Kv, Kh, eta);
  }
  if constexpr ((MASK & (OUT_VEL | OUT_DVORT)) != 0) {
    // c = d x a_s
    const float cx = __builtin_fmaf(dy, sa.z, -dz * sa.y);
    const float cy = __builtin_fmaf(dz, sa.x, -dx * sa.z);
    const float cz = __builtin_fmaf(dx, sa.y, -dy * sa.x);
    if constexpr ((MASK & OUT_VEL) != 0) {
      acc[R::F_VEL + 0] = __builtin_fmaf(Kv, cx, acc[R::F_VEL + 0]);
      acc[R::F_VEL + 1] = __builtin_fmaf(Kv, cy, acc[R::F_VEL + 1]);
      acc[R::F_VEL + 2] = __builtin_fmaf(Kv, cz, acc[R::F_VEL + 2]);
    }
    if constexpr ((MASK & OUT_DVORT) != 0) {
      if constexpr (R::SPLIT) {
        // the far range's stretching sum was added by eval_split's callback
      } else if constexpr (R::DBL) {
        const double2 sxy = *pxy, szv = *pzv;
        const double k = (double)Kv;
        dacc[R::D_A + 0] = __builtin_fma(k, sxy.x, dacc[R::D_A + 0]);
        dacc[R::D_A + 1] = __builtin_fma(k, sxy.y, dacc[R::D_A + 1]);
        dacc[R::D_A + 2] = __builtin_fma(k, szv.x, dacc[R::D_A + 2]);
      } else {
        acc[R::F_A + 0] = __builtin_fmaf(Kv, sa.x, acc[R::F_A + 0]);
        acc[R::F_A + 1] = __builtin_fmaf(Kv, sa.y, acc[R::F_A + 1]);
        acc[R::F_A + 2] = __builtin_fmaf(Kv, sa.z, acc[R::F_A + 2]);
      }
      const float t = Kh * __builtin_fmaf(at.x, cx, __builtin_fmaf(at.y, cy, at.z * cz));
      acc[R::F_H + 0] = __builtin_fmaf(t, dx, acc[R::F_H + 0]);
      acc[R::F_H + 1] = __builtin_fmaf(t, dy, acc[R::F_H + 1]);
      acc[R::F_H + 2] = __builtin_fmaf(t, dz, acc[R::F_H + 2]);
    }
  }
  if constexpr ((MASK & OUT_VISC) != 0 && !R::SPLIT) {
    if constexpr (R::DBL) {
      const double2 sxy = *pxy, szv = *pzv;
      const double e = (double)eta;
      dacc[R::D_E + 0] = __builtin_fma(e, sxy.x, dacc[R::D_E + 0]);
      dacc[R::D_E + 1] = __builtin_fma(e, sxy.y, dacc[R::D_E + 1]);
      dacc[R::D_E + 2] = __builtin_fma(e, szv.x, dacc[R::D_E + 2]);
      dacc[R::D_E + 3] = __builtin_fma(e, szv.y, dacc[R::D_E + 3]);
    } else {
      acc[R::F_E + 0] = __builtin_fmaf(eta, sa.x, acc[R::F_E + 0]);
      acc[R::F_E + 1] = __builtin_fmaf(eta, sa.y, acc[R::F_E + 1]);
      acc[R::F_E + 2] = __builtin_fmaf(eta, sa.z, acc[R::F_E + 2]);
      acc[R::F_E + 3] = __builtin_fmaf(eta, sp.w, acc[R::F_E + 3]);
    }
  }
}

// Final combination of the reduced sums for one target (FP64).
template <int MASK>
__device__ __forceinline__ void p3d_finish(const double* a, float3 at, float vt, const Scales& sc,
                                           float* out_vel, float* out_dvort, float* out_visc,
                                           int i) {
  using L = P3DLayout<MASK>;
  if constexpr ((MASK & OUT_VEL) != 0) {
    out_vel[3 * (size_t)i + 0] = (float)(sc.vel * a[L::VEL + 0]);
    out_vel[3 * (size_t)i + 1] = (float)(sc.vel * a[L::VEL + 1]);
    out_vel[3 * (size_t)i + 2] = (float)(sc.vel * a[L::VEL + 2]);
  }
  if constexpr ((MASK & OUT_DVORT) != 0) {
    const double ax = at.x, ay = at.y, az = at.z;
    const double Ax = a[L::A + 0], Ay = a[L::A + 1], Az = a[L::A + 2];
    const double wx = ay * Az - az * Ay;
    const double wy = az * Ax - ax * Az;
    const double wz = ax * Ay - ay * Ax;
    out_dvort[3 * (size_t)i + 0] = (float)(sc.dG * wx + sc.dH * a[L::H + 0]);
    out_dvort[3 * (size_t)i + 1] = (float)(sc.dG * wy + sc.dH * a[L::H + 1]);
    out_dvort[3 * (size_t)i + 2] = (float)(sc.dG * wz + sc.dH * a[L::H + 2]);
  }
  if constexpr ((MASK & OUT_VISC) != 0) {
    const double v = vt, e = a[L::EV];
    out_visc[3 * (size_t)i + 0] = (float)(sc.visc * (v * a[L::E + 0] - (double)at.x * e));
    out_visc[3 * (size_t)i + 1] = (float)(sc.visc * (v * a[L::E + 1] - (double)at.y * e));
    out_visc[3 * (size_t)i + 2] = (float)(sc.visc * (v * a[L::E + 2] - (double)at.z * e));
  }
}

// Shared-memory staging of one workgroup.  A source is read from LDS on
// every pair: staged as {x, y, z, ax} and {ay, az} (with the viscous term
// {ay, az, volume, -}), it is one 16-byte and one 8-byte read, where
// {x, y, z, volume} and {ax, ay, az, -} took two 12-byte reads.  Measured on
// the 2^20 ring, same box: fused vel + dvort 671.4 -> 659.7 ms, dvort alone
// 595.0 -> 585.6 ms, vel + dvort + visc 695.4 -> 673.0 ms, identical results.
template <int REG, int MASK, int ACC, int WG>
struct P3DSmem {
  static constexpr bool DBL = P3DRegs<REG, MASK, ACC>::ND > 0;  // FP64 source strengths
  static constexpr bool VOL = (MASK & OUT_VISC) != 0;
  using Tail = typename std::conditional<VOL, float4, float2>::type;
  float4 xa[WG];   // x, y, z, ax
  Tail tl[WG];     // ay, az (, volume)
  double2 xy[DBL ? WG : 1];
  double2 zv[DBL ? WG : 1];
  float4 tab[REG == REG_GAUSSIAN ? 2 * GTAB_N : 1];
};

// One workgroup: targets base + k*WG + lane (k < T) that are < tgt_end,
// against sources [s0, s1).  Writes the outputs (n_split == 1) or the FP64
// partials of partition `split`.  The caller has loaded sm.tab (Gaussian)
// and issues no barrier-sensitive work on sm around the call.
template <int REG, int MASK, int ACC, int T, int WG>
__device__ __forceinline__ void p3d_block(const P3DKParams& p, const RegConsts& rc,
                                          const Scales& sc, int base, int tgt_end, int s0,
                                          int s1, int split, P3DSmem<REG, MASK, ACC, WG>& sm) {
  using L = P3DLayout<MASK>;
  using R = P3DRegs<REG, MASK, ACC>;
  constexpr int NF = R::NF;
  constexpr int ND = R::ND > 0 ? R::ND : 1;
  constexpr bool DBL_ = P3DSmem<REG, MASK, ACC, WG>::DBL;
  const int tid = threadIdx.x;
  float3 xt[T], at[T];
  float vt[T];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    int i = base + k * WG + tid;
    i = i < tgt_end ? i : tgt_end - 1;
    if (p.tgt_perm) i = p.tgt_perm[i];
    const float4 q = p.tgt_pv[i];
    xt[k] = make_float3(q.x, q.y, q.z);
    vt[k] = q.w;
    if constexpr (L::NEED_TA) {
      const float4 w = p.tgt_a[i];
      at[k] = make_float3(w.x, w.y, w.z);
    } else {
      at[k] = make_float3(0.f, 0.f, 0.f);
    }
  }

  float acc[T][NF];
  float lo[T][NF];
  float hi[T][NF];
  double dacc[T][ND];
#pragma unroll
  for (int k = 0; k < T; ++k) {
#pragma unroll
    for (int a = 0; a < NF; ++a) { acc[k][a] = 0.f; lo[k][a] = 0.f; hi[k][a] = 0.f; }
#pragma unroll
    for (int a = 0; a < ND; ++a) dacc[k][a] = 0.0;
  }

  const RegConsts rv = loop_consts(rc);
  for (int ts = s0; ts < s1; ts += WG) {
    const int cnt = min(WG, s1 - ts);
    __syncthreads();
    if (tid < cnt) {
      const float4 q = p.src_pv[ts + tid];
      const float4 w = p.src_a[ts + tid];
      sm.xa[tid] = make_float4(q.x, q.y, q.z, w.x);
      if constexpr (P3DSmem<REG, MASK, ACC, WG>::VOL)
        sm.tl[tid] = make_float4(w.y, w.z, q.w, 0.f);
      else
        sm.tl[tid] = make_float2(w.y, w.z);
      if constexpr (DBL_) {
        sm.xy[tid] = make_double2((double)w.x, (double)w.y);
        sm.zv[tid] = make_double2((double)w.z, (double)q.w);
      }
    }
    __syncthreads();
    if constexpr (ACC != ACC_PLAIN) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < NF; ++a) acc[k][a] = 0.f;
    }
    // the LDS index runs in a VGPR too: an SGPR index would cost an
    // SGPR-reading move per source to form the ds_read address
    int jv = in_vgpr(0);
    // two sources per iteration, except around the velocity-only Gaussian's
    // pinned near-range branch, which cannot be unrolled, and four for the
    // short velocity-only pair with one target per lane (uniform cube of
    // 8192, Winckelmans: 41.0 -> 39.0 us per evaluation; 79.7 -> 63.0 us
    // with one wave per SIMD; eight sources: 40.9 us)
    constexpr int kUnroll = (R::SPLIT && (MASK & (OUT_DVORT | OUT_VISC)) == 0)
                                ? 1
                                : ((MASK == OUT_VEL && T == 1) ? 4 : 2);
#pragma unroll kUnroll
    for (int j = 0; j < cnt; ++j, ++jv) {
      const float4 xa = sm.xa[jv];
      const auto tl = sm.tl[jv];
      float vol = 0.f;
      if constexpr (P3DSmem<REG, MASK, ACC, WG>::VOL) vol = tl.z;
      const float4 sp = make_float4(xa.x, xa.y, xa.z, vol);
      const float4 sa = make_float4(xa.w, tl.x, tl.y, 0.f);
#pragma unroll
      for (int k = 0; k < T; ++k)
        p3d_pair<REG, MASK, ACC>(sp, sa, &sm.xy[DBL_ ? jv : 0], &sm.zv[DBL_ ? jv : 0], xt[k], at[k],
                                 rv, sm.tab, acc[k], dacc[k]);
    }
    if constexpr (ACC != ACC_PLAIN) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < NF; ++a) two_sum_add(hi[k][a], lo[k][a], acc[k][a]);
    }
  }

#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * WG + tid;
    if (i >= tgt_end) continue;
    // gather the canonical layout
    double d[L::N];
    auto fval = [&](int a) -> double {
      return (ACC != ACC_PLAIN) ? (double)hi[k][a] + (double)lo[k][a] : (double)acc[k][a];
    };
    if constexpr ((MASK & OUT_VEL) != 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) d[L::VEL + c] = fval(R::F_VEL + c);
    }
    if constexpr ((MASK & OUT_DVORT) != 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if constexpr (R::SPLIT)
          d[L::A + c] = dacc[k][R::D_A + c] + fval(R::F_A + c);
        else
          d[L::A + c] = R::DBL ? dacc[k][R::D_A + c] : fval(R::F_A + c);
        d[L::H + c] = fval(R::F_H + c);
      }
    }
    if constexpr ((MASK & OUT_VISC) != 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) d[L::E + c] = R::DBL ? dacc[k][R::D_E + c] : fval(R::F_E + c);
    }
    if (p.n_split == 1) {
      p3d_finish<MASK>(d, at[k], vt[k], sc, p.out_vel, p.out_dvort, p.out_visc,
                       p.tgt_perm ? p.tgt_perm[i] : i);
    } else {
#pragma unroll
      for (int a = 0; a < L::N; ++a)
        p.partial[((size_t)split * L::N + a) * p.n_tgt + i] = d[a];
    }
  }
}


}  // namespace cvtx
