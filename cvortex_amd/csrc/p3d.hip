// p3d.hip -- 3D vortex particle direct-sum kernels for gfx950.
//
// One kernel evaluates any subset of {vel, dvort, visc} for the targets of a
// workgroup against one partition of the sources:
//   * each lane owns T targets in registers (targets k*WG + lane of the
//     workgroup's block, so target loads are coalesced);
//   * sources stream through LDS in tiles of WG particles, loaded coalesced
//     by the whole workgroup and then read by every lane at the same address
//     (an LDS broadcast), so each source is fetched from memory once per
//     workgroup and shared by WG*T targets;
//   * per pair: d, r2, one regularisation evaluation shared by all outputs,
//     then FMA accumulation into the reduced sums of device_common.hpp;
//   * ACC_COMP: per-tile FP32 sums folded into a TwoSum (hi, lo) pair per
//     accumulator; ACC_PLAIN: one FP32 running sum;
//   * n_split > 1: each grid row handles one source partition and writes FP64
//     partial sums; p3d_reduce adds the partitions in FP64 and applies the
//     epilogue.  Otherwise the epilogue runs in the main kernel.
// Reference semantics: src/P3D.cpp:51-144 (pairs), :230-296 (sums),
// :324-456 (M2M); GPU scaling ocl_P3D.cpp:164,299-304,337,482-487,698-702.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

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
// pair in FP64 (full rate on CDNA4): their near-field terms are large and
// nearly parallel to a_t, and only their small component normal to a_t (or
// the small difference V_t E - a_t e) survives, so FP32 tile sums would
// lose several digits there.  With ACC_PLAIN everything is FP32.
template <int MASK, int ACC>
struct P3DRegs {
  static constexpr bool DBL = (ACC == ACC_COMP);
  static constexpr int F_VEL = 0;
  static constexpr int F_A = (MASK & OUT_VEL) ? 3 : 0;                        // plain only
  static constexpr int F_H = F_A + ((!DBL && (MASK & OUT_DVORT)) ? 3 : 0);
  static constexpr int F_E = F_H + ((MASK & OUT_DVORT) ? 3 : 0);              // plain only
  static constexpr int NF = F_E + ((!DBL && (MASK & OUT_VISC)) ? 4 : 0);
  static constexpr int D_A = 0;
  static constexpr int D_E = (MASK & OUT_DVORT) ? 3 : 0;
  static constexpr int ND = DBL ? (D_E + ((MASK & OUT_VISC) ? 4 : 0)) : 0;
};

int p3d_nacc(int mask) {
  return ((mask & OUT_VEL) ? 3 : 0) + ((mask & OUT_DVORT) ? 6 : 0) + ((mask & OUT_VISC) ? 4 : 0);
}

struct P3DKParams {
  const float4* __restrict__ src_pv;
  const float4* __restrict__ src_a;
  const float4* __restrict__ tgt_pv;
  const float4* __restrict__ tgt_a;
  const float4* __restrict__ gtab;
  float* __restrict__ out_vel;
  float* __restrict__ out_dvort;
  float* __restrict__ out_visc;
  double* __restrict__ partial;
  int n_src, n_tgt, split_len, n_split;
  RegConsts rc;
  Scales sc;
};

template <int REG, int MASK, int ACC>
__device__ __forceinline__ void p3d_pair(const float4 sp, const float4 sa, const double2 sxy,
                                         const double2 szv, const float3 xt, const float3 at,
                                         const RegConsts& rc, const float4* tab, float* acc,
                                         double* dacc) {
  using R = P3DRegs<MASK, ACC>;
  const float dx = xt.x - sp.x;
  const float dy = xt.y - sp.y;
  const float dz = xt.z - sp.z;
  const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
  float Kv, Kh, eta;
  RegKernel<REG>::eval(r2, rc, tab, Kv, Kh, eta);
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
      if constexpr (R::DBL) {
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
  if constexpr ((MASK & OUT_VISC) != 0) {
    if constexpr (R::DBL) {
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
    out_vel[3 * i + 0] = (float)(sc.vel * a[L::VEL + 0]);
    out_vel[3 * i + 1] = (float)(sc.vel * a[L::VEL + 1]);
    out_vel[3 * i + 2] = (float)(sc.vel * a[L::VEL + 2]);
  }
  if constexpr ((MASK & OUT_DVORT) != 0) {
    const double ax = at.x, ay = at.y, az = at.z;
    const double Ax = a[L::A + 0], Ay = a[L::A + 1], Az = a[L::A + 2];
    const double wx = ay * Az - az * Ay;
    const double wy = az * Ax - ax * Az;
    const double wz = ax * Ay - ay * Ax;
    out_dvort[3 * i + 0] = (float)(sc.dG * wx + sc.dH * a[L::H + 0]);
    out_dvort[3 * i + 1] = (float)(sc.dG * wy + sc.dH * a[L::H + 1]);
    out_dvort[3 * i + 2] = (float)(sc.dG * wz + sc.dH * a[L::H + 2]);
  }
  if constexpr ((MASK & OUT_VISC) != 0) {
    const double v = vt, e = a[L::EV];
    out_visc[3 * i + 0] = (float)(sc.visc * (v * a[L::E + 0] - (double)at.x * e));
    out_visc[3 * i + 1] = (float)(sc.visc * (v * a[L::E + 1] - (double)at.y * e));
    out_visc[3 * i + 2] = (float)(sc.visc * (v * a[L::E + 2] - (double)at.z * e));
  }
}

template <int REG, int MASK, int ACC, int T, int WG>
__global__ __launch_bounds__(WG) void p3d_kernel(P3DKParams p) {
  using L = P3DLayout<MASK>;
  using R = P3DRegs<MASK, ACC>;
  constexpr int NF = R::NF;
  constexpr int ND = R::ND > 0 ? R::ND : 1;
  __shared__ float4 s_pv[WG];
  __shared__ float4 s_a[WG];
  __shared__ double2 s_xy[R::ND > 0 ? WG : 1];
  __shared__ double2 s_zv[R::ND > 0 ? WG : 1];
  __shared__ float4 s_tab[REG == REG_GAUSSIAN ? 2 * GTAB_N : 1];

  const int tid = threadIdx.x;
  if constexpr (REG == REG_GAUSSIAN) {
    for (int i = tid; i < 2 * GTAB_N; i += WG) s_tab[i] = p.gtab[i];
  }

  const int base = blockIdx.x * (WG * T);
  float3 xt[T], at[T];
  float vt[T];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    int i = base + k * WG + tid;
    i = i < p.n_tgt ? i : p.n_tgt - 1;
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

  const int s0 = blockIdx.y * p.split_len;
  const int s1 = min(s0 + p.split_len, p.n_src);
  for (int ts = s0; ts < s1; ts += WG) {
    const int cnt = min(WG, s1 - ts);
    __syncthreads();
    if (tid < cnt) {
      const float4 q = p.src_pv[ts + tid];
      const float4 w = p.src_a[ts + tid];
      s_pv[tid] = q;
      s_a[tid] = w;
      if constexpr (R::ND > 0) {
        s_xy[tid] = make_double2((double)w.x, (double)w.y);
        s_zv[tid] = make_double2((double)w.z, (double)q.w);
      }
    }
    __syncthreads();
    if constexpr (ACC == ACC_COMP) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < NF; ++a) acc[k][a] = 0.f;
    }
#pragma unroll 2
    for (int j = 0; j < cnt; ++j) {
      const float4 sp = s_pv[j];
      const float4 sa = s_a[j];
      double2 sxy = make_double2(0.0, 0.0), szv = make_double2(0.0, 0.0);
      if constexpr (R::ND > 0) {
        sxy = s_xy[j];
        szv = s_zv[j];
      }
#pragma unroll
      for (int k = 0; k < T; ++k)
        p3d_pair<REG, MASK, ACC>(sp, sa, sxy, szv, xt[k], at[k], p.rc, s_tab, acc[k], dacc[k]);
    }
    if constexpr (ACC == ACC_COMP) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < NF; ++a) two_sum_add(hi[k][a], lo[k][a], acc[k][a]);
    }
  }

#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * WG + tid;
    if (i >= p.n_tgt) continue;
    // gather the canonical layout
    double d[L::N];
    auto fval = [&](int a) -> double {
      return (ACC == ACC_COMP) ? (double)hi[k][a] + (double)lo[k][a] : (double)acc[k][a];
    };
    if constexpr ((MASK & OUT_VEL) != 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) d[L::VEL + c] = fval(R::F_VEL + c);
    }
    if constexpr ((MASK & OUT_DVORT) != 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        d[L::A + c] = R::DBL ? dacc[k][R::D_A + c] : fval(R::F_A + c);
        d[L::H + c] = fval(R::F_H + c);
      }
    }
    if constexpr ((MASK & OUT_VISC) != 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) d[L::E + c] = R::DBL ? dacc[k][R::D_E + c] : fval(R::F_E + c);
    }
    if (p.n_split == 1) {
      p3d_finish<MASK>(d, at[k], vt[k], p.sc, p.out_vel, p.out_dvort, p.out_visc, i);
    } else {
#pragma unroll
      for (int a = 0; a < L::N; ++a)
        p.partial[((size_t)blockIdx.y * L::N + a) * p.n_tgt + i] = d[a];
    }
  }
}

template <int MASK>
__global__ __launch_bounds__(256) void p3d_reduce(P3DKParams p) {
  using L = P3DLayout<MASK>;
  constexpr int NA = L::N;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n_tgt) return;
  double d[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) d[a] = 0.0;
  for (int s = 0; s < p.n_split; ++s)
#pragma unroll
    for (int a = 0; a < NA; ++a) d[a] += p.partial[((size_t)s * NA + a) * p.n_tgt + i];
  const float4 q = p.tgt_pv[i];
  float3 at = make_float3(0.f, 0.f, 0.f);
  if constexpr (L::NEED_TA) {
    const float4 w = p.tgt_a[i];
    at = make_float3(w.x, w.y, w.z);
  }
  p3d_finish<MASK>(d, at, q.w, p.sc, p.out_vel, p.out_dvort, p.out_visc, i);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

bool p3d_supported(int reg, int mask) {
  if (reg < REG_SINGULAR || reg > REG_GAUSSIAN) return false;
  if (mask <= 0 || mask > 7) return false;
  if ((mask & OUT_VISC) && (reg == REG_SINGULAR || reg == REG_PLANETARY)) return false;
  return true;
}

RegConsts p3d_consts(int reg, float sigma) {
  RegConsts k{};
  const double s = std::fabs((double)sigma);
  k.s2 = (float)(s * s);
  k.inv_s2 = (float)(1.0 / (s * s));
  if (reg == REG_GAUSSIAN) {
    k.c1 = (float)(-0.5 / std::log(2.0));
    k.c2 = (float)std::sqrt(2.0 / M_PI);
  }
  return k;
}

Scales p3d_scales(int reg, float sigma, float nu) {
  Scales sc;
  const double s = std::fabs((double)sigma);
  const double sgn = sigma < 0 ? -1.0 : 1.0;
  const double k4pi = 1.0 / (4.0 * M_PI);
  const double s3 = s * s * s, s5 = s3 * s * s;
  switch (reg) {
    case REG_GAUSSIAN:
      sc.vel = -k4pi / s3;
      sc.dG = sgn * k4pi / s3;
      sc.dH = sgn * k4pi / s5;
      sc.visc = 2.0 * (double)nu / (s * s);
      break;
    case REG_WINCKELMANS:
      sc.vel = -k4pi / s3;
      sc.dG = sgn * k4pi / s3;
      sc.dH = 3.0 * sgn * k4pi / s5;
      sc.visc = 2.0 * (double)nu / (s * s) * 52.5;
      break;
    case REG_PLANETARY:
      sc.vel = -k4pi / s3;
      sc.dG = sgn * k4pi / s3;
      sc.dH = 3.0 * sgn * k4pi / s5;
      break;
    default:  // singular: Kv = r^-3, Kh = r^-5 are dimensional
      sc.vel = -k4pi;
      sc.dG = sgn * k4pi;
      sc.dH = 3.0 * sgn * k4pi;
      break;
  }
  return sc;
}

size_t p3d_workspace_bytes(const P3DJob& j, const LaunchCfg& c) {
  if (c.n_split <= 1) return 0;
  return (size_t)c.n_split * p3d_nacc(j.mask) * (size_t)j.n_tgt * sizeof(double);
}

namespace {

using KernelFn = void (*)(P3DKParams);

template <int REG, int MASK, int ACC>
KernelFn pick_T(int T) {
  switch (T) {
    case 1: return p3d_kernel<REG, MASK, ACC, 1, 256>;
    case 4: return p3d_kernel<REG, MASK, ACC, 4, 256>;
    default: return p3d_kernel<REG, MASK, ACC, 2, 256>;
  }
}

template <int REG, int MASK>
KernelFn pick_acc(int acc, int T) {
  return acc == ACC_PLAIN ? pick_T<REG, MASK, ACC_PLAIN>(T) : pick_T<REG, MASK, ACC_COMP>(T);
}

template <int REG>
KernelFn pick_mask_inviscid(int mask, int acc, int T) {
  switch (mask) {
    case OUT_VEL: return pick_acc<REG, OUT_VEL>(acc, T);
    case OUT_DVORT: return pick_acc<REG, OUT_DVORT>(acc, T);
    case OUT_VEL | OUT_DVORT: return pick_acc<REG, OUT_VEL | OUT_DVORT>(acc, T);
    default: return nullptr;
  }
}

template <int REG>
KernelFn pick_mask_viscous(int mask, int acc, int T) {
  switch (mask) {
    case OUT_VEL: return pick_acc<REG, OUT_VEL>(acc, T);
    case OUT_DVORT: return pick_acc<REG, OUT_DVORT>(acc, T);
    case OUT_VISC: return pick_acc<REG, OUT_VISC>(acc, T);
    case OUT_VEL | OUT_DVORT: return pick_acc<REG, OUT_VEL | OUT_DVORT>(acc, T);
    case OUT_DVORT | OUT_VISC: return pick_acc<REG, OUT_DVORT | OUT_VISC>(acc, T);
    case OUT_VEL | OUT_DVORT | OUT_VISC: return pick_acc<REG, OUT_VEL | OUT_DVORT | OUT_VISC>(acc, T);
    default: return nullptr;
  }
}

KernelFn pick_kernel(int reg, int mask, int acc, int T) {
  switch (reg) {
    case REG_SINGULAR: return pick_mask_inviscid<REG_SINGULAR>(mask, acc, T);
    case REG_PLANETARY: return pick_mask_inviscid<REG_PLANETARY>(mask, acc, T);
    case REG_WINCKELMANS: return pick_mask_viscous<REG_WINCKELMANS>(mask, acc, T);
    case REG_GAUSSIAN: return pick_mask_viscous<REG_GAUSSIAN>(mask, acc, T);
    default: return nullptr;
  }
}

KernelFn pick_reduce(int mask) {
  switch (mask) {
    case OUT_VEL: return p3d_reduce<OUT_VEL>;
    case OUT_DVORT: return p3d_reduce<OUT_DVORT>;
    case OUT_VISC: return p3d_reduce<OUT_VISC>;
    case OUT_VEL | OUT_DVORT: return p3d_reduce<OUT_VEL | OUT_DVORT>;
    case OUT_DVORT | OUT_VISC: return p3d_reduce<OUT_DVORT | OUT_VISC>;
    case OUT_VEL | OUT_DVORT | OUT_VISC: return p3d_reduce<OUT_VEL | OUT_DVORT | OUT_VISC>;
    default: return nullptr;
  }
}

}  // namespace

hipError_t launch_p3d(const P3DJob& j, const LaunchCfg& c, hipStream_t s) {
  if (j.n_tgt <= 0) return hipSuccess;
  if (!p3d_supported(j.reg, j.mask)) return hipErrorInvalidValue;
  if (j.n_src < 0 || c.wg != 256) return hipErrorInvalidValue;
  if ((j.mask & (OUT_DVORT | OUT_VISC)) && j.tgt_a == nullptr) return hipErrorInvalidValue;
  if (j.reg == REG_GAUSSIAN && j.gtab == nullptr) return hipErrorInvalidValue;
  KernelFn k = pick_kernel(j.reg, j.mask, j.acc, c.T);
  if (!k) return hipErrorInvalidValue;

  P3DKParams p{};
  p.src_pv = j.src_pv;
  p.src_a = j.src_a;
  p.tgt_pv = j.tgt_pv;
  p.tgt_a = j.tgt_a;
  p.gtab = j.gtab;
  p.out_vel = j.out_vel;
  p.out_dvort = j.out_dvort;
  p.out_visc = j.out_visc;
  p.partial = j.partial;
  p.n_src = j.n_src;
  p.n_tgt = j.n_tgt;
  p.rc = p3d_consts(j.reg, j.sigma);
  p.sc = p3d_scales(j.reg, j.sigma, j.nu);

  int n_split = c.n_split;
  int split_len = c.split_len;
  if (n_split <= 1 || j.n_src == 0) {
    n_split = 1;
    split_len = j.n_src > 0 ? j.n_src : 1;
  } else {
    if (split_len <= 0 || split_len % c.wg != 0) return hipErrorInvalidValue;
    if ((long long)split_len * n_split < j.n_src) return hipErrorInvalidValue;
    if (j.partial == nullptr) return hipErrorInvalidValue;
  }
  p.n_split = n_split;
  p.split_len = split_len;

  const int tgt_per_block = c.wg * c.T;
  dim3 grid((j.n_tgt + tgt_per_block - 1) / tgt_per_block, n_split);
  hipLaunchKernelGGL(k, grid, dim3(c.wg), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_split == 1) return e;
  KernelFn r = pick_reduce(j.mask);
  hipLaunchKernelGGL(r, dim3((j.n_tgt + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace cvtx
