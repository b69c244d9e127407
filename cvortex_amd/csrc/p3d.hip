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
//     accumulator, FP64 stretching and PSE sums per pair; ACC_TILES: the
//     folded tile sums for every accumulator; ACC_PLAIN: one FP32 running sum;
//   * n_split > 1: each grid row handles one source partition and writes FP64
//     partial sums; p3d_reduce adds the partitions in FP64 and applies the
//     epilogue.  Otherwise the epilogue runs in the main kernel.
// Reference semantics: src/P3D.cpp:51-144 (pairs), :230-296 (sums),
// :324-456 (M2M); GPU scaling ocl_P3D.cpp:164,299-304,337,482-487,698-702.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "kernels.hpp"
#include "p3d_body.hpp"

namespace cvtx {

int p3d_nacc(int mask) {
  return ((mask & OUT_VEL) ? 3 : 0) + ((mask & OUT_DVORT) ? 6 : 0) + ((mask & OUT_VISC) ? 4 : 0);
}

template <int REG, int MASK, int ACC, int T, int WG>
__global__ __launch_bounds__(WG) void p3d_kernel(P3DKParams p) {
  __shared__ P3DSmem<REG, MASK, ACC, WG> sm;
  if constexpr (REG == REG_GAUSSIAN) {
    for (int i = threadIdx.x; i < 2 * GTAB_N; i += WG) sm.tab[i] = p.gtab[i];
  }
  const int s0 = blockIdx.y * p.split_len;
  const int s1 = min(s0 + p.split_len, p.n_src);
  p3d_block<REG, MASK, ACC, T, WG>(p, p.rc, p.sc, blockIdx.x * (WG * T), p.n_tgt, s0, s1,
                                   blockIdx.y, sm);
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
  // partitions added in order; unrolled so that the loads of several
  // partitions are in flight at once
#pragma unroll 4
  for (int s = 0; s < p.n_split; ++s)
#pragma unroll
    for (int a = 0; a < NA; ++a) d[a] += p.partial[((size_t)s * NA + a) * p.n_tgt + i];
  const int t = p.tgt_perm ? p.tgt_perm[i] : i;
  const float4 q = p.tgt_pv[t];
  float3 at = make_float3(0.f, 0.f, 0.f);
  if constexpr (L::NEED_TA) {
    const float4 w = p.tgt_a[t];
    at = make_float3(w.x, w.y, w.z);
  }
  p3d_finish<MASK>(d, at, q.w, p.sc, p.out_vel, p.out_dvort, p.out_visc, t);
}

// Velocity alone: one thread per target and component, so three times the
// threads of p3d_reduce (a small problem's reduction has few targets), each
// with sixteen partitions' loads in flight; the partitions are added in
// order, as p3d_reduce does.  Cube of 8192 (32 partitions): 44.0 -> 43.6 us
// per evaluation.
__global__ __launch_bounds__(64) void p3d_reduce_vel(P3DKParams p) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= p.n_tgt) return;
  double d = 0.0;
#pragma unroll 16
  for (int s = 0; s < p.n_split; ++s) d += p.partial[((size_t)s * 3 + c) * p.n_tgt + i];
  const int t = p.tgt_perm ? p.tgt_perm[i] : i;
  p.out_vel[3 * (size_t)t + c] = (float)(p.sc.vel * d);
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
      sc.dH = 3.0 * sgn * k4pi / s5;
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
  if (acc == ACC_PLAIN) return pick_T<REG, MASK, ACC_PLAIN>(T);
  if (acc == ACC_TILES) return pick_T<REG, MASK, ACC_TILES>(T);
  return pick_T<REG, MASK, ACC_COMP>(T);
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
  // every output the mask selects must exist, and the inputs it reads
  if (((j.mask & OUT_VEL) && !j.out_vel) || ((j.mask & OUT_DVORT) && !j.out_dvort) ||
      ((j.mask & OUT_VISC) && !j.out_visc) || !j.tgt_pv || (j.n_src > 0 && (!j.src_pv || !j.src_a)))
    return hipErrorInvalidValue;
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
  p.tgt_perm = j.tgt_perm;
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
  if (j.mask == OUT_VEL) {
    hipLaunchKernelGGL(p3d_reduce_vel, dim3((j.n_tgt + 63) / 64, 3), dim3(64), 0, s, p);
  } else {
    KernelFn r = pick_reduce(j.mask);
    hipLaunchKernelGGL(r, dim3((j.n_tgt + 255) / 256), dim3(256), 0, s, p);
  }
  return hipGetLastError();
}

}  // namespace cvtx
