// vort.hip -- vorticity field of 3D vortex particles (cvtx_P3D_M2M_vort) and
// the Pedrizzetti relaxation update, for gfx950.
//
// Semantics are those of the reference CPU path (src/P3D.cpp:298-322,
// 458-498, 669-707):
//   omega(x) = sum over sources s with |dx|, |dy|, |dz| < 5 sigma of
//              zeta(|x_s - x| / sigma) alpha_s,   then / (4 pi sigma^3)
// with d = x_s - x, the cut-off box tested component-wise in FP32 exactly as
// the reference tests it, and the self-pair included (zeta(0) is finite).
// (The reference's OpenCL kernel has neither the cut-off nor the self-pair;
// the CPU path is the one followed here.)
//
// Structure: as the pair kernels -- 256-source LDS tiles shared by the
// workgroup, T targets per lane, FP32 tile sums combined by TwoSum, optional
// source partitions with an FP64 reduction.  Because the cut-off makes the
// interaction short-ranged, large problems run on Morton-sorted sources and
// targets: each 256-source tile and each of its 64-source chunks carries its
// bounding box; a workgroup skips a tile, and a wave a chunk, whose box is
// separated from the box of its own targets by at least the cut-off along
// any axis.  FP32 subtraction is monotonic, so every pair of a skipped tile
// or chunk would fail the reference's own box test: skipping changes no
// result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "kernels.hpp"

namespace cvtx {

namespace {

// zeta(rho) up to the per-call constant the epilogue applies (vort_scale).
template <int REG>
struct Zeta;

// eval(dx, dy, dz, rsigma, sigma, divide, consts): rsigma = 1/sigma and
// sigma feed only the planetary core test (rho = |d| rsigma, or |d| / sigma
// when `divide`, as the reference's two vorticity forms compute it).
template <>
struct Zeta<REG_SINGULAR> {  // zeta = 0
  static __device__ __forceinline__ float eval(float, float, float, float, float, bool,
                                               const RegConsts&) {
    return 0.f;
  }
};

template <>
struct Zeta<REG_WINCKELMANS> {  // 7.5 (rho^2 + 1)^-3.5 -> y^7, y = (q + 1)^-1/2
  static __device__ __forceinline__ float eval(float dx, float dy, float dz, float, float, bool,
                                               const RegConsts& k) {
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    const float y = __builtin_amdgcn_rsqf(__builtin_fmaf(r2, k.inv_s2, 1.f));
    const float y2 = y * y;
    const float y3 = y2 * y;
    return (y3 * y3) * y;
  }
};

template <>
struct Zeta<REG_PLANETARY> {  // 3 inside the core, 0 outside -> 1 / 0
  // The core test is a discontinuity, so rho is formed with the reference's
  // FP32 operations: |d| = sqrt(dx dx + dy dy + dz dz) rounded at every
  // step, then rho = |d| * (1/sigma) (src/P3D.cpp:314-315) or |d| / sigma
  // (src/P3D.cpp:154-155).
  static __device__ __forceinline__ float eval(float dx, float dy, float dz, float rsigma,
                                               float sigma, bool divide, const RegConsts&) {
    const float r2 = rn_add(rn_add(rn_mul(dx, dx), rn_mul(dy, dy)), rn_mul(dz, dz));
    const float r = __fsqrt_rn(r2);
    const float rho = divide ? __fdiv_rn(r, sigma) : rn_mul(r, rsigma);
    return rho < 1.f ? 1.f : 0.f;
  }
};

template <>
struct Zeta<REG_GAUSSIAN> {  // sqrt(2/pi) exp(-rho^2 / 2) -> 2^(q c1)
  static __device__ __forceinline__ float eval(float dx, float dy, float dz, float, float, bool,
                                               const RegConsts& k) {
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    return __builtin_amdgcn_exp2f(r2 * k.inv_s2 * k.c1);
  }
};

struct VortKParams {
  const float4* __restrict__ src_pv;
  const float4* __restrict__ src_a;
  const float4* __restrict__ tgt;
  const int* __restrict__ tgt_perm;   // slot -> target (nullptr: identity)
  const float4* __restrict__ box_lo;  // per 256-source tile (nullptr: no skipping)
  const float4* __restrict__ box_hi;
  const float4* __restrict__ sub_lo;  // per 64-source chunk
  const float4* __restrict__ sub_hi;
  float* __restrict__ out;
  double* __restrict__ partial;
  int n_src, n_tgt, split_len, n_split;
  float cutoff;  // 5 sigma in FP32, as the reference forms it (+inf: no cut-off)
  float rsigma;  // 1 / sigma in FP32
  float sigma;
  int divide;    // planetary rho = |d| / sigma instead of |d| * rsigma
  RegConsts rc;
  double scale;  // zeta constant / (4 pi sigma^3)
};

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

template <int REG, int ACC, int T, int WG>
__global__ __launch_bounds__(WG) void vort_kernel(VortKParams p) {
  __shared__ float4 s_pv[WG];
  __shared__ float4 s_a[WG];
  __shared__ float s_box[WG / 64][6];
  const int tid = threadIdx.x;
  const int base = blockIdx.x * (WG * T);
  float3 xt[T];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    int i = base + k * WG + tid;
    i = i < p.n_tgt ? i : p.n_tgt - 1;
    if (p.tgt_perm) i = p.tgt_perm[i];
    const float4 q = p.tgt[i];
    xt[k] = make_float3(q.x, q.y, q.z);
  }
  // bounding boxes of this wave's and this workgroup's targets (used only
  // with tile boxes)
  float3 wlo = make_float3(0.f, 0.f, 0.f), whi = wlo;
  float3 lo = wlo, hi = wlo;
  const bool skipping = p.box_lo != nullptr;
  if (skipping) {
    lo = xt[0];
    hi = xt[0];
#pragma unroll
    for (int k = 1; k < T; ++k) {
      lo.x = fminf(lo.x, xt[k].x); lo.y = fminf(lo.y, xt[k].y); lo.z = fminf(lo.z, xt[k].z);
      hi.x = fmaxf(hi.x, xt[k].x); hi.y = fmaxf(hi.y, xt[k].y); hi.z = fmaxf(hi.z, xt[k].z);
    }
    lo.x = wave_min(lo.x); lo.y = wave_min(lo.y); lo.z = wave_min(lo.z);
    hi.x = wave_max(hi.x); hi.y = wave_max(hi.y); hi.z = wave_max(hi.z);
    if ((tid & 63) == 0) {
      float* b = s_box[tid >> 6];
      b[0] = lo.x; b[1] = lo.y; b[2] = lo.z; b[3] = hi.x; b[4] = hi.y; b[5] = hi.z;
    }
    __syncthreads();
    wlo = make_float3(s_box[0][0], s_box[0][1], s_box[0][2]);
    whi = make_float3(s_box[0][3], s_box[0][4], s_box[0][5]);
#pragma unroll
    for (int w = 1; w < WG / 64; ++w) {
      wlo.x = fminf(wlo.x, s_box[w][0]); wlo.y = fminf(wlo.y, s_box[w][1]);
      wlo.z = fminf(wlo.z, s_box[w][2]);
      whi.x = fmaxf(whi.x, s_box[w][3]); whi.y = fmaxf(whi.y, s_box[w][4]);
      whi.z = fmaxf(whi.z, s_box[w][5]);
    }
  }

  float acc[T][3], sh[T][3], sl[T][3];
#pragma unroll
  for (int k = 0; k < T; ++k)
#pragma unroll
    for (int a = 0; a < 3; ++a) { acc[k][a] = 0.f; sh[k][a] = 0.f; sl[k][a] = 0.f; }

  const float c = p.cutoff;
  // per-pair uniform values in VGPRs (an SGPR operand blocks dual issue)
  const float cv = in_vgpr(c);
  const RegConsts rv = loop_consts(p.rc);
  const int s0 = blockIdx.y * p.split_len;
  const int s1 = min(s0 + p.split_len, p.n_src);
  for (int ts = s0; ts < s1; ts += WG) {
    if (skipping) {
      // identical in every lane: the whole workgroup skips together
      const float4 bl = p.box_lo[ts / WG];
      const float4 bh = p.box_hi[ts / WG];
      const bool apart = (bl.x - whi.x >= c) || (wlo.x - bh.x >= c) || (bl.y - whi.y >= c) ||
                         (wlo.y - bh.y >= c) || (bl.z - whi.z >= c) || (wlo.z - bh.z >= c);
      if (apart) continue;
    }
    const int cnt = min(WG, s1 - ts);
    __syncthreads();
    if (tid < cnt) {
      s_pv[tid] = p.src_pv[ts + tid];
      s_a[tid] = p.src_a[ts + tid];
    }
    __syncthreads();
    if constexpr (ACC == ACC_COMP) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) acc[k][a] = 0.f;
    }
    // 64-source chunks; with boxes, a wave skips a chunk whose box is apart
    // from the wave's own targets (uniform within the wave)
    for (int j0 = 0; j0 < cnt; j0 += 64) {
      if (skipping) {
        const float4 bl = p.sub_lo[(ts + j0) / 64];
        const float4 bh = p.sub_hi[(ts + j0) / 64];
        const bool apart = (bl.x - hi.x >= c) || (lo.x - bh.x >= c) || (bl.y - hi.y >= c) ||
                           (lo.y - bh.y >= c) || (bl.z - hi.z >= c) || (lo.z - bh.z >= c);
        if (apart) continue;
      }
      const int j1 = min(cnt, j0 + 64);
      int jv = in_vgpr(j0);  // LDS index in a VGPR
#pragma unroll 2
      for (int j = j0; j < j1; ++j, ++jv) {
        const float4 sp = s_pv[jv];
        const float4 sa = s_a[jv];
#pragma unroll
        for (int k = 0; k < T; ++k) {
          const float dx = sp.x - xt[k].x;
          const float dy = sp.y - xt[k].y;
          const float dz = sp.z - xt[k].z;
          const bool in = (__builtin_fabsf(dx) < cv) && (__builtin_fabsf(dy) < cv) &&
                          (__builtin_fabsf(dz) < cv);
          float z = Zeta<REG>::eval(dx, dy, dz, p.rsigma, p.sigma, p.divide != 0, rv);
          z = in ? z : 0.f;
          acc[k][0] = __builtin_fmaf(z, sa.x, acc[k][0]);
          acc[k][1] = __builtin_fmaf(z, sa.y, acc[k][1]);
          acc[k][2] = __builtin_fmaf(z, sa.z, acc[k][2]);
        }
      }
    }
    if constexpr (ACC == ACC_COMP) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) two_sum_add(sh[k][a], sl[k][a], acc[k][a]);
    }
  }

#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * WG + tid;
    if (i >= p.n_tgt) continue;
    double d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
      d[a] = (ACC == ACC_COMP) ? (double)sh[k][a] + (double)sl[k][a] : (double)acc[k][a];
    if (p.n_split == 1) {
      const int o = p.tgt_perm ? p.tgt_perm[i] : i;
#pragma unroll
      for (int a = 0; a < 3; ++a) p.out[3 * (size_t)o + a] = (float)(p.scale * d[a]);
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) p.partial[((size_t)blockIdx.y * 3 + a) * p.n_tgt + i] = d[a];
    }
  }
}

__global__ __launch_bounds__(256) void vort_reduce(VortKParams p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n_tgt) return;
  double d[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
  for (int s = 0; s < p.n_split; ++s)
#pragma unroll
    for (int a = 0; a < 3; ++a) d[a] += p.partial[((size_t)s * 3 + a) * p.n_tgt + i];
  const int o = p.tgt_perm ? p.tgt_perm[i] : i;
#pragma unroll
  for (int a = 0; a < 3; ++a) p.out[3 * (size_t)o + a] = (float)(p.scale * d[a]);
}

// Sorted copies of the sources: pv_out[i] = pv[perm[i]], a_out[i] = a[perm[i]].
__global__ __launch_bounds__(256) void gather_sorted(const float4* __restrict__ pv,
                                                     const float4* __restrict__ a,
                                                     const int* __restrict__ perm, int n,
                                                     float4* __restrict__ pv_out,
                                                     float4* __restrict__ a_out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  pv_out[i] = pv[j];
  a_out[i] = a[j];
}

// Bounding box of each 256-point tile and of each of its 64-point chunks
// (one workgroup per tile, one wave per chunk).  Lanes past n repeat the
// last point, which lies in the same tile.
__global__ __launch_bounds__(256) void tile_boxes(const float4* __restrict__ pv, int n,
                                                  float4* __restrict__ lo,
                                                  float4* __restrict__ hi,
                                                  float4* __restrict__ sub_lo,
                                                  float4* __restrict__ sub_hi) {
  __shared__ float s[4][6];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const float4 q = pv[i < n ? i : n - 1];
  float a0 = wave_min(q.x), a1 = wave_min(q.y), a2 = wave_min(q.z);
  float b0 = wave_max(q.x), b1 = wave_max(q.y), b2 = wave_max(q.z);
  if ((threadIdx.x & 63) == 0) {
    float* w = s[threadIdx.x >> 6];
    w[0] = a0; w[1] = a1; w[2] = a2; w[3] = b0; w[4] = b1; w[5] = b2;
    const int chunk = blockIdx.x * 4 + (threadIdx.x >> 6);
    sub_lo[chunk] = make_float4(a0, a1, a2, 0.f);
    sub_hi[chunk] = make_float4(b0, b1, b2, 0.f);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // a*, b* hold wave 0's box; fold in waves 1..3
    for (int w = 1; w < 4; ++w) {
      a0 = fminf(a0, s[w][0]); a1 = fminf(a1, s[w][1]); a2 = fminf(a2, s[w][2]);
      b0 = fmaxf(b0, s[w][3]); b1 = fmaxf(b1, s[w][4]); b2 = fmaxf(b2, s[w][5]);
    }
    lo[blockIdx.x] = make_float4(a0, a1, a2, 0.f);
    hi[blockIdx.x] = make_float4(b0, b1, b2, 0.f);
  }
}

// Pedrizzetti update (reference src/P3D.cpp:690-702), FP32 in the
// reference's operation order: new = a (1 - fdt) + omega (|a| / |omega|) fdt,
// zero where |omega| == 0.  a is float4 {alpha, 0}; omega float[3 n].
__global__ __launch_bounds__(256) void relax_kernel(float4* __restrict__ a,
                                                    const float* __restrict__ omega, int n,
                                                    float fdt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 ov = a[i];
  const float wx = omega[3 * (size_t)i + 0], wy = omega[3 * (size_t)i + 1], wz = omega[3 * (size_t)i + 2];
  const float aw =
      __fsqrt_rn(rn_add(rn_add(rn_mul(wx, wx), rn_mul(wy, wy)), rn_mul(wz, wz)));
  const float aa = __fsqrt_rn(
      rn_add(rn_add(rn_mul(ov.x, ov.x), rn_mul(ov.y, ov.y)), rn_mul(ov.z, ov.z)));
  const float coeff = __fdiv_rn(aa, aw);
  const float tmp = rn_sub(1.f, fdt);
  const float cf = rn_mul(coeff, fdt);
  float4 nv;
  nv.x = rn_add(rn_mul(ov.x, tmp), rn_mul(wx, cf));
  nv.y = rn_add(rn_mul(ov.y, tmp), rn_mul(wy, cf));
  nv.z = rn_add(rn_mul(ov.z, tmp), rn_mul(wz, cf));
  nv.w = ov.w;
  if (aw == 0.f) nv = make_float4(0.f, 0.f, 0.f, ov.w);
  a[i] = nv;
}

using VFn = void (*)(VortKParams);

template <int REG, int ACC>
VFn vort_T(int T) {
  switch (T) {
    case 1: return vort_kernel<REG, ACC, 1, 256>;
    case 4: return vort_kernel<REG, ACC, 4, 256>;
    default: return vort_kernel<REG, ACC, 2, 256>;
  }
}
template <int REG>
VFn vort_acc(int acc, int T) {
  return acc == ACC_PLAIN ? vort_T<REG, ACC_PLAIN>(T) : vort_T<REG, ACC_COMP>(T);
}
VFn vort_pick(int reg, int acc, int T) {
  switch (reg) {
    case REG_SINGULAR: return vort_acc<REG_SINGULAR>(acc, T);
    case REG_WINCKELMANS: return vort_acc<REG_WINCKELMANS>(acc, T);
    case REG_PLANETARY: return vort_acc<REG_PLANETARY>(acc, T);
    case REG_GAUSSIAN: return vort_acc<REG_GAUSSIAN>(acc, T);
    default: return nullptr;
  }
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

constexpr int VORT_SORTED_T = 1;

}  // namespace

double vort_scale(int reg, float sigma) {
  // zeta constants: Winckelmans 7.5, planetary 3, Gaussian sqrt(2/pi).  The
  // reference divides by 4 pi sigma^3 with the signed sigma
  // (src/P3D.cpp:320-321); sigma = 0 gives 0 * inf = NaN as there.
  const double s = (double)sigma;
  const double k = reg == REG_WINCKELMANS ? 7.5
                   : reg == REG_PLANETARY ? 3.0
                   : reg == REG_GAUSSIAN  ? std::sqrt(2.0 / M_PI)
                                          : 0.0;
  return k * (1.0 / (4.0 * M_PI * s * s * s));
}

LaunchCfg vort_default_cfg(int n_src, int n_tgt, int n_cu, bool sorted) {
  if (!sorted) return default_cfg(n_src, n_tgt, n_cu);
  // Sorted sources with tile skipping: one target per lane (a smaller target
  // block has a smaller bounding box, so it skips more source tiles) and 64
  // source partitions.  The work left after skipping is very uneven --
  // target blocks in dense regions keep most of their tiles -- and small
  // partitions spread it over many workgroups.  Measured on the ring (ms,
  // no partitions -> 64):
  //   2^14: 0.14 -> 0.11,  2^16: 1.25 -> 0.40,  2^18: 6.57 -> 2.49,
  //   2^20: 37.1 -> 22.6 (32 partitions: 23.2, 128: 23.1, 256: 25.1).
  // The FP64 partial sums (24 B per target and partition) are held to 2^26
  // target-partitions, 1.5 GiB.
  LaunchCfg c;
  c.T = VORT_SORTED_T;
  const long long tiles = ((long long)n_src + 255) / 256;
  long long k = std::min<long long>(64, tiles);
  k = std::min<long long>(k, std::max<long long>(1, (1LL << 26) / std::max(n_tgt, 1)));
  if (k <= 1) return c;
  c.split_len = (int)((tiles + k - 1) / k * 256);
  c.n_split = (int)(((long long)n_src + c.split_len - 1) / c.split_len);
  if (c.n_split <= 1) {
    c.n_split = 1;
    c.split_len = 0;
  }
  return c;
}

size_t vort_workspace_bytes(const VortJob& j, const LaunchCfg& c, bool sorted) {
  size_t b = 0;
  if (c.n_split > 1) b += align256((size_t)c.n_split * 3 * (size_t)j.n_tgt * sizeof(double));
  if (sorted) {
    const size_t tiles = ((size_t)j.n_src + 255) / 256;
    b += 2 * align256(sizeof(float4) * (size_t)j.n_src) + 2 * align256(sizeof(float4) * tiles) +
         2 * align256(sizeof(float4) * tiles * 4);
  }
  return b;
}

hipError_t launch_vort(const VortJob& j, const LaunchCfg& c, void* ws, size_t ws_bytes,
                       hipStream_t s) {
  if (j.n_tgt <= 0) return hipSuccess;
  if (j.n_src < 0 || c.wg != 256 || j.reg < REG_SINGULAR || j.reg > REG_GAUSSIAN)
    return hipErrorInvalidValue;
  const bool sorted = j.src_perm != nullptr;
  if (ws_bytes < vort_workspace_bytes(j, c, sorted)) return hipErrorInvalidValue;
  VFn k = vort_pick(j.reg, j.acc, c.T);
  if (!k) return hipErrorInvalidValue;
  VortKParams p{};
  p.src_pv = j.src_pv;
  p.src_a = j.src_a;
  p.tgt = j.tgt;
  p.tgt_perm = j.tgt_perm;
  p.out = j.out;
  p.n_src = j.n_src;
  p.n_tgt = j.n_tgt;
  // the reference's single-source forms (S2S, S2M) have no cut-off and
  // divide by sigma (src/P3D.cpp:146-161); M2S and M2M cut at 5 sigma
  p.cutoff = j.cutoff ? 5.f * j.sigma : __builtin_inff();
  p.rsigma = 1.f / j.sigma;
  p.sigma = j.sigma;
  p.divide = j.cutoff ? 0 : 1;
  const double sa = (double)j.sigma;
  p.rc.s2 = (float)(sa * sa);
  p.rc.inv_s2 = (float)(1.0 / (sa * sa));
  p.rc.c1 = (float)(-0.5 / std::log(2.0));
  p.scale = vort_scale(j.reg, j.sigma);
  char* w = (char*)ws;
  int n_split = c.n_split, split_len = c.split_len;
  if (n_split <= 1 || j.n_src <= 0) {
    n_split = 1;
    split_len = j.n_src > 0 ? j.n_src : 1;
  } else {
    p.partial = (double*)w;
    w += align256((size_t)n_split * 3 * (size_t)j.n_tgt * sizeof(double));
  }
  p.n_split = n_split;
  p.split_len = split_len;
  if (sorted && j.n_src > 0) {
    const int tiles = (j.n_src + 255) / 256;
    float4* spv = (float4*)w;
    w += align256(sizeof(float4) * (size_t)j.n_src);
    float4* sa4 = (float4*)w;
    w += align256(sizeof(float4) * (size_t)j.n_src);
    float4* blo = (float4*)w;
    w += align256(sizeof(float4) * (size_t)tiles);
    float4* bhi = (float4*)w;
    w += align256(sizeof(float4) * (size_t)tiles);
    float4* clo = (float4*)w;
    w += align256(sizeof(float4) * (size_t)tiles * 4);
    float4* chi = (float4*)w;
    hipLaunchKernelGGL(gather_sorted, dim3((j.n_src + 255) / 256), dim3(256), 0, s, j.src_pv,
                       j.src_a, j.src_perm, j.n_src, spv, sa4);
    hipLaunchKernelGGL(tile_boxes, dim3(tiles), dim3(256), 0, s, spv, j.n_src, blo, bhi, clo, chi);
    p.src_pv = spv;
    p.src_a = sa4;
    p.box_lo = blo;
    p.box_hi = bhi;
    p.sub_lo = clo;
    p.sub_hi = chi;
  }
  const int gx = (j.n_tgt + 256 * c.T - 1) / (256 * c.T);
  hipLaunchKernelGGL(k, dim3(gx, n_split), dim3(256), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_split == 1) return e;
  hipLaunchKernelGGL(vort_reduce, dim3((j.n_tgt + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_relax(float4* a, const float* omega, int n, float fdt, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(relax_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, omega, n, fdt);
  return hipGetLastError();
}

}  // namespace cvtx
