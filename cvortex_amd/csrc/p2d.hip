// p2d.hip -- 2D vortex particle direct-sum kernels for gfx950.
//
// Same structure as p3d.hip (targets in registers, LDS source tiles shared by
// the workgroup, TwoSum or plain accumulation, optional source partitions
// with an FP64 reduction).  Sources and targets are float4 {x, y, alpha, area}.
//
// Reduced sums per target (d = x_t - x_s, q = r^2 / sigma^2):
//   vel  = (1/2pi) * scale_K * ( sum K a_s d_y, -sum K a_s d_x )
//   visc = (2 nu / sigma^2) * scale_eta * sum eta (A_t a_s - a_t A_s)
// with K = g_2D(rho)/r^2 in the units below.  The exchange is summed pair by
// pair, as the reference does: its two products nearly cancel for many
// targets, and two separate sums (A_t sum eta a_s and a_t sum eta A_s)
// would carry the rounding of the larger sums into the small difference.
// ACC_COMP sums it in FP64 (each pair's difference of two float products is
// exact there up to one rounding); ACC_TILES and ACC_PLAIN in FP32.  Reference semantics:
// src/P2D.cpp:49-69 (vel pair), :99-162 (M2M), :167-276 (visc);
// src/VortFunc.cpp:78-80,117-131,150-153,191-199 (g_2D, eta_2D).
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"
#include "tile_cull.hpp"

namespace cvtx {

int p2d_nacc(int mask) { return ((mask & OUT_VEL) ? 2 : 0) + ((mask & OUT_VISC) ? 1 : 0); }

template <int REG>
struct Reg2D;

// singular: g = 1 -> K = 1/r^2 (dimensional)
template <>
struct Reg2D<REG_SINGULAR> {
  static __device__ __forceinline__ void eval(float r2, const RegConsts&, float& K, float& eta) {
    float k = __builtin_amdgcn_rcpf(r2);
    K = (r2 == 0.f) ? 0.f : k;
    eta = 0.f;
  }
};

// Winckelmans: a = q + 1, y = 1/a
//   g_2D / r^2 = (q + 2) / ((q + 1)^2 sigma^2) = (y + y^2) / sigma^2
//            -> K = y + y^2  (x 1/sigma^2)
//   eta_2D = 24 exp(4 / a^3) / a^4  (reference VortFunc.cpp:124-131, kept as is)
//            -> eta = exp(4 y^3) y^4  (x 24)
template <>
struct Reg2D<REG_WINCKELMANS> {
  static __device__ __forceinline__ void eval(float r2, const RegConsts& k, float& K, float& eta) {
    float a = __builtin_fmaf(r2, k.inv_s2, 1.f);
    float y = __builtin_amdgcn_rcpf(a);
    K = __builtin_fmaf(y, y, y);  // finite at r2 = 0 and multiplied by d = 0 there
    float y2 = y * y;
    float e = __builtin_amdgcn_exp2f((y2 * y) * k.c1);
    float et = e * (y2 * y2);
    eta = (r2 == 0.f) ? 0.f : et;
  }
};

// planetary: g = rho^2 inside the core, 1 outside, so g / r^2 is 1/sigma^2
// inside and 1/r^2 outside: K = 1 / max(r^2, sigma^2) (dimensional).  The
// host floors sigma^2 at 1e-30, so K stays finite for a coincident pair
// (where it multiplies d = 0) also when sigma = 0.
template <>
struct Reg2D<REG_PLANETARY> {
  static __device__ __forceinline__ void eval(float r2, const RegConsts& k, float& K, float& eta) {
    K = __builtin_amdgcn_rcpf(__builtin_fmaxf(r2, k.s2));
    eta = 0.f;
  }
};

// Gaussian: g = 1 - exp(-q/2) -> K = (1 - exp(-q/2)) / q (x 1/sigma^2),
// eta_2D = exp(-q/2).  Three nested ranges of q, so that a wave runs only
// the ranges its lanes need (a wave of nearby targets is mostly in the first):
//   q >= 208: exp(-q/2) < 2^-150 rounds to 0 in FP32, so K = 1/q and eta = 0
//             (what the general form gives there);
//   q >= 1:   K = (1 - e)/q, eta = e;
//   q < 1:    the series 1/2 sum_{m<9} (-q/2)^m / (m+1)! avoids the
//             cancellation of 1 - e; eta is zeroed at r2 = 0.
// The empty asm keeps the first test a branch: if-converted, every pair
// would pay the exp and the series.
template <>
struct Reg2D<REG_GAUSSIAN> {
  static __device__ __forceinline__ void eval(float r2, const RegConsts& k, float& K, float& eta) {
    const float q = r2 * k.inv_s2;
    const float y = __builtin_amdgcn_rcpf(q);
    K = y;
    eta = 0.f;
    if (q < k.q_far) {
      asm volatile("");
      const float e = __builtin_amdgcn_exp2f(q * k.c1);
      K = (1.f - e) * y;
      eta = e;
      if (q < 1.f) {
        const float x = -0.5f * q;
        float s = 1.f / 362880.f;
        s = __builtin_fmaf(s, x, 1.f / 40320.f);
        s = __builtin_fmaf(s, x, 1.f / 5040.f);
        s = __builtin_fmaf(s, x, 1.f / 720.f);
        s = __builtin_fmaf(s, x, 1.f / 120.f);
        s = __builtin_fmaf(s, x, 1.f / 24.f);
        s = __builtin_fmaf(s, x, 1.f / 6.f);
        s = __builtin_fmaf(s, x, 0.5f);
        s = __builtin_fmaf(s, x, 1.f);
        K = 0.5f * s;
        eta = (r2 == 0.f) ? 0.f : e;
      }
    }
  }
};

struct P2DKParams {
  const float4* __restrict__ src;
  const float4* __restrict__ tgt;
  float* __restrict__ out_vel;
  float* __restrict__ out_visc;
  double* __restrict__ partial;
  int n_src, n_tgt, split_len, n_split;
  RegConsts rc;
  double sc_vel, sc_visc;
};

template <int MASK>
__device__ __forceinline__ void p2d_finish(const double* a, const P2DKParams& p, int i) {
  constexpr int V = 0;
  constexpr int E = (MASK & OUT_VEL) ? 2 : 0;
  if constexpr ((MASK & OUT_VEL) != 0) {
    p.out_vel[2 * (size_t)i + 0] = (float)(p.sc_vel * a[V + 0]);
    p.out_vel[2 * (size_t)i + 1] = (float)(-p.sc_vel * a[V + 1]);
  }
  if constexpr ((MASK & OUT_VISC) != 0) {
    p.out_visc[i] = (float)(p.sc_visc * a[E]);
  }
}

// Far tiles (tile_cull.hpp): for planetary and Gaussian, a pair at
// r^2 >= near2 has K = 1/r^2 in the kernel's units (planetary: r^2 >=
// sigma^2, outside the core; Gaussian: rho^2 >= 208, where exp(-rho^2/2) is
// 0 in FP32 and K = 1/rho^2 = sigma^2/r^2, eta = 0), so a tile whose box is
// at least that far from the block's targets needs no range test: its
// pairs take K from one shared reciprocal per four pairs.
template <int REG>
struct P2DFar {
  static constexpr bool ENABLED = REG == REG_PLANETARY || REG == REG_GAUSSIAN;
  static __device__ __forceinline__ float near2(const RegConsts& k) {
    return REG == REG_PLANETARY ? k.s2 : k.q_far * k.s2;
  }
  // factor of 1/r^2 in K
  static __device__ __forceinline__ float kscale(const RegConsts& k) {
    return REG == REG_PLANETARY ? 1.f : k.s2;
  }
};

template <int REG, int MASK, int ACC, int T, int WG>
__global__ __launch_bounds__(WG) void p2d_kernel(P2DKParams p) {
  constexpr int NA = ((MASK & OUT_VEL) ? 2 : 0) + ((MASK & OUT_VISC) ? 1 : 0);
  constexpr int E = (MASK & OUT_VEL) ? 2 : 0;
  constexpr bool CULL = P2DFar<REG>::ENABLED && (MASK & OUT_VEL) != 0;
  constexpr bool TS = ACC != ACC_PLAIN;                           // tile sums folded by TwoSum
  constexpr bool V64 = ACC == ACC_COMP && (MASK & OUT_VISC) != 0;  // exchange summed in FP64
  constexpr int WAVES = WG / 64;
  __shared__ float4 s_src[WG];
  __shared__ BoxPartials<2, WAVES> s_box;
  const int tid = threadIdx.x;
  const int base = blockIdx.x * (WG * T);
  float4 xt[T];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    int i = base + k * WG + tid;
    i = i < p.n_tgt ? i : p.n_tgt - 1;
    xt[k] = p.tgt[i];
  }
  float acc[T][NA], hi[T][NA], lo[T][NA];
#pragma unroll
  for (int k = 0; k < T; ++k)
#pragma unroll
    for (int a = 0; a < NA; ++a) { acc[k][a] = 0.f; hi[k][a] = 0.f; lo[k][a] = 0.f; }
  double vacc[T], tA[T], ta[T];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    vacc[k] = 0.0;
    tA[k] = (double)xt[k].w;
    ta[k] = (double)xt[k].z;
  }

  // box of the block's targets (the clamped duplicates are real targets)
  Box<2> tbox;
  if constexpr (CULL) {
    float lo2[2] = {xt[0].x, xt[0].y}, hi2[2] = {xt[0].x, xt[0].y};
#pragma unroll
    for (int k = 1; k < T; ++k) {
      lo2[0] = fminf(lo2[0], xt[k].x); lo2[1] = fminf(lo2[1], xt[k].y);
      hi2[0] = fmaxf(hi2[0], xt[k].x); hi2[1] = fmaxf(hi2[1], xt[k].y);
    }
    bool fin = true;
#pragma unroll
    for (int k = 0; k < T; ++k) fin = fin && __builtin_isfinite(xt[k].x) && __builtin_isfinite(xt[k].y);
    tbox = block_box<2, WAVES>(lo2, hi2, fin, s_box);
  }

  const int s0 = blockIdx.y * p.split_len;
  const int s1 = min(s0 + p.split_len, p.n_src);
  const RegConsts rv = loop_consts(p.rc);
  for (int ts = s0; ts < s1; ts += WG) {
    const int cnt = min(WG, s1 - ts);
    __syncthreads();
    float4 src = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid < cnt) {
      src = p.src[ts + tid];
      s_src[tid] = src;
    }
    if constexpr (CULL) {
      const float x2[2] = {src.x, src.y};
      box_partial<2, WAVES>(x2, tid < cnt, s_box);
    }
    __syncthreads();
    if constexpr (TS) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < NA; ++a) acc[k][a] = 0.f;
    }
    int jv = in_vgpr(0);  // LDS index in a VGPR (no SGPR operand per source)
    bool far = false;
    if constexpr (CULL) far = boxes_far<2>(tbox, box_combine<2, WAVES>(s_box), P2DFar<REG>::near2(p.rc));
    if (far) {
      if constexpr (CULL) {
        // U sources x T targets = 4 pairs per shared reciprocal; eta = 0
        constexpr int U = 4 / T;
        const float ks = in_vgpr(P2DFar<REG>::kscale(p.rc));
        // a_s times the factor of 1/r^2 in K (planetary: 1, so no multiply)
        auto sk = [=](float a) {
          if constexpr (REG == REG_PLANETARY) return a;
          else return a * ks;
        };
        int j = 0;
        for (; j + U <= cnt; j += U, jv += U) {
          float4 s[U];
#pragma unroll
          for (int u = 0; u < U; ++u) s[u] = s_src[jv + u];
          float dx[4], dy[4], r2[4], K[4];
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < T; ++k) {
              const int q = u * T + k;
              dx[q] = xt[k].x - s[u].x;
              dy[q] = xt[k].y - s[u].y;
              r2[q] = __builtin_fmaf(dx[q], dx[q], dy[q] * dy[q]);
            }
          // K[q] a_s for the four pairs
          if constexpr (T == 1) {
            shared_rcp4(r2, K);
#pragma unroll
            for (int u = 0; u < U; ++u) K[u] *= sk(s[u].z);
          } else {
            // pairs 0, 1 and 2, 3 share a source: its strength scales the
            // shared reciprocal once
            shared_rcp4_scaled(r2, sk(s[0].z), sk(s[U - 1].z), K);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int k = q % T;
            acc[k][0] = __builtin_fmaf(K[q], dy[q], acc[k][0]);
            acc[k][1] = __builtin_fmaf(K[q], dx[q], acc[k][1]);
          }
        }
        for (; j < cnt; ++j, ++jv) {
          const float4 s = s_src[jv];
          const float as = sk(s.z);
#pragma unroll
          for (int k = 0; k < T; ++k) {
            const float dx = xt[k].x - s.x;
            const float dy = xt[k].y - s.y;
            const float ka = __builtin_amdgcn_rcpf(__builtin_fmaf(dx, dx, dy * dy)) * as;
            acc[k][0] = __builtin_fmaf(ka, dy, acc[k][0]);
            acc[k][1] = __builtin_fmaf(ka, dx, acc[k][1]);
          }
        }
      }
    } else {
      // (the Gaussian's pinned branch cannot be unrolled around)
      constexpr int kUnroll = REG == REG_GAUSSIAN ? 1 : 4;
#pragma unroll kUnroll
      for (int j = 0; j < cnt; ++j, ++jv) {
        const float4 s = s_src[jv];
#pragma unroll
        for (int k = 0; k < T; ++k) {
          const float dx = xt[k].x - s.x;
          const float dy = xt[k].y - s.y;
          const float r2 = __builtin_fmaf(dx, dx, dy * dy);
          float K, eta;
          Reg2D<REG>::eval(r2, rv, K, eta);
          if constexpr ((MASK & OUT_VEL) != 0) {
            const float ka = K * s.z;
            acc[k][0] = __builtin_fmaf(ka, dy, acc[k][0]);
            acc[k][1] = __builtin_fmaf(ka, dx, acc[k][1]);
          }
          if constexpr (V64) {
            const double d = __builtin_fma(tA[k], (double)s.z, -(ta[k] * (double)s.w));
            vacc[k] = __builtin_fma(d, (double)eta, vacc[k]);
          } else if constexpr ((MASK & OUT_VISC) != 0) {
            const float d = __builtin_fmaf(xt[k].w, s.z, -(xt[k].z * s.w));
            acc[k][E] = __builtin_fmaf(d, eta, acc[k][E]);
          }
        }
      }
    }
    if constexpr (TS) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < (V64 ? E : NA); ++a) two_sum_add(hi[k][a], lo[k][a], acc[k][a]);
    }
  }

#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * WG + tid;
    if (i >= p.n_tgt) continue;
    double d[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a)
      d[a] = TS ? (double)hi[k][a] + (double)lo[k][a] : (double)acc[k][a];
    if constexpr (V64) d[E] = vacc[k];
    if (p.n_split == 1) {
      p2d_finish<MASK>(d, p, i);
    } else {
#pragma unroll
      for (int a = 0; a < NA; ++a) p.partial[((size_t)blockIdx.y * NA + a) * p.n_tgt + i] = d[a];
    }
  }
}

template <int MASK>
__global__ __launch_bounds__(256) void p2d_reduce(P2DKParams p) {
  constexpr int NA = ((MASK & OUT_VEL) ? 2 : 0) + ((MASK & OUT_VISC) ? 1 : 0);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n_tgt) return;
  double d[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) d[a] = 0.0;
#pragma unroll 4
  for (int s = 0; s < p.n_split; ++s)
#pragma unroll
    for (int a = 0; a < NA; ++a) d[a] += p.partial[((size_t)s * NA + a) * p.n_tgt + i];
  p2d_finish<MASK>(d, p, i);
}

bool p2d_supported(int reg, int mask) {
  if (reg < REG_SINGULAR || reg > REG_GAUSSIAN) return false;
  if (mask != OUT_VEL && mask != OUT_VISC && mask != (OUT_VEL | OUT_VISC)) return false;
  if ((mask & OUT_VISC) && (reg == REG_SINGULAR || reg == REG_PLANETARY)) return false;
  return true;
}

size_t p2d_workspace_bytes(const P2DJob& j, const LaunchCfg& c) {
  if (c.n_split <= 1) return 0;
  return (size_t)c.n_split * p2d_nacc(j.mask) * (size_t)j.n_tgt * sizeof(double);
}

namespace {
using K2Fn = void (*)(P2DKParams);

template <int REG, int MASK, int ACC>
K2Fn p2_T(int T) {
  switch (T) {
    case 1: return p2d_kernel<REG, MASK, ACC, 1, 256>;
    case 4: return p2d_kernel<REG, MASK, ACC, 4, 256>;
    default: return p2d_kernel<REG, MASK, ACC, 2, 256>;
  }
}
// ACC_TILES differs from ACC_COMP only in the viscous exchange's sum
template <int REG, int MASK>
K2Fn p2_acc(int acc, int T) {
  if (acc == ACC_PLAIN) return p2_T<REG, MASK, ACC_PLAIN>(T);
  if constexpr ((MASK & OUT_VISC) != 0)
    if (acc == ACC_TILES) return p2_T<REG, MASK, ACC_TILES>(T);
  return p2_T<REG, MASK, ACC_COMP>(T);
}
template <int REG>
K2Fn p2_mask_inviscid(int mask, int acc, int T) {
  return mask == OUT_VEL ? p2_acc<REG, OUT_VEL>(acc, T) : nullptr;
}
template <int REG>
K2Fn p2_mask_viscous(int mask, int acc, int T) {
  switch (mask) {
    case OUT_VEL: return p2_acc<REG, OUT_VEL>(acc, T);
    case OUT_VISC: return p2_acc<REG, OUT_VISC>(acc, T);
    case OUT_VEL | OUT_VISC: return p2_acc<REG, OUT_VEL | OUT_VISC>(acc, T);
    default: return nullptr;
  }
}
K2Fn p2_pick(int reg, int mask, int acc, int T) {
  switch (reg) {
    case REG_SINGULAR: return p2_mask_inviscid<REG_SINGULAR>(mask, acc, T);
    case REG_PLANETARY: return p2_mask_inviscid<REG_PLANETARY>(mask, acc, T);
    case REG_WINCKELMANS: return p2_mask_viscous<REG_WINCKELMANS>(mask, acc, T);
    case REG_GAUSSIAN: return p2_mask_viscous<REG_GAUSSIAN>(mask, acc, T);
    default: return nullptr;
  }
}
K2Fn p2_reduce(int mask) {
  switch (mask) {
    case OUT_VEL: return p2d_reduce<OUT_VEL>;
    case OUT_VISC: return p2d_reduce<OUT_VISC>;
    case OUT_VEL | OUT_VISC: return p2d_reduce<OUT_VEL | OUT_VISC>;
    default: return nullptr;
  }
}
}  // namespace

hipError_t launch_p2d(const P2DJob& j, const LaunchCfg& c, hipStream_t s) {
  if (j.n_tgt <= 0) return hipSuccess;
  if (!p2d_supported(j.reg, j.mask) || j.n_src < 0 || c.wg != 256) return hipErrorInvalidValue;
  // every output the mask selects must exist, and the inputs it reads
  if (((j.mask & OUT_VEL) && !j.out_vel) || ((j.mask & OUT_VISC) && !j.out_visc) || !j.tgt ||
      (j.n_src > 0 && !j.src))
    return hipErrorInvalidValue;
  K2Fn k = p2_pick(j.reg, j.mask, j.acc, c.T);
  if (!k) return hipErrorInvalidValue;
  P2DKParams p{};
  p.src = j.src;
  p.tgt = j.tgt;
  p.out_vel = j.out_vel;
  p.out_visc = j.out_visc;
  p.partial = j.partial;
  p.n_src = j.n_src;
  p.n_tgt = j.n_tgt;
  const double sa = std::fabs((double)j.sigma);
  p.rc.s2 = (float)(sa * sa);
  p.rc.inv_s2 = (float)(1.0 / (sa * sa));
  const double k2pi = 1.0 / (2.0 * M_PI);
  switch (j.reg) {
    case REG_SINGULAR:
      p.sc_vel = k2pi;
      break;
    case REG_WINCKELMANS:
      p.rc.c1 = (float)(4.0 / std::log(2.0));
      p.sc_vel = k2pi / (sa * sa);
      p.sc_visc = 2.0 * (double)j.nu / (sa * sa) * 24.0;
      break;
    case REG_PLANETARY:
      p.rc.s2 = std::fmax((float)(sa * sa), 1e-30f);
      p.sc_vel = k2pi;
      break;
    default:
      p.rc.c1 = (float)(-0.5 / std::log(2.0));
      p.sc_vel = k2pi / (sa * sa);
      p.sc_visc = 2.0 * (double)j.nu / (sa * sa);
      break;
  }
  int n_split = c.n_split, split_len = c.split_len;
  if (n_split <= 1 || j.n_src == 0) {
    n_split = 1;
    split_len = j.n_src > 0 ? j.n_src : 1;
  } else {
    if (split_len <= 0 || split_len % c.wg != 0) return hipErrorInvalidValue;
    if ((long long)split_len * n_split < j.n_src || j.partial == nullptr) return hipErrorInvalidValue;
  }
  p.n_split = n_split;
  p.split_len = split_len;
  const int per = c.wg * c.T;
  hipLaunchKernelGGL(k, dim3((j.n_tgt + per - 1) / per, n_split), dim3(c.wg), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_split == 1) return e;
  hipLaunchKernelGGL(p2_reduce(j.mask), dim3((j.n_tgt + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace cvtx
