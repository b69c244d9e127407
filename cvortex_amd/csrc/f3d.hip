// f3d.hip -- straight singular vortex filament kernels for gfx950.
//
// Filaments stream through LDS in tiles; when a tile is staged each
// filament is expanded once per workgroup into
//   { a (start), Gamma/(4pi) },  { b (end), 3 Gamma/(4pi |r0|) },  { r0 = b - a },
// so the per-pair work only involves the target-dependent quantities.
//
// Per pair (x the target): r1 = x - a, r2 = x - b, c = r1 x r2 (= r0 x r1),
// P = |r1||r2|, s = |r1| + |r2|, d = r1.r2.  The reference's factors
// (F3D.cpp:34-85)
//   t2 = r0.r1/|r1| - r0.r2/|r2|            and   1/|r1| - 1/|r2|
// cancel in FP32 for targets far from the segment or near its end points,
// so they are evaluated in the algebraically identical forms
//   k = (Gamma/4pi) t2 / |c|^2 = (Gamma/4pi) s / (P (P + d))          (d >= 0)
//                              = (Gamma/4pi) s (P - d) / (P |c|^2)    (d < 0)
//   1/|r1| - 1/|r2| = -r0.(r1 + r2) / (P s)
// (using |c|^2 = P^2 - d^2 and |r2|^2 - |r1|^2 = -r0.(r1 + r2)).  Then
//   vel:   v += k c
//   dvort: B += 3 Gamma/(4pi|r0|) |c| (1/|r1| - 1/|r2|),  S += k r0,
//          dvort = a_t B - S x a_t
// A pair whose terms are not finite (target on a filament's line or an end
// point) contributes zero, as in the reference's GPU kernels
// (nbody.cl:637,718); on the line's extension the velocity term is exactly
// zero there (c = 0) and the stretching term is dropped.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"
#include "tile_cull.hpp"

namespace cvtx {

int f3d_nacc(int mask) { return ((mask & OUT_VEL) ? 3 : 0) + ((mask & OUT_DVORT) ? 4 : 0); }

struct F3DKParams {
  const float4* __restrict__ fa;
  const float4* __restrict__ fb;
  const float4* __restrict__ tgt_pv;
  const float4* __restrict__ tgt_a;
  float* __restrict__ out_vel;
  float* __restrict__ out_dvort;
  double* __restrict__ partial;
  const int* __restrict__ tgt_perm;  // slot -> target (nullptr: identity)
  int n_fil, n_tgt, split_len, n_split;
};

template <int MASK>
__device__ __forceinline__ void f3d_finish(const double* a, float3 at, const F3DKParams& p, int i) {
  constexpr int D = (MASK & OUT_VEL) ? 3 : 0;
  if constexpr ((MASK & OUT_VEL) != 0) {
    p.out_vel[3 * (size_t)i + 0] = (float)a[0];
    p.out_vel[3 * (size_t)i + 1] = (float)a[1];
    p.out_vel[3 * (size_t)i + 2] = (float)a[2];
  }
  if constexpr ((MASK & OUT_DVORT) != 0) {
    const double B = a[D], Sx = a[D + 1], Sy = a[D + 2], Sz = a[D + 3];
    const double ax = at.x, ay = at.y, az = at.z;
    p.out_dvort[3 * (size_t)i + 0] = (float)(ax * B - (Sy * az - Sz * ay));
    p.out_dvort[3 * (size_t)i + 1] = (float)(ay * B - (Sz * ax - Sx * az));
    p.out_dvort[3 * (size_t)i + 2] = (float)(az * B - (Sx * ay - Sy * ax));
  }
}

// One pair by the general formula (any target position): accumulates into
// acc[0..NA) of one target.  sa = {a, Gamma/(4pi)}, sb = {b, 3 Gamma/(4pi|r0|)},
// f0 = {r0}.
template <int MASK>
__device__ __forceinline__ void f3d_pair(const float4 fa, const float4 fb, const float4 f0,
                                         const float3 xt, const float fin, float* acc) {
  constexpr int D = (MASK & OUT_VEL) ? 3 : 0;
  const float r1x = xt.x - fa.x, r1y = xt.y - fa.y, r1z = xt.z - fa.z;
  const float r2x = xt.x - fb.x, r2y = xt.y - fb.y, r2z = xt.z - fb.z;
  // c = r1 x r2 (= r0 x r1); for targets beside the segment its
  // products are smaller than those of r0 x r1, so it loses less to
  // cancellation near the filament
  const float cx = __builtin_fmaf(r1y, r2z, -r1z * r2y);
  const float cy = __builtin_fmaf(r1z, r2x, -r1x * r2z);
  const float cz = __builtin_fmaf(r1x, r2y, -r1y * r2x);
  const float cc = __builtin_fmaf(cx, cx, __builtin_fmaf(cy, cy, cz * cz));
  const float n1s = __builtin_fmaf(r1x, r1x, __builtin_fmaf(r1y, r1y, r1z * r1z));
  const float n2s = __builtin_fmaf(r2x, r2x, __builtin_fmaf(r2y, r2y, r2z * r2z));
  const float n1 = n1s * __builtin_amdgcn_rsqf(n1s);
  const float n2 = n2s * __builtin_amdgcn_rsqf(n2s);
  const float P = n1 * n2;
  const float sum = n1 + n2;
  const float d12 = __builtin_fmaf(r1x, r2x, __builtin_fmaf(r1y, r2y, r1z * r2z));
  const bool same_side = d12 >= 0.f;
  const float den = same_side ? P * (P + d12) : P * cc;
  const float num = same_side ? sum : sum * (P - d12);
  float kk = fa.w * num * __builtin_amdgcn_rcpf(den);
  if constexpr ((MASK & OUT_DVORT) != 0) {
    const float rcc = __builtin_amdgcn_rsqf(cc);
    const float cn = cc * rcc;  // |c|; 0 * inf = NaN when c = 0
    const float p12 = __builtin_fmaf(f0.x, r1x + r2x, __builtin_fmaf(f0.y, r1y + r2y, f0.z * (r1z + r2z)));
    float B = -fb.w * cn * p12 * __builtin_amdgcn_rcpf(P * sum);
    // B is NaN whenever c = 0 (target on the filament's line), so the
    // finiteness of B also drops those pairs
    const bool ok = (__builtin_fabsf(kk) < __builtin_inff()) &&
                    (__builtin_fabsf(B) < __builtin_inff());
    B = ok ? B : 0.f;
    const float kd = ok ? kk : 0.f;
    if constexpr ((MASK & OUT_VEL) != 0) {
      kk = (__builtin_fabsf(kk) < __builtin_inff()) ? kk : 0.f;
      acc[0] = __builtin_fmaf(kk, cx, acc[0]);
      acc[1] = __builtin_fmaf(kk, cy, acc[1]);
      acc[2] = __builtin_fmaf(kk, cz, acc[2]);
    }
    acc[D + 0] += B;
    acc[D + 1] = __builtin_fmaf(kd, f0.x, acc[D + 1]);
    acc[D + 2] = __builtin_fmaf(kd, f0.y, acc[D + 2]);
    acc[D + 3] = __builtin_fmaf(kd, f0.z, acc[D + 3]);
  } else {
    kk = (__builtin_fabsf(kk) < fin) ? kk : 0.f;
    acc[0] = __builtin_fmaf(kk, cx, acc[0]);
    acc[1] = __builtin_fmaf(kk, cy, acc[1]);
    acc[2] = __builtin_fmaf(kk, cz, acc[2]);
  }
}

// Far tiles (tile_cull.hpp).  When every target of the block is at least
// half the longest filament of the tile away from the tile's box, every
// pair has r1.r2 = |x - m|^2 - |r0|^2/4 >= 0 (m the filament's midpoint),
// so the same-side form holds for all of them, with y_i = 1/|r_i|:
//   k = (Gamma/4pi) (|r1| + |r2|) / (|r1||r2| (|r1||r2| + r1.r2))
//     = (Gamma/4pi) (y1 + y2) / (P + r1.r2),           P = |r1||r2|,
//   1/|r1| - 1/|r2| = -r0.(r1 + r2) y1 y2 / (|r1| + |r2|),
// and no pair's terms can be infinite.  The reciprocals of P + r1.r2 (and,
// for the stretching, of |r1| + |r2|) of the group's pairs come from one
// shared reciprocal (batch_rcp).  The stretching terms of a pair with c = 0
// (a target on the extension of a filament's line) are dropped, as the
// general form does (there |c| = 0 and the reference's terms are NaN).
template <int MASK, int T>
struct F3DFar {
  // pairs per group and reciprocals per pair
  static constexpr int G = (MASK == OUT_VEL) ? 4 : 2;
  static constexpr int U = G >= T ? G / T : 1;  // filaments per group
  static constexpr int NP = U * T;              // pairs per group
  static constexpr int NR = (MASK & OUT_DVORT) ? 2 : 1;
  static constexpr int N = NP * NR;             // values per shared reciprocal
};

// Waves per SIMD the register allocation must leave room for.  The fused
// kernel with one target per lane is held to 5 (at most 96 VGPRs; it takes
// 103 and 4 waves unbounded, and spills 16 bytes held to 5): 16384
// filaments x 2^20 targets, 21.29 -> 19.69 ms.  The velocity alone is held
// to 6 (80 VGPRs instead of 89; 20 bytes spilled outside the source loops):
// 10.37 -> 10.20 ms (three A/B pairs on one box); held to 7 it was slower.
// The stretching alone runs best unbounded (5 waves; 17.24 ms against 17.73
// held to 6).
template <int MASK, int T>
constexpr int f3d_min_waves() {
  if constexpr (T != 1) return 1;
  return MASK == (OUT_VEL | OUT_DVORT) ? 5 : (MASK == OUT_VEL ? 6 : 1);
}

template <int MASK, int ACC, int T, int WG>
__global__ __launch_bounds__(WG, (f3d_min_waves<MASK, T>())) void f3d_kernel(F3DKParams p) {
  constexpr int NA = ((MASK & OUT_VEL) ? 3 : 0) + ((MASK & OUT_DVORT) ? 4 : 0);
  constexpr int D = (MASK & OUT_VEL) ? 3 : 0;
  constexpr int WAVES = WG / 64;
  using FF = F3DFar<MASK, T>;
  __shared__ float4 s_a[WG];
  __shared__ float4 s_b[WG];
  __shared__ float4 s_r0[WG];
  __shared__ BoxPartials<3, WAVES> s_box;
  const int tid = threadIdx.x;
  const int base = blockIdx.x * (WG * T);
  float3 xt[T], at[T];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    int i = base + k * WG + tid;
    i = i < p.n_tgt ? i : p.n_tgt - 1;
    if (p.tgt_perm) i = p.tgt_perm[i];
    const float4 q = p.tgt_pv[i];
    xt[k] = make_float3(q.x, q.y, q.z);
    if constexpr ((MASK & OUT_DVORT) != 0) {
      const float4 w = p.tgt_a[i];
      at[k] = make_float3(w.x, w.y, w.z);
    } else {
      at[k] = make_float3(0.f, 0.f, 0.f);
    }
  }
  float acc[T][NA], hi[T][NA], lo[T][NA];
#pragma unroll
  for (int k = 0; k < T; ++k)
#pragma unroll
    for (int a = 0; a < NA; ++a) { acc[k][a] = 0.f; hi[k][a] = 0.f; lo[k][a] = 0.f; }

  // box of the block's targets (the clamped duplicates are real targets)
  Box<3> tbox;
  {
    float l3[3] = {xt[0].x, xt[0].y, xt[0].z}, h3[3] = {xt[0].x, xt[0].y, xt[0].z};
    bool fin = true;
#pragma unroll
    for (int k = 0; k < T; ++k) {
      l3[0] = fminf(l3[0], xt[k].x); l3[1] = fminf(l3[1], xt[k].y); l3[2] = fminf(l3[2], xt[k].z);
      h3[0] = fmaxf(h3[0], xt[k].x); h3[1] = fmaxf(h3[1], xt[k].y); h3[2] = fmaxf(h3[2], xt[k].z);
      fin = fin && __builtin_isfinite(xt[k].x) && __builtin_isfinite(xt[k].y) &&
            __builtin_isfinite(xt[k].z);
    }
    tbox = block_box<3, WAVES>(l3, h3, fin, s_box);
  }

  const float k4pi = (float)(1.0 / (4.0 * M_PI));
  const int s0 = blockIdx.y * p.split_len;
  const int s1 = min(s0 + p.split_len, p.n_fil);
  for (int ts = s0; ts < s1; ts += WG) {
    const int cnt = min(WG, s1 - ts);
    __syncthreads();
    {
      float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
      if (tid < cnt) {
        a = p.fa[ts + tid];
        b = p.fb[ts + tid];
        const float rx = b.x - a.x, ry = b.y - a.y, rz = b.z - a.z;
        const float g = a.w * k4pi;
        const float len = __builtin_sqrtf(rx * rx + ry * ry + rz * rz);
        // link: this filament ends where the next one of the tile starts,
        // bit for bit (a polyline), so the two can share that point's
        // distance terms in the far loop
        int link = 0;
        if (tid + 1 < cnt) {
          const float4 an = p.fa[ts + tid + 1];
          link = (int)(__float_as_uint(an.x) == __float_as_uint(b.x)) &
                 (int)(__float_as_uint(an.y) == __float_as_uint(b.y)) &
                 (int)(__float_as_uint(an.z) == __float_as_uint(b.z));
        }
        s_a[tid] = make_float4(a.x, a.y, a.z, g);
        s_b[tid] = make_float4(b.x, b.y, b.z, 3.f * g / len);
        s_r0[tid] = make_float4(rx, ry, rz, __int_as_float(link));
      }
      // the tile's box holds both ends of every filament; extra: the
      // longest filament's squared length
      const float l3[3] = {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)};
      const float h3[3] = {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)};
      const float rx = b.x - a.x, ry = b.y - a.y, rz = b.z - a.z;
      const bool fin = __builtin_isfinite(a.x) && __builtin_isfinite(a.y) &&
                       __builtin_isfinite(a.z) && __builtin_isfinite(b.x) &&
                       __builtin_isfinite(b.y) && __builtin_isfinite(b.z) &&
                       __builtin_isfinite(a.w);
      box_partial<3, WAVES>(l3, h3, rx * rx + ry * ry + rz * rz, fin, tid < cnt, s_box);
    }
    __syncthreads();
    if constexpr (ACC == ACC_COMP) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < NA; ++a) acc[k][a] = 0.f;
    }
    int jv = in_vgpr(0);  // LDS index in a VGPR (no SGPR operand per filament)
    // velocity only: +inf in a VGPR, so |kk| < fin is a VGPR compare (against
    // the literal the compiler emits a class test whose mask operand is an
    // SGPR).  Measured on 16384 filaments x 2^20 targets: vel 16.79 -> 16.54
    // ms; the dvort and fused kernels were slower with it (23.06 -> 23.27,
    // 24.46 -> 24.91) and keep the literal.
    const float fin = [] {
      if constexpr ((MASK & OUT_DVORT) == 0) return in_vgpr(__builtin_inff());
      else return __builtin_inff();
    }();
    const float c126 = in_vgpr(0x1p126f);
    float l2max;
    const Box<3> fbox = box_combine<3, WAVES>(s_box, &l2max);
    // every pair same-side and finite, and each of the N values that share
    // a reciprocal (P + r1.r2 in [dmin^2, 2 dmax^2], |r1| + |r2| in
    // [2 dmin, 2 dmax]) within 2^(+-120/N), so their products stay normal
    bool far = false;
    if (tbox.finite && fbox.finite) {
      float dmin2, dmax2;
      box_dist2<3>(tbox, fbox, dmin2, dmax2);
      const float lo_ok = FF::N <= 4 ? 9.3e-10f : 3.1e-5f;   // 2^-30, 2^-15
      const float hi_ok = FF::N <= 4 ? 5.3e8f : 1.6e4f;      // 2^29, 2^14
      far = dmin2 >= 0.2501f * l2max && dmin2 >= lo_ok && dmax2 <= hi_ok;
    }
    if (far) {
      constexpr int U = FF::U;
      int j = 0;
      // the last end point of a group, kept for the next group when the
      // chain runs on into it (its first filament starts there, bit for bit,
      // so the values are those the next group would compute).  With the
      // stretching (two filaments per group) this saves a quarter of the end
      // points' rsq: lattice dvort 18.67 -> 17.94 ms, fused 19.74 -> 19.22;
      // both paths leave the point, which keeps the register allocator from
      // copying it around the branch (dvort 18.07 -> 17.23, fused 19.14 ->
      // 18.62), and then the velocity alone (four filaments per group, a
      // fifth of its rsq saved) gains too: 10.87 -> 10.58 ms.
      float lpx[T], lpy[T], lpz[T], lpns[T], lpy1[T];
      bool carry = false;
      for (; j + U <= cnt; j += U, jv += U) {
        float4 fa[U], fb[U], f0[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          fa[u] = s_a[jv + u];
          fb[u] = s_b[jv + u];
          f0[u] = s_r0[jv + u];
        }
        float cx[FF::NP], cy[FF::NP], cz[FF::NP], ys[FF::NP], y12[FF::NP], p12[FF::NP],
            cn[FF::NP], den[FF::N], inv[FF::N];
        // the terms of pair q (filament u, target k) from r1 = x - a,
        // r2 = x - b, their squared norms and reciprocal norms
        auto terms = [&](int q, int u, float r1x, float r1y, float r1z, float r2x, float r2y,
                         float r2z, float n1s, float n2s, float y1, float y2) {
          cx[q] = __builtin_fmaf(r1y, r2z, -r1z * r2y);
          cy[q] = __builtin_fmaf(r1z, r2x, -r1x * r2z);
          cz[q] = __builtin_fmaf(r1x, r2y, -r1y * r2x);
          const float n1 = n1s * y1, n2 = n2s * y2;
          const float d12 = __builtin_fmaf(r1x, r2x, __builtin_fmaf(r1y, r2y, r1z * r2z));
          ys[q] = y1 + y2;
          den[q] = __builtin_fmaf(n1, n2, d12);  // P + r1.r2
          if constexpr ((MASK & OUT_DVORT) != 0) {
            y12[q] = y1 * y2;
            den[FF::NP + q] = n1 + n2;
            const float cc = __builtin_fmaf(cx[q], cx[q], __builtin_fmaf(cy[q], cy[q], cz[q] * cz[q]));
            cn[q] = __builtin_amdgcn_sqrtf(cc);
            p12[q] = __builtin_fmaf(f0[u].x, r1x + r2x,
                                    __builtin_fmaf(f0[u].y, r1y + r2y, f0[u].z * (r1z + r2z)));
          }
        };
        // a chained group (each filament ends where the next begins) has
        // U + 1 distinct end points: each point's distance and rsq serve
        // the two filaments that meet there (same formulas as below).
        // Lattice of 16384 filaments in chains of 128 x 2^20 targets:
        // vel 13.22 -> 10.65 ms, dvort 19.23 -> 18.67, fused 20.41 -> 19.69
        bool chained = false;
        if constexpr (U > 1) {
          int f = __float_as_int(f0[0].w);
#pragma unroll
          for (int u = 1; u + 1 < U; ++u) f &= __float_as_int(f0[u].w);
          chained = __builtin_amdgcn_readfirstlane(f) != 0;
        }
        if (chained) {
          if (!carry) {
#pragma unroll
            for (int k = 0; k < T; ++k) {
              lpx[k] = xt[k].x - fa[0].x;
              lpy[k] = xt[k].y - fa[0].y;
              lpz[k] = xt[k].z - fa[0].z;
              lpns[k] = __builtin_fmaf(lpx[k], lpx[k], __builtin_fmaf(lpy[k], lpy[k], lpz[k] * lpz[k]));
              lpy1[k] = __builtin_amdgcn_rsqf(lpns[k]);
            }
          }
#pragma unroll
          for (int k = 0; k < T; ++k) {
            // walk the points: each filament's terms as soon as its end is known
            float px = lpx[k], py = lpy[k], pz = lpz[k], pns = lpns[k], py1 = lpy1[k];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const float4 pt = u + 1 < U ? fa[u + 1] : fb[U - 1];
              const float qx = xt[k].x - pt.x, qy = xt[k].y - pt.y, qz = xt[k].z - pt.z;
              const float qns = __builtin_fmaf(qx, qx, __builtin_fmaf(qy, qy, qz * qz));
              const float qy1 = __builtin_amdgcn_rsqf(qns);
              terms(u * T + k, u, px, py, pz, qx, qy, qz, pns, qns, py1, qy1);
              px = qx; py = qy; pz = qz; pns = qns; py1 = qy1;
            }
            lpx[k] = px; lpy[k] = py; lpz[k] = pz; lpns[k] = pns; lpy1[k] = py1;
          }
          carry = __builtin_amdgcn_readfirstlane(__float_as_int(f0[U - 1].w)) != 0;
        } else {
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < T; ++k) {
              const float r1x = xt[k].x - fa[u].x, r1y = xt[k].y - fa[u].y, r1z = xt[k].z - fa[u].z;
              const float r2x = xt[k].x - fb[u].x, r2y = xt[k].y - fb[u].y, r2z = xt[k].z - fb[u].z;
              const float n1s = __builtin_fmaf(r1x, r1x, __builtin_fmaf(r1y, r1y, r1z * r1z));
              const float n2s = __builtin_fmaf(r2x, r2x, __builtin_fmaf(r2y, r2y, r2z * r2z));
              const float y2 = __builtin_amdgcn_rsqf(n2s);
              terms(u * T + k, u, r1x, r1y, r1z, r2x, r2y, r2z, n1s, n2s,
                    __builtin_amdgcn_rsqf(n1s), y2);
              if (u == U - 1) {
                // the group's last end point, as the chained walk leaves it
                lpx[k] = r2x; lpy[k] = r2y; lpz[k] = r2z; lpns[k] = n2s; lpy1[k] = y2;
              }
            }
          carry = __builtin_amdgcn_readfirstlane(__float_as_int(f0[U - 1].w)) != 0;
        }
        batch_rcp<FF::N>(den, inv);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int k = 0; k < T; ++k) {
            const int q = u * T + k;
            const float kk = fa[u].w * ys[q] * inv[q];
            if constexpr ((MASK & OUT_VEL) != 0) {
              acc[k][0] = __builtin_fmaf(kk, cx[q], acc[k][0]);
              acc[k][1] = __builtin_fmaf(kk, cy[q], acc[k][1]);
              acc[k][2] = __builtin_fmaf(kk, cz[q], acc[k][2]);
            }
            if constexpr ((MASK & OUT_DVORT) != 0) {
              const float B = -fb[u].w * cn[q] * p12[q] * y12[q] * inv[FF::NP + q];
              // kk, or 0 when c = 0: |c| is 0 or at least 2^-75 (the root
              // of a float), so |c| 2^126 clamped to [0, 1] is 0 or 1 (one
              // multiplication with the clamp modifier, where a compare does
              // not dual-issue): lattice fused 19.18 -> 19.03 ms
              // (kk is finite in far tiles, so kk * 0 is 0)
              const float kd = kk * __builtin_amdgcn_fmed3f(cn[q] * c126, 0.f, 1.f);
              acc[k][D + 0] += B;
              acc[k][D + 1] = __builtin_fmaf(kd, f0[u].x, acc[k][D + 1]);
              acc[k][D + 2] = __builtin_fmaf(kd, f0[u].y, acc[k][D + 2]);
              acc[k][D + 3] = __builtin_fmaf(kd, f0[u].z, acc[k][D + 3]);
            }
          }
      }
      for (; j < cnt; ++j, ++jv) {
        const float4 fa = s_a[jv], fb = s_b[jv], f0 = s_r0[jv];
#pragma unroll
        for (int k = 0; k < T; ++k) f3d_pair<MASK>(fa, fb, f0, xt[k], fin, acc[k]);
      }
    } else {
#pragma unroll 2
      for (int j = 0; j < cnt; ++j, ++jv) {
        const float4 fa = s_a[jv];
        const float4 fb = s_b[jv];
        const float4 f0 = s_r0[jv];
#pragma unroll
        for (int k = 0; k < T; ++k) f3d_pair<MASK>(fa, fb, f0, xt[k], fin, acc[k]);
      }
    }
    if constexpr (ACC == ACC_COMP) {
#pragma unroll
      for (int k = 0; k < T; ++k)
#pragma unroll
        for (int a = 0; a < NA; ++a) two_sum_add(hi[k][a], lo[k][a], acc[k][a]);
    }
  }

#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int i = base + k * WG + tid;
    if (i >= p.n_tgt) continue;
    double d[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a)
      d[a] = (ACC == ACC_COMP) ? (double)hi[k][a] + (double)lo[k][a] : (double)acc[k][a];
    if (p.n_split == 1) {
      f3d_finish<MASK>(d, at[k], p, p.tgt_perm ? p.tgt_perm[i] : i);
    } else {
#pragma unroll
      for (int a = 0; a < NA; ++a) p.partial[((size_t)blockIdx.y * NA + a) * p.n_tgt + i] = d[a];
    }
  }
}

template <int MASK>
__global__ __launch_bounds__(256) void f3d_reduce(F3DKParams p) {
  constexpr int NA = ((MASK & OUT_VEL) ? 3 : 0) + ((MASK & OUT_DVORT) ? 4 : 0);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n_tgt) return;
  double d[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) d[a] = 0.0;
#pragma unroll 4
  for (int s = 0; s < p.n_split; ++s)
#pragma unroll
    for (int a = 0; a < NA; ++a) d[a] += p.partial[((size_t)s * NA + a) * p.n_tgt + i];
  const int t = p.tgt_perm ? p.tgt_perm[i] : i;
  float3 at = make_float3(0.f, 0.f, 0.f);
  if constexpr ((MASK & OUT_DVORT) != 0) {
    const float4 w = p.tgt_a[t];
    at = make_float3(w.x, w.y, w.z);
  }
  f3d_finish<MASK>(d, at, p, t);
}

size_t f3d_workspace_bytes(const F3DJob& j, const LaunchCfg& c) {
  if (c.n_split <= 1) return 0;
  return (size_t)c.n_split * f3d_nacc(j.mask) * (size_t)j.n_tgt * sizeof(double);
}

namespace {
using F3Fn = void (*)(F3DKParams);
template <int MASK, int ACC>
F3Fn f3_T(int T) {
  switch (T) {
    case 1: return f3d_kernel<MASK, ACC, 1, 256>;
    case 4: return f3d_kernel<MASK, ACC, 4, 256>;
    default: return f3d_kernel<MASK, ACC, 2, 256>;
  }
}
template <int MASK>
F3Fn f3_acc(int acc, int T) {
  return acc == ACC_PLAIN ? f3_T<MASK, ACC_PLAIN>(T) : f3_T<MASK, ACC_COMP>(T);
}
F3Fn f3_pick(int mask, int acc, int T) {
  switch (mask) {
    case OUT_VEL: return f3_acc<OUT_VEL>(acc, T);
    case OUT_DVORT: return f3_acc<OUT_DVORT>(acc, T);
    case OUT_VEL | OUT_DVORT: return f3_acc<OUT_VEL | OUT_DVORT>(acc, T);
    default: return nullptr;
  }
}
F3Fn f3_reduce(int mask) {
  switch (mask) {
    case OUT_VEL: return f3d_reduce<OUT_VEL>;
    case OUT_DVORT: return f3d_reduce<OUT_DVORT>;
    case OUT_VEL | OUT_DVORT: return f3d_reduce<OUT_VEL | OUT_DVORT>;
    default: return nullptr;
  }
}
// Influence matrix (cvtx_F3D_inf_mtrx, reference src/F3D.cpp:204-227):
//   out[i * n_fil + j] = vel(filament j, point i) . dir_i
// with the filament's own strength, the velocity term of f3d_kernel (same
// well-conditioned formula and the same flush of non-finite terms).  Each
// lane owns INF_NJ consecutive filaments (columns) of its workgroup's
// block and the workgroup walks `rows` consecutive points, whose
// coordinates and directions are the same in every lane.  Where a lane's
// filaments form a polyline (each starts, bit for bit, where the previous
// one ends, as the spanwise segments of a vortex lattice do) and every lane
// of the wave has that, the shared end point's distance terms are computed
// once per row: the same values the unshared form gives, with one rsq and
// seven other instructions fewer per element.  16384 x 16384 lattice:
// 0.362 -> 0.348 ms (three A/B pairs on one box).  Four filaments per lane
// were slower (0.41 ms: a wave's stores then cover 1 KiB at a 16-byte
// stride), and two columns written as one 8-byte store gained nothing.
constexpr int INF_NJ = 2;

__global__ __launch_bounds__(256) void f3d_inf_kernel(const float4* __restrict__ fa,
                                                      const float4* __restrict__ fb, int n_fil,
                                                      const float4* __restrict__ mes,
                                                      const float4* __restrict__ dir, int n_mes,
                                                      int rows, float* __restrict__ out) {
  const int j0 = (blockIdx.x * 256 + threadIdx.x) * INF_NJ;
  float3 a[INF_NJ], b[INF_NJ];
  float g[INF_NJ];
  bool linked = true;
#pragma unroll
  for (int k = 0; k < INF_NJ; ++k) {
    const int j = min(j0 + k, n_fil - 1);
    const float4 fa4 = fa[j], fb4 = fb[j];
    a[k] = make_float3(fa4.x, fa4.y, fa4.z);
    b[k] = make_float3(fb4.x, fb4.y, fb4.z);
    g[k] = fa4.w * (float)(1.0 / (4.0 * M_PI));
    if (k > 0)
      linked = linked && __float_as_uint(a[k].x) == __float_as_uint(b[k - 1].x) &&
               __float_as_uint(a[k].y) == __float_as_uint(b[k - 1].y) &&
               __float_as_uint(a[k].z) == __float_as_uint(b[k - 1].z);
  }
  // one answer for the whole wave, so the row loop branches uniformly
  const bool share = __all(linked);
  const int i0 = blockIdx.y * rows;
  const int i1 = min(i0 + rows, n_mes);
  for (int i = i0; i < i1; ++i) {
    const float4 x = mes[i];
    const float4 e = dir[i];
    float r1x = x.x - a[0].x, r1y = x.y - a[0].y, r1z = x.z - a[0].z;
    float n1s = __builtin_fmaf(r1x, r1x, __builtin_fmaf(r1y, r1y, r1z * r1z));
    float n1 = n1s * __builtin_amdgcn_rsqf(n1s);
#pragma unroll
    for (int k = 0; k < INF_NJ; ++k) {
      if (k > 0 && !share) {
        r1x = x.x - a[k].x;
        r1y = x.y - a[k].y;
        r1z = x.z - a[k].z;
        n1s = __builtin_fmaf(r1x, r1x, __builtin_fmaf(r1y, r1y, r1z * r1z));
        n1 = n1s * __builtin_amdgcn_rsqf(n1s);
      }
      const float r2x = x.x - b[k].x, r2y = x.y - b[k].y, r2z = x.z - b[k].z;
      const float cx = __builtin_fmaf(r1y, r2z, -r1z * r2y);
      const float cy = __builtin_fmaf(r1z, r2x, -r1x * r2z);
      const float cz = __builtin_fmaf(r1x, r2y, -r1y * r2x);
      const float cc = __builtin_fmaf(cx, cx, __builtin_fmaf(cy, cy, cz * cz));
      const float n2s = __builtin_fmaf(r2x, r2x, __builtin_fmaf(r2y, r2y, r2z * r2z));
      const float n2 = n2s * __builtin_amdgcn_rsqf(n2s);
      const float P = n1 * n2;
      const float sum = n1 + n2;
      const float d12 = __builtin_fmaf(r1x, r2x, __builtin_fmaf(r1y, r2y, r1z * r2z));
      const bool same_side = d12 >= 0.f;
      const float den = same_side ? P * (P + d12) : P * cc;
      const float num = same_side ? sum : sum * (P - d12);
      float kk = g[k] * num * __builtin_amdgcn_rcpf(den);
      kk = (__builtin_fabsf(kk) < __builtin_inff()) ? kk : 0.f;
      const float v = kk * __builtin_fmaf(cx, e.x, __builtin_fmaf(cy, e.y, cz * e.z));
      if (j0 + k < n_fil) out[(size_t)i * (size_t)n_fil + j0 + k] = v;
      // the next filament starts at this one's end point
      r1x = r2x;
      r1y = r2y;
      r1z = r2z;
      n1 = n2;
    }
  }
}

}  // namespace

hipError_t launch_f3d_inf(const float4* fa, const float4* fb, int n_fil, const float4* mes,
                          const float4* dir, int n_mes, float* out, hipStream_t s) {
  if (n_fil <= 0 || n_mes <= 0) return hipSuccess;
  if (!fa || !fb || !mes || !dir || !out) return hipErrorInvalidValue;
  // rows per workgroup: enough workgroups to fill the chip several times
  const int gx = (n_fil + 256 * INF_NJ - 1) / (256 * INF_NJ);
  int rows = 1;
  while (rows < 64 && (long long)gx * ((n_mes + 2 * rows - 1) / (2 * rows)) >= 4096) rows *= 2;
  hipLaunchKernelGGL(f3d_inf_kernel, dim3(gx, (n_mes + rows - 1) / rows), dim3(256), 0, s, fa, fb,
                     n_fil, mes, dir, n_mes, rows, out);
  return hipGetLastError();
}

hipError_t launch_f3d(const F3DJob& j, const LaunchCfg& c, hipStream_t s) {
  if (j.n_tgt <= 0) return hipSuccess;
  if (j.n_fil < 0 || c.wg != 256) return hipErrorInvalidValue;
  if ((j.mask & OUT_DVORT) && j.tgt_a == nullptr) return hipErrorInvalidValue;
  // every output the mask selects must exist, and the inputs it reads
  if (((j.mask & OUT_VEL) && !j.out_vel) || ((j.mask & OUT_DVORT) && !j.out_dvort) || !j.tgt_pv ||
      (j.n_fil > 0 && (!j.fa || !j.fb)))
    return hipErrorInvalidValue;
  F3Fn k = f3_pick(j.mask, j.acc, c.T);
  if (!k) return hipErrorInvalidValue;
  F3DKParams p{};
  p.fa = j.fa;
  p.fb = j.fb;
  p.tgt_pv = j.tgt_pv;
  p.tgt_a = j.tgt_a;
  p.out_vel = j.out_vel;
  p.out_dvort = j.out_dvort;
  p.partial = j.partial;
  p.tgt_perm = j.tgt_perm;
  p.n_fil = j.n_fil;
  p.n_tgt = j.n_tgt;
  int n_split = c.n_split, split_len = c.split_len;
  if (n_split <= 1 || j.n_fil == 0) {
    n_split = 1;
    split_len = j.n_fil > 0 ? j.n_fil : 1;
  } else {
    if (split_len <= 0 || split_len % c.wg != 0) return hipErrorInvalidValue;
    if ((long long)split_len * n_split < j.n_fil || j.partial == nullptr) return hipErrorInvalidValue;
  }
  p.n_split = n_split;
  p.split_len = split_len;
  const int per = c.wg * c.T;
  hipLaunchKernelGGL(k, dim3((j.n_tgt + per - 1) / per, n_split), dim3(c.wg), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || n_split == 1) return e;
  hipLaunchKernelGGL(f3_reduce(j.mask), dim3((j.n_tgt + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace cvtx
