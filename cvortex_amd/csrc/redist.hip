// redist.hip -- redistribution of vortex particles onto a regular grid
// (remeshing): cvtx_P3D_redistribute_on_grid, cvtx_P2D_redistribute_on_grid.
//
// Semantics of the reference CPU path (src/P3D.cpp:509-665,
// src/P2D.cpp:283-436, src/RedistFunc.cpp:36-96,
// src/redistribution_helper_funcs.cpp:33-80):
//   1. grid origin from the particles' minimum and mean (one cell layout for
//      the whole call), node key = round((x - origin) / h) per axis;
//   2. every particle gives strength alpha * f(U) f(W) (f(V)) to the
//      (2R+1)^dim nodes around its nearest node, R = round(radius), with
//      U = |x - x_node| / h per axis; exactly-zero contributions make no node;
//   3. the nodes' summed strengths become particles of volume h^dim, in the
//      ascending Morton order of the node keys (octant index x + 2y + 4z,
//      quadrant x + 2y), the order the reference's tree emits them in;
//   4. particles weaker than negligible x mean strength are removed and their
//      strength shared equally by the rest;
//   5. with an output array and more particles than it holds, a histogram
//      search picks a strength threshold for about max_output survivors, and
//      of the first max_output particles those above it are kept, again
//      sharing the removed strength.
// The weights use the reference's FP32 operations in its order (no fused
// multiply-adds), so the set of nodes, their positions and the weights of
// every contribution are the reference's; node sums are formed in FP64 in
// particle order (the reference sums in FP32).
//
// GPU structure: one thread per (particle, neighbour) writes a (Morton key,
// contribution) entry; rocPRIM's stable radix sort groups the entries by
// node; a thread per node sums its run; removal and truncation are
// order-preserving compactions (rocPRIM select) with the removed strength
// reduced in FP64.  Scalar control (grid origin, thresholds, the histogram
// search's loop) runs on the host between kernels.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"

namespace cvtx {

namespace {

constexpr int kPartBlocks = 512;  // fixed grid of the reductions (deterministic)

// Redistribution functions, reference src/RedistFunc.cpp:36-96, in FP32
// with the reference's operation order.
template <int RID>
__device__ __forceinline__ float rd_f(float U) {
  if constexpr (RID == 0) {  // lambda0
    return U < 0.5f ? 1.f : 0.f;
  } else if constexpr (RID == 1) {  // lambda1
    return U <= 1.f ? rn_sub(1.f, U) : 0.f;
  } else if constexpr (RID == 2) {  // lambda2
    return U < 0.5f ? rn_sub(1.f, rn_mul(U, U))
                    : (U < 1.5f ? rn_mul(rn_mul(0.5f, rn_sub(1.f, U)), rn_sub(2.f, U))
                                : 0.f);
  } else if constexpr (RID == 3) {  // lambda3
    return U < 1.f ? rn_mul(rn_mul(0.5f, rn_sub(1.f, rn_mul(U, U))), rn_sub(2.f, U))
                   : (U < 2.f ? rn_mul(rn_mul(rn_mul(1.f / 6.f, rn_sub(1.f, U)),
                                                    rn_sub(2.f, U)),
                                          rn_sub(3.f, U))
                              : 0.f);
  } else {  // m4'
    return U < 1.f ? rn_add(rn_sub(1.f, rn_mul(rn_mul(2.5f, U), U)),
                               rn_mul(rn_mul(rn_mul(1.5f, U), U), U))
                   : (U < 2.f ? rn_mul(rn_mul(rn_mul(0.5f, rn_sub(1.f, U)),
                                                    rn_sub(2.f, U)),
                                          rn_sub(2.f, U))
                              : 0.f);
  }
}

__host__ __device__ __forceinline__ unsigned long long spread3(unsigned v) {
  unsigned long long x = v & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__host__ __device__ __forceinline__ unsigned compact3(unsigned long long x) {
  x &= 0x1249249249249249ull;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
  x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
  x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
  x = (x ^ (x >> 32)) & 0x1fffffull;
  return (unsigned)x;
}
__host__ __device__ __forceinline__ unsigned long long spread2(unsigned v) {
  unsigned long long x = v;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
__host__ __device__ __forceinline__ unsigned compact2(unsigned long long x) {
  x &= 0x5555555555555555ull;
  x = (x ^ (x >> 1)) & 0x3333333333333333ull;
  x = (x ^ (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
  x = (x ^ (x >> 4)) & 0x00ff00ff00ff00ffull;
  x = (x ^ (x >> 8)) & 0x0000ffff0000ffffull;
  x = (x ^ (x >> 16)) & 0x00000000ffffffffull;
  return (unsigned)x;
}

struct Grid {
  float mn[3];  // origin
  float h, rh;  // spacing and its reciprocal (FP32 as the reference)
  int R;        // neighbourhood half-width
};

// Per-block min (FP32) and sum (FP64) of each coordinate.
template <int DIM>
__global__ __launch_bounds__(256) void rd_bounds(const float4* __restrict__ pv, int n,
                                                 float* __restrict__ pmin,
                                                 float* __restrict__ pmax,
                                                 double* __restrict__ psum) {
  __shared__ float smin[DIM][256];
  __shared__ float smax[DIM][256];
  __shared__ double ssum[DIM][256];
  float m[DIM], M[DIM];
  double s[DIM];
  for (int c = 0; c < DIM; ++c) { m[c] = __builtin_inff(); M[c] = -__builtin_inff(); s[c] = 0.0; }
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float4 p = pv[i];
    const float x[3] = {p.x, p.y, p.z};
    for (int c = 0; c < DIM; ++c) {
      m[c] = fminf(m[c], x[c]);
      M[c] = fmaxf(M[c], x[c]);
      s[c] += (double)x[c];
    }
  }
  for (int c = 0; c < DIM; ++c) {
    smin[c][threadIdx.x] = m[c];
    smax[c][threadIdx.x] = M[c];
    ssum[c][threadIdx.x] = s[c];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < DIM; ++c) {
        smin[c][threadIdx.x] = fminf(smin[c][threadIdx.x], smin[c][threadIdx.x + w]);
        smax[c][threadIdx.x] = fmaxf(smax[c][threadIdx.x], smax[c][threadIdx.x + w]);
        ssum[c][threadIdx.x] += ssum[c][threadIdx.x + w];
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < DIM; ++c) {
      pmin[blockIdx.x * DIM + c] = smin[c][0];
      pmax[blockIdx.x * DIM + c] = smax[c][0];
      psum[blockIdx.x * DIM + c] = ssum[c][0];
    }
}

// Entry e = i * W^DIM + j: the contribution of particle i to its j-th
// neighbour node (a window of W = 2R + 1 nodes per axis around the nearest
// node).  3D: pv = {x, y, z, -}, a = {alpha, -}; 2D: pv = {x, y, alpha, -}.
// Returns false for an exactly-zero contribution (the reference makes no
// node for it).
template <int DIM, int RID>
__device__ __forceinline__ bool rd_entry(const float4* __restrict__ pv,
                                         const float4* __restrict__ a, int e, const Grid& g,
                                         unsigned long long& key, float4& v) {
  const int W = 2 * g.R + 1;
  const int WD = DIM == 3 ? W * W * W : W * W;
  const int i = e / WD;
  const int j = e - i * WD;
  const float4 p = pv[i];
  const float x[3] = {p.x, p.y, p.z};
  unsigned q[3];
  int off[3];
  if constexpr (DIM == 3) {
    off[0] = j / (W * W) - g.R;
    off[1] = (j / W) % W - g.R;
    off[2] = j % W - g.R;
  } else {
    off[0] = j / W - g.R;
    off[1] = j % W - g.R;
  }
  float frac = 1.f;
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    float k;
    if constexpr (DIM == 3) {  // reference UIntKey96.cpp: FP32 offset
      k = roundf(rn_mul(rn_sub(x[c], g.mn[c]), g.rh));
    } else {  // reference UIntKey64.cpp:79-96: the offset is formed in FP64
      k = roundf((float)(((double)x[c] - (double)g.mn[c]) * (double)g.rh));
    }
    q[c] = (unsigned)k + (unsigned)off[c];
    const float node = rn_add(g.mn[c], rn_mul(g.h, (float)q[c]));
    const float U = fabsf(rn_mul(rn_sub(x[c], node), g.rh));
    frac = c == 0 ? rd_f<RID>(U) : rn_mul(frac, rd_f<RID>(U));
  }
  if constexpr (DIM == 3) {
    const float4 al = a[i];
    v = make_float4(rn_mul(al.x, frac), rn_mul(al.y, frac), rn_mul(al.z, frac), 0.f);
    key = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    return !(v.x == 0.f && v.y == 0.f && v.z == 0.f);
  } else {
    v = make_float4(rn_mul(p.z, frac), 0.f, 0.f, 0.f);
    key = spread2(q[0]) | (spread2(q[1]) << 1);
    return v.x != 0.f;
  }
}

// Chunked stable compaction (used for the nonzero entries and for the node
// starts): a first kernel counts the selected items of each chunk of kChunk
// = 16 rounds of 256, an inclusive scan of the counts gives each chunk's end
// offset, and a second kernel writes the chunk's items from its start
// offset in their original order, a wave ballot ranking them within a round.
constexpr int kChunk = 4096;

__device__ __forceinline__ int block_sum(int c, int* s_w) {
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
  __syncthreads();
  return s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Within one round of a chunk: the output slot of this lane's item (if it is
// selected), given the round's first slot; advances `off` past the round.
__device__ __forceinline__ int round_slot(bool sel, int& off, int* s_w) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long m = __ballot(sel);
  const int rank = __popcll(m & ((1ull << lane) - 1ull));
  __syncthreads();  // s_w of the previous round has been read
  if (lane == 0) s_w[w] = __popcll(m);
  __syncthreads();
  int wo = 0;
  for (int q = 0; q < w; ++q) wo += s_w[q];
  const int slot = off + wo + rank;
  off += s_w[0] + s_w[1] + s_w[2] + s_w[3];
  return slot;
}

template <int DIM, int RID>
__global__ __launch_bounds__(256) void rd_contrib_count(const float4* __restrict__ pv,
                                                        const float4* __restrict__ a, long ne,
                                                        Grid g, int* __restrict__ counts) {
  __shared__ int s_w[4];
  const long base = (long)blockIdx.x * kChunk;
  int c = 0;
  for (int r = 0; r < kChunk / 256; ++r) {
    const long e = base + r * 256 + threadIdx.x;
    unsigned long long key;
    float4 v;
    if (e < ne && rd_entry<DIM, RID>(pv, a, (int)e, g, key, v)) ++c;
  }
  c = block_sum(c, s_w);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// The nonzero entries, compacted in entry order (ne < 2^31, checked on the
// host: 32-bit entry indices).
template <int DIM, int RID>
__global__ __launch_bounds__(256) void rd_contrib_write(const float4* __restrict__ pv,
                                                        const float4* __restrict__ a, long ne,
                                                        Grid g, const int* __restrict__ incl,
                                                        unsigned long long* __restrict__ keys,
                                                        float4* __restrict__ vals) {
  __shared__ int s_w[4];
  const long base = (long)blockIdx.x * kChunk;
  int off = blockIdx.x > 0 ? incl[blockIdx.x - 1] : 0;
  for (int r = 0; r < kChunk / 256; ++r) {
    const long e = base + r * 256 + threadIdx.x;
    unsigned long long key = 0;
    float4 v;
    const bool sel = e < ne && rd_entry<DIM, RID>(pv, a, (int)e, g, key, v);
    const int slot = round_slot(sel, off, s_w);
    if (sel) {
      keys[slot] = key;
      vals[slot] = v;
    }
  }
}

// Node starts: the heads of the runs of equal keys among the sorted nonzero
// entries, in increasing order, by the chunked compaction above.  Heads are
// sparse (one per node, a few per hundred entries), so no flag array is
// written.
__device__ __forceinline__ bool run_head(const unsigned long long* __restrict__ keys, long i,
                                         long n) {
  return i < n && (i == 0 || keys[i - 1] != keys[i]);
}

__global__ __launch_bounds__(256) void rd_head_counts(const unsigned long long* __restrict__ keys,
                                                      long n, int* __restrict__ counts) {
  __shared__ int s_w[4];
  const long base = (long)blockIdx.x * kChunk;
  int c = 0;
  for (int r = 0; r < kChunk / 256; ++r) c += run_head(keys, base + r * 256 + threadIdx.x, n);
  c = block_sum(c, s_w);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void rd_head_write(const unsigned long long* __restrict__ keys,
                                                     long n, const int* __restrict__ incl,
                                                     long* __restrict__ starts) {
  __shared__ int s_w[4];
  const long base = (long)blockIdx.x * kChunk;
  int off = blockIdx.x > 0 ? incl[blockIdx.x - 1] : 0;
  for (int r = 0; r < kChunk / 256; ++r) {
    const long i = base + r * 256 + threadIdx.x;
    const bool h = run_head(keys, i, n);
    const int slot = round_slot(h, off, s_w);
    if (h) starts[slot] = i;
  }
}

// One wave per node: the wave reads the node's run of entries 64 at a time
// (coalesced) and reduces it in FP64 with a fixed shuffle tree, so the sum
// is the same on every call; the run ends at the first different key.
// pos = {x, y, z, volume} (2D: {x, y, -, area}); w = strength; str = |w|.
template <int DIM>
__global__ __launch_bounds__(256) void rd_nodes(const unsigned long long* __restrict__ keys,
                                                const float4* __restrict__ vals, long n_entries,
                                                const long* __restrict__ starts, int n_nodes,
                                                Grid g, float4* __restrict__ pos,
                                                float4* __restrict__ w, float* __restrict__ str) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n_nodes) return;  // whole waves leave together
  const long s = starts[k];
  const unsigned long long key = keys[s];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (long base = s;; base += 64) {
    const long i = base + lane;
    const bool in = i < n_entries && keys[i] == key;
    if (in) {
      const float4 v = vals[i];
      a0 += (double)v.x;
      a1 += (double)v.y;
      a2 += (double)v.z;
    }
    // the run is contiguous: it continues only if every lane was inside it
    if (__ballot(in) != ~0ull) break;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o);
    a1 += __shfl_xor(a1, o);
    a2 += __shfl_xor(a2, o);
  }
  if (lane != 0) return;
  const float vol = DIM == 3 ? rn_mul(rn_mul(g.h, g.h), g.h) : rn_mul(g.h, g.h);
  float4 p;
  if constexpr (DIM == 3) {
    p = make_float4(rn_add(g.mn[0], rn_mul(g.h, (float)compact3(key))),
                    rn_add(g.mn[1], rn_mul(g.h, (float)compact3(key >> 1))),
                    rn_add(g.mn[2], rn_mul(g.h, (float)compact3(key >> 2))), vol);
  } else {
    p = make_float4(rn_add(g.mn[0], rn_mul(g.h, (float)compact2(key))),
                    rn_add(g.mn[1], rn_mul(g.h, (float)compact2(key >> 1))), 0.f, vol);
  }
  const float4 ww = make_float4((float)a0, (float)a1, (float)a2, 0.f);
  pos[k] = p;
  w[k] = ww;
  str[k] = DIM == 3 ? __fsqrt_rn(rn_add(rn_add(rn_mul(ww.x, ww.x), rn_mul(ww.y, ww.y)),
                                        rn_mul(ww.z, ww.z)))
                    : fabsf(ww.x);
}

// Per-block FP64 sum and FP32 min / max of the strengths.
__global__ __launch_bounds__(256) void rd_str_stats(const float* __restrict__ str, int n,
                                                    double* __restrict__ psum,
                                                    float* __restrict__ pmin,
                                                    float* __restrict__ pmax) {
  __shared__ double ss[256];
  __shared__ float sl[256], sh[256];
  double s = 0.0;
  float lo = __builtin_inff(), hi = -__builtin_inff();
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float v = str[i];
    s += (double)v;
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
  }
  ss[threadIdx.x] = s;
  sl[threadIdx.x] = lo;
  sh[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      ss[threadIdx.x] += ss[threadIdx.x + w];
      sl[threadIdx.x] = fminf(sl[threadIdx.x], sl[threadIdx.x + w]);
      sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    psum[blockIdx.x] = ss[0];
    pmin[blockIdx.x] = sl[0];
    pmax[blockIdx.x] = sh[0];
  }
}

// keep[i] = str[i] > thr && i < max_keep; per-block FP64 sums of the
// removed strength (3 components).
__global__ __launch_bounds__(256) void rd_flags(const float* __restrict__ str,
                                                const float4* __restrict__ w, int n, float thr,
                                                int max_keep, unsigned char* __restrict__ keep,
                                                double* __restrict__ pdef) {
  __shared__ double sd[3][256];
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const bool k = str[i] > thr && i < max_keep;
    keep[i] = k ? 1 : 0;
    if (!k) {
      const float4 v = w[i];
      d0 += (double)v.x;
      d1 += (double)v.y;
      d2 += (double)v.z;
    }
  }
  sd[0][threadIdx.x] = d0;
  sd[1][threadIdx.x] = d1;
  sd[2][threadIdx.x] = d2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 3; ++c) sd[c][threadIdx.x] += sd[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) pdef[blockIdx.x * 3 + c] = sd[c][0];
}

// Gather the kept particles, add the shared deficit, refresh |w|.
template <int DIM>
__global__ __launch_bounds__(256) void rd_keep(const int* __restrict__ idx, int n_kept,
                                               const float4* __restrict__ pos_in,
                                               const float4* __restrict__ w_in, float4 def,
                                               float4* __restrict__ pos_out,
                                               float4* __restrict__ w_out,
                                               float* __restrict__ str_out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n_kept) return;
  const int i = idx[j];
  pos_out[j] = pos_in[i];
  const float4 v = w_in[i];
  const float4 ww = make_float4(rn_add(v.x, def.x), rn_add(v.y, def.y),
                                rn_add(v.z, def.z), 0.f);
  w_out[j] = ww;
  str_out[j] = DIM == 3 ? __fsqrt_rn(rn_add(rn_add(rn_mul(ww.x, ww.x), rn_mul(ww.y, ww.y)),
                                               rn_mul(ww.z, ww.z)))
                        : fabsf(ww.x);
}

// Histogram of the strength-threshold search (reference
// src/redistribution_helper_funcs.cpp:54-60): bin = floor(1023 (s - minv) / range).
__global__ __launch_bounds__(256) void rd_hist(const float* __restrict__ str, int n, double minv,
                                               double range, int* __restrict__ counts) {
  __shared__ int sc[1024];
  for (int b = threadIdx.x; b < 1024; b += 256) sc[b] = 0;
  __syncthreads();
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    int b = (int)floor((double)1023 * ((double)str[i] - minv) / range);
    b = b < 0 ? 0 : (b > 1023 ? 1023 : b);
    atomicAdd(&sc[b], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 1024; b += 256)
    if (sc[b]) atomicAdd(&counts[b], sc[b]);
}

// AoS output records: 3D {coord[3], vorticity[3], volume} (7 floats),
// 2D {coord[2], vorticity, area} (4 floats).
template <int DIM>
__global__ __launch_bounds__(256) void rd_pack(const float4* __restrict__ pos,
                                               const float4* __restrict__ w, int n,
                                               float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 p = pos[i];
  const float4 v = w[i];
  if constexpr (DIM == 3) {
    float* r = out + 7 * (size_t)i;
    r[0] = p.x; r[1] = p.y; r[2] = p.z; r[3] = v.x; r[4] = v.y; r[5] = v.z; r[6] = p.w;
  } else {
    float* r = out + 4 * (size_t)i;
    r[0] = p.x; r[1] = p.y; r[2] = v.x; r[3] = p.w;
  }
}

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

// Bump allocator over one runtime workspace slot.
struct Arena {
  char* base = nullptr;
  size_t used = 0, cap = 0;
  template <class T>
  T* take(size_t n) {
    T* p = (T*)(base + used);
    used += al256(sizeof(T) * (n > 0 ? n : 1));
    return used <= cap ? p : nullptr;
  }
};

template <int DIM>
hipError_t launch_contrib_count(int rid, const float4* pv, const float4* a, long ne, const Grid& g,
                                int* counts, hipStream_t s) {
  const dim3 grid((unsigned)((ne + kChunk - 1) / kChunk));
  switch (rid) {
    case 0: hipLaunchKernelGGL((rd_contrib_count<DIM, 0>), grid, dim3(256), 0, s, pv, a, ne, g, counts); break;
    case 1: hipLaunchKernelGGL((rd_contrib_count<DIM, 1>), grid, dim3(256), 0, s, pv, a, ne, g, counts); break;
    case 2: hipLaunchKernelGGL((rd_contrib_count<DIM, 2>), grid, dim3(256), 0, s, pv, a, ne, g, counts); break;
    case 3: hipLaunchKernelGGL((rd_contrib_count<DIM, 3>), grid, dim3(256), 0, s, pv, a, ne, g, counts); break;
    default: hipLaunchKernelGGL((rd_contrib_count<DIM, 4>), grid, dim3(256), 0, s, pv, a, ne, g, counts); break;
  }
  return hipGetLastError();
}

template <int DIM>
hipError_t launch_contrib_write(int rid, const float4* pv, const float4* a, long ne, const Grid& g,
                                const int* incl, unsigned long long* keys, float4* vals,
                                hipStream_t s) {
  const dim3 grid((unsigned)((ne + kChunk - 1) / kChunk));
  switch (rid) {
    case 0: hipLaunchKernelGGL((rd_contrib_write<DIM, 0>), grid, dim3(256), 0, s, pv, a, ne, g, incl, keys, vals); break;
    case 1: hipLaunchKernelGGL((rd_contrib_write<DIM, 1>), grid, dim3(256), 0, s, pv, a, ne, g, incl, keys, vals); break;
    case 2: hipLaunchKernelGGL((rd_contrib_write<DIM, 2>), grid, dim3(256), 0, s, pv, a, ne, g, incl, keys, vals); break;
    case 3: hipLaunchKernelGGL((rd_contrib_write<DIM, 3>), grid, dim3(256), 0, s, pv, a, ne, g, incl, keys, vals); break;
    default: hipLaunchKernelGGL((rd_contrib_write<DIM, 4>), grid, dim3(256), 0, s, pv, a, ne, g, incl, keys, vals); break;
  }
  return hipGetLastError();
}

}  // namespace

float redist_radius(int rid) {
  static const float r[5] = {0.5f, 1.0f, 1.5f, 2.0f, 2.0f};
  return (rid >= 0 && rid < 5) ? r[rid] : -1.f;
}

template <int DIM>
static int redistribute_dim(Runtime& rt, Device& d, const RedistJob& j, float** out_dev,
                            hipStream_t s) {
  auto fail = [&](const char* what, hipError_t e) {
    if (e != hipSuccess) check_hip(rt, e, what);
    else set_error(rt, what);
    return -1;
  };
  const int n = j.n;
  *out_dev = nullptr;
  if (n == 0) return 0;
  // ---- grid origin (reference src/P3D.cpp:538-548) ----
  float* pmin = nullptr;
  float* pmax = nullptr;
  double* psum = nullptr;
  {
    Arena small;
    const size_t bytes = 2 * al256(sizeof(float) * kPartBlocks * 3) + al256(sizeof(double) * kPartBlocks * 3);
    small.base = (char*)device_buffer(rt, d, SLOT_RD_SMALL, bytes);
    small.cap = bytes;
    if (!small.base) return -1;
    pmin = small.take<float>(kPartBlocks * 3);
    pmax = small.take<float>(kPartBlocks * 3);
    psum = small.take<double>(kPartBlocks * 3);
  }
  hipLaunchKernelGGL((rd_bounds<DIM>), dim3(kPartBlocks), dim3(256), 0, s, j.pv, n, pmin, pmax, psum);
  std::vector<float> hmin(kPartBlocks * DIM), hmax(kPartBlocks * DIM);
  std::vector<double> hsum(kPartBlocks * DIM);
  hipError_t e = hipMemcpyAsync(hmin.data(), pmin, sizeof(float) * hmin.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hmax.data(), pmax, sizeof(float) * hmax.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hsum.data(), psum, sizeof(double) * hsum.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail("redistribution bounds", e);
  Grid g{};
  g.h = j.h;
  g.rh = 1.f / j.h;
  g.R = (int)roundf(redist_radius(j.rid));
  int axis_bits = 1;  // bits of the largest node index on any axis
  for (int c = 0; c < DIM; ++c) {
    float m = hmin[c], M = hmax[c];
    double sum = 0.0;
    for (int b = 0; b < kPartBlocks; ++b) {
      m = hmin[b * DIM + c] < m ? hmin[b * DIM + c] : m;
      M = hmax[b * DIM + c] > M ? hmax[b * DIM + c] : M;
      sum += hsum[b * DIM + c];
    }
    if (!std::isfinite(m) || !std::isfinite(M)) return fail("non-finite particle coordinates", hipSuccess);
    const float mean = (float)(sum / (double)n);
    float mn = m - 1.f * ((float)g.R * j.h);
    float dc = (mean - mn) / j.h;
    dc = roundf(dc) + 5;
    g.mn[c] = mean - dc * j.h;
    // node indices must fit the Morton key: 21 bits per axis in 3D, 32 in 2D
    const double span = ((double)M - (double)g.mn[c]) / (double)j.h + g.R + 2;
    if (span >= (DIM == 3 ? 2097152.0 : 4294967295.0))
      return fail("grid too fine for the particles' extent", hipSuccess);
    while (axis_bits < 32 && (double)(1ull << axis_bits) <= span) ++axis_bits;
  }
  // the sort only needs the interleaved bits of the node indices
  const int key_bits = DIM * axis_bits < 64 ? DIM * axis_bits : 64;
  const int W = 2 * g.R + 1;
  const long WD = DIM == 3 ? (long)W * W * W : (long)W * W;
  const long ne = (long)n * WD;
  if (ne >= 2147483647L) return fail("too many grid contributions for one call", hipSuccess);
  // ---- nonzero entries: count per chunk, scan, compacting write ----
  const long nc = (ne + kChunk - 1) / kChunk;
  size_t scan_tmp = 0;
  e = rocprim::inclusive_scan(nullptr, scan_tmp, (int*)nullptr, (int*)nullptr, (size_t)nc,
                              rocprim::plus<int>(), s);
  if (e != hipSuccess) return fail("scan size", e);
  Arena ca;
  const size_t cneed = al256(sizeof(int) * nc) * 2 + al256(scan_tmp);
  ca.base = (char*)device_buffer(rt, d, SLOT_RD_CHUNKS, cneed);
  ca.cap = cneed;
  if (!ca.base) return -1;
  auto* ccount = ca.take<int>(nc);
  auto* cincl = ca.take<int>(nc);
  void* ct = ca.take<char>(scan_tmp);
  // the count of the chunks' items (in cincl's last element) on the host
  auto scan_total = [&](long chunks, int* total) -> bool {
    size_t tb = scan_tmp;
    hipError_t ee = rocprim::inclusive_scan(ct, tb, ccount, cincl, (size_t)chunks,
                                            rocprim::plus<int>(), s);
    if (ee == hipSuccess)
      ee = hipMemcpyAsync(total, cincl + (chunks - 1), sizeof(int), hipMemcpyDeviceToHost, s);
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    if (ee != hipSuccess) {
      fail("redistribution scan", ee);
      return false;
    }
    return true;
  };
  e = launch_contrib_count<DIM>(j.rid, j.pv, j.a, ne, g, ccount, s);
  if (e != hipSuccess) return fail("redistribution contributions", e);
  int nz = 0;
  if (!scan_total(nc, &nz)) return -1;
  if (nz == 0) return 0;
  // ---- workspace: the nonzero entries and the sort ----
  size_t sort_tmp = 0;
  e = rocprim::radix_sort_pairs(nullptr, sort_tmp, (unsigned long long*)nullptr,
                                (unsigned long long*)nullptr, (float4*)nullptr, (float4*)nullptr,
                                (size_t)nz, 0, 64, s);
  if (e != hipSuccess) return fail("radix sort size", e);
  const size_t need = al256(sizeof(unsigned long long) * nz) * 2 + al256(sizeof(float4) * nz) * 2 +
                      al256(sort_tmp);
  Arena ar;
  ar.base = (char*)device_buffer(rt, d, SLOT_RD, need);
  ar.cap = need;
  if (!ar.base) return -1;
  auto* k_in = ar.take<unsigned long long>(nz);
  auto* k_out = ar.take<unsigned long long>(nz);
  auto* v_in = ar.take<float4>(nz);
  auto* v_out = ar.take<float4>(nz);
  void* t = ar.take<char>(sort_tmp);
  if (!t) return fail("redistribution workspace", hipSuccess);
  e = launch_contrib_write<DIM>(j.rid, j.pv, j.a, ne, g, cincl, k_in, v_in, s);
  if (e != hipSuccess) return fail("redistribution contributions", e);
  size_t tb = sort_tmp;
  e = rocprim::radix_sort_pairs(t, tb, k_in, k_out, v_in, v_out, (size_t)nz, 0, key_bits, s);
  if (e != hipSuccess) return fail("redistribution sort", e);
  // ---- node starts ----
  const long nh = (nz + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(rd_head_counts, dim3((unsigned)nh), dim3(256), 0, s, k_out, (long)nz, ccount);
  int nn = 0;
  if (!scan_total(nh, &nn)) return -1;
  if (nn == 0) return 0;
  // ---- workspace: nodes (sized by the node count) ----
  size_t sel_tmp = 0;
  e = rocprim::select(nullptr, sel_tmp, rocprim::counting_iterator<int>(0), (unsigned char*)nullptr,
                      (int*)nullptr, (int*)nullptr, (size_t)nn, s);
  if (e != hipSuccess) return fail("select size", e);
  const size_t nneed = al256(sel_tmp) + al256(nn) + al256(sizeof(long) * nn) +
                       al256(sizeof(float4) * nn) * 4 + al256(sizeof(float) * nn) * 2 +
                       al256(sizeof(int) * nn) + al256(sizeof(int)) +
                       al256(sizeof(double) * kPartBlocks * 3) + al256(sizeof(float) * kPartBlocks) * 2 +
                       al256(sizeof(int) * 1024) + al256(sizeof(float) * 7 * (size_t)nn);
  Arena na;
  na.base = (char*)device_buffer(rt, d, SLOT_RD_NODES, nneed);
  na.cap = nneed;
  if (!na.base) return -1;
  void* t2 = na.take<char>(sel_tmp);
  auto* head = na.take<unsigned char>(nn);
  auto* starts = na.take<long>(nn);
  auto* pos0 = na.take<float4>(nn);
  auto* w0 = na.take<float4>(nn);
  auto* pos1 = na.take<float4>(nn);
  auto* w1 = na.take<float4>(nn);
  auto* str0 = na.take<float>(nn);
  auto* str1 = na.take<float>(nn);
  auto* kidx = na.take<int>(nn);
  auto* nkept = na.take<int>(1);
  auto* pdef = na.take<double>(kPartBlocks * 3);
  auto* psmin = na.take<float>(kPartBlocks);
  auto* psmax = na.take<float>(kPartBlocks);
  auto* counts = na.take<int>(1024);
  auto* packed = na.take<float>(7 * (size_t)nn);
  if (!packed) return fail("redistribution workspace", hipSuccess);
  hipLaunchKernelGGL(rd_head_write, dim3((unsigned)nh), dim3(256), 0, s, k_out, (long)nz, cincl,
                     starts);
  hipLaunchKernelGGL((rd_nodes<DIM>), dim3((nn + 3) / 4), dim3(256), 0, s, k_out, v_out, (long)nz,
                     starts, nn, g, pos0, w0, str0);
  // ---- removal: strength <= threshold and index >= max_keep ----
  auto remove = [&](float thr, int max_keep) -> int {
    hipLaunchKernelGGL(rd_flags, dim3(kPartBlocks), dim3(256), 0, s, str0, w0, nn, thr, max_keep,
                       head, pdef);
    size_t tb2 = sel_tmp;
    hipError_t ee = rocprim::select(t2, tb2, rocprim::counting_iterator<int>(0), head, kidx, nkept,
                                    (size_t)nn, s);
    if (ee != hipSuccess) return fail("redistribution compaction", ee);
    std::vector<double> hdef(kPartBlocks * 3);
    int hk = 0;
    ee = hipMemcpyAsync(hdef.data(), pdef, sizeof(double) * hdef.size(), hipMemcpyDeviceToHost, s);
    if (ee == hipSuccess) ee = hipMemcpyAsync(&hk, nkept, sizeof(int), hipMemcpyDeviceToHost, s);
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    if (ee != hipSuccess) return fail("redistribution compaction", ee);
    double dd[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < kPartBlocks; ++b)
      for (int c = 0; c < 3; ++c) dd[c] += hdef[b * 3 + c];
    // the reference divides the removed strength by the kept count in FP32
    const float4 def = make_float4((float)dd[0] / (float)hk, (float)dd[1] / (float)hk,
                                   (float)dd[2] / (float)hk, 0.f);
    if (hk > 0)
      hipLaunchKernelGGL((rd_keep<DIM>), dim3((hk + 255) / 256), dim3(256), 0, s, kidx, hk, pos0,
                         w0, def, pos1, w1, str1);
    std::swap(pos0, pos1);
    std::swap(w0, w1);
    std::swap(str0, str1);
    nn = hk;
    return 0;
  };
  auto stats = [&](double* sum, float* lo, float* hi) -> bool {
    hipLaunchKernelGGL(rd_str_stats, dim3(kPartBlocks), dim3(256), 0, s, str0, nn, pdef, psmin, psmax);
    std::vector<double> hs(kPartBlocks);
    std::vector<float> hl(kPartBlocks), hh(kPartBlocks);
    hipError_t ee = hipMemcpyAsync(hs.data(), pdef, sizeof(double) * kPartBlocks, hipMemcpyDeviceToHost, s);
    if (ee == hipSuccess) ee = hipMemcpyAsync(hl.data(), psmin, sizeof(float) * kPartBlocks, hipMemcpyDeviceToHost, s);
    if (ee == hipSuccess) ee = hipMemcpyAsync(hh.data(), psmax, sizeof(float) * kPartBlocks, hipMemcpyDeviceToHost, s);
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    if (ee != hipSuccess) {
      fail("redistribution strengths", ee);
      return false;
    }
    double a = 0.0;
    float l = __builtin_inff(), u = -__builtin_inff();
    for (int b = 0; b < kPartBlocks; ++b) {
      a += hs[b];
      l = hl[b] < l ? hl[b] : l;
      u = hh[b] > u ? hh[b] : u;
    }
    *sum = a;
    *lo = l;
    *hi = u;
    return true;
  };
  double ssum;
  float smin, smax;
  if (!stats(&ssum, &smin, &smax)) return -1;
  const float mean = (float)ssum / (float)nn;
  if (remove(mean * j.negligible, nn) != 0) return -1;
  if (j.max_out >= 0 && nn > j.max_out) {
    // strength threshold, reference src/redistribution_helper_funcs.cpp:33-80
    if (!stats(&ssum, &smin, &smax)) return -1;
    double minv = smin, maxv = smax, range;
    const int n_guesses = 1024;
    std::vector<float> guesses(n_guesses);
    std::vector<int> gc(n_guesses);
    int k = 0;
    const float tol = 0.01f;
    for (int iter = 0; iter < 4096; ++iter) {
      range = (maxv - minv) * 1.05;
      for (int i = 0; i < n_guesses; ++i)
        guesses[i] = (float)(minv + i * range / (float)(n_guesses - 1));
      e = hipMemsetAsync(counts, 0, sizeof(int) * 1024, s);
      if (e != hipSuccess) return fail("redistribution histogram", e);
      hipLaunchKernelGGL(rd_hist, dim3(kPartBlocks), dim3(256), 0, s, str0, nn, minv, range, counts);
      e = hipMemcpyAsync(gc.data(), counts, sizeof(int) * 1024, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) return fail("redistribution histogram", e);
      int interp = gc[n_guesses - 1];
      for (int i = n_guesses - 2; i >= 0; --i) {
        maxv = guesses[i + 1];
        minv = guesses[i];
        interp += gc[i];
        gc[i] = interp;
        if (interp > j.max_out) {
          k = i + 1;
          break;
        }
      }
      if (minv == maxv || gc[k] == gc[k - 1] ||
          std::fabs((float)(j.max_out - gc[k]) / ((float)j.max_out)) < tol * 0.6)
        break;
    }
    if (remove(guesses[k], j.max_out) != 0) return -1;
  }
  if (j.max_out < 0 || nn == 0) return nn;
  hipLaunchKernelGGL((rd_pack<DIM>), dim3((nn + 255) / 256), dim3(256), 0, s, pos0, w0, nn, packed);
  e = hipGetLastError();
  if (e != hipSuccess) return fail("redistribution output", e);
  *out_dev = packed;
  return nn;
}

int redistribute(Runtime& rt, Device& d, const RedistJob& j, float** out_dev, hipStream_t s) {
  if (j.rid < 0 || j.rid > 4 || !(j.h > 0.f) || j.n < 0 || (j.n > 0 && !j.pv) ||
      (j.dim == 3 && j.n > 0 && !j.a)) {
    set_error(rt, "invalid redistribution job");
    return -1;
  }
  return j.dim == 3 ? redistribute_dim<3>(rt, d, j, out_dev, s)
                    : redistribute_dim<2>(rt, d, j, out_dev, s);
}

}  // namespace cvtx
