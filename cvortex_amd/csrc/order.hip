// order.hip -- spatial ordering of targets (Morton / Z-order).
//
// The pair kernels assign 256*T consecutive targets to a workgroup and 64
// to a wave.  When consecutive targets are close in space, a wave's targets
// are either all near a given source or all far from it, so a kernel can
// skip a near-field-only code path for the whole wave (the Gaussian table
// lookup).  This file computes a permutation that sorts targets by the
// 30-bit Morton code of their position inside the targets' bounding box:
// a two-pass min/max reduction, a key kernel, and rocPRIM's device radix
// sort of (key, index) pairs.  The kernels then read target perm[i] and
// write its outputs there; source order -- and hence every per-target sum
// -- is unchanged.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "kernels.hpp"

namespace cvtx {

namespace {

constexpr int kBoxBlocks = 256;

__global__ __launch_bounds__(256) void box_partial(const float4* __restrict__ pts, int n,
                                                   float4* __restrict__ lo, float4* __restrict__ hi) {
  __shared__ float3 slo[256], shi[256];
  float3 a = make_float3(__builtin_inff(), __builtin_inff(), __builtin_inff());
  float3 b = make_float3(-__builtin_inff(), -__builtin_inff(), -__builtin_inff());
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float4 p = pts[i];
    a.x = fminf(a.x, p.x); a.y = fminf(a.y, p.y); a.z = fminf(a.z, p.z);
    b.x = fmaxf(b.x, p.x); b.y = fmaxf(b.y, p.y); b.z = fmaxf(b.z, p.z);
  }
  slo[threadIdx.x] = a;
  shi[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const float3 c = slo[threadIdx.x + s], d = shi[threadIdx.x + s];
      float3& x = slo[threadIdx.x];
      float3& y = shi[threadIdx.x];
      x.x = fminf(x.x, c.x); x.y = fminf(x.y, c.y); x.z = fminf(x.z, c.z);
      y.x = fmaxf(y.x, d.x); y.y = fmaxf(y.y, d.y); y.z = fmaxf(y.z, d.z);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    lo[blockIdx.x] = make_float4(slo[0].x, slo[0].y, slo[0].z, 0.f);
    hi[blockIdx.x] = make_float4(shi[0].x, shi[0].y, shi[0].z, 0.f);
  }
}

// Spread the low 10 bits of v to every third bit.
__device__ __forceinline__ unsigned spread3(unsigned v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ __launch_bounds__(256) void morton_keys(const float4* __restrict__ pts, int n,
                                                   const float4* __restrict__ lo,
                                                   const float4* __restrict__ hi, int nparts,
                                                   unsigned* __restrict__ keys,
                                                   int* __restrict__ idx) {
  // every thread folds the (few) partial boxes itself
  float3 a = make_float3(__builtin_inff(), __builtin_inff(), __builtin_inff());
  float3 b = make_float3(-__builtin_inff(), -__builtin_inff(), -__builtin_inff());
  for (int k = 0; k < nparts; ++k) {
    const float4 l = lo[k], h = hi[k];
    a.x = fminf(a.x, l.x); a.y = fminf(a.y, l.y); a.z = fminf(a.z, l.z);
    b.x = fmaxf(b.x, h.x); b.y = fmaxf(b.y, h.y); b.z = fmaxf(b.z, h.z);
  }
  const float ext = fmaxf(fmaxf(b.x - a.x, b.y - a.y), fmaxf(b.z - a.z, 1e-30f));
  const float sc = 1023.f / ext;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 p = pts[i];
  // non-finite coordinates map to cell 0
  const float fx = fminf(fmaxf((p.x - a.x) * sc, 0.f), 1023.f);
  const float fy = fminf(fmaxf((p.y - a.y) * sc, 0.f), 1023.f);
  const float fz = fminf(fmaxf((p.z - a.z) * sc, 0.f), 1023.f);
  const unsigned ux = (fx == fx) ? (unsigned)fx : 0u;
  const unsigned uy = (fy == fy) ? (unsigned)fy : 0u;
  const unsigned uz = (fz == fz) ? (unsigned)fz : 0u;
  keys[i] = (spread3(ux) << 2) | (spread3(uy) << 1) | spread3(uz);
  idx[i] = i;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

size_t target_order_workspace(int n) {
  if (n <= 0) return 0;
  size_t temp = 0;
  (void)rocprim::radix_sort_pairs(nullptr, temp, (unsigned*)nullptr, (unsigned*)nullptr,
                                  (int*)nullptr, (int*)nullptr, (unsigned)n, 0, 30, (hipStream_t)0,
                                  false);
  const size_t nn = (size_t)n;
  return align256(sizeof(float4) * kBoxBlocks) * 2 + align256(sizeof(unsigned) * nn) * 2 +
         align256(sizeof(int) * nn) + align256(temp);
}

hipError_t target_order(const float4* pts, int n, int* perm, void* ws, size_t ws_bytes,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (ws_bytes < target_order_workspace(n) || ws == nullptr || perm == nullptr)
    return hipErrorInvalidValue;
  char* w = (char*)ws;
  float4* lo = (float4*)w;
  w += align256(sizeof(float4) * kBoxBlocks);
  float4* hi = (float4*)w;
  w += align256(sizeof(float4) * kBoxBlocks);
  unsigned* k_in = (unsigned*)w;
  w += align256(sizeof(unsigned) * (size_t)n);
  unsigned* k_out = (unsigned*)w;
  w += align256(sizeof(unsigned) * (size_t)n);
  int* v_in = (int*)w;
  w += align256(sizeof(int) * (size_t)n);
  size_t temp = 0;
  (void)rocprim::radix_sort_pairs(nullptr, temp, k_in, k_out, v_in, perm, (unsigned)n, 0, 30, s,
                                  false);
  const int nb = n < 256 * kBoxBlocks ? (n + 255) / 256 : kBoxBlocks;
  hipLaunchKernelGGL(box_partial, dim3(nb), dim3(256), 0, s, pts, n, lo, hi);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(morton_keys, dim3((n + 255) / 256), dim3(256), 0, s, pts, n, lo, hi, nb, k_in,
                     v_in);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return rocprim::radix_sort_pairs((void*)w, temp, k_in, k_out, v_in, perm, (unsigned)n, 0, 30, s,
                                   false);
}

}  // namespace cvtx
