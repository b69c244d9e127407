// kernels.hpp -- host-visible launch interface of the gfx950 pair kernels.
//
// All buffers are device pointers in structure-of-arrays form:
//   P3D particles: pv = float4 {x, y, z, volume}, a = float4 {ax, ay, az, 0}
//   P2D particles: float4 {x, y, vorticity, area}
//   F3D filaments: fa = float4 {start x, y, z, strength}, fb = float4 {end x, y, z, 0}
//   points (measurement positions): float4 {x, y, z, 0} or {x, y, 0, 0}
// Outputs are float[3*n] (P3D/F3D), float[2*n] (P2D vel) or float[n] (P2D visc),
// i.e. exactly the caller's bsv_V3f / bsv_V2f / float result layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "device_common.hpp"

namespace cvtx {

// Launch shape chosen by the tuner.
struct LaunchCfg {
  int T = 2;          // targets per lane (1, 2 or 4)
  int wg = 256;       // workgroup size (= source tile length); 256 only for now
  int n_split = 1;    // number of source partitions (each a grid row)
  int split_len = 0;  // sources per partition, a multiple of wg
};

// FP64 epilogue scale factors.
struct Scales {
  double vel = 0, dG = 0, dH = 0, visc = 0;
};

struct P3DJob {
  const float4* src_pv = nullptr;
  const float4* src_a = nullptr;
  int n_src = 0;
  const float4* tgt_pv = nullptr;
  const float4* tgt_a = nullptr;  // may be null when mask == OUT_VEL
  int n_tgt = 0;
  float* out_vel = nullptr;
  float* out_dvort = nullptr;
  float* out_visc = nullptr;
  int reg = REG_WINCKELMANS;
  int mask = OUT_VEL;
  int acc = ACC_COMP;
  float sigma = 1.f;
  float nu = 0.f;
  const float4* gtab = nullptr;  // Gaussian table (device), 2*GTAB_N float4
  double* partial = nullptr;     // workspace for n_split > 1
  const int* tgt_perm = nullptr; // optional target order: slot i handles target tgt_perm[i]
};

struct P2DJob {
  const float4* src = nullptr;
  int n_src = 0;
  const float4* tgt = nullptr;
  int n_tgt = 0;
  float* out_vel = nullptr;   // float[2*n_tgt]
  float* out_visc = nullptr;  // float[n_tgt]
  int reg = REG_WINCKELMANS;
  int mask = OUT_VEL;         // OUT_VEL and/or OUT_VISC
  int acc = ACC_COMP;
  float sigma = 1.f;
  float nu = 0.f;
  double* partial = nullptr;
};

struct F3DJob {
  const float4* fa = nullptr;
  const float4* fb = nullptr;
  int n_fil = 0;
  const float4* tgt_pv = nullptr;
  const float4* tgt_a = nullptr;  // needed for OUT_DVORT
  int n_tgt = 0;
  float* out_vel = nullptr;
  float* out_dvort = nullptr;
  int mask = OUT_VEL;             // OUT_VEL and/or OUT_DVORT
  int acc = ACC_COMP;
  double* partial = nullptr;
  const int* tgt_perm = nullptr;  // spatial order of the targets (order.hip), or identity
};

// Vorticity field (cvtx_P3D_M2M_vort semantics: 5 sigma box cut-off).
struct VortJob {
  const float4* src_pv = nullptr;  // {x, y, z, volume}
  const float4* src_a = nullptr;   // {alpha, 0}
  const int* src_perm = nullptr;   // Morton order of the sources: enables tile skipping
  int n_src = 0;
  const float4* tgt = nullptr;     // {x, y, z, -}
  const int* tgt_perm = nullptr;   // Morton order of the targets (nullptr: identity)
  int n_tgt = 0;
  float* out = nullptr;            // float[3 n_tgt]
  int reg = REG_WINCKELMANS;
  int acc = ACC_COMP;
  float sigma = 1.f;
  bool cutoff = true;              // false: no 5 sigma box (single-source forms)
};

// Redistribution on a grid (remeshing).  dim 3: pv = {x, y, z, volume},
// a = {alpha, -}; dim 2: pv = {x, y, alpha, area}, a unused.  rid 0..4 =
// lambda0, lambda1, lambda2, lambda3, m4'.  max_out < 0: count only.
struct RedistJob {
  int dim = 3;
  int rid = 4;
  const float4* pv = nullptr;
  const float4* a = nullptr;
  int n = 0;
  float h = 1.f;
  float negligible = 0.f;
  int max_out = -1;
};
float redist_radius(int rid);

// number of FP64 partial accumulators per target for a mask
int p3d_nacc(int mask);
int p2d_nacc(int mask);
int f3d_nacc(int mask);

// Workspace bytes needed for a launch shape (0 when n_split == 1).
size_t p3d_workspace_bytes(const P3DJob& j, const LaunchCfg& c);
size_t p2d_workspace_bytes(const P2DJob& j, const LaunchCfg& c);
size_t f3d_workspace_bytes(const F3DJob& j, const LaunchCfg& c);

// Default launch shape for a problem size (the tuner may override).
LaunchCfg default_cfg(int n_src, int n_tgt, int n_cu);
// The default source partitions for T targets per lane.
LaunchCfg partitioned_cfg(int T, int n_src, int n_tgt, int n_cu);

// Enqueue the kernels on `s`.  Return hipSuccess or the launch error.
// Valid combinations: P3D visc only with Winckelmans or Gaussian.
hipError_t launch_p3d(const P3DJob& j, const LaunchCfg& c, hipStream_t s);
hipError_t launch_p2d(const P2DJob& j, const LaunchCfg& c, hipStream_t s);
hipError_t launch_f3d(const F3DJob& j, const LaunchCfg& c, hipStream_t s);
// F3D influence matrix: out[i * n_fil + j] = vel(filament j, mes_i) . dir_i
// (fa, fb as in F3DJob; mes, dir float4 {x, y, z, -}).
hipError_t launch_f3d_inf(const float4* fa, const float4* fb, int n_fil, const float4* mes,
                          const float4* dir, int n_mes, float* out, hipStream_t s);
// Vorticity field: workspace of vort_workspace_bytes (source partitions;
// sorted source copies and tile boxes when j.src_perm is set).
size_t vort_workspace_bytes(const VortJob& j, const LaunchCfg& c, bool sorted);
hipError_t launch_vort(const VortJob& j, const LaunchCfg& c, void* ws, size_t ws_bytes,
                       hipStream_t s);
double vort_scale(int reg, float sigma);
// Default shape of a vorticity-field launch (sorted: with tile skipping).
LaunchCfg vort_default_cfg(int n_src, int n_tgt, int n_cu, bool sorted);
// Pedrizzetti update of a[n] (float4 {alpha, w}) from omega[3 n], in place.
hipError_t launch_relax(float4* a, const float* omega, int n, float fdt, hipStream_t s);

// Host-side helpers.
Scales p3d_scales(int reg, float sigma, float nu);
RegConsts p3d_consts(int reg, float sigma);
bool p3d_supported(int reg, int mask);
bool p2d_supported(int reg, int mask);

// Gaussian near-field table (2*GTAB_N float4): cubic pieces of G and H/3,
// computed in FP64 on the host.
void build_gaussian_table(float* out /* 8*GTAB_N floats */);

// Spatial (Morton) order of n points: perm[i] = index of the i-th point.
// Workspace: target_order_workspace(n) bytes (device).
size_t target_order_workspace(int n);
hipError_t target_order(const float4* pts, int n, int* perm, void* ws, size_t ws_bytes,
                        hipStream_t s);

// Whether ordering the targets pays for a problem (a kernel with a
// near-field-only path and enough targets to amortise the sort).
// Sort sources and targets of a vorticity-field evaluation (tile skipping).
inline bool vort_wants_sort(int n_src, int n_tgt) {
  return n_src >= 4096 && n_tgt >= 1024 && (long long)n_src * n_tgt >= (1LL << 26);
}

// Gaussian ring, ms per vel+dvort evaluation, ordered -> unordered:
// 8192: 0.113 -> 0.101; 16384: 0.255 -> 0.304; 32768: 0.85 -> 1.12.
inline bool p3d_wants_order(int reg, int n_src, int n_tgt) {
  return reg == REG_GAUSSIAN && n_tgt >= 4096 && (long long)n_src * n_tgt > (1LL << 26);
}

// Filament kernels: spatially ordered targets make a block's box small, so
// most filament tiles are far from it and take the far-tile loop (f3d.hip).
inline bool f3d_wants_order(int n_fil, int n_tgt) {
  return n_tgt >= 4096 && (long long)n_fil * n_tgt >= (1LL << 26);
}

}  // namespace cvtx
