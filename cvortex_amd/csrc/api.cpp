// api.cpp -- the C ABI of libcvortex.so: accelerator controls, the public
// cvtx_* M2M functions, the int-returning HIP seam and the device-resident
// API.  Declarations: include/cvortex/libcvtx.h, include/cvortex_hip.h.
//
// Host-array calls follow one path: gather the caller's pointer arrays into
// pinned structure-of-arrays staging (parallel over host threads), copy to
// the device on the library stream, run the kernels, copy the results back
// and synchronise.  The seam never evaluates a pair on the CPU: the public
// void functions (dispatch.cpp) do that when the seam declines.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>

#include <cvortex/libcvtx.h>
#include <cvortex_hip.h>

#include "kernels.hpp"
#include "runtime.hpp"
#include "api_common.hpp"
#include "tune.hpp"

using namespace cvtx;

namespace cvtx {

// ---------------- host gathers: AoP records -> SoA float4 ----------------

void gather_p3d(const cvtx_P3D* const* src, int n, float4* pv, float4* a) {
  parallel_for(n, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      const cvtx_P3D* p = src[i];
      pv[i] = make_float4(p->coord.x[0], p->coord.x[1], p->coord.x[2], p->volume);
      a[i] = make_float4(p->vorticity.x[0], p->vorticity.x[1], p->vorticity.x[2], 0.f);
    }
  });
}

void gather_pts3(const bsv_V3f* pts, int n, float4* pv) {
  parallel_for(n, [&](long b, long e) {
    for (long i = b; i < e; ++i) pv[i] = make_float4(pts[i].x[0], pts[i].x[1], pts[i].x[2], 0.f);
  });
}

void gather_p2d(const cvtx_P2D* const* src, int n, float4* out) {
  parallel_for(n, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      const cvtx_P2D* p = src[i];
      out[i] = make_float4(p->coord.x[0], p->coord.x[1], p->vorticity, p->area);
    }
  });
}

void gather_pts2(const bsv_V2f* pts, int n, float4* out) {
  parallel_for(n, [&](long b, long e) {
    for (long i = b; i < e; ++i) out[i] = make_float4(pts[i].x[0], pts[i].x[1], 0.f, 0.f);
  });
}

void gather_f3d(const cvtx_F3D* const* src, int n, float4* fa, float4* fb) {
  parallel_for(n, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      const cvtx_F3D* f = src[i];
      fa[i] = make_float4(f->start.x[0], f->start.x[1], f->start.x[2], f->strength);
      fb[i] = make_float4(f->end.x[0], f->end.x[1], f->end.x[2], 0.f);
    }
  });
}

void fill_nan(float* out, long n) {
  if (!out) return;
  const float q = std::numeric_limits<float>::quiet_NaN();
  for (long i = 0; i < n; ++i) out[i] = q;
}

void warn_no_eta_once() {
  static bool warned = false;
  if (!warned) {
    std::fprintf(stderr,
                 "cvortex: viscous vorticity exchange requested with a regularisation that has "
                 "no eta function (singular or planetary); modelled as inviscid (zero).\n");
    warned = true;
  }
}

bool bad_counts(Runtime& rt, int ns, int nt, const void* src, const void* tgt) {
  if (ns < 0 || nt < 0) {
    set_error(rt, "negative element count");
    return true;
  }
  if ((ns > 0 && src == nullptr) || (nt > 0 && tgt == nullptr)) {
    set_error(rt, "null input array");
    return true;
  }
  return false;
}

// Copy `bytes` from pinned staging to device and back helpers.
bool h2d(Runtime& rt, Device& d, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return true;
  return check_hip(rt, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, d.stream), "H2D copy");
}
bool d2h(Runtime& rt, Device& d, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return true;
  return check_hip(rt, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, d.stream), "D2H copy");
}

void report_failure(const char* fn, float* out, long n) {
  std::fprintf(stderr, "cvortex: %s failed: %s (result set to NaN)\n", fn,
               runtime().last_error.c_str());
  fill_nan(out, n);
}

hipStream_t pick_stream(Device* d, void* s) { return s ? (hipStream_t)s : d->stream; }

bool make_launch_shape(int T, int n_split, int n_src, LaunchCfg* c) {
  if (T != 1 && T != 2 && T != 4) return false;
  if (n_split < 1) return false;
  *c = launch_shape(T, n_split, n_src);
  return true;
}

int* order_points(Runtime& rt, Device& d, const float4* pts, int n, int slot, hipStream_t s) {
  const size_t wb = target_order_workspace(n);
  void* ws = device_buffer(rt, d, SLOT_ORDER, wb);
  int* perm = (int*)device_buffer(rt, d, slot, sizeof(int) * (size_t)n);
  if (!ws || !perm) return nullptr;
  if (!check_hip(rt, target_order(pts, n, perm, ws, wb, s), "spatial order")) return nullptr;
  return perm;
}

}  // namespace cvtx

namespace {

using Lock = std::lock_guard<std::recursive_mutex>;

bool maybe_order(Runtime& rt, Device& d, P3DJob& j, hipStream_t s);

// ---------------------------- P3D host path ----------------------------

int p3d_host(const cvtx_P3D** src, int ns, const bsv_V3f* mes, const cvtx_P3D** ind, int nt,
             float* vel, float* dvort, float* visc, const cvtx_VortFunc* vf, float sigma, float nu) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (vf == nullptr) {
    set_error(rt, "null regularisation");
    return -1;
  }
  const int reg = reg_from_name(vf->cl_kernel_name_ext);
  if (reg < 0) {
    set_error(rt, "regularisation name is not one of singular, winckelmans, planetary, gaussian");
    return -1;
  }
  if (bad_counts(rt, ns, nt, src, mes ? (const void*)mes : (const void*)ind)) return -1;
  int mask = (vel ? OUT_VEL : 0) | (dvort ? OUT_DVORT : 0) | (visc ? OUT_VISC : 0);
  if ((mask & OUT_VISC) && !p3d_supported(reg, OUT_VISC)) {
    warn_no_eta_once();
    std::memset(visc, 0, sizeof(float) * 3 * (size_t)nt);
    mask &= ~OUT_VISC;
  }
  if (mask == 0 || nt == 0) return 0;
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);

  const size_t ob = sizeof(float) * 3 * (size_t)nt;
  const bool need_ta = (mask & (OUT_DVORT | OUT_VISC)) != 0;
  // sources {pv, a} and targets {pv, a} in one staging buffer, so that one
  // copy moves them (each copy costs microseconds of latency, which is most
  // of a small problem's call)
  const size_t ns1 = (size_t)(ns > 0 ? ns : 1);
  const size_t n4 = 2 * ns1 + (size_t)nt * (need_ta ? 2 : 1);
  float4* h_in = (float4*)host_buffer(rt, *d, SLOT_SRC0, sizeof(float4) * n4);
  float4* d_in = (float4*)device_buffer(rt, *d, SLOT_SRC0, sizeof(float4) * n4);
  if (!h_in || !d_in) return -1;
  float4* h_spv = h_in;
  float4* h_sa = h_in + ns1;
  float4* h_tpv = h_in + 2 * ns1;
  float4* h_ta = need_ta ? h_tpv + nt : nullptr;
  float4* d_spv = d_in;
  float4* d_sa = d_in + ns1;
  float4* d_tpv = d_in + 2 * ns1;
  float4* d_ta = need_ta ? d_tpv + nt : nullptr;
  // the selected outputs side by side, so that one copy brings them back
  const int n_out = ((mask & OUT_VEL) ? 1 : 0) + ((mask & OUT_DVORT) ? 1 : 0) + ((mask & OUT_VISC) ? 1 : 0);
  float* h_outs = (float*)host_buffer(rt, *d, SLOT_OUT0, ob * n_out);
  float* d_outs = (float*)device_buffer(rt, *d, SLOT_OUT0, ob * n_out);
  if (!h_outs || !d_outs) return -1;
  float* h_out[3] = {nullptr, nullptr, nullptr};
  float* d_out[3] = {nullptr, nullptr, nullptr};
  float* user_out[3] = {vel, dvort, visc};
  for (int o = 0, k = 0; o < 3; ++o) {
    if (!(mask & (1 << o))) continue;
    h_out[o] = h_outs + (size_t)k * 3 * nt;
    d_out[o] = d_outs + (size_t)k * 3 * nt;
    ++k;
  }

  gather_p3d(src, ns, h_spv, h_sa);
  if (mes) {
    gather_pts3(mes, nt, h_tpv);
  } else {
    gather_p3d(ind, nt, h_tpv, h_ta);
  }
  if (!h2d(rt, *d, d_in, h_in, sizeof(float4) * n4)) return -1;

  P3DJob j;
  j.src_pv = d_spv;
  j.src_a = d_sa;
  j.n_src = ns;
  j.tgt_pv = d_tpv;
  j.tgt_a = d_ta;
  j.n_tgt = nt;
  j.out_vel = d_out[0];
  j.out_dvort = d_out[1];
  j.out_visc = d_out[2];
  j.reg = reg;
  j.mask = mask;
  j.acc = rt.acc_mode;
  j.sigma = sigma;
  j.nu = nu;
  j.gtab = d->gtab;
  LaunchCfg c = choose_cfg(KIND_P3D, reg, mask, j.acc, ns, nt, d->n_cu);
  const size_t wb = p3d_workspace_bytes(j, c);
  if (wb) {
    j.partial = (double*)device_buffer(rt, *d, SLOT_WORK, wb);
    if (!j.partial) return -1;
  }
  if (!maybe_order(rt, *d, j, d->stream)) return -1;
  if (!check_hip(rt, launch_p3d(j, c, d->stream), "P3D kernel launch")) return -1;
  if (!d2h(rt, *d, h_outs, d_outs, ob * n_out)) return -1;
  if (!check_hip(rt, hipStreamSynchronize(d->stream), "P3D kernels")) return -1;
  for (int o = 0; o < 3; ++o)
    if (d_out[o]) std::memcpy(user_out[o], h_out[o], ob);
  return 0;
}

// ---------------------------- F3D host path ----------------------------

int f3d_host(const cvtx_F3D** src, int nf, const bsv_V3f* mes, const cvtx_P3D** ind, int nt,
             float* vel, float* dvort) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (bad_counts(rt, nf, nt, src, mes ? (const void*)mes : (const void*)ind)) return -1;
  const int mask = (vel ? OUT_VEL : 0) | (dvort ? OUT_DVORT : 0);
  if (mask == 0 || nt == 0) return 0;
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);
  const size_t ob = sizeof(float) * 3 * (size_t)nt;
  const bool need_ta = (mask & OUT_DVORT) != 0;
  // filaments and targets in one staging buffer, the outputs side by side:
  // one copy each way (as the P3D path)
  const size_t nf1 = (size_t)(nf > 0 ? nf : 1);
  const size_t n4 = 2 * nf1 + (size_t)nt * (need_ta ? 2 : 1);
  const int n_out = ((mask & OUT_VEL) ? 1 : 0) + ((mask & OUT_DVORT) ? 1 : 0);
  float4* h_in = (float4*)host_buffer(rt, *d, SLOT_SRC0, sizeof(float4) * n4);
  float4* d_in = (float4*)device_buffer(rt, *d, SLOT_SRC0, sizeof(float4) * n4);
  float* h_outs = (float*)host_buffer(rt, *d, SLOT_OUT0, ob * n_out);
  float* d_outs = (float*)device_buffer(rt, *d, SLOT_OUT0, ob * n_out);
  if (!h_in || !d_in || !h_outs || !d_outs) return -1;
  float4* h_fa = h_in;
  float4* h_fb = h_in + nf1;
  float4* h_tpv = h_in + 2 * nf1;
  float4* h_ta = need_ta ? h_tpv + nt : nullptr;
  float4* d_fa = d_in;
  float4* d_fb = d_in + nf1;
  float4* d_tpv = d_in + 2 * nf1;
  float4* d_ta = need_ta ? d_tpv + nt : nullptr;
  float* h_out[2] = {nullptr, nullptr};
  float* d_out[2] = {nullptr, nullptr};
  float* user_out[2] = {vel, dvort};
  for (int o = 0, k = 0; o < 2; ++o) {
    if (!(mask & (1 << o))) continue;
    h_out[o] = h_outs + (size_t)k * 3 * nt;
    d_out[o] = d_outs + (size_t)k * 3 * nt;
    ++k;
  }
  gather_f3d(src, nf, h_fa, h_fb);
  if (mes) {
    gather_pts3(mes, nt, h_tpv);
  } else {
    gather_p3d(ind, nt, h_tpv, h_ta);
  }
  if (!h2d(rt, *d, d_in, h_in, sizeof(float4) * n4)) return -1;
  F3DJob j;
  j.fa = d_fa;
  j.fb = d_fb;
  j.n_fil = nf;
  j.tgt_pv = d_tpv;
  j.tgt_a = d_ta;
  j.n_tgt = nt;
  j.out_vel = d_out[0];
  j.out_dvort = d_out[1];
  j.mask = mask;
  j.acc = rt.acc_mode;
  LaunchCfg c = choose_cfg(KIND_F3D, 0, mask, j.acc, nf, nt, d->n_cu);
  const size_t wb = f3d_workspace_bytes(j, c);
  if (wb) {
    j.partial = (double*)device_buffer(rt, *d, SLOT_WORK, wb);
    if (!j.partial) return -1;
  }
  if (rt.order_targets && f3d_wants_order(nf, nt)) {
    j.tgt_perm = order_points(rt, *d, d_tpv, nt, SLOT_PERM, d->stream);
    if (!j.tgt_perm) return -1;
  }
  if (!check_hip(rt, launch_f3d(j, c, d->stream), "F3D kernel launch")) return -1;
  if (!d2h(rt, *d, h_outs, d_outs, ob * n_out)) return -1;
  if (!check_hip(rt, hipStreamSynchronize(d->stream), "F3D kernels")) return -1;
  for (int o = 0; o < 2; ++o)
    if (d_out[o]) std::memcpy(user_out[o], h_out[o], ob);
  return 0;
}

// ---------------------------- P2D host path ----------------------------

int p2d_host(const cvtx_P2D** src, int ns, const bsv_V2f* mes, const cvtx_P2D** ind, int nt,
             float* vel, float* visc, const cvtx_VortFunc* vf, float sigma, float nu) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (vf == nullptr) {
    set_error(rt, "null regularisation");
    return -1;
  }
  const int reg = reg_from_name(vf->cl_kernel_name_ext);
  if (reg < 0) {
    set_error(rt, "regularisation name is not one of singular, winckelmans, planetary, gaussian");
    return -1;
  }
  if (bad_counts(rt, ns, nt, src, mes ? (const void*)mes : (const void*)ind)) return -1;
  int mask = (vel ? OUT_VEL : 0) | (visc ? OUT_VISC : 0);
  if ((mask & OUT_VISC) && !p2d_supported(reg, OUT_VISC)) {
    warn_no_eta_once();
    std::memset(visc, 0, sizeof(float) * (size_t)nt);
    mask &= ~OUT_VISC;
  }
  if (mask == 0 || nt == 0) return 0;
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);
  // sources and targets in one staging buffer, the outputs side by side:
  // one copy each way (as the P3D path)
  const size_t ns1 = (size_t)(ns > 0 ? ns : 1);
  const size_t n4 = ns1 + (size_t)nt;
  const size_t vb = (mask & OUT_VEL) ? sizeof(float) * 2 * (size_t)nt : 0;
  const size_t eb = (mask & OUT_VISC) ? sizeof(float) * (size_t)nt : 0;
  float4* h_in = (float4*)host_buffer(rt, *d, SLOT_SRC0, sizeof(float4) * n4);
  float4* d_in = (float4*)device_buffer(rt, *d, SLOT_SRC0, sizeof(float4) * n4);
  float* h_outs = (float*)host_buffer(rt, *d, SLOT_OUT0, vb + eb);
  float* d_outs = (float*)device_buffer(rt, *d, SLOT_OUT0, vb + eb);
  if (!h_in || !d_in || !h_outs || !d_outs) return -1;
  float4* h_s = h_in;
  float4* h_t = h_in + ns1;
  float4* d_s = d_in;
  float4* d_t = d_in + ns1;
  float* h_v = vb ? h_outs : nullptr;
  float* d_v = vb ? d_outs : nullptr;
  float* h_e = eb ? h_outs + vb / sizeof(float) : nullptr;
  float* d_e = eb ? d_outs + vb / sizeof(float) : nullptr;
  gather_p2d(src, ns, h_s);
  if (mes) {
    gather_pts2(mes, nt, h_t);
  } else {
    gather_p2d(ind, nt, h_t);
  }
  if (!h2d(rt, *d, d_in, h_in, sizeof(float4) * n4)) return -1;
  P2DJob j;
  j.src = d_s;
  j.n_src = ns;
  j.tgt = d_t;
  j.n_tgt = nt;
  j.out_vel = d_v;
  j.out_visc = d_e;
  j.reg = reg;
  j.mask = mask;
  j.acc = rt.acc_mode;
  j.sigma = sigma;
  j.nu = nu;
  LaunchCfg c = choose_cfg(KIND_P2D, reg, mask, j.acc, ns, nt, d->n_cu);
  const size_t wb = p2d_workspace_bytes(j, c);
  if (wb) {
    j.partial = (double*)device_buffer(rt, *d, SLOT_WORK, wb);
    if (!j.partial) return -1;
  }
  if (!check_hip(rt, launch_p2d(j, c, d->stream), "P2D kernel launch")) return -1;
  if (!d2h(rt, *d, h_outs, d_outs, vb + eb)) return -1;
  if (!check_hip(rt, hipStreamSynchronize(d->stream), "P2D kernels")) return -1;
  if (d_v) std::memcpy(vel, h_v, vb);
  if (d_e) std::memcpy(visc, h_e, eb);
  return 0;
}

// Spatially order a P3D job's targets when its kernel has a near-field-only
// path that coherent waves can skip (see order.hip).
bool maybe_order(Runtime& rt, Device& d, P3DJob& j, hipStream_t s) {
  if (!rt.order_targets || !p3d_wants_order(j.reg, j.n_src, j.n_tgt)) return true;
  j.tgt_perm = order_points(rt, d, j.tgt_pv, j.n_tgt, SLOT_PERM, s);
  return j.tgt_perm != nullptr;
}

}  // namespace

// =========================== accelerator controls ===========================

extern "C" {

void cvtx_initialise() {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  ready_enabled_devices(rt);
}

void cvtx_finalise() {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  finalise_locked(rt);
}

const char* cvtx_information() {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  return rt.info.c_str();
}

int cvtx_num_accelerators() {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  return (int)rt.devices.size();
}

int cvtx_num_enabled_accelerators() {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  int n = 0;
  for (auto& d : rt.devices) n += d.enabled ? 1 : 0;
  return n;
}

const char* cvtx_accelerator_name(int id) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  if (id < 0 || id >= (int)rt.devices.size()) return nullptr;
  return rt.devices[id].name.c_str();
}

int cvtx_accelerator_enabled(int id) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  if (id < 0 || id >= (int)rt.devices.size()) return 0;
  return rt.devices[id].enabled ? 1 : 0;
}

void cvtx_accelerator_enable(int id) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  if (id >= 0 && id < (int)rt.devices.size()) rt.devices[id].enabled = true;
  ready_enabled_devices(rt);
}

void cvtx_accelerator_disable(int id) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  initialise_locked(rt);
  if (id >= 0 && id < (int)rt.devices.size()) rt.devices[id].enabled = false;
}

// ================================ seam ================================

int cvtx_hip_P3D_M2M_vel(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nm, bsv_V3f* out,
                         const cvtx_VortFunc* k, float sigma) {
  if (nm > 0 && out == nullptr) return -1;
  return p3d_host(src, ns, mes, nullptr, nm, (float*)out, nullptr, nullptr, k, sigma, 0.f);
}

int cvtx_hip_P3D_M2M_dvort(const cvtx_P3D** src, int ns, const cvtx_P3D** ind, int ni,
                           bsv_V3f* out, const cvtx_VortFunc* k, float sigma) {
  if (ni > 0 && out == nullptr) return -1;
  return p3d_host(src, ns, nullptr, ind, ni, nullptr, (float*)out, nullptr, k, sigma, 0.f);
}

int cvtx_hip_P3D_M2M_visc_dvort(const cvtx_P3D** src, int ns, const cvtx_P3D** ind, int ni,
                                bsv_V3f* out, const cvtx_VortFunc* k, float sigma, float nu) {
  if (ni > 0 && out == nullptr) return -1;
  return p3d_host(src, ns, nullptr, ind, ni, nullptr, nullptr, (float*)out, k, sigma, nu);
}

int cvtx_hip_P3D_M2M_fused(const cvtx_P3D** src, int ns, const cvtx_P3D** ind, int ni,
                           bsv_V3f* vel, bsv_V3f* dvort, bsv_V3f* visc, const cvtx_VortFunc* k,
                           float sigma, float nu) {
  return p3d_host(src, ns, nullptr, ind, ni, (float*)vel, (float*)dvort, (float*)visc, k, sigma, nu);
}

int cvtx_hip_F3D_M2M_vel(const cvtx_F3D** src, int nf, const bsv_V3f* mes, int nm, bsv_V3f* out) {
  if (nm > 0 && out == nullptr) return -1;
  return f3d_host(src, nf, mes, nullptr, nm, (float*)out, nullptr);
}

int cvtx_hip_F3D_M2M_dvort(const cvtx_F3D** src, int nf, const cvtx_P3D** ind, int ni,
                           bsv_V3f* out) {
  if (ni > 0 && out == nullptr) return -1;
  return f3d_host(src, nf, nullptr, ind, ni, nullptr, (float*)out);
}

int cvtx_hip_F3D_M2M_fused(const cvtx_F3D** src, int nf, const cvtx_P3D** ind, int ni,
                           bsv_V3f* vel, bsv_V3f* dvort) {
  return f3d_host(src, nf, nullptr, ind, ni, (float*)vel, (float*)dvort);
}

int cvtx_hip_P2D_M2M_vel(const cvtx_P2D** src, int ns, const bsv_V2f* mes, int nm, bsv_V2f* out,
                         const cvtx_VortFunc* k, float sigma) {
  if (nm > 0 && out == nullptr) return -1;
  return p2d_host(src, ns, mes, nullptr, nm, (float*)out, nullptr, k, sigma, 0.f);
}

int cvtx_hip_P2D_M2M_visc_dvort(const cvtx_P2D** src, int ns, const cvtx_P2D** ind, int ni,
                                float* out, const cvtx_VortFunc* k, float sigma, float nu) {
  if (ni > 0 && out == nullptr) return -1;
  return p2d_host(src, ns, nullptr, ind, ni, nullptr, out, k, sigma, nu);
}

int cvtx_hip_P2D_M2M_fused(const cvtx_P2D** src, int ns, const cvtx_P2D** ind, int ni,
                           bsv_V2f* vel, float* visc, const cvtx_VortFunc* k, float sigma,
                           float nu) {
  return p2d_host(src, ns, nullptr, ind, ni, (float*)vel, visc, k, sigma, nu);
}

// ============================== controls ==============================

const char* cvtx_hip_last_error(void) { return runtime().last_error.c_str(); }

int cvtx_hip_set_accumulation(int mode) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (!valid_acc(mode)) return -1;
  const int prev = rt.acc_mode;
  rt.acc_mode = mode;
  return prev;
}

int cvtx_hip_get_accumulation(void) { return runtime().acc_mode; }

int cvtx_hip_set_target_order(int enable) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  const int prev = rt.order_targets ? 1 : 0;
  rt.order_targets = enable != 0;
  return prev;
}

void* cvtx_hip_stream(void) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  Device* d = active_device(rt);
  return d ? (void*)d->stream : nullptr;
}

int cvtx_hip_synchronize(void) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  Device* d = active_device(rt);
  if (!d) return -1;
  return check_hip(rt, hipStreamSynchronize(d->stream), "synchronize") ? 0 : -1;
}

// ========================= device-resident API =========================

}  // extern "C"

namespace cvtx {

bool p3d_job_from_problem(Runtime& rt, const Device& d, const cvtx_hip_p3d_problem* p, P3DJob* out) {
  if (!p) {
    set_error(rt, "null problem");
    return false;
  }
  P3DJob j;
  j.src_pv = (const float4*)p->src_pv;
  j.src_a = (const float4*)p->src_a;
  j.n_src = p->n_src;
  j.tgt_pv = (const float4*)p->tgt_pv;
  j.tgt_a = (const float4*)p->tgt_a;
  j.n_tgt = p->n_tgt;
  j.out_vel = p->vel;
  j.out_dvort = p->dvort;
  j.out_visc = p->visc;
  j.reg = p->regularisation;
  j.mask = (p->vel ? OUT_VEL : 0) | (p->dvort ? OUT_DVORT : 0) | (p->visc ? OUT_VISC : 0);
  j.acc = p->accumulate;
  j.sigma = p->sigma;
  j.nu = p->nu;
  j.gtab = d.gtab;
  if (j.n_src < 0 || j.n_tgt < 0 || (j.n_src > 0 && (!j.src_pv || !j.src_a)) ||
      (j.n_tgt > 0 && !j.tgt_pv)) {
    set_error(rt, "invalid device problem");
    return false;
  }
  if (!valid_acc(j.acc)) {
    set_error(rt, "invalid accumulation mode");
    return false;
  }
  if (!p3d_supported(j.reg, j.mask)) {
    set_error(rt, "unsupported regularisation/output combination");
    return false;
  }
  *out = j;
  return true;
}

}  // namespace cvtx

extern "C" {

static int p3d_dev(const cvtx_hip_p3d_problem* p, const LaunchCfg* forced, void* stream) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  Device* d = active_device(rt);
  if (!d) return -1;
  P3DJob j;
  if (!p3d_job_from_problem(rt, *d, p, &j)) return -1;
  StreamUse use(*d, pick_stream(d, stream));
  LaunchCfg c = forced ? *forced : choose_cfg(KIND_P3D, j.reg, j.mask, j.acc, j.n_src, j.n_tgt, d->n_cu);
  const size_t wb = p3d_workspace_bytes(j, c);
  if (wb) {
    j.partial = (double*)device_buffer(rt, *d, SLOT_WORK, wb);
    if (!j.partial) return -1;
  }
  if (!maybe_order(rt, *d, j, pick_stream(d, stream))) return -1;
  return check_hip(rt, launch_p3d(j, c, pick_stream(d, stream)), "P3D launch") ? 0 : -1;
}

static int p2d_dev(const cvtx_hip_p2d_problem* p, const LaunchCfg* forced, void* stream) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (!p) {
    set_error(rt, "null problem");
    return -1;
  }
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, pick_stream(d, stream));
  P2DJob j;
  j.src = (const float4*)p->src;
  j.n_src = p->n_src;
  j.tgt = (const float4*)p->tgt;
  j.n_tgt = p->n_tgt;
  j.out_vel = p->vel;
  j.out_visc = p->visc;
  j.reg = p->regularisation;
  j.mask = (p->vel ? OUT_VEL : 0) | (p->visc ? OUT_VISC : 0);
  j.acc = p->accumulate;
  j.sigma = p->sigma;
  j.nu = p->nu;
  if (j.n_src < 0 || j.n_tgt < 0 || (j.n_src > 0 && !j.src) || (j.n_tgt > 0 && !j.tgt)) {
    set_error(rt, "invalid device problem");
    return -1;
  }
  if (!valid_acc(j.acc) || !p2d_supported(j.reg, j.mask)) {
    set_error(rt, "unsupported regularisation/output/accumulation combination");
    return -1;
  }
  LaunchCfg c = forced ? *forced : choose_cfg(KIND_P2D, j.reg, j.mask, j.acc, j.n_src, j.n_tgt, d->n_cu);
  const size_t wb = p2d_workspace_bytes(j, c);
  if (wb) {
    j.partial = (double*)device_buffer(rt, *d, SLOT_WORK, wb);
    if (!j.partial) return -1;
  }
  return check_hip(rt, launch_p2d(j, c, pick_stream(d, stream)), "P2D launch") ? 0 : -1;
}

static int f3d_dev(const cvtx_hip_f3d_problem* p, const LaunchCfg* forced, void* stream) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (!p) {
    set_error(rt, "null problem");
    return -1;
  }
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, pick_stream(d, stream));
  F3DJob j;
  j.fa = (const float4*)p->fa;
  j.fb = (const float4*)p->fb;
  j.n_fil = p->n_fil;
  j.tgt_pv = (const float4*)p->tgt_pv;
  j.tgt_a = (const float4*)p->tgt_a;
  j.n_tgt = p->n_tgt;
  j.out_vel = p->vel;
  j.out_dvort = p->dvort;
  j.mask = (p->vel ? OUT_VEL : 0) | (p->dvort ? OUT_DVORT : 0);
  j.acc = p->accumulate;
  if (j.n_fil < 0 || j.n_tgt < 0 || (j.n_fil > 0 && (!j.fa || !j.fb)) || (j.n_tgt > 0 && !j.tgt_pv) ||
      j.mask == 0 || !valid_acc(j.acc)) {
    set_error(rt, "invalid device problem");
    return -1;
  }
  LaunchCfg c = forced ? *forced : choose_cfg(KIND_F3D, 0, j.mask, j.acc, j.n_fil, j.n_tgt, d->n_cu);
  const size_t wb = f3d_workspace_bytes(j, c);
  if (wb) {
    j.partial = (double*)device_buffer(rt, *d, SLOT_WORK, wb);
    if (!j.partial) return -1;
  }
  if (rt.order_targets && f3d_wants_order(j.n_fil, j.n_tgt)) {
    j.tgt_perm = order_points(rt, *d, j.tgt_pv, j.n_tgt, SLOT_PERM, pick_stream(d, stream));
    if (!j.tgt_perm) return -1;
  }
  return check_hip(rt, launch_f3d(j, c, pick_stream(d, stream)), "F3D launch") ? 0 : -1;
}


int cvtx_hip_p3d_eval(const cvtx_hip_p3d_problem* p, void* stream) { return p3d_dev(p, nullptr, stream); }
int cvtx_hip_p2d_eval(const cvtx_hip_p2d_problem* p, void* stream) { return p2d_dev(p, nullptr, stream); }
int cvtx_hip_f3d_eval(const cvtx_hip_f3d_problem* p, void* stream) { return f3d_dev(p, nullptr, stream); }

int cvtx_hip_p3d_eval_shape(const cvtx_hip_p3d_problem* p, int T, int n_split, void* stream) {
  LaunchCfg c;
  if (!p || !make_launch_shape(T, n_split, p->n_src, &c)) {
    set_error(runtime(), "invalid launch shape");
    return -1;
  }
  return p3d_dev(p, &c, stream);
}

int cvtx_hip_p2d_eval_shape(const cvtx_hip_p2d_problem* p, int T, int n_split, void* stream) {
  LaunchCfg c;
  if (!p || !make_launch_shape(T, n_split, p->n_src, &c)) {
    set_error(runtime(), "invalid launch shape");
    return -1;
  }
  return p2d_dev(p, &c, stream);
}

int cvtx_hip_f3d_eval_shape(const cvtx_hip_f3d_problem* p, int T, int n_split, void* stream) {
  LaunchCfg c;
  if (!p || !make_launch_shape(T, n_split, p->n_fil, &c)) {
    set_error(runtime(), "invalid launch shape");
    return -1;
  }
  return f3d_dev(p, &c, stream);
}

}  // extern "C"
