// dispatch.cpp -- the public void cvtx_* functions: device or host leg.
//
// Like the reference (src/P3D.cpp:343-498, src/F3D.cpp:162-202,
// src/P2D.cpp:141-276), each public function first offers the call to the
// accelerator seam (cvtx_hip_*, api.cpp) and evaluates it on the CPU
// (host_leg.cpp) when the seam is not asked or declines:
//   * no device is enabled (src/accelerators.cpp:107-118);
//   * the regularisation's cl_kernel_name_ext is not one of the four kernel
//     names, e.g. a caller's own cvtx_VortFunc, which then runs through its
//     function pointers (include/cvortex/libcvtx.h:74-94);
//   * the problem is below the reference's size guards:
//       P3D vel, dvort, visc_dvort, vort; F3D dvort; P2D visc_dvort:
//         both counts >= 256 (src/P3D.cpp:353-355, src/F3D.cpp:190-191,
//         src/P2D.cpp:263-265);
//       F3D vel: n_f n_m >= n_f n_m / 20 + 5000 + 300 n_m (src/ocl_F3D.cpp:49-51);
//       P2D vel: n_s n_m > 20000 + n_s n_m / 20 (src/ocl_P2D.cpp:54-57);
//   * the device call fails (then a line goes to stderr).
// cvtx_hip_set_dispatch() replaces the size guards: DEVICE sends every call
// the seam can take to the device, HOST sends every call to the CPU.
// cvtx_hip_last_leg() tells which leg evaluated the last call.
#include <cstdio>
#include <cstring>
#include <string>

#include <cvortex/libcvtx.h>
#include <cvortex_hip.h>

#include "api_common.hpp"
#include "host_leg.hpp"
#include "runtime.hpp"

using namespace cvtx;

namespace cvtx {

namespace {

using Lock = std::lock_guard<std::recursive_mutex>;

bool any_enabled(Runtime& rt) {
  initialise_locked(rt);
  for (auto& d : rt.devices)
    if (d.enabled) return true;
  return false;
}

}  // namespace

bool want_device(bool guard_met, const cvtx_VortFunc* vf) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  rt.last_leg = -1;
  if (rt.dispatch == DISPATCH_HOST) return false;
  if (!any_enabled(rt)) return false;
  if (vf != nullptr && reg_from_name(vf->cl_kernel_name_ext) < 0) return false;
  return rt.dispatch == DISPATCH_DEVICE || guard_met;
}

void device_done() { runtime().last_leg = CVTX_HIP_LEG_DEVICE; }

void device_declined(const char* fn) {
  std::fprintf(stderr, "cvortex: %s on the device failed (%s); evaluating on the host\n", fn,
               runtime().last_error.c_str());
}

bool host_args_ok(const char* fn, bool ok, float* out, long n_out) {
  if (ok) return true;
  set_error(runtime(), "invalid arguments");
  report_failure(fn, out, n_out);
  return false;
}

void host_done(const char* fn, int status, const std::string& err, float* out, long n_out) {
  Runtime& rt = runtime();
  if (status == 0) {
    rt.last_leg = CVTX_HIP_LEG_HOST;
    return;
  }
  set_error(rt, err);
  report_failure(fn, out, n_out);
}

}  // namespace cvtx

namespace {

inline long cnt(int n) { return n > 0 ? n : 0; }

bool guard256(int a, int b) { return a >= 256 && b >= 256; }

template <class T>
bool arrays_ok(const T* arr, int n) {
  return n >= 0 && (n == 0 || arr != nullptr);
}

}  // namespace

namespace cvtx {

// Each run_* offers the call to the device when want_device() agrees
// (guard: the caller's size rule), else or on failure evaluates it on the
// host leg; invalid arguments give NaN results and a line on stderr.

void run_p3d_vel(const char* fn, bool guard, const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                 int nm, bsv_V3f* out, const cvtx_VortFunc* k, float sigma) {
  if (want_device(guard, k)) {
    if (cvtx_hip_P3D_M2M_vel(src, ns, mes, nm, out, k, sigma) == 0) return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, k && arrays_ok(src, ns) && arrays_ok(mes, nm) && arrays_ok(out, nm),
                    (float*)out, 3 * cnt(nm)))
    return;
  std::string err;
  host_done(fn, host::p3d_vel(src, ns, mes, nm, out, k, sigma, &err), err, (float*)out, 3 * cnt(nm));
}

void run_p3d_dvort(const char* fn, bool guard, const cvtx_P3D** src, int ns,
                   const cvtx_P3D** ind, int ni, bsv_V3f* out, const cvtx_VortFunc* k,
                   float sigma) {
  if (want_device(guard, k)) {
    if (cvtx_hip_P3D_M2M_dvort(src, ns, ind, ni, out, k, sigma) == 0) return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, k && arrays_ok(src, ns) && arrays_ok(ind, ni) && arrays_ok(out, ni),
                    (float*)out, 3 * cnt(ni)))
    return;
  std::string err;
  host_done(fn, host::p3d_dvort(src, ns, ind, ni, out, k, sigma, &err), err, (float*)out,
            3 * cnt(ni));
}

void run_p3d_visc(const char* fn, bool guard, const cvtx_P3D** src, int ns,
                  const cvtx_P3D** ind, int ni, bsv_V3f* out, const cvtx_VortFunc* k,
                  float sigma, float nu) {
  if (want_device(guard, k)) {
    if (cvtx_hip_P3D_M2M_visc_dvort(src, ns, ind, ni, out, k, sigma, nu) == 0)
      return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, k && arrays_ok(src, ns) && arrays_ok(ind, ni) && arrays_ok(out, ni),
                    (float*)out, 3 * cnt(ni)))
    return;
  std::string err;
  host_done(fn, host::p3d_visc(src, ns, ind, ni, out, k, sigma, nu, &err), err, (float*)out,
            3 * cnt(ni));
}

void run_p3d_vort(const char* fn, bool guard, const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                  int nm, bsv_V3f* out, const cvtx_VortFunc* k, float sigma, bool cutoff) {
  if (want_device(guard, k)) {
    if (vort_host(src, ns, mes, nm, (float*)out, k, sigma, cutoff) == 0) return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, k && arrays_ok(src, ns) && arrays_ok(mes, nm) && arrays_ok(out, nm),
                    (float*)out, 3 * cnt(nm)))
    return;
  std::string err;
  host_done(fn, host::p3d_vort(src, ns, mes, nm, out, k, sigma, cutoff, &err), err, (float*)out,
            3 * cnt(nm));
}

void run_f3d_vel(const char* fn, bool guard, const cvtx_F3D** src, int nf, const bsv_V3f* mes,
                 int nm, bsv_V3f* out) {
  if (want_device(guard, nullptr)) {
    if (cvtx_hip_F3D_M2M_vel(src, nf, mes, nm, out) == 0) return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, arrays_ok(src, nf) && arrays_ok(mes, nm) && arrays_ok(out, nm),
                    (float*)out, 3 * cnt(nm)))
    return;
  host::f3d_vel(src, nf, mes, nm, out);
  runtime().last_leg = CVTX_HIP_LEG_HOST;
}

void run_f3d_dvort(const char* fn, bool guard, const cvtx_F3D** src, int nf,
                   const cvtx_P3D** ind, int ni, bsv_V3f* out) {
  if (want_device(guard, nullptr)) {
    if (cvtx_hip_F3D_M2M_dvort(src, nf, ind, ni, out) == 0) return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, arrays_ok(src, nf) && arrays_ok(ind, ni) && arrays_ok(out, ni),
                    (float*)out, 3 * cnt(ni)))
    return;
  host::f3d_dvort(src, nf, ind, ni, out);
  runtime().last_leg = CVTX_HIP_LEG_HOST;
}

void run_p2d_vel(const char* fn, bool guard, const cvtx_P2D** src, int ns, const bsv_V2f* mes,
                 int nm, bsv_V2f* out, const cvtx_VortFunc* k, float sigma) {
  if (want_device(guard, k)) {
    if (cvtx_hip_P2D_M2M_vel(src, ns, mes, nm, out, k, sigma) == 0) return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, k && arrays_ok(src, ns) && arrays_ok(mes, nm) && arrays_ok(out, nm),
                    (float*)out, 2 * cnt(nm)))
    return;
  std::string err;
  host_done(fn, host::p2d_vel(src, ns, mes, nm, out, k, sigma, &err), err, (float*)out,
            2 * cnt(nm));
}

void run_p2d_visc(const char* fn, bool guard, const cvtx_P2D** src, int ns,
                  const cvtx_P2D** ind, int ni, float* out, const cvtx_VortFunc* k, float sigma,
                  float nu) {
  if (want_device(guard, k)) {
    if (cvtx_hip_P2D_M2M_visc_dvort(src, ns, ind, ni, out, k, sigma, nu) == 0)
      return device_done();
    device_declined(fn);
  }
  if (!host_args_ok(fn, k && arrays_ok(src, ns) && arrays_ok(ind, ni) && arrays_ok(out, ni), out,
                    cnt(ni)))
    return;
  std::string err;
  host_done(fn, host::p2d_visc(src, ns, ind, ni, out, k, sigma, nu, &err), err, out, cnt(ni));
}

}  // namespace cvtx

namespace {

// the reference's cost models for the filament and 2D velocities
bool f3d_vel_guard(int nf, int nm) {
  const long long w = (long long)cnt(nf) * cnt(nm);
  return !(w < w / 20 + 5000 + 300LL * cnt(nm));
}

bool p2d_vel_guard(int ns, int nm) {
  const long long w = (long long)cnt(ns) * cnt(nm);
  return !(20000 + w / 20 >= w);
}

}  // namespace

extern "C" {

// ---------------------------------- P3D ----------------------------------

void cvtx_P3D_M2M_vel(const cvtx_P3D** src, const int ns, const bsv_V3f* mes, const int nm,
                      bsv_V3f* out, const cvtx_VortFunc* k, float sigma) {
  run_p3d_vel("cvtx_P3D_M2M_vel", guard256(ns, nm), src, ns, mes, nm, out, k, sigma);
}

void cvtx_P3D_M2M_dvort(const cvtx_P3D** src, const int ns, const cvtx_P3D** ind, const int ni,
                        bsv_V3f* out, const cvtx_VortFunc* k, float sigma) {
  run_p3d_dvort("cvtx_P3D_M2M_dvort", guard256(ns, ni), src, ns, ind, ni, out, k, sigma);
}

void cvtx_P3D_M2M_visc_dvort(const cvtx_P3D** src, const int ns, const cvtx_P3D** ind,
                             const int ni, bsv_V3f* out, const cvtx_VortFunc* k, float sigma,
                             float nu) {
  run_p3d_visc("cvtx_P3D_M2M_visc_dvort", guard256(ns, ni), src, ns, ind, ni, out, k, sigma, nu);
}

void cvtx_P3D_M2M_vort(const cvtx_P3D** src, const int ns, const bsv_V3f* mes, const int nm,
                       bsv_V3f* out, const cvtx_VortFunc* k, float sigma) {
  run_p3d_vort("cvtx_P3D_M2M_vort", guard256(ns, nm), src, ns, mes, nm, out, k, sigma, true);
}

// reference src/P3D.cpp:669-707: the vorticity field at the particles
// (cvtx_P3D_M2M_vort's dispatch) and the strength update
void cvtx_P3D_pedrizzetti_relaxation(cvtx_P3D** particles, const int n, float fdt,
                                     const cvtx_VortFunc* k, float sigma) {
  const char* fn = "cvtx_P3D_pedrizzetti_relaxation";
  if (want_device(guard256(n, n), k)) {
    if (cvtx_hip_P3D_pedrizzetti_relaxation(particles, n, fdt, k, sigma) == 0)
      return device_done();
    device_declined(fn);
  }
  if (!k || !arrays_ok(particles, n)) {
    set_error(runtime(), "invalid arguments");
    report_failure(fn, nullptr, 0);
    return;
  }
  std::string err;
  if (host::p3d_relax(particles, n, fdt, k, sigma, &err) != 0) {
    // the strengths are the result here: report and set them to NaN
    set_error(runtime(), err);
    report_failure(fn, nullptr, 0);
    for (int i = 0; i < n; ++i)
      if (particles[i]) fill_nan(particles[i]->vorticity.x, 3);
    return;
  }
  runtime().last_leg = CVTX_HIP_LEG_HOST;
}

// ---------------------------------- F3D ----------------------------------

void cvtx_F3D_M2M_vel(const cvtx_F3D** src, const int nf, const bsv_V3f* mes, const int nm,
                      bsv_V3f* out) {
  run_f3d_vel("cvtx_F3D_M2M_vel", f3d_vel_guard(nf, nm), src, nf, mes, nm, out);
}

void cvtx_F3D_M2M_dvort(const cvtx_F3D** src, const int nf, const cvtx_P3D** ind, const int ni,
                        bsv_V3f* out) {
  run_f3d_dvort("cvtx_F3D_M2M_dvort", guard256(nf, ni), src, nf, ind, ni, out);
}

// The reference computes the influence matrix on the CPU only
// (src/F3D.cpp:204-227); here it goes to the device under the filament
// velocity's size guard.
void cvtx_F3D_inf_mtrx(const cvtx_F3D** src, const int nf, const bsv_V3f* mes,
                       const bsv_V3f* dir, const int nm, float* out) {
  const char* fn = "cvtx_F3D_inf_mtrx";
  if (want_device(f3d_vel_guard(nf, nm), nullptr)) {
    if (cvtx_hip_F3D_inf_mtrx(src, nf, mes, dir, nm, out) == 0) return device_done();
    device_declined(fn);
  }
  const long w = cnt(nf) * cnt(nm);
  if (!host_args_ok(fn, arrays_ok(src, nf) && arrays_ok(mes, nm) && arrays_ok(dir, nm) &&
                            (w == 0 || out != nullptr),
                    out, w))
    return;
  host::f3d_inf(src, nf, mes, dir, nm, out);
  runtime().last_leg = CVTX_HIP_LEG_HOST;
}

// ---------------------------------- P2D ----------------------------------

void cvtx_P2D_M2M_vel(const cvtx_P2D** src, const int ns, const bsv_V2f* mes, const int nm,
                      bsv_V2f* out, const cvtx_VortFunc* k, float sigma) {
  run_p2d_vel("cvtx_P2D_M2M_vel", p2d_vel_guard(ns, nm), src, ns, mes, nm, out, k, sigma);
}

void cvtx_P2D_M2M_visc_dvort(const cvtx_P2D** src, const int ns, const cvtx_P2D** ind,
                             const int ni, float* out, const cvtx_VortFunc* k, float sigma,
                             float nu) {
  run_p2d_visc("cvtx_P2D_M2M_visc_dvort", guard256(ns, ni), src, ns, ind, ni, out, k, sigma, nu);
}

// -------------------------------- controls --------------------------------

int cvtx_hip_set_dispatch(int mode) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (mode != DISPATCH_AUTO && mode != DISPATCH_DEVICE && mode != DISPATCH_HOST) return -1;
  const int prev = rt.dispatch;
  rt.dispatch = mode;
  return prev;
}

int cvtx_hip_get_dispatch(void) { return runtime().dispatch; }

int cvtx_hip_last_leg(void) { return runtime().last_leg; }

int cvtx_hip_host_threads(void) { return host::threads(); }

}  // extern "C"
