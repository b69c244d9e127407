// redist_api.cpp -- C ABI of the redistribution on a grid (remeshing):
// the cvtx_RedistFunc factories (reference src/RedistFunc.cpp:36-96) and
// cvtx_P3D_redistribute_on_grid / cvtx_P2D_redistribute_on_grid (reference
// libcvtx.h:250-257, 372-380).  The device pipeline is in redist.hip.
//
// cvtx_RedistFunc carries a host function pointer and a radius but no name,
// so the device kernel is chosen by recognising the library's own five
// functions; a redistributor built around any other function is rejected
// (return -1, message on stderr), as is an unknown regularisation name.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include <cvortex/libcvtx.h>
#include <cvortex_hip.h>

#include "api_common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace cvtx;

namespace {

using Lock = std::lock_guard<std::recursive_mutex>;

// reference src/RedistFunc.cpp:36-96 (FP32 host functions of the plugin)
float f_lambda0(float U) { return U < 0.5f ? 1.f : 0.f; }
float f_lambda1(float U) { return U <= 1.f ? 1.f - U : 0.f; }
float f_lambda2(float U) {
  return U < 0.5f ? 1.f - U * U : (U < 1.5f ? 0.5f * (1.f - U) * (2.f - U) : 0.f);
}
float f_lambda3(float U) {
  return U < 1.f ? 0.5f * (1.f - U * U) * (2.f - U)
                 : (U < 2.f ? (1.f / 6.f) * (1.f - U) * (2.f - U) * (3.f - U) : 0.f);
}
float f_m4p(float U) {
  return U < 1.f ? 1.f - 2.5f * U * U + 1.5f * U * U * U
                 : (U < 2.f ? 0.5f * (1.f - U) * (2.f - U) * (2.f - U) : 0.f);
}

float (*const kFuncs[5])(float) = {f_lambda0, f_lambda1, f_lambda2, f_lambda3, f_m4p};

int redistributor_id(Runtime& rt, const cvtx_RedistFunc* r) {
  if (r == nullptr) {
    set_error(rt, "null redistributor");
    return -1;
  }
  for (int k = 0; k < 5; ++k)
    if (r->func == kFuncs[k] && r->radius == redist_radius(k)) return k;
  set_error(rt, "redistributor is not one of the library's lambda0..3 / m4p");
  return -1;
}

template <int DIM, class Rec>
int redistribute_host(const Rec** in, int n, Rec* out, int max_out, const cvtx_RedistFunc* r,
                      float h, float negligible) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  const int rid = redistributor_id(rt, r);
  if (rid < 0) return -1;
  if (n < 0 || max_out < 0 || !(h > 0.f) || !(negligible >= 0.f) || !(negligible < 1.f) ||
      (n > 0 && in == nullptr)) {
    set_error(rt, "invalid redistribution arguments");
    return -1;
  }
  if (n == 0) return 0;
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);
  const size_t pb = sizeof(float4) * (size_t)n;
  float4* h_pv = (float4*)host_buffer(rt, *d, SLOT_SRC0, pb);
  float4* d_pv = (float4*)device_buffer(rt, *d, SLOT_SRC0, pb);
  float4* h_a = nullptr;
  float4* d_a = nullptr;
  if (!h_pv || !d_pv) return -1;
  if constexpr (DIM == 3) {
    h_a = (float4*)host_buffer(rt, *d, SLOT_SRC1, pb);
    d_a = (float4*)device_buffer(rt, *d, SLOT_SRC1, pb);
    if (!h_a || !d_a) return -1;
    gather_p3d(in, n, h_pv, h_a);
    if (!h2d(rt, *d, d_a, h_a, pb)) return -1;
  } else {
    gather_p2d(in, n, h_pv);
  }
  if (!h2d(rt, *d, d_pv, h_pv, pb)) return -1;
  RedistJob j;
  j.dim = DIM;
  j.rid = rid;
  j.pv = d_pv;
  j.a = d_a;
  j.n = n;
  j.h = h;
  j.negligible = negligible;
  j.max_out = out ? max_out : -1;
  float* dev_out = nullptr;
  const int nn = redistribute(rt, *d, j, &dev_out, d->stream);
  if (nn < 0) return -1;
  if (out && nn > 0) {
    static_assert(sizeof(Rec) == sizeof(float) * (DIM == 3 ? 7 : 4), "record layout");
    if (!d2h(rt, *d, out, dev_out, sizeof(Rec) * (size_t)nn)) return -1;
  }
  if (!check_hip(rt, hipStreamSynchronize(d->stream), "redistribution")) return -1;
  return nn;
}

}  // namespace

extern "C" {

const cvtx_RedistFunc cvtx_RedistFunc_lambda0(void) { return cvtx_RedistFunc{f_lambda0, 0.5f}; }
const cvtx_RedistFunc cvtx_RedistFunc_lambda1(void) { return cvtx_RedistFunc{f_lambda1, 1.0f}; }
const cvtx_RedistFunc cvtx_RedistFunc_lambda2(void) { return cvtx_RedistFunc{f_lambda2, 1.5f}; }
const cvtx_RedistFunc cvtx_RedistFunc_lambda3(void) { return cvtx_RedistFunc{f_lambda3, 2.0f}; }
const cvtx_RedistFunc cvtx_RedistFunc_m4p(void) { return cvtx_RedistFunc{f_m4p, 2.0f}; }

int cvtx_P3D_redistribute_on_grid(const cvtx_P3D** input_array_start, const int n_input_particles,
                                  cvtx_P3D* output_particles, int max_output_particles,
                                  const cvtx_RedistFunc* redistributor, const float grid_density,
                                  float negligible_vort) {
  const int n = redistribute_host<3>(input_array_start, n_input_particles, output_particles,
                                     max_output_particles, redistributor, grid_density,
                                     negligible_vort);
  if (n < 0) report_failure("cvtx_P3D_redistribute_on_grid", nullptr, 0);
  return n;
}

int cvtx_P2D_redistribute_on_grid(const cvtx_P2D** input_array_start, const int n_input_particles,
                                  cvtx_P2D* output_particles, int max_output_particles,
                                  const cvtx_RedistFunc* redistributor, const float grid_density,
                                  float negligible_vort) {
  const int n = redistribute_host<2>(input_array_start, n_input_particles, output_particles,
                                     max_output_particles, redistributor, grid_density,
                                     negligible_vort);
  if (n < 0) report_failure("cvtx_P2D_redistribute_on_grid", nullptr, 0);
  return n;
}

}  // extern "C"
