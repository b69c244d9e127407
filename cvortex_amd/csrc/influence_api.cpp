// influence_api.cpp -- C ABI of the vortex-filament influence matrix
// (cvtx_F3D_inf_mtrx, reference libcvtx.h:299-305, src/F3D.cpp:204-227):
// the host-array entry point, its int-returning seam and the
// device-resident form.  Kernel: f3d_inf_kernel in f3d.hip.
#include <hip/hip_runtime.h>

#include <mutex>

#include <cvortex/libcvtx.h>
#include <cvortex_hip.h>

#include "api_common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace cvtx;

namespace {

using Lock = std::lock_guard<std::recursive_mutex>;

int inf_host(const cvtx_F3D** fils, int nf, const bsv_V3f* mes, const bsv_V3f* dir, int nm,
             float* out) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (bad_counts(rt, nf, nm, fils, mes) || (nm > 0 && dir == nullptr)) {
    if (nm > 0 && dir == nullptr) set_error(rt, "null direction array");
    return -1;
  }
  if (nm == 0 || nf == 0) return 0;  // an empty matrix: nothing to write
  if (out == nullptr) {
    set_error(rt, "null result matrix");
    return -1;
  }
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);
  const size_t fb = sizeof(float4) * (size_t)nf;
  const size_t mb = sizeof(float4) * (size_t)nm;
  const size_t ob = sizeof(float) * (size_t)nm * (size_t)nf;
  float4* h_fa = (float4*)host_buffer(rt, *d, SLOT_SRC0, fb);
  float4* h_fb = (float4*)host_buffer(rt, *d, SLOT_SRC1, fb);
  float4* h_m = (float4*)host_buffer(rt, *d, SLOT_TGT0, mb);
  float4* h_e = (float4*)host_buffer(rt, *d, SLOT_TGT1, mb);
  float4* d_fa = (float4*)device_buffer(rt, *d, SLOT_SRC0, fb);
  float4* d_fb = (float4*)device_buffer(rt, *d, SLOT_SRC1, fb);
  float4* d_m = (float4*)device_buffer(rt, *d, SLOT_TGT0, mb);
  float4* d_e = (float4*)device_buffer(rt, *d, SLOT_TGT1, mb);
  float* d_out = (float*)device_buffer(rt, *d, SLOT_OUT0, ob);
  if (!h_fa || !h_fb || !h_m || !h_e || !d_fa || !d_fb || !d_m || !d_e || !d_out) return -1;
  gather_f3d(fils, nf, h_fa, h_fb);
  gather_pts3(mes, nm, h_m);
  gather_pts3(dir, nm, h_e);
  if (!h2d(rt, *d, d_fa, h_fa, fb) || !h2d(rt, *d, d_fb, h_fb, fb) || !h2d(rt, *d, d_m, h_m, mb) ||
      !h2d(rt, *d, d_e, h_e, mb))
    return -1;
  if (!check_hip(rt, launch_f3d_inf(d_fa, d_fb, nf, d_m, d_e, nm, d_out, d->stream),
                 "influence matrix kernel"))
    return -1;
  // the matrix can be large: copy it straight into the caller's array
  if (!d2h(rt, *d, out, d_out, ob)) return -1;
  return check_hip(rt, hipStreamSynchronize(d->stream), "influence matrix") ? 0 : -1;
}

}  // namespace

extern "C" {

// Seam form (no reference counterpart: the reference computes the matrix
// on the CPU only, src/F3D.cpp:204-227).
int cvtx_hip_F3D_inf_mtrx(const cvtx_F3D** array_start, int num_filaments,
                          const bsv_V3f* mes_start, const bsv_V3f* dir_start, int num_mes,
                          float* result_matrix) {
  return inf_host(array_start, num_filaments, mes_start, dir_start, num_mes, result_matrix);
}

int cvtx_hip_f3d_inf_eval(const float* fa, const float* fb, int n_fil, const float* mes,
                          const float* dir, int n_mes, float* out, void* stream) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (n_fil < 0 || n_mes < 0 || (n_fil > 0 && n_mes > 0 && (!fa || !fb || !mes || !dir || !out))) {
    set_error(rt, "invalid influence-matrix problem");
    return -1;
  }
  Device* d = active_device(rt);
  if (!d) return -1;
  return check_hip(rt,
                   launch_f3d_inf((const float4*)fa, (const float4*)fb, n_fil,
                                  (const float4*)mes, (const float4*)dir, n_mes, out,
                                  pick_stream(d, stream)),
                   "influence matrix kernel")
             ? 0
             : -1;
}

}  // extern "C"
