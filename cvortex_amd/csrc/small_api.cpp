// small_api.cpp -- the reference's single-source and single-target entry
// points (S2S, S2M, M2S; reference libcvtx.h:125-211, 266-283, 306-360).
//
// The reference evaluates these on the CPU only (src/P3D.cpp:74-296,
// src/F3D.cpp:87-160, src/P2D.cpp:71-231).  Here a single pair (S2S) and
// any S2M/M2S call with fewer than CVTX_HIP_SMALL_DEVICE_MIN elements run
// on the library's CPU leg: one pair, or a few thousand, cost far less on
// the host than a device round trip (gather, copies, launch, sync).  Larger
// S2M calls are one-source M2M launches and larger M2S calls one-target M2M
// launches, whose source partitions fill the GPU; the CPU leg remains their
// fallback (dispatch.cpp).  The values are those of the reference's pair
// and sum, except for the vorticity field, whose single-source forms have
// no cut-off (reference src/P3D.cpp:146-161, 213-228).  Null records give
// NaN results and a line on stderr (the reference asserts).
#include <cmath>
#include <limits>

#include <cvortex/libcvtx.h>
#include <cvortex_hip.h>

#include "api_common.hpp"
#include "runtime.hpp"

using namespace cvtx;

namespace {

bsv_V3f nan3() {
  const float q = std::numeric_limits<float>::quiet_NaN();
  return bsv_V3f{{q, q, q}};
}

bsv_V2f nan2() {
  const float q = std::numeric_limits<float>::quiet_NaN();
  return bsv_V2f{{q, q}};
}

// a null single record is an invalid argument: NaN results
bool null_rec(const char* fn, const void* p, float* out, long n) {
  if (p) return false;
  set_error(runtime(), "null record");
  report_failure(fn, out, n);
  return true;
}

// whether a single-source / single-target call of n elements is big enough
// to be worth a device round trip
bool big(int n) { return n >= CVTX_HIP_SMALL_DEVICE_MIN; }

}  // namespace

extern "C" {

// ------------------------------- P3D -------------------------------

bsv_V3f cvtx_P3D_S2S_vel(const cvtx_P3D* self, const bsv_V3f mes_point,
                         const cvtx_VortFunc* kernel, float regularisation_radius) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_P3D_S2S_vel", self, nullptr, 0)) return out;
  const cvtx_P3D* src[1] = {self};
  run_p3d_vel("cvtx_P3D_S2S_vel", false, src, 1, &mes_point, 1, &out, kernel,
              regularisation_radius);
  return out;
}

bsv_V3f cvtx_P3D_S2S_dvort(const cvtx_P3D* self, const cvtx_P3D* induced_particle,
                           const cvtx_VortFunc* kernel, float regularisation_radius) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_P3D_S2S_dvort", self, nullptr, 0) ||
      null_rec("cvtx_P3D_S2S_dvort", induced_particle, nullptr, 0))
    return out;
  const cvtx_P3D* src[1] = {self};
  const cvtx_P3D* ind[1] = {induced_particle};
  run_p3d_dvort("cvtx_P3D_S2S_dvort", false, src, 1, ind, 1, &out, kernel,
                regularisation_radius);
  return out;
}

bsv_V3f cvtx_P3D_S2S_visc_dvort(const cvtx_P3D* self, const cvtx_P3D* induced_particle,
                                const cvtx_VortFunc* kernel, float regularisation_radius,
                                float kinematic_visc) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_P3D_S2S_visc_dvort", self, nullptr, 0) ||
      null_rec("cvtx_P3D_S2S_visc_dvort", induced_particle, nullptr, 0))
    return out;
  const cvtx_P3D* src[1] = {self};
  const cvtx_P3D* ind[1] = {induced_particle};
  run_p3d_visc("cvtx_P3D_S2S_visc_dvort", false, src, 1, ind, 1, &out, kernel,
               regularisation_radius, kinematic_visc);
  return out;
}

bsv_V3f cvtx_P3D_S2S_vort(const cvtx_P3D* self, const bsv_V3f mes_point,
                          const cvtx_VortFunc* kernel, float regularisation_radius) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_P3D_S2S_vort", self, nullptr, 0)) return out;
  const cvtx_P3D* src[1] = {self};
  run_p3d_vort("cvtx_P3D_S2S_vort", false, src, 1, &mes_point, 1, &out, kernel,
               regularisation_radius, false);
  return out;
}

void cvtx_P3D_S2M_vel(const cvtx_P3D* self, const bsv_V3f* mes_start, const int num_mes,
                      bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                      float regularisation_radius) {
  const long n = 3L * (num_mes > 0 ? num_mes : 0);
  if (null_rec("cvtx_P3D_S2M_vel", self, (float*)result_array, n)) return;
  const cvtx_P3D* src[1] = {self};
  run_p3d_vel("cvtx_P3D_S2M_vel", big(num_mes), src, 1, mes_start, num_mes, result_array, kernel,
              regularisation_radius);
}

void cvtx_P3D_S2M_dvort(const cvtx_P3D* self, const cvtx_P3D** induced_start,
                        const int num_induced, bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                        float regularisation_radius) {
  const long n = 3L * (num_induced > 0 ? num_induced : 0);
  if (null_rec("cvtx_P3D_S2M_dvort", self, (float*)result_array, n)) return;
  const cvtx_P3D* src[1] = {self};
  run_p3d_dvort("cvtx_P3D_S2M_dvort", big(num_induced), src, 1, induced_start, num_induced,
                result_array, kernel, regularisation_radius);
}

void cvtx_P3D_S2M_visc_dvort(const cvtx_P3D* self, const cvtx_P3D** induced_start,
                             const int num_induced, bsv_V3f* result_array,
                             const cvtx_VortFunc* kernel, float regularisation_radius,
                             float kinematic_visc) {
  const long n = 3L * (num_induced > 0 ? num_induced : 0);
  if (null_rec("cvtx_P3D_S2M_visc_dvort", self, (float*)result_array, n)) return;
  const cvtx_P3D* src[1] = {self};
  run_p3d_visc("cvtx_P3D_S2M_visc_dvort", big(num_induced), src, 1, induced_start, num_induced,
               result_array, kernel, regularisation_radius, kinematic_visc);
}

void cvtx_P3D_S2M_vort(const cvtx_P3D* self, const bsv_V3f* mes_start, const int num_mes,
                       bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                       float regularisation_radius) {
  const long n = 3L * (num_mes > 0 ? num_mes : 0);
  if (null_rec("cvtx_P3D_S2M_vort", self, (float*)result_array, n)) return;
  const cvtx_P3D* src[1] = {self};
  run_p3d_vort("cvtx_P3D_S2M_vort", big(num_mes), src, 1, mes_start, num_mes, result_array,
               kernel, regularisation_radius, false);
}

bsv_V3f cvtx_P3D_M2S_vel(const cvtx_P3D** array_start, const int num_particles,
                         const bsv_V3f mes_point, const cvtx_VortFunc* kernel,
                         float regularisation_radius) {
  bsv_V3f out = nan3();
  run_p3d_vel("cvtx_P3D_M2S_vel", big(num_particles), array_start, num_particles, &mes_point, 1,
              &out, kernel, regularisation_radius);
  return out;
}

bsv_V3f cvtx_P3D_M2S_dvort(const cvtx_P3D** array_start, const int num_particles,
                           const cvtx_P3D* induced_particle, const cvtx_VortFunc* kernel,
                           float regularisation_radius) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_P3D_M2S_dvort", induced_particle, nullptr, 0)) return out;
  const cvtx_P3D* ind[1] = {induced_particle};
  run_p3d_dvort("cvtx_P3D_M2S_dvort", big(num_particles), array_start, num_particles, ind, 1,
                &out, kernel, regularisation_radius);
  return out;
}

bsv_V3f cvtx_P3D_M2S_visc_dvort(const cvtx_P3D** array_start, const int num_particles,
                                const cvtx_P3D* induced_particle, const cvtx_VortFunc* kernel,
                                float regularisation_radius, float kinematic_visc) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_P3D_M2S_visc_dvort", induced_particle, nullptr, 0)) return out;
  const cvtx_P3D* ind[1] = {induced_particle};
  run_p3d_visc("cvtx_P3D_M2S_visc_dvort", big(num_particles), array_start, num_particles, ind, 1,
               &out, kernel, regularisation_radius, kinematic_visc);
  return out;
}

bsv_V3f cvtx_P3D_M2S_vort(const cvtx_P3D** array_start, const int num_particles,
                          const bsv_V3f mes_point, const cvtx_VortFunc* kernel,
                          float regularisation_radius) {
  bsv_V3f out = nan3();
  run_p3d_vort("cvtx_P3D_M2S_vort", big(num_particles), array_start, num_particles, &mes_point, 1,
               &out, kernel, regularisation_radius, true);
  return out;
}

// ------------------------------- F3D -------------------------------

bsv_V3f cvtx_F3D_S2S_vel(const cvtx_F3D* self, const bsv_V3f mes_point) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_F3D_S2S_vel", self, nullptr, 0)) return out;
  const cvtx_F3D* src[1] = {self};
  run_f3d_vel("cvtx_F3D_S2S_vel", false, src, 1, &mes_point, 1, &out);
  return out;
}

bsv_V3f cvtx_F3D_S2S_dvort(const cvtx_F3D* self, const cvtx_P3D* induced_particle) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_F3D_S2S_dvort", self, nullptr, 0) ||
      null_rec("cvtx_F3D_S2S_dvort", induced_particle, nullptr, 0))
    return out;
  const cvtx_F3D* src[1] = {self};
  const cvtx_P3D* ind[1] = {induced_particle};
  run_f3d_dvort("cvtx_F3D_S2S_dvort", false, src, 1, ind, 1, &out);
  return out;
}

bsv_V3f cvtx_F3D_M2S_vel(const cvtx_F3D** array_start, const int num_filaments,
                         const bsv_V3f mes_point) {
  bsv_V3f out = nan3();
  run_f3d_vel("cvtx_F3D_M2S_vel", big(num_filaments), array_start, num_filaments, &mes_point, 1,
              &out);
  return out;
}

bsv_V3f cvtx_F3D_M2S_dvort(const cvtx_F3D** array_start, const int num_filaments,
                           const cvtx_P3D* induced_particle) {
  bsv_V3f out = nan3();
  if (null_rec("cvtx_F3D_M2S_dvort", induced_particle, nullptr, 0)) return out;
  const cvtx_P3D* ind[1] = {induced_particle};
  run_f3d_dvort("cvtx_F3D_M2S_dvort", big(num_filaments), array_start, num_filaments, ind, 1,
                &out);
  return out;
}

// ------------------------------- P2D -------------------------------

bsv_V2f cvtx_P2D_S2S_vel(const cvtx_P2D* self, const bsv_V2f mes_point,
                         const cvtx_VortFunc* kernel, float regularisation_radius) {
  bsv_V2f out = nan2();
  if (null_rec("cvtx_P2D_S2S_vel", self, nullptr, 0)) return out;
  const cvtx_P2D* src[1] = {self};
  run_p2d_vel("cvtx_P2D_S2S_vel", false, src, 1, &mes_point, 1, &out, kernel,
              regularisation_radius);
  return out;
}

void cvtx_P2D_S2M_vel(const cvtx_P2D* self, const bsv_V2f* mes_start, const int num_mes,
                      bsv_V2f* result_array, const cvtx_VortFunc* kernel,
                      float regularisation_radius) {
  const long n = 2L * (num_mes > 0 ? num_mes : 0);
  if (null_rec("cvtx_P2D_S2M_vel", self, (float*)result_array, n)) return;
  const cvtx_P2D* src[1] = {self};
  run_p2d_vel("cvtx_P2D_S2M_vel", big(num_mes), src, 1, mes_start, num_mes, result_array, kernel,
              regularisation_radius);
}

bsv_V2f cvtx_P2D_M2S_vel(const cvtx_P2D** array_start, const int num_particles,
                         const bsv_V2f mes_point, const cvtx_VortFunc* kernel,
                         float regularisation_radius) {
  bsv_V2f out = nan2();
  run_p2d_vel("cvtx_P2D_M2S_vel", big(num_particles), array_start, num_particles, &mes_point, 1,
              &out, kernel, regularisation_radius);
  return out;
}

float cvtx_P2D_S2S_visc_dvort(const cvtx_P2D* self, const cvtx_P2D* induced_particle,
                              const cvtx_VortFunc* kernel, float regularisation_radius,
                              float kinematic_visc) {
  float out = std::numeric_limits<float>::quiet_NaN();
  if (null_rec("cvtx_P2D_S2S_visc_dvort", self, nullptr, 0) ||
      null_rec("cvtx_P2D_S2S_visc_dvort", induced_particle, nullptr, 0))
    return out;
  const cvtx_P2D* src[1] = {self};
  const cvtx_P2D* ind[1] = {induced_particle};
  run_p2d_visc("cvtx_P2D_S2S_visc_dvort", false, src, 1, ind, 1, &out, kernel,
               regularisation_radius, kinematic_visc);
  return out;
}

void cvtx_P2D_S2M_visc_dvort(const cvtx_P2D* self, const cvtx_P2D** induced_start,
                             const int num_induced, float* result_array,
                             const cvtx_VortFunc* kernel, float regularisation_radius,
                             float kinematic_visc) {
  const long n = num_induced > 0 ? num_induced : 0;
  if (null_rec("cvtx_P2D_S2M_visc_dvort", self, result_array, n)) return;
  const cvtx_P2D* src[1] = {self};
  run_p2d_visc("cvtx_P2D_S2M_visc_dvort", big(num_induced), src, 1, induced_start, num_induced,
               result_array, kernel, regularisation_radius, kinematic_visc);
}

float cvtx_P2D_M2S_visc_dvort(const cvtx_P2D** array_start, const int num_particles,
                              const cvtx_P2D* induced_particle, const cvtx_VortFunc* kernel,
                              float regularisation_radius, float kinematic_visc) {
  float out = std::numeric_limits<float>::quiet_NaN();
  if (null_rec("cvtx_P2D_M2S_visc_dvort", induced_particle, nullptr, 0)) return out;
  const cvtx_P2D* ind[1] = {induced_particle};
  run_p2d_visc("cvtx_P2D_M2S_visc_dvort", big(num_particles), array_start, num_particles, ind, 1,
               &out, kernel, regularisation_radius, kinematic_visc);
  return out;
}

}  // extern "C"
