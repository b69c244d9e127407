// small_api.cpp -- the reference's single-source and single-target entry
// points (S2S, S2M, M2S; reference libcvtx.h:125-211, 266-283, 306-360),
// evaluated on the device through the M2M paths: S2M is a one-source M2M,
// M2S a one-target M2M (whose source partitions fill the GPU), S2S a 1 x 1
// M2M.  The values are those of the reference's pair and sum (the single
// forms differ from M2M only in the final rounding of the one-term sum),
// except for the vorticity field, whose single-source forms have no cut-off
// and divide by the signed radius (reference src/P3D.cpp:146-161, 213-228).
//
// Each call is one device round trip; many small evaluations are better
// expressed as one M2M call.  On failure the result is NaN and a line goes
// to stderr, as for the M2M entry points.
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

// null single-record arguments are errors (the reference asserts)
bool null_arg(const void* p) {
  if (p) return false;
  set_error(runtime(), "null record");
  return true;
}

bsv_V3f fail3(const char* fn) {
  report_failure(fn, nullptr, 0);
  return nan3();
}

bsv_V2f fail2(const char* fn) {
  report_failure(fn, nullptr, 0);
  return nan2();
}

float fail1(const char* fn) {
  report_failure(fn, nullptr, 0);
  return std::numeric_limits<float>::quiet_NaN();
}

}  // namespace

extern "C" {

// ------------------------------- P3D -------------------------------

bsv_V3f cvtx_P3D_S2S_vel(const cvtx_P3D* self, const bsv_V3f mes_point,
                         const cvtx_VortFunc* kernel, float regularisation_radius) {
  if (null_arg(self)) return fail3("cvtx_P3D_S2S_vel");
  const cvtx_P3D* src[1] = {self};
  bsv_V3f out;
  if (cvtx_hip_P3D_M2M_vel(src, 1, &mes_point, 1, &out, kernel, regularisation_radius) != 0) {
    report_failure("cvtx_P3D_S2S_vel", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_P3D_S2S_dvort(const cvtx_P3D* self, const cvtx_P3D* induced_particle,
                           const cvtx_VortFunc* kernel, float regularisation_radius) {
  if (null_arg(self) || null_arg(induced_particle)) return fail3("cvtx_P3D_S2S_dvort");
  const cvtx_P3D* src[1] = {self};
  const cvtx_P3D* ind[1] = {induced_particle};
  bsv_V3f out;
  if (cvtx_hip_P3D_M2M_dvort(src, 1, ind, 1, &out, kernel, regularisation_radius) != 0) {
    report_failure("cvtx_P3D_S2S_dvort", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_P3D_S2S_visc_dvort(const cvtx_P3D* self, const cvtx_P3D* induced_particle,
                                const cvtx_VortFunc* kernel, float regularisation_radius,
                                float kinematic_visc) {
  if (null_arg(self) || null_arg(induced_particle)) return fail3("cvtx_P3D_S2S_visc_dvort");
  const cvtx_P3D* src[1] = {self};
  const cvtx_P3D* ind[1] = {induced_particle};
  bsv_V3f out;
  if (cvtx_hip_P3D_M2M_visc_dvort(src, 1, ind, 1, &out, kernel, regularisation_radius,
                                  kinematic_visc) != 0) {
    report_failure("cvtx_P3D_S2S_visc_dvort", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_P3D_S2S_vort(const cvtx_P3D* self, const bsv_V3f mes_point,
                          const cvtx_VortFunc* kernel, float regularisation_radius) {
  if (null_arg(self)) return fail3("cvtx_P3D_S2S_vort");
  const cvtx_P3D* src[1] = {self};
  bsv_V3f out;
  if (vort_host(src, 1, &mes_point, 1, out.x, kernel, regularisation_radius, false) != 0) {
    report_failure("cvtx_P3D_S2S_vort", nullptr, 0);
    return nan3();
  }
  return out;
}

void cvtx_P3D_S2M_vel(const cvtx_P3D* self, const bsv_V3f* mes_start, const int num_mes,
                      bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                      float regularisation_radius) {
  const cvtx_P3D* src[1] = {self};
  if (null_arg(self) ||
      cvtx_hip_P3D_M2M_vel(src, 1, mes_start, num_mes, result_array, kernel,
                           regularisation_radius) != 0)
    report_failure("cvtx_P3D_S2M_vel", (float*)result_array, 3L * (num_mes > 0 ? num_mes : 0));
}

void cvtx_P3D_S2M_dvort(const cvtx_P3D* self, const cvtx_P3D** induced_start,
                        const int num_induced, bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                        float regularisation_radius) {
  const cvtx_P3D* src[1] = {self};
  if (null_arg(self) ||
      cvtx_hip_P3D_M2M_dvort(src, 1, induced_start, num_induced, result_array, kernel,
                             regularisation_radius) != 0)
    report_failure("cvtx_P3D_S2M_dvort", (float*)result_array,
                   3L * (num_induced > 0 ? num_induced : 0));
}

void cvtx_P3D_S2M_visc_dvort(const cvtx_P3D* self, const cvtx_P3D** induced_start,
                             const int num_induced, bsv_V3f* result_array,
                             const cvtx_VortFunc* kernel, float regularisation_radius,
                             float kinematic_visc) {
  const cvtx_P3D* src[1] = {self};
  if (null_arg(self) ||
      cvtx_hip_P3D_M2M_visc_dvort(src, 1, induced_start, num_induced, result_array, kernel,
                                  regularisation_radius, kinematic_visc) != 0)
    report_failure("cvtx_P3D_S2M_visc_dvort", (float*)result_array,
                   3L * (num_induced > 0 ? num_induced : 0));
}

void cvtx_P3D_S2M_vort(const cvtx_P3D* self, const bsv_V3f* mes_start, const int num_mes,
                       bsv_V3f* result_array, const cvtx_VortFunc* kernel,
                       float regularisation_radius) {
  const cvtx_P3D* src[1] = {self};
  if (null_arg(self) ||
      vort_host(src, 1, mes_start, num_mes, (float*)result_array, kernel, regularisation_radius,
                false) != 0)
    report_failure("cvtx_P3D_S2M_vort", (float*)result_array, 3L * (num_mes > 0 ? num_mes : 0));
}

bsv_V3f cvtx_P3D_M2S_vel(const cvtx_P3D** array_start, const int num_particles,
                         const bsv_V3f mes_point, const cvtx_VortFunc* kernel,
                         float regularisation_radius) {
  bsv_V3f out;
  if (cvtx_hip_P3D_M2M_vel(array_start, num_particles, &mes_point, 1, &out, kernel,
                           regularisation_radius) != 0) {
    report_failure("cvtx_P3D_M2S_vel", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_P3D_M2S_dvort(const cvtx_P3D** array_start, const int num_particles,
                           const cvtx_P3D* induced_particle, const cvtx_VortFunc* kernel,
                           float regularisation_radius) {
  if (null_arg(induced_particle)) return fail3("cvtx_P3D_M2S_dvort");
  const cvtx_P3D* ind[1] = {induced_particle};
  bsv_V3f out;
  if (cvtx_hip_P3D_M2M_dvort(array_start, num_particles, ind, 1, &out, kernel,
                             regularisation_radius) != 0) {
    report_failure("cvtx_P3D_M2S_dvort", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_P3D_M2S_visc_dvort(const cvtx_P3D** array_start, const int num_particles,
                                const cvtx_P3D* induced_particle, const cvtx_VortFunc* kernel,
                                float regularisation_radius, float kinematic_visc) {
  if (null_arg(induced_particle)) return fail3("cvtx_P3D_M2S_visc_dvort");
  const cvtx_P3D* ind[1] = {induced_particle};
  bsv_V3f out;
  if (cvtx_hip_P3D_M2M_visc_dvort(array_start, num_particles, ind, 1, &out, kernel,
                                  regularisation_radius, kinematic_visc) != 0) {
    report_failure("cvtx_P3D_M2S_visc_dvort", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_P3D_M2S_vort(const cvtx_P3D** array_start, const int num_particles,
                          const bsv_V3f mes_point, const cvtx_VortFunc* kernel,
                          float regularisation_radius) {
  bsv_V3f out;
  if (vort_host(array_start, num_particles, &mes_point, 1, out.x, kernel, regularisation_radius,
                true) != 0) {
    report_failure("cvtx_P3D_M2S_vort", nullptr, 0);
    return nan3();
  }
  return out;
}

// ------------------------------- F3D -------------------------------

bsv_V3f cvtx_F3D_S2S_vel(const cvtx_F3D* self, const bsv_V3f mes_point) {
  if (null_arg(self)) return fail3("cvtx_F3D_S2S_vel");
  const cvtx_F3D* src[1] = {self};
  bsv_V3f out;
  if (cvtx_hip_F3D_M2M_vel(src, 1, &mes_point, 1, &out) != 0) {
    report_failure("cvtx_F3D_S2S_vel", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_F3D_S2S_dvort(const cvtx_F3D* self, const cvtx_P3D* induced_particle) {
  if (null_arg(self) || null_arg(induced_particle)) return fail3("cvtx_F3D_S2S_dvort");
  const cvtx_F3D* src[1] = {self};
  const cvtx_P3D* ind[1] = {induced_particle};
  bsv_V3f out;
  if (cvtx_hip_F3D_M2M_dvort(src, 1, ind, 1, &out) != 0) {
    report_failure("cvtx_F3D_S2S_dvort", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_F3D_M2S_vel(const cvtx_F3D** array_start, const int num_filaments,
                         const bsv_V3f mes_point) {
  bsv_V3f out;
  if (cvtx_hip_F3D_M2M_vel(array_start, num_filaments, &mes_point, 1, &out) != 0) {
    report_failure("cvtx_F3D_M2S_vel", nullptr, 0);
    return nan3();
  }
  return out;
}

bsv_V3f cvtx_F3D_M2S_dvort(const cvtx_F3D** array_start, const int num_filaments,
                           const cvtx_P3D* induced_particle) {
  if (null_arg(induced_particle)) return fail3("cvtx_F3D_M2S_dvort");
  const cvtx_P3D* ind[1] = {induced_particle};
  bsv_V3f out;
  if (cvtx_hip_F3D_M2M_dvort(array_start, num_filaments, ind, 1, &out) != 0) {
    report_failure("cvtx_F3D_M2S_dvort", nullptr, 0);
    return nan3();
  }
  return out;
}

// ------------------------------- P2D -------------------------------

bsv_V2f cvtx_P2D_S2S_vel(const cvtx_P2D* self, const bsv_V2f mes_point,
                         const cvtx_VortFunc* kernel, float regularisation_radius) {
  if (null_arg(self)) return fail2("cvtx_P2D_S2S_vel");
  const cvtx_P2D* src[1] = {self};
  bsv_V2f out;
  if (cvtx_hip_P2D_M2M_vel(src, 1, &mes_point, 1, &out, kernel, regularisation_radius) != 0) {
    report_failure("cvtx_P2D_S2S_vel", nullptr, 0);
    return nan2();
  }
  return out;
}

void cvtx_P2D_S2M_vel(const cvtx_P2D* self, const bsv_V2f* mes_start, const int num_mes,
                      bsv_V2f* result_array, const cvtx_VortFunc* kernel,
                      float regularisation_radius) {
  const cvtx_P2D* src[1] = {self};
  if (null_arg(self) ||
      cvtx_hip_P2D_M2M_vel(src, 1, mes_start, num_mes, result_array, kernel,
                           regularisation_radius) != 0)
    report_failure("cvtx_P2D_S2M_vel", (float*)result_array, 2L * (num_mes > 0 ? num_mes : 0));
}

bsv_V2f cvtx_P2D_M2S_vel(const cvtx_P2D** array_start, const int num_particles,
                         const bsv_V2f mes_point, const cvtx_VortFunc* kernel,
                         float regularisation_radius) {
  bsv_V2f out;
  if (cvtx_hip_P2D_M2M_vel(array_start, num_particles, &mes_point, 1, &out, kernel,
                           regularisation_radius) != 0) {
    report_failure("cvtx_P2D_M2S_vel", nullptr, 0);
    return nan2();
  }
  return out;
}

float cvtx_P2D_S2S_visc_dvort(const cvtx_P2D* self, const cvtx_P2D* induced_particle,
                              const cvtx_VortFunc* kernel, float regularisation_radius,
                              float kinematic_visc) {
  if (null_arg(self) || null_arg(induced_particle)) return fail1("cvtx_P2D_S2S_visc_dvort");
  const cvtx_P2D* src[1] = {self};
  const cvtx_P2D* ind[1] = {induced_particle};
  float out;
  if (cvtx_hip_P2D_M2M_visc_dvort(src, 1, ind, 1, &out, kernel, regularisation_radius,
                                  kinematic_visc) != 0) {
    report_failure("cvtx_P2D_S2S_visc_dvort", nullptr, 0);
    return std::numeric_limits<float>::quiet_NaN();
  }
  return out;
}

void cvtx_P2D_S2M_visc_dvort(const cvtx_P2D* self, const cvtx_P2D** induced_start,
                             const int num_induced, float* result_array,
                             const cvtx_VortFunc* kernel, float regularisation_radius,
                             float kinematic_visc) {
  const cvtx_P2D* src[1] = {self};
  if (null_arg(self) ||
      cvtx_hip_P2D_M2M_visc_dvort(src, 1, induced_start, num_induced, result_array, kernel,
                                  regularisation_radius, kinematic_visc) != 0)
    report_failure("cvtx_P2D_S2M_visc_dvort", result_array,
                   (long)(num_induced > 0 ? num_induced : 0));
}

float cvtx_P2D_M2S_visc_dvort(const cvtx_P2D** array_start, const int num_particles,
                              const cvtx_P2D* induced_particle, const cvtx_VortFunc* kernel,
                              float regularisation_radius, float kinematic_visc) {
  if (null_arg(induced_particle)) return fail1("cvtx_P2D_M2S_visc_dvort");
  const cvtx_P2D* ind[1] = {induced_particle};
  float out;
  if (cvtx_hip_P2D_M2M_visc_dvort(array_start, num_particles, ind, 1, &out, kernel,
                                  regularisation_radius, kinematic_visc) != 0) {
    report_failure("cvtx_P2D_M2S_visc_dvort", nullptr, 0);
    return std::numeric_limits<float>::quiet_NaN();
  }
  return out;
}

}  // extern "C"
