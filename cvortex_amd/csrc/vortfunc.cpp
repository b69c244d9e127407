// vortfunc.cpp -- the regularisation plugin structs (cvtx_VortFunc).
//
// The GPU kernels are selected by cl_kernel_name_ext; the function pointers
// are scalar host functions kept for callers that evaluate g, zeta or eta
// themselves (the reference's own tests do, test/testvortfunc.h:34-69).
// They are evaluated in FP64 and rounded to float, so that they agree with
// the GPU kernels (whose Gaussian uses exact erf values) rather than with the
// reference's FP32 Abramowitz & Stegun approximation, which loses accuracy
// below rho ~ 0.5.  Formulas: reference src/VortFunc.cpp:64-199.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>

#include <cvortex/libcvtx.h>

namespace {

constexpr double kSqrt2OverPi = 0.79788456080286535588;

float eta_missing(float) {
  // reference src/VortFunc.cpp:52-62: one warning, then zero (the CPU leg
  // calls this from several threads at once)
  static std::atomic<bool> warned{false};
  if (!warned.exchange(true))
    std::fprintf(stderr,
                 "cvortex: eta requested from a regularisation without a viscous kernel "
                 "(singular or planetary); returning zero.\n");
  return 0.f;
}

// ---- singular ----
float sing_g3(float) { return 1.f; }
float sing_g2(float) { return 1.f; }
float sing_zeta3(float) { return 0.f; }
void sing_comb3(float, float* g, float* z) {
  *g = 1.f;
  *z = 0.f;
}

// ---- Winckelmans ----
float winck_g3(float rho) {
  const double r = rho, a = r * r + 1.0;
  return (float)((r * r + 2.5) * r * r * r * std::pow(a, -2.5));
}
float winck_zeta3(float rho) {
  const double r = rho;
  return (float)(7.5 * std::pow(r * r + 1.0, -3.5));
}
float winck_eta3(float rho) {
  const double r = rho;
  return (float)(52.5 * std::pow(r * r + 1.0, -4.5));
}
void winck_comb3(float rho, float* g, float* z) {
  *g = winck_g3(rho);
  *z = winck_zeta3(rho);
}
float winck_g2(float rho) {
  const double q = (double)rho * rho;
  return (float)((q * q + 2.0 * q) / (q * q + 2.0 * q + 1.0));
}
float winck_eta2(float rho) {
  // 24 exp(4 / a^3) / a^4 with a = rho^2 + 1, as defined by the reference
  // (src/VortFunc.cpp:124-131 and nbody.cl:597-599)
  const double a = (double)rho * rho + 1.0;
  return (float)(24.0 * std::exp(4.0 / (a * a * a)) / (a * a * a * a));
}

// ---- planetary ----
float plan_g3(float rho) { return rho < 1.f ? (float)((double)rho * rho * rho) : 1.f; }
float plan_zeta3(float rho) { return rho < 1.f ? 3.f : 0.f; }
void plan_comb3(float rho, float* g, float* z) {
  *g = plan_g3(rho);
  *z = plan_zeta3(rho);
}
float plan_g2(float rho) { return rho < 1.f ? (float)((double)rho * rho) : 1.f; }

// ---- Gaussian ----
double gauss_g3_d(double r) {
  if (r < 0.5) {
    // g = rho^3 sqrt(2/pi) sum_k (-rho^2/2)^k / (k! (2k+3))
    const double x = -0.5 * r * r;
    double term = 1.0, sum = 0.0;
    for (int k = 0; k < 30; ++k) {
      sum += term / (2.0 * k + 3.0);
      term *= x / (k + 1.0);
    }
    return kSqrt2OverPi * sum * r * r * r;
  }
  return std::erf(r / std::sqrt(2.0)) - r * kSqrt2OverPi * std::exp(-0.5 * r * r);
}
float gauss_g3(float rho) { return (float)gauss_g3_d(rho); }
float gauss_zeta3(float rho) {
  const double r = rho;
  return (float)(kSqrt2OverPi * std::exp(-0.5 * r * r));
}
void gauss_comb3(float rho, float* g, float* z) {
  *g = gauss_g3(rho);
  *z = gauss_zeta3(rho);
}
float gauss_g2(float rho) {
  const double r = rho;
  return (float)(-std::expm1(-0.5 * r * r));
}
float gauss_eta2(float rho) {
  const double r = rho;
  return (float)std::exp(-0.5 * r * r);
}

cvtx_VortFunc make(float (*g3)(float), float (*g2)(float), float (*z3)(float),
                   void (*c3)(float, float*, float*), float (*e3)(float), float (*e2)(float),
                   const char* name) {
  cvtx_VortFunc f;
  std::memset(&f, 0, sizeof(f));
  f.g_3D = g3;
  f.g_2D = g2;
  f.zeta_3D = z3;
  f.combined_3D = c3;
  f.eta_3D = e3;
  f.eta_2D = e2;
  std::strncpy(f.cl_kernel_name_ext, name, sizeof(f.cl_kernel_name_ext) - 1);
  return f;
}

}  // namespace

extern "C" {

// reference src/VortFunc.cpp:201-212
const cvtx_VortFunc cvtx_VortFunc_singular(void) {
  return make(sing_g3, sing_g2, sing_zeta3, sing_comb3, eta_missing, eta_missing, "singular");
}

// reference src/VortFunc.cpp:214-225
const cvtx_VortFunc cvtx_VortFunc_winckelmans(void) {
  return make(winck_g3, winck_g2, winck_zeta3, winck_comb3, winck_eta3, winck_eta2, "winckelmans");
}

// reference src/VortFunc.cpp:227-238
const cvtx_VortFunc cvtx_VortFunc_planetary(void) {
  return make(plan_g3, plan_g2, plan_zeta3, plan_comb3, eta_missing, eta_missing, "planetary");
}

// reference src/VortFunc.cpp:240-251 (eta_3D is zeta_3D for the Gaussian)
const cvtx_VortFunc cvtx_VortFunc_gaussian(void) {
  return make(gauss_g3, gauss_g2, gauss_zeta3, gauss_comb3, gauss_zeta3, gauss_eta2, "gaussian");
}

}  // extern "C"
