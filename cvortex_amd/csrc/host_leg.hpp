// host_leg.hpp -- the library's CPU evaluation of the direct sums.
//
// The reference library runs its own OpenMP loops whenever its accelerator
// seam declines a call: no device enabled, a regularisation without a GPU
// kernel name (a caller's own cvtx_VortFunc), a problem below the size
// guards, or a device error (reference src/P3D.cpp:343-456,
// src/F3D.cpp:162-202, src/P2D.cpp:141-276, src/accelerators.cpp:107-118).
// These functions are that leg of this library.  Regularisations are
// evaluated through the caller's cvtx_VortFunc function pointers, once per
// pair, as the reference does; geometry and per-target sums are FP64 and
// each output is rounded to FP32 once.  Work is spread over host threads:
// by target when there are enough targets, else by fixed-size source
// chunks whose partial sums are added in chunk order, so results do not
// depend on the thread count.
//
// Each function returns 0, or -1 with a message in *err when a function
// pointer it needs is missing (the reference asserts on that).
#pragma once

#include <string>

#include <cvortex/libcvtx.h>

namespace cvtx {
namespace host {

// Host threads the leg uses: CVTX_NUM_THREADS, else OMP_NUM_THREADS, else
// the hardware concurrency (at most 64).
int threads();

int p3d_vel(const cvtx_P3D* const* src, long ns, const bsv_V3f* mes, long nm, bsv_V3f* out,
            const cvtx_VortFunc* vf, float sigma, std::string* err);
int p3d_dvort(const cvtx_P3D* const* src, long ns, const cvtx_P3D* const* ind, long ni,
              bsv_V3f* out, const cvtx_VortFunc* vf, float sigma, std::string* err);
int p3d_visc(const cvtx_P3D* const* src, long ns, const cvtx_P3D* const* ind, long ni,
             bsv_V3f* out, const cvtx_VortFunc* vf, float sigma, float nu, std::string* err);
// box_cutoff: the M2S/M2M semantics (only sources inside the 5 sigma box
// around the point); false: the single-source forms (every source).
int p3d_vort(const cvtx_P3D* const* src, long ns, const bsv_V3f* mes, long nm, bsv_V3f* out,
             const cvtx_VortFunc* vf, float sigma, bool box_cutoff, std::string* err);
int p3d_relax(cvtx_P3D* const* p, long n, float fdt, const cvtx_VortFunc* vf, float sigma,
              std::string* err);

void f3d_vel(const cvtx_F3D* const* f, long nf, const bsv_V3f* mes, long nm, bsv_V3f* out);
void f3d_dvort(const cvtx_F3D* const* f, long nf, const cvtx_P3D* const* ind, long ni,
               bsv_V3f* out);
void f3d_inf(const cvtx_F3D* const* f, long nf, const bsv_V3f* mes, const bsv_V3f* dir, long nm,
             float* out);

int p2d_vel(const cvtx_P2D* const* src, long ns, const bsv_V2f* mes, long nm, bsv_V2f* out,
            const cvtx_VortFunc* vf, float sigma, std::string* err);
int p2d_visc(const cvtx_P2D* const* src, long ns, const cvtx_P2D* const* ind, long ni,
             float* out, const cvtx_VortFunc* vf, float sigma, float nu, std::string* err);

}  // namespace host
}  // namespace cvtx
