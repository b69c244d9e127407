// api_common.hpp -- host staging helpers shared by the C-ABI translation
// units (api.cpp, vort_api.cpp): gathers of the caller's pointer arrays into
// pinned structure-of-arrays buffers, copies on the device stream, failure
// reporting and spatial ordering.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include <cvortex/libcvtx.h>
#include <cvortex_hip.h>

#include "kernels.hpp"
#include "runtime.hpp"

namespace cvtx {

// AoP records -> SoA float4 (parallel over host threads)
void gather_p3d(const cvtx_P3D* const* src, int n, float4* pv, float4* a);
void gather_pts3(const bsv_V3f* pts, int n, float4* pv);
void gather_p2d(const cvtx_P2D* const* src, int n, float4* out);
void gather_pts2(const bsv_V2f* pts, int n, float4* out);
void gather_f3d(const cvtx_F3D* const* src, int n, float4* fa, float4* fb);

void fill_nan(float* out, long n);
void warn_no_eta_once();
// true (and the error recorded) for negative counts or missing arrays
bool bad_counts(Runtime& rt, int ns, int nt, const void* src, const void* tgt);
// asynchronous copies on the device's stream
bool h2d(Runtime& rt, Device& d, void* dst, const void* src, size_t bytes);
bool d2h(Runtime& rt, Device& d, void* dst, const void* src, size_t bytes);
// stderr message with the last error, and NaN results
void report_failure(const char* fn, float* out, long n);
hipStream_t pick_stream(Device* d, void* s);
// Morton permutation of n points into runtime slot `slot` (nullptr on failure)
int* order_points(Runtime& rt, Device& d, const float4* pts, int n, int slot, hipStream_t s);
// Vorticity field of host records at host points (vort_api.cpp); cutoff =
// false for the reference's single-source forms (no 5 sigma box).
int vort_host(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nt, float* out,
              const cvtx_VortFunc* vf, float sigma, bool cutoff);
// Redistribution on a grid (redist.hip): returns the number of output
// particles (>= 0) or -1 with the error recorded.  With j.max_out >= 0,
// *out_dev holds the packed records (7 floats per P3D, 4 per P2D) in a
// library buffer, valid until the next call.
int redistribute(Runtime& rt, Device& d, const RedistJob& j, float** out_dev, hipStream_t s);
// Public-function dispatch (dispatch.cpp).  want_device: whether to offer a
// call to the seam (the dispatch mode, an enabled device, a known kernel
// name when vf is given, and the reference's size guard when the mode is
// automatic); device_done / host_done record the leg; device_declined
// reports a failed device call that falls back to the host; host_args_ok
// reports invalid arguments (NaN results) and returns false.
bool want_device(bool guard_met, const cvtx_VortFunc* vf);
void device_done();
void device_declined(const char* fn);
bool host_args_ok(const char* fn, bool ok, float* out, long n_out);
void host_done(const char* fn, int status, const std::string& err, float* out, long n_out);
// One public evaluation: the device when want_device(guard, ...) agrees and
// the seam succeeds, else the host leg (dispatch.cpp).
void run_p3d_vel(const char* fn, bool guard, const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                 int nm, bsv_V3f* out, const cvtx_VortFunc* k, float sigma);
void run_p3d_dvort(const char* fn, bool guard, const cvtx_P3D** src, int ns,
                   const cvtx_P3D** ind, int ni, bsv_V3f* out, const cvtx_VortFunc* k,
                   float sigma);
void run_p3d_visc(const char* fn, bool guard, const cvtx_P3D** src, int ns,
                  const cvtx_P3D** ind, int ni, bsv_V3f* out, const cvtx_VortFunc* k,
                  float sigma, float nu);
void run_p3d_vort(const char* fn, bool guard, const cvtx_P3D** src, int ns, const bsv_V3f* mes,
                  int nm, bsv_V3f* out, const cvtx_VortFunc* k, float sigma, bool cutoff);
void run_f3d_vel(const char* fn, bool guard, const cvtx_F3D** src, int nf, const bsv_V3f* mes,
                 int nm, bsv_V3f* out);
void run_f3d_dvort(const char* fn, bool guard, const cvtx_F3D** src, int nf,
                   const cvtx_P3D** ind, int ni, bsv_V3f* out);
void run_p2d_vel(const char* fn, bool guard, const cvtx_P2D** src, int ns, const bsv_V2f* mes,
                 int nm, bsv_V2f* out, const cvtx_VortFunc* k, float sigma);
void run_p2d_visc(const char* fn, bool guard, const cvtx_P2D** src, int ns,
                  const cvtx_P2D** ind, int ni, float* out, const cvtx_VortFunc* k, float sigma,
                  float nu);
// A device-resident problem as a P3D job on device d (validated; false and
// the error recorded when it is invalid).
bool p3d_job_from_problem(Runtime& rt, const Device& d, const cvtx_hip_p3d_problem* p, P3DJob* j);
// explicit launch shape: T in {1, 2, 4}, n_split >= 1 source partitions
// (rounded to whole 256-source tiles); false for an invalid shape
bool make_launch_shape(int T, int n_split, int n_src, LaunchCfg* c);

}  // namespace cvtx
