// noalloc.hip -- the P3D device-resident evaluation that allocates nothing.
//
// cvtx_hip_p3d_eval_shape() takes the FP64 partial sums of a
// source-partitioned launch from the library's grow-only workspace, which
// may call hipMalloc/hipFree when a call needs more than any call before it,
// and orders Gaussian targets with a library sort that takes temporary
// storage from the same workspace.  Callers that cannot allocate (code
// under HIP graph capture, tight serving loops) use
// cvtx_hip_p3d_eval_shape_noalloc() instead: the partial sums live in a
// fixed array of kScratchBytes in device memory, defined below, which every
// GPU the library runs on holds from the time the library's code is loaded
// until the process ends.  The call
//   * takes the caller's launch shape as it is and rejects, before any
//     launch, a shape whose partial sums need more than kScratchBytes;
//   * does not reorder targets (the results are the same either way);
//   * allocates nothing and needs no clearing: the main kernel writes every
//     partial sum it owns before the reduction kernel reads it.
// All calls on one GPU share its array.  Calls that are not captured into
// a graph are ordered across streams (an event per call); calls that run
// concurrently anyway -- replays of captured graphs on different streams --
// race on the array.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <string>

#include <cvortex_hip.h>

#include "api_common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

// The scratch array.  In host code this name is a stand-in for the symbol,
// not a device address: host code uses the address attach_scratch() looks up
// for each device (Device::scratch) and never the name itself.
static __device__ __align__(256) unsigned char scratch[CVTX_HIP_NOALLOC_SCRATCH_BYTES];

namespace cvtx {

namespace {

constexpr size_t kScratchBytes = (size_t)CVTX_HIP_NOALLOC_SCRATCH_BYTES;

}  // namespace

bool attach_scratch(Runtime& rt, Device& d) {
  void* ptr = nullptr;
  if (!check_hip(rt, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch symbol address"))
    return false;
  if (ptr == nullptr) {
    set_error(rt, "scratch symbol address is null");
    return false;
  }
  d.scratch = ptr;
  d.scratch_bytes = kScratchBytes;
  return true;
}

}  // namespace cvtx

using namespace cvtx;

extern "C" {

int cvtx_hip_p3d_eval_shape_noalloc(const cvtx_hip_p3d_problem* p, int targets_per_lane,
                                    int n_split, void* stream) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  Device* d = active_device(rt);
  if (!d) return -1;
  if (!d->scratch) {
    set_error(rt, "no scratch array on this device");
    return -1;
  }
  P3DJob j;
  if (!p3d_job_from_problem(rt, *d, p, &j)) return -1;
  LaunchCfg c;
  if (!make_launch_shape(targets_per_lane, n_split, j.n_src, &c)) {
    set_error(rt, "invalid launch shape");
    return -1;
  }
  const size_t need = p3d_workspace_bytes(j, c);
  if (need > d->scratch_bytes) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "the launch needs %zu bytes of partial sums (%d partitions x %d targets), "
                  "more than the %zu-byte scratch array; use fewer partitions",
                  need, c.n_split, j.n_tgt, d->scratch_bytes);
    set_error(rt, msg);
    return -1;
  }
  j.partial = need ? (double*)d->scratch : nullptr;
  const hipStream_t s = pick_stream(d, stream);
  StreamUse use(d->scratch_use, s);
  return check_hip(rt, launch_p3d(j, c, s), "P3D launch") ? 0 : -1;
}

}  // extern "C"
