// runtime.hpp -- device registry, per-device streams and buffers, host gather.
//
// Replaces the role of the reference's OpenCL device runtime
// (src/opencl_acc.cpp, src/OclPlatformState.cpp, src/OclDeviceState.cpp):
// an accelerator is a HIP device; cvtx_initialise enables the first one;
// each device gets one stream (default priority) and grow-only device and
// pinned host buffers that are reused across calls (the reference allocated
// and freed OpenCL buffers per call and per 256-source chunk,
// src/ocl_P3D.cpp:193,260-271,289).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace cvtx {

enum Slot : int {
  SLOT_SRC0 = 0,
  SLOT_SRC1,
  SLOT_TGT0,
  SLOT_TGT1,
  SLOT_OUT0,
  SLOT_OUT1,
  SLOT_OUT2,
  SLOT_WORK,
  SLOT_ORDER,  // target-order workspace
  SLOT_PERM,   // target permutation
  SLOT_PERM2,  // source permutation (vorticity field)
  SLOT_RD,     // redistribution workspace: grid entries
  SLOT_RD_NODES,  // redistribution workspace: nodes
  SLOT_RD_CHUNKS,  // redistribution workspace: per-chunk counts and scan
  SLOT_RD_SMALL,
  SLOT_COUNT
};

// Which leg the public void functions use (cvtx_hip_set_dispatch).
enum Dispatch : int { DISPATCH_AUTO = 0, DISPATCH_DEVICE = 1, DISPATCH_HOST = 2 };

struct Buffer {
  void* p = nullptr;
  size_t bytes = 0;
};

// The last use of a shared device buffer, on `stream`: either an event
// recorded after it (pending), or, when that stream is the device's own
// (`own`, which lives as long as the device), nothing yet (lazy): the event
// is recorded on `own` only when a call on another stream needs to wait.
struct UseOrder {
  hipEvent_t ev = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t own = nullptr;
  bool pending = false;
  bool lazy = false;
};

struct Device {
  int id = -1;
  std::string name;
  bool enabled = false;
  bool ready = false;
  int n_cu = 0;
  hipStream_t stream = nullptr;
  float4* gtab = nullptr;  // Gaussian near-field table
  // The slot buffers are shared by every call on the device, whatever its
  // stream; so is the fixed scratch array of the no-allocation entry point
  // (noalloc.hip).  Each UseOrder marks the end of the last call's use of
  // one of them (see StreamUse).
  UseOrder slot_use;
  UseOrder scratch_use;
  void* scratch = nullptr;  // this device's address of the scratch array
  size_t scratch_bytes = 0;
  Buffer dev[SLOT_COUNT];
  Buffer host[SLOT_COUNT];  // pinned
};

struct Runtime {
  bool initialised = false;
  std::vector<Device> devices;
  std::string info;
  std::string last_error;
  int acc_mode = ACC_COMP;
  bool order_targets = true;
  int dispatch = DISPATCH_AUTO;  // public void functions: which leg (dispatch.cpp)
  int last_leg = -1;             // leg of the last public call (0 device, 1 host)
  unsigned generation = 0;       // counts cvtx_finalise calls (objects made before are stale)
  std::recursive_mutex mu;
};

Runtime& runtime();

// Idempotent initialisation (device enumeration, device 0 enabled).
void initialise_locked(Runtime& rt);
void finalise_locked(Runtime& rt);

// First enabled device, made ready (stream, table).  nullptr and
// last_error set when there is none or setup fails.
Device* active_device(Runtime& rt);

// Grow-only buffers.  nullptr (and last_error) on allocation failure.
void* device_buffer(Runtime& rt, Device& d, int slot, size_t bytes);
void* host_buffer(Runtime& rt, Device& d, int slot, size_t bytes);

void set_error(Runtime& rt, const std::string& msg);

// Orders the calls that use a shared device buffer across streams: made at
// the start of a call on stream s, it makes s wait for the last such call
// if that ran on another stream; at its end it marks this call's work on s.
// A call on the device's own stream records nothing (an event record costs
// the GPU a few microseconds between launches); its work is marked when,
// and if, a call on another stream comes.  Nothing is recorded or waited
// for while s is being captured into a graph (the graph orders its own
// nodes).  StreamUse(d, s) guards the slot buffers.
class StreamUse {
 public:
  StreamUse(UseOrder& o, hipStream_t s);
  StreamUse(Device& d, hipStream_t s) : StreamUse(d.slot_use, s) {}
  ~StreamUse();
  StreamUse(const StreamUse&) = delete;
  StreamUse& operator=(const StreamUse&) = delete;

 private:
  UseOrder& o_;
  hipStream_t s_;
  bool capturing_ = false;
};

// Looks up this device's address of the no-allocation scratch array
// (noalloc.hip); called with the device current when it is made ready.
bool attach_scratch(Runtime& rt, Device& d);
// Makes every enabled device ready (stream, table, scratch address): the
// set-up cvtx_initialise does so that later calls need not.
void ready_enabled_devices(Runtime& rt);
bool check_hip(Runtime& rt, hipError_t e, const char* what);

// Run body(begin, end) over [0, n) on up to 16 host threads.
template <class F>
void parallel_for(long n, F body);

// Regularisation id from cl_kernel_name_ext, -1 if unknown.
int reg_from_name(const char* name);

}  // namespace cvtx

#include <thread>

namespace cvtx {
template <class F>
void parallel_for(long n, F body) {
  if (n <= 0) return;
  unsigned hw = std::thread::hardware_concurrency();
  long nt = hw == 0 ? 1 : (long)(hw > 16 ? 16 : hw);
  if (n < 65536 || nt <= 1) {
    body(0L, n);
    return;
  }
  if (nt > n / 16384) nt = n / 16384;
  if (nt < 1) nt = 1;
  std::vector<std::thread> th;
  th.reserve(nt);
  const long chunk = (n + nt - 1) / nt;
  for (long t = 0; t < nt; ++t) {
    const long b = t * chunk, e = b + chunk < n ? b + chunk : n;
    if (b >= e) break;
    th.emplace_back([=]() { body(b, e); });
  }
  for (auto& x : th) x.join();
}
}  // namespace cvtx
