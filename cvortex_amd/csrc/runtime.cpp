// runtime.cpp -- device registry and buffers (see runtime.hpp).
#include "runtime.hpp"

#include <cstdio>
#include <cstring>
#include <sstream>

namespace cvtx {

Runtime& runtime() {
  static Runtime rt;
  return rt;
}

void set_error(Runtime& rt, const std::string& msg) { rt.last_error = msg; }

bool check_hip(Runtime& rt, hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::ostringstream os;
  os << what << ": " << hipGetErrorString(e);
  set_error(rt, os.str());
  return false;
}

int reg_from_name(const char* name) {
  if (name == nullptr) return -1;
  char buf[33];
  std::memcpy(buf, name, 32);
  buf[32] = 0;
  if (!std::strcmp(buf, "singular")) return REG_SINGULAR;
  if (!std::strcmp(buf, "winckelmans")) return REG_WINCKELMANS;
  if (!std::strcmp(buf, "planetary")) return REG_PLANETARY;
  if (!std::strcmp(buf, "gaussian")) return REG_GAUSSIAN;
  return -1;
}

void initialise_locked(Runtime& rt) {
  if (rt.initialised) return;
  rt.devices.clear();
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  for (int i = 0; i < n; ++i) {
    Device d;
    d.id = i;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, i) == hipSuccess) {
      d.name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
      d.n_cu = prop.multiProcessorCount;
    } else {
      d.name = "HIP device " + std::to_string(i);
    }
    rt.devices.push_back(d);
  }
  if (!rt.devices.empty()) rt.devices[0].enabled = true;
  std::ostringstream os;
  os << "cvortex-mi355x 0.1.0 (HIP, gfx950 direct-sum Biot-Savart)\n";
  os << "compiler: " << __VERSION__ << "\n";
  os << "HIP runtime: " << HIP_VERSION_MAJOR << "." << HIP_VERSION_MINOR << "\n";
  os << "devices: " << rt.devices.size() << "\n";
  for (auto& d : rt.devices) os << "  [" << d.id << "] " << d.name << " CUs=" << d.n_cu << "\n";
  rt.info = os.str();
  rt.initialised = true;
}

static void release_device(Device& d) {
  if (!d.ready) return;
  (void)hipSetDevice(d.id);
  if (d.stream) (void)hipStreamSynchronize(d.stream);
  for (int s = 0; s < SLOT_COUNT; ++s) {
    if (d.dev[s].p) (void)hipFree(d.dev[s].p);
    if (d.host[s].p) (void)hipHostFree(d.host[s].p);
    d.dev[s] = Buffer{};
    d.host[s] = Buffer{};
  }
  if (d.gtab) (void)hipFree(d.gtab);
  d.gtab = nullptr;
  for (UseOrder* o : {&d.slot_use, &d.scratch_use}) {
    if (o->ev) (void)hipEventDestroy(o->ev);
    *o = UseOrder{};
  }
  d.scratch = nullptr;
  d.scratch_bytes = 0;
  if (d.stream) (void)hipStreamDestroy(d.stream);
  d.stream = nullptr;
  d.ready = false;
}

void finalise_locked(Runtime& rt) {
  for (auto& d : rt.devices) release_device(d);
  rt.devices.clear();
  rt.initialised = false;
  ++rt.generation;
}

static bool make_ready(Runtime& rt, Device& d) {
  if (d.ready) return true;
  if (!check_hip(rt, hipSetDevice(d.id), "hipSetDevice")) return false;
  if (!check_hip(rt, hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking), "hipStreamCreate"))
    return false;
  // ordering events: waited on by other streams of this device only, so no
  // timing and no system-scope fence (a device-scope release suffices)
  for (UseOrder* o : {&d.slot_use, &d.scratch_use}) {
    if (!check_hip(rt,
                   hipEventCreateWithFlags(&o->ev,
                                           hipEventDisableTiming | hipEventDisableSystemFence),
                   "hipEventCreate"))
      return false;
    o->own = d.stream;
  }
  if (!attach_scratch(rt, d)) return false;
  float tab[8 * GTAB_N];
  build_gaussian_table(tab);
  if (!check_hip(rt, hipMalloc(&d.gtab, sizeof(tab)), "hipMalloc(table)")) return false;
  if (!check_hip(rt, hipMemcpy(d.gtab, tab, sizeof(tab), hipMemcpyHostToDevice), "hipMemcpy(table)"))
    return false;
  d.ready = true;
  return true;
}

Device* active_device(Runtime& rt) {
  initialise_locked(rt);
  for (auto& d : rt.devices) {
    if (!d.enabled) continue;
    if (!make_ready(rt, d)) return nullptr;
    if (!check_hip(rt, hipSetDevice(d.id), "hipSetDevice")) return nullptr;
    return &d;
  }
  set_error(rt, rt.devices.empty() ? "no HIP device present" : "no accelerator enabled");
  return nullptr;
}

void* device_buffer(Runtime& rt, Device& d, int slot, size_t bytes) {
  Buffer& b = d.dev[slot];
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return b.p;
  if (b.p) {
    // queued work on any stream (device-resident calls take the caller's
    // stream) may still use the old buffer
    (void)hipDeviceSynchronize();
    (void)hipFree(b.p);
    b = Buffer{};
  }
  size_t want = bytes + bytes / 4;
  if (!check_hip(rt, hipMalloc(&b.p, want), "hipMalloc")) {
    b = Buffer{};
    return nullptr;
  }
  b.bytes = want;
  return b.p;
}

void* host_buffer(Runtime& rt, Device& d, int slot, size_t bytes) {
  Buffer& b = d.host[slot];
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return b.p;
  if (b.p) {
    (void)hipStreamSynchronize(d.stream);
    (void)hipHostFree(b.p);
    b = Buffer{};
  }
  size_t want = bytes + bytes / 4;
  if (!check_hip(rt, hipHostMalloc(&b.p, want, hipHostMallocDefault), "hipHostMalloc")) {
    b = Buffer{};
    return nullptr;
  }
  b.bytes = want;
  return b.p;
}

StreamUse::StreamUse(UseOrder& o, hipStream_t s) : o_(o), s_(s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  capturing_ = hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
  if (capturing_ || !o_.stream || o_.stream == s) return;
  if (o_.lazy) {
    // the last use ran on the device's own stream: mark it now
    o_.pending = hipEventRecord(o_.ev, o_.stream) == hipSuccess;
    o_.lazy = false;
  }
  if (o_.pending) (void)hipStreamWaitEvent(s, o_.ev, 0);
}

StreamUse::~StreamUse() {
  if (capturing_ || !o_.ev) return;
  if (s_ == o_.own) {
    o_.stream = s_;
    o_.lazy = true;
    o_.pending = false;
    return;
  }
  if (hipEventRecord(o_.ev, s_) == hipSuccess) {
    o_.stream = s_;
    o_.pending = true;
    o_.lazy = false;
  }
}

void ready_enabled_devices(Runtime& rt) {
  for (auto& d : rt.devices)
    if (d.enabled) (void)make_ready(rt, d);
}

}  // namespace cvtx
