// batched.hip -- many independent P3D particle clouds in one persistent
// launch, and RK4 time steps captured as one HIP graph per step.
//
// A batch is a set of clouds stored back to back in the same SoA buffers;
// cloud c owns particles [begin_c, end_c) and its own regularisation radius.
// Each cloud interacts only with itself (sources = targets = the cloud).
//
// Work units are (cloud, block of WG*T targets), sorted by cost (cloud size)
// in decreasing order -- the longest-processing-time-first rule -- so that
// the 64x spread of cloud sizes (1024 .. 65536 particles) leaves no long
// tail.  A persistent grid of a few workgroups per CU pulls units from a
// global atomic counter; every workgroup leaves the loop once the counter
// passes the number of units.  The counter is reset by a one-thread kernel
// launched before each evaluation, so an evaluation can be captured in a
// graph (no memset node).
//
// RK4 (classical, low storage):
//   k1 = f(y0);  S = k1;   ys = y0 + dt/2 k1
//   k2 = f(ys);  S += 2k2; ys = y0 + dt/2 k2
//   k3 = f(ys);  S += 2k3; ys = y0 + dt k3
//   k4 = f(ys);  y0 += dt/6 (S + k4)
// with y = (positions, vorticity), f = (vel, dvort) from the batched kernel;
// particle volumes are carried unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include <cvortex_hip.h>

#include "p3d_body.hpp"
#include "runtime.hpp"
#include "tune.hpp"

namespace cvtx {

struct BatchParams {
  P3DKParams base;  // concatenated particle arrays and outputs; n_split = 1
  const int2* __restrict__ units;          // (cloud, first target)
  const int2* __restrict__ range;          // cloud -> (begin, end)
  const RegConsts* __restrict__ rc;        // per cloud
  const Scales* __restrict__ sc;           // per cloud
  unsigned int* __restrict__ counter;
  int n_units;
};

__global__ void batch_reset_counter(unsigned int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
}

template <int REG, int MASK, int ACC, int T, int WG>
__global__ __launch_bounds__(WG) void p3d_batched_kernel(BatchParams bp) {
  __shared__ P3DSmem<REG, MASK, ACC, WG> sm;
  __shared__ int s_unit;
  if constexpr (REG == REG_GAUSSIAN) {
    for (int i = threadIdx.x; i < 2 * GTAB_N; i += WG) sm.tab[i] = bp.base.gtab[i];
  }
  for (;;) {
    __syncthreads();  // everyone has read s_unit (and the table is staged)
    if (threadIdx.x == 0) s_unit = (int)atomicAdd(bp.counter, 1u);
    __syncthreads();
    const int u = s_unit;
    if (u >= bp.n_units) break;  // uniform over the workgroup
    const int2 un = bp.units[u];
    const int2 r = bp.range[un.x];
    p3d_block<REG, MASK, ACC, T, WG>(bp.base, bp.rc[un.x], bp.sc[un.x], un.y, r.y, r.x, r.y, 0, sm);
  }
}

// ys = y0 + c * k;  S = w*k (first) or S + w*k
__global__ void rk4_stage(int n, const float4* __restrict__ y0p, const float4* __restrict__ y0a,
                          const float* __restrict__ kv, const float* __restrict__ kd,
                          float4* __restrict__ ysp, float4* __restrict__ ysa,
                          float4* __restrict__ Sp, float4* __restrict__ Sa, float c, float w,
                          int first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float vx = kv[3 * (size_t)i], vy = kv[3 * (size_t)i + 1], vz = kv[3 * (size_t)i + 2];
  const float dx = kd[3 * (size_t)i], dy = kd[3 * (size_t)i + 1], dz = kd[3 * (size_t)i + 2];
  const float4 p = y0p[i], a = y0a[i];
  ysp[i] = make_float4(__builtin_fmaf(c, vx, p.x), __builtin_fmaf(c, vy, p.y),
                       __builtin_fmaf(c, vz, p.z), p.w);
  ysa[i] = make_float4(__builtin_fmaf(c, dx, a.x), __builtin_fmaf(c, dy, a.y),
                       __builtin_fmaf(c, dz, a.z), 0.f);
  if (first) {
    Sp[i] = make_float4(w * vx, w * vy, w * vz, 0.f);
    Sa[i] = make_float4(w * dx, w * dy, w * dz, 0.f);
  } else {
    const float4 sp = Sp[i], sa = Sa[i];
    Sp[i] = make_float4(__builtin_fmaf(w, vx, sp.x), __builtin_fmaf(w, vy, sp.y),
                        __builtin_fmaf(w, vz, sp.z), 0.f);
    Sa[i] = make_float4(__builtin_fmaf(w, dx, sa.x), __builtin_fmaf(w, dy, sa.y),
                        __builtin_fmaf(w, dz, sa.z), 0.f);
  }
}

// y0 += h * (S + k4)
__global__ void rk4_final(int n, float4* __restrict__ y0p, float4* __restrict__ y0a,
                          const float* __restrict__ kv, const float* __restrict__ kd,
                          const float4* __restrict__ Sp, const float4* __restrict__ Sa, float h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 sp = Sp[i], sa = Sa[i];
  float4 p = y0p[i], a = y0a[i];
  p.x = __builtin_fmaf(h, sp.x + kv[3 * (size_t)i], p.x);
  p.y = __builtin_fmaf(h, sp.y + kv[3 * (size_t)i + 1], p.y);
  p.z = __builtin_fmaf(h, sp.z + kv[3 * (size_t)i + 2], p.z);
  a.x = __builtin_fmaf(h, sa.x + kd[3 * (size_t)i], a.x);
  a.y = __builtin_fmaf(h, sa.y + kd[3 * (size_t)i + 1], a.y);
  a.z = __builtin_fmaf(h, sa.z + kd[3 * (size_t)i + 2], a.z);
  y0p[i] = p;
  y0a[i] = a;
}

namespace {

using BFn = void (*)(BatchParams);

template <int REG, int MASK, int ACC, int WG>
BFn b_T(int T) {
  switch (T) {
    case 1: return p3d_batched_kernel<REG, MASK, ACC, 1, WG>;
    case 4: return p3d_batched_kernel<REG, MASK, ACC, 4, WG>;
    default: return p3d_batched_kernel<REG, MASK, ACC, 2, WG>;
  }
}
template <int REG, int MASK, int ACC>
BFn b_wg(int wg, int T) {
  return wg == 128 ? b_T<REG, MASK, ACC, 128>(T) : b_T<REG, MASK, ACC, 256>(T);
}
template <int REG, int MASK>
BFn b_acc(int acc, int wg, int T) {
  if (acc == ACC_PLAIN) return b_wg<REG, MASK, ACC_PLAIN>(wg, T);
  if (acc == ACC_TILES) return b_wg<REG, MASK, ACC_TILES>(wg, T);
  return b_wg<REG, MASK, ACC_COMP>(wg, T);
}
template <int REG>
BFn b_mask(int mask, int wg, int T, int acc) {
  if (mask == (OUT_VEL | OUT_DVORT)) return b_acc<REG, OUT_VEL | OUT_DVORT>(acc, wg, T);
  if (mask == (OUT_VEL | OUT_DVORT | OUT_VISC))
    return b_acc<REG, OUT_VEL | OUT_DVORT | OUT_VISC>(acc, wg, T);
  return nullptr;
}
BFn b_pick(int reg, int mask, int acc, int wg, int T) {
  if (wg != 128 && wg != 256) return nullptr;
  if (reg == REG_WINCKELMANS) return b_mask<REG_WINCKELMANS>(mask, wg, T, acc);
  if (reg == REG_GAUSSIAN) return b_mask<REG_GAUSSIAN>(mask, wg, T, acc);
  return nullptr;
}

}  // namespace

}  // namespace cvtx

using namespace cvtx;

struct cvtx_hip_batch {
  Device* dev = nullptr;  // valid while gen equals the runtime's generation
  unsigned gen = 0;
  int n_clouds = 0;
  int n_particles = 0;
  int reg = REG_WINCKELMANS;
  int mask = OUT_VEL | OUT_DVORT;
  int acc = ACC_COMP;
  float nu = 0.f;
  int T = 2;
  int wg = 256;  // workgroup size = source tile length
  int wpc = 4;   // resident workgroups per CU of the persistent grid
  int n_units = 0;
  int grid = 0;
  std::vector<int> begin, end;
  std::vector<float> sigma;
  // device
  int2* d_units = nullptr;
  int2* d_range = nullptr;
  RegConsts* d_rc = nullptr;
  Scales* d_sc = nullptr;
  unsigned int* d_counter = nullptr;
  double* d_work = nullptr;
  size_t work_bytes = 0;
  // RK4
  float4 *ysp = nullptr, *ysa = nullptr, *Sp = nullptr, *Sa = nullptr;
  float *kv = nullptr, *kd = nullptr;
  float4 *y0p = nullptr, *y0a = nullptr;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// RK4 stage buffers, capture stream and graph (also after a failed init)
void rk4_release(cvtx_hip_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  b->exec = nullptr;
  b->graph = nullptr;
  b->cap_stream = nullptr;
  float4** f4[] = {&b->ysp, &b->ysa, &b->Sp, &b->Sa};
  for (float4** p : f4) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  float** f1[] = {&b->kv, &b->kd};
  for (float** p : f1) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

void batch_free(cvtx_hip_batch* b) {
  if (!b) return;
  rk4_release(b);
  void* ptrs[] = {b->d_units, b->d_range, b->d_rc, b->d_sc, b->d_counter, b->d_work};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete b;
}

// A batch holds a pointer into the runtime's device list and a graph that
// names the device's Gaussian table: both end with cvtx_finalise.
bool batch_live(Runtime& rt, const cvtx_hip_batch* b) {
  if (!b) {
    set_error(rt, "batch: null batch");
    return false;
  }
  if (b->gen != rt.generation || !rt.initialised) {
    set_error(rt, "batch: made before cvtx_finalise; destroy it and make a new one");
    return false;
  }
  return true;
}

bool build_units(cvtx_hip_batch* b, int T, int wg) {
  const int per = wg * T;
  std::vector<int> order(b->n_clouds);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return (b->end[x] - b->begin[x]) > (b->end[y] - b->begin[y]);
  });
  std::vector<int2> units;
  for (int c : order)
    for (int t = b->begin[c]; t < b->end[c]; t += per) units.push_back(make_int2(c, t));
  b->n_units = (int)units.size();
  b->T = T;
  b->wg = wg;
  if (b->d_units) (void)hipFree(b->d_units);
  b->d_units = nullptr;
  if (units.empty()) return true;
  if (hipMalloc(&b->d_units, sizeof(int2) * units.size()) != hipSuccess) return false;
  return hipMemcpy(b->d_units, units.data(), sizeof(int2) * units.size(), hipMemcpyHostToDevice) ==
         hipSuccess;
}

P3DKParams base_params(cvtx_hip_batch* b, const float* pv, const float* a, float* vel, float* dvort,
                       float* visc) {
  P3DKParams p{};
  p.src_pv = (const float4*)pv;
  p.src_a = (const float4*)a;
  p.tgt_pv = (const float4*)pv;
  p.tgt_a = (const float4*)a;
  p.gtab = b->dev->gtab;
  p.out_vel = vel;
  p.out_dvort = dvort;
  p.out_visc = visc;
  p.n_src = b->n_particles;
  p.n_tgt = b->n_particles;
  p.n_split = 1;
  p.split_len = b->n_particles;
  return p;
}

hipError_t enqueue_eval(cvtx_hip_batch* b, const float* pv, const float* a, float* vel,
                        float* dvort, float* visc, hipStream_t s) {
  if (b->n_units == 0) return hipSuccess;
  BFn k = b_pick(b->reg, b->mask, b->acc, b->wg, b->T);
  if (!k) return hipErrorInvalidValue;
  BatchParams bp{};
  bp.base = base_params(b, pv, a, vel, dvort, visc);
  bp.units = b->d_units;
  bp.range = b->d_range;
  bp.rc = b->d_rc;
  bp.sc = b->d_sc;
  bp.counter = b->d_counter;
  bp.n_units = b->n_units;
  hipLaunchKernelGGL(batch_reset_counter, dim3(1), dim3(64), 0, s, b->d_counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(b->grid), dim3(b->wg), 0, s, bp);
  return hipGetLastError();
}

// What the autotuner's cache keys a batch by: regularisation, outputs,
// accumulation, and the buckets (floor log2) of its total pair count and of
// its number of clouds.
BatchSignature batch_signature(const cvtx_hip_batch* b) {
  long long pairs = 0;
  for (int c = 0; c < b->n_clouds; ++c) {
    const long long n = b->end[c] - b->begin[c];
    pairs += n * n;
  }
  BatchSignature g;
  g.reg = b->reg;
  g.mask = b->mask;
  g.acc = b->acc;
  g.pairs = pairs;
  g.n_clouds = b->n_clouds;
  return g;
}

}  // namespace

extern "C" {

cvtx_hip_batch* cvtx_hip_batch_create(int n_clouds, const int* offsets, const float* sigmas,
                                      int regularisation, int outputs, int accumulate, float nu) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (n_clouds < 0 || !offsets || (n_clouds > 0 && !sigmas)) {
    set_error(rt, "batch: invalid clouds");
    return nullptr;
  }
  if (regularisation != REG_WINCKELMANS && regularisation != REG_GAUSSIAN) {
    set_error(rt, "batch: regularisation must be winckelmans or gaussian");
    return nullptr;
  }
  if (outputs != (OUT_VEL | OUT_DVORT) && outputs != (OUT_VEL | OUT_DVORT | OUT_VISC)) {
    set_error(rt, "batch: outputs must be vel|dvort or vel|dvort|visc");
    return nullptr;
  }
  if (!valid_acc(accumulate)) {
    set_error(rt, "batch: invalid accumulation mode");
    return nullptr;
  }
  for (int c = 0; c < n_clouds; ++c)
    if (offsets[c + 1] < offsets[c] || offsets[c] < 0) {
      set_error(rt, "batch: offsets must be non-decreasing");
      return nullptr;
    }
  Device* d = active_device(rt);
  if (!d) return nullptr;
  cvtx_hip_batch* b = new cvtx_hip_batch;
  b->dev = d;
  b->gen = rt.generation;
  b->n_clouds = n_clouds;
  b->n_particles = offsets[n_clouds];
  b->reg = regularisation;
  b->mask = outputs;
  b->acc = accumulate;
  b->nu = nu;
  b->begin.assign(offsets, offsets + n_clouds);
  b->end.assign(offsets + 1, offsets + n_clouds + 1);
  b->sigma.assign(sigmas, sigmas + n_clouds);
  std::vector<int2> range(n_clouds);
  std::vector<RegConsts> rc(n_clouds);
  std::vector<Scales> sc(n_clouds);
  for (int c = 0; c < n_clouds; ++c) {
    range[c] = make_int2(offsets[c], offsets[c + 1]);
    rc[c] = p3d_consts(regularisation, sigmas[c]);
    sc[c] = p3d_scales(regularisation, sigmas[c], nu);
  }
  bool ok = true;
  const size_t nc = std::max(1, n_clouds);
  ok = ok && hipMalloc(&b->d_range, sizeof(int2) * nc) == hipSuccess;
  ok = ok && hipMalloc(&b->d_rc, sizeof(RegConsts) * nc) == hipSuccess;
  ok = ok && hipMalloc(&b->d_sc, sizeof(Scales) * nc) == hipSuccess;
  ok = ok && hipMalloc(&b->d_counter, 256) == hipSuccess;
  if (ok && n_clouds > 0) {
    ok = ok && hipMemcpy(b->d_range, range.data(), sizeof(int2) * n_clouds, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b->d_rc, rc.data(), sizeof(RegConsts) * n_clouds, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b->d_sc, sc.data(), sizeof(Scales) * n_clouds, hipMemcpyHostToDevice) == hipSuccess;
  }
  // the autotuner's shape for this batch's signature, if one is cached;
  // else two targets per lane, 256-wide workgroups, 4 resident per CU
  BatchShape shp{2, 256, 4, 0.f};
  (void)batch_shape_lookup(batch_signature(b), &shp);
  b->wpc = shp.wpc;
  ok = ok && build_units(b, shp.T, shp.wg);
  // a few resident workgroups per CU; the loop exits when the units run out
  b->grid = std::max(1, std::min(b->n_units, b->wpc * std::max(1, d->n_cu)));
  if (!ok) {
    set_error(rt, "batch: device allocation failed");
    batch_free(b);
    return nullptr;
  }
  return b;
}

void cvtx_hip_batch_destroy(cvtx_hip_batch* b) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  batch_free(b);
}

int cvtx_hip_batch_set_shape(cvtx_hip_batch* b, int targets_per_lane, int workgroups_per_cu) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (!batch_live(rt, b)) return -1;
  if ((targets_per_lane != 1 && targets_per_lane != 2 && targets_per_lane != 4) ||
      workgroups_per_cu < 1 || workgroups_per_cu > 64 || b->exec) {
    set_error(rt, "batch: invalid shape (or RK4 graph already captured)");
    return -1;
  }
  if (!build_units(b, targets_per_lane, b->wg)) {
    set_error(rt, "batch: allocation failed");
    return -1;
  }
  b->wpc = workgroups_per_cu;
  b->grid = std::max(1, std::min(b->n_units, workgroups_per_cu * std::max(1, b->dev->n_cu)));
  return 0;
}

int cvtx_hip_batch_eval(cvtx_hip_batch* b, const float* pv, const float* a, float* vel,
                        float* dvort, float* visc, void* stream) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (!batch_live(rt, b)) return -1;
  if ((b->n_particles > 0 && (!pv || !a || !vel || !dvort)) ||
      ((b->mask & OUT_VISC) && b->n_particles > 0 && !visc)) {
    set_error(rt, "batch: null buffer");
    return -1;
  }
  hipStream_t s = stream ? (hipStream_t)stream : b->dev->stream;
  return check_hip(rt, enqueue_eval(b, pv, a, vel, dvort, visc, s), "batch eval") ? 0 : -1;
}

// One regular launch per cloud (the unbatched baseline, same kernels and
// accumulation, launch shape from the tuner for each cloud's size).
int cvtx_hip_batch_eval_per_cloud(cvtx_hip_batch* b, const float* pv, const float* a, float* vel,
                                  float* dvort, float* visc, int plain, void* stream) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (!batch_live(rt, b)) return -1;
  if ((b->n_particles > 0 && (!pv || !a || !vel || !dvort)) ||
      ((b->mask & OUT_VISC) && b->n_particles > 0 && !visc)) {
    set_error(rt, "batch: null buffer");
    return -1;
  }
  hipStream_t s = stream ? (hipStream_t)stream : b->dev->stream;
  StreamUse use(*b->dev, s);  // the partitioned launches share the device's workspace
  const float4* P = (const float4*)pv;
  const float4* A = (const float4*)a;
  for (int c = 0; c < b->n_clouds; ++c) {
    const int n = b->end[c] - b->begin[c];
    if (n == 0) continue;
    const int o = b->begin[c];
    P3DJob j;
    j.src_pv = P + o;
    j.src_a = A + o;
    j.n_src = n;
    j.tgt_pv = P + o;
    j.tgt_a = A + o;
    j.n_tgt = n;
    j.out_vel = vel + 3 * (size_t)o;
    j.out_dvort = dvort + 3 * (size_t)o;
    j.out_visc = (b->mask & OUT_VISC) ? visc + 3 * (size_t)o : nullptr;
    j.reg = b->reg;
    j.mask = b->mask;
    j.acc = b->acc;
    j.sigma = b->sigma[c];
    j.nu = b->nu;
    j.gtab = b->dev->gtab;
    LaunchCfg cfg = default_cfg(n, n, b->dev->n_cu);
    if (plain) {  // one target per lane, no source partitions
      cfg.T = 1;
      cfg.n_split = 1;
      cfg.split_len = 0;
    }
    const size_t wb = p3d_workspace_bytes(j, cfg);
    if (wb) {
      j.partial = (double*)device_buffer(rt, *b->dev, SLOT_WORK, wb);
      if (!j.partial) return -1;
    }
    if (!check_hip(rt, launch_p3d(j, cfg, s), "per-cloud launch")) return -1;
    if (plain && !check_hip(rt, hipStreamSynchronize(s), "per-cloud synchronize")) return -1;
  }
  return 0;
}

// Prepare RK4 stepping of the particles in pv/a (updated in place by each
// step) with time step dt: allocates the stage buffers and captures one
// step -- 4 evaluations and 4 update kernels -- into a HIP graph.
int cvtx_hip_batch_rk4_init(cvtx_hip_batch* b, float* pv, float* a, float dt) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (!batch_live(rt, b)) return -1;
  if (b->exec || b->mask != (OUT_VEL | OUT_DVORT) || (b->n_particles > 0 && (!pv || !a))) {
    set_error(rt, "batch: RK4 needs a vel|dvort batch, buffers, and no earlier init");
    return -1;
  }
  const size_t n = (size_t)std::max(1, b->n_particles);
  bool ok = true;
  ok = ok && hipMalloc(&b->ysp, sizeof(float4) * n) == hipSuccess;
  ok = ok && hipMalloc(&b->ysa, sizeof(float4) * n) == hipSuccess;
  ok = ok && hipMalloc(&b->Sp, sizeof(float4) * n) == hipSuccess;
  ok = ok && hipMalloc(&b->Sa, sizeof(float4) * n) == hipSuccess;
  ok = ok && hipMalloc(&b->kv, sizeof(float) * 3 * n) == hipSuccess;
  ok = ok && hipMalloc(&b->kd, sizeof(float) * 3 * n) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) == hipSuccess;
  if (!ok) {
    set_error(rt, "batch: RK4 allocation failed");
    rk4_release(b);
    return -1;
  }
  b->y0p = (float4*)pv;
  b->y0a = (float4*)a;
  const int N = b->n_particles;
  const dim3 g((N + 255) / 256 > 0 ? (N + 255) / 256 : 1), blk(256);
  hipStream_t s = b->cap_stream;
  if (!check_hip(rt, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture")) {
    rk4_release(b);
    return -1;
  }
  hipError_t e = hipSuccess;
  auto eval = [&](const float4* p_, const float4* a_) {
    if (e == hipSuccess) e = enqueue_eval(b, (const float*)p_, (const float*)a_, b->kv, b->kd, nullptr, s);
  };
  auto stage = [&](float c, float w, int first) {
    if (e != hipSuccess || N == 0) return;
    hipLaunchKernelGGL(rk4_stage, g, blk, 0, s, N, b->y0p, b->y0a, b->kv, b->kd, b->ysp, b->ysa,
                       b->Sp, b->Sa, c, w, first);
    e = hipGetLastError();
  };
  eval(b->y0p, b->y0a);
  stage(0.5f * dt, 1.f, 1);
  eval(b->ysp, b->ysa);
  stage(0.5f * dt, 2.f, 0);
  eval(b->ysp, b->ysa);
  stage(dt, 2.f, 0);
  eval(b->ysp, b->ysa);
  if (e == hipSuccess && N > 0) {
    hipLaunchKernelGGL(rk4_final, g, blk, 0, s, N, b->y0p, b->y0a, b->kv, b->kd, b->Sp, b->Sa,
                       dt / 6.f);
    e = hipGetLastError();
  }
  hipError_t e2 = hipStreamEndCapture(s, &b->graph);
  if (!check_hip(rt, e, "RK4 capture") || !check_hip(rt, e2, "end capture") ||
      !check_hip(rt, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0),
                 "graph instantiate")) {
    b->exec = nullptr;
    rk4_release(b);
    return -1;
  }
  return 0;
}

// Launch `n_steps` captured RK4 steps on `stream` (NULL = library stream).
int cvtx_hip_batch_rk4_step(cvtx_hip_batch* b, int n_steps, void* stream) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (!batch_live(rt, b)) return -1;
  if (!b->exec || n_steps < 0) {
    set_error(rt, "batch: RK4 not initialised");
    return -1;
  }
  hipStream_t s = stream ? (hipStream_t)stream : b->dev->stream;
  for (int i = 0; i < n_steps; ++i)
    if (!check_hip(rt, hipGraphLaunch(b->exec, s), "graph launch")) return -1;
  return 0;
}

// Times the persistent launch of this batch (its outputs are overwritten)
// for every candidate shape -- targets per lane {1, 2, 4} x workgroup size
// {128, 256} x resident workgroups per CU {2, 4, 8} -- with HIP events on
// the library's stream, keeps the fastest, and records it in the shape cache
// under the batch's signature, so later batches of the same signature start
// with it.  Returns 0 and the winner, or -1.
int cvtx_hip_batch_autotune(cvtx_hip_batch* b, const float* pv, const float* a, float* vel,
                            float* dvort, float* visc, int reps, int* targets_per_lane,
                            int* workgroup_size, int* workgroups_per_cu, float* best_ms) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (!batch_live(rt, b)) return -1;
  if (reps < 1 || b->exec || (b->n_particles > 0 && (!pv || !a || !vel || !dvort)) ||
      ((b->mask & OUT_VISC) && b->n_particles > 0 && !visc)) {
    set_error(rt, "batch autotune: null buffer, reps < 1, or RK4 graph already captured");
    return -1;
  }
  hipStream_t s = b->dev->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!check_hip(rt, hipEventCreate(&e0), "event") || !check_hip(rt, hipEventCreate(&e1), "event")) {
    if (e0) (void)hipEventDestroy(e0);
    return -1;
  }
  BatchShape best{b->T, b->wg, b->wpc, -1.f};
  bool ok = true;
  for (int wg : {128, 256}) {
    for (int T : {1, 2, 4}) {
      if (!build_units(b, T, wg)) {
        ok = false;
        break;
      }
      for (int wpc : {2, 4, 8}) {
        b->grid = std::max(1, std::min(b->n_units, wpc * std::max(1, b->dev->n_cu)));
        float ms = -1.f;
        if (enqueue_eval(b, pv, a, vel, dvort, visc, s) == hipSuccess &&  // warm-up
            hipEventRecord(e0, s) == hipSuccess) {
          bool run = true;
          for (int r = 0; r < reps && run; ++r)
            run = enqueue_eval(b, pv, a, vel, dvort, visc, s) == hipSuccess;
          if (run && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
              hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
            ms /= (float)reps;
          else
            ms = -1.f;
        }
        if (ms > 0.f && (best.ms < 0.f || ms < best.ms)) best = BatchShape{T, wg, wpc, ms};
      }
    }
    if (!ok) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // leave the batch in the winning shape (or its previous one)
  if (!build_units(b, best.T, best.wg)) {
    set_error(rt, "batch autotune: allocation failed");
    return -1;
  }
  b->wpc = best.wpc;
  b->grid = std::max(1, std::min(b->n_units, b->wpc * std::max(1, b->dev->n_cu)));
  if (!ok || best.ms < 0.f) {
    set_error(rt, "batch autotune: no candidate ran");
    return -1;
  }
  batch_shape_record(batch_signature(b), best);
  if (targets_per_lane) *targets_per_lane = best.T;
  if (workgroup_size) *workgroup_size = best.wg;
  if (workgroups_per_cu) *workgroups_per_cu = best.wpc;
  if (best_ms) *best_ms = best.ms;
  return 0;
}

int cvtx_hip_batch_shape(const cvtx_hip_batch* b, int* targets_per_lane, int* workgroup_size,
                         int* workgroups_per_cu) {
  if (!b) return -1;
  if (targets_per_lane) *targets_per_lane = b->T;
  if (workgroup_size) *workgroup_size = b->wg;
  if (workgroups_per_cu) *workgroups_per_cu = b->wpc;
  return 0;
}

int cvtx_hip_batch_info(const cvtx_hip_batch* b, int* n_units, int* grid, int* targets_per_lane) {
  if (!b) return -1;
  if (n_units) *n_units = b->n_units;
  if (grid) *grid = b->grid;
  if (targets_per_lane) *targets_per_lane = b->T;
  return 0;
}

}  // extern "C"
