// tune.cpp -- launch-shape cache and autotuner.
//
// The autotuner times candidate launch shapes (targets per lane T, number of
// source partitions) for a caller's device-resident problem with HIP events
// and records the fastest in a cache keyed by
//   (kind, regularisation, outputs, accumulation, log2 n_src, log2 n_tgt).
// choose_cfg() consults the cache before falling back to default_cfg().
// The cache can be saved to and loaded from a small text file, one entry
// per line: "kind reg mask acc bucket_src bucket_tgt T n_split ms wg" (kind 3:
// a multi-cloud batch, keyed by its pair count and cloud count; n_split is
// then the resident workgroups per CU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include <cvortex_hip.h>

#include "api_common.hpp"
#include "runtime.hpp"
#include "tune.hpp"

namespace cvtx {

namespace {

using Key = std::tuple<int, int, int, int, int, int>;
// n_split is the number of source partitions, or for KIND_BATCH the
// resident workgroups per CU; wg the workgroup size.
struct Entry {
  int T;
  int n_split;
  float ms;
  int wg = 256;
};

std::map<Key, Entry>& cache() {
  static std::map<Key, Entry> c;
  return c;
}

}  // namespace

LaunchCfg launch_shape(int T, int n_split, int n_src) {
  LaunchCfg c;
  c.T = T;
  c.wg = 256;
  const long tiles = ((long)n_src + 255) / 256;
  if (n_split <= 1 || tiles <= 1) return c;
  c.split_len = (int)((tiles + n_split - 1) / n_split * 256);
  c.n_split = (int)((n_src + c.split_len - 1) / c.split_len);
  if (c.n_split <= 1) {
    c.n_split = 1;
    c.split_len = 0;
  }
  return c;
}

int size_bucket(int n) {
  int b = 0;
  while (n > 1) {
    n >>= 1;
    ++b;
  }
  return b;
}

// Whether a launch of this kind orders its targets spatially (api.cpp), which
// changes the cost of its launch shapes: part of the cache key, folded into
// the kind field (kind + 16).
int ordered_kind(int kind, int reg, int n_src, int n_tgt) {
  const Runtime& rt = runtime();
  const bool ordered =
      rt.order_targets && ((kind == KIND_P3D && p3d_wants_order(reg, n_src, n_tgt)) ||
                           (kind == KIND_F3D && f3d_wants_order(n_src, n_tgt)));
  return ordered ? kind + 16 : kind;
}

LaunchCfg choose_cfg(int kind, int reg, int mask, int acc, int n_src, int n_tgt, int n_cu) {
  auto it = cache().find(Key{ordered_kind(kind, reg, n_src, n_tgt), reg, mask, acc,
                             size_bucket(n_src), size_bucket(n_tgt)});
  if (it != cache().end()) return launch_shape(it->second.T, it->second.n_split, n_src);
  const int cus = n_cu > 0 ? n_cu : 256;
  // Spatially ordered targets (Gaussian P3D, compensated sums).  One target
  // per lane: a 64-target wave of Morton-ordered targets is more compact
  // than a 128-target one, so fewer of its source rounds take the FP64
  // near-field branch.  Source partitions: the near-field work is
  // concentrated in the workgroups whose targets sit in dense regions, and
  // partitioning spreads each of them over several workgroups instead of
  // leaving them as the launch's tail.  Enough partitions for about 64
  // workgroups per CU, and at least one per 65536 sources, but no more than
  // about 256 workgroups per CU.  Measured on the ring (vel+dvort, ms; one
  // partition -> this rule, on different boxes):
  //   2^14: 0.29 -> 0.26 (64 partitions),  2^16: 3.73 -> 2.85 (64),
  //   2^18: 58.0 -> 42.4 (16; 32 and 64 the same),
  //   2^20: 710.6 -> 665.5 (16; 4 partitions: 673.6, 32: 665.3),
  //   2^22: 13742 -> 10564 (4; 1 partition on that box: 10703).
  // vel alone and vel+dvort+visc at 2^20 also gain 1.2-1.4 % from 4 to 16.
  // The plain FP32 kernel has no FP64 branch and keeps the general shape
  // (two targets per lane: 678 ms against 861 ms with one at 2^20).
  if (kind == KIND_P3D && acc == ACC_COMP && p3d_wants_order(reg, n_src, n_tgt)) {
    // The velocity alone has no FP64 near sums, so the near branch costs
    // little and four targets per lane share each source's LDS reads: four
    // targets per lane and partitions for about 256 workgroups per CU (ring
    // 2^20 vel: 475.7 ms with the rule below, 419.6 ms with this).
    const int T = mask == OUT_VEL ? 4 : 1;
    const long long blocks = ((long long)n_tgt + 256LL * T - 1) / (256LL * T);
    const long long few = (64LL * cus + blocks - 1) / blocks;
    const long long per = ((long long)n_src + 65535) / 65536;
    const long long many = (256LL * cus + blocks - 1) / blocks;
    long long k = T == 4 ? many : std::min(std::max(few, per), many);
    // partitions of one tile spend much of a workgroup on its prologue and
    // epilogue: two tiles each while that still leaves 32 workgroups per CU
    // (ring 32768: 128 partitions 0.85 ms, 64 partitions 0.77 ms)
    const long long tiles = ((long long)n_src + 255) / 256;
    if (k > tiles / 2 && blocks * (tiles / 2) >= 32LL * cus) k = tiles / 2;
    return launch_shape(T, (int)std::min<long long>(k, 1 << 20), n_src);
  }
  LaunchCfg c = default_cfg(n_src, n_tgt, n_cu);
  const bool many_blocks = (long long)(n_tgt + 255) / 256 >= 4LL * cus;
  // Filaments: one target per lane.  It fits in fewer VGPRs (more waves per
  // SIMD to pair instructions for dual issue) and, with ordered targets,
  // makes a block's box smaller, so more filament tiles take the far-tile
  // loop.  16384 filaments x 2^20 targets, 4 partitions (ms, two targets
  // per lane -> one): vel 13.7 -> 13.1, dvort 21.1 -> 19.2, fused 21.4 ->
  // 20.4.
  if (kind == KIND_F3D && c.T > 1 && many_blocks) c = partitioned_cfg(1, n_src, n_tgt, n_cu);
  // 2D Gaussian with the viscous term: two targets per lane (2^22 sheet,
  // 4 partitions: 3331 ms with four, 3226 with two); planetary and the
  // velocity alone keep four (2720 ms against 3042 with two).
  if (kind == KIND_P2D && reg == REG_GAUSSIAN && (mask & OUT_VISC) && c.T == 4)
    c = partitioned_cfg(2, n_src, n_tgt, n_cu);
  return c;
}

void record_cfg(int kind, int reg, int mask, int acc, int n_src, int n_tgt, int T, int n_split) {
  cache()[Key{ordered_kind(kind, reg, n_src, n_tgt), reg, mask, acc, size_bucket(n_src),
              size_bucket(n_tgt)}] = Entry{T, n_split, 0.f};
}

namespace {
int bucket64(long long n) {
  int b = 0;
  while (n > 1) {
    n >>= 1;
    ++b;
  }
  return b;
}
Key batch_key(const BatchSignature& g) {
  return Key{KIND_BATCH, g.reg, g.mask, g.acc, bucket64(g.pairs), size_bucket(g.n_clouds)};
}
}  // namespace

bool batch_shape_lookup(const BatchSignature& g, BatchShape* out) {
  auto it = cache().find(batch_key(g));
  if (it == cache().end()) return false;
  *out = BatchShape{it->second.T, it->second.wg, it->second.n_split, it->second.ms};
  return true;
}

void batch_shape_record(const BatchSignature& g, const BatchShape& s) {
  cache()[batch_key(g)] = Entry{s.T, s.wpc, s.ms, s.wg};
}

}  // namespace cvtx

using namespace cvtx;

namespace {

// Time `reps` launches of one shape; returns ms per launch or -1 on error.
template <class Launch>
float time_shape(hipStream_t s, int reps, Launch launch) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return -1.f;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return -1.f;
  }
  float ms = -1.f;
  if (launch() == hipSuccess && hipEventRecord(a, s) == hipSuccess) {
    bool ok = true;
    for (int r = 0; r < reps && ok; ++r) ok = launch() == hipSuccess;
    if (ok && hipEventRecord(b, s) == hipSuccess && hipEventSynchronize(b) == hipSuccess) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, a, b) == hipSuccess) ms = t / (float)reps;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

std::vector<LaunchCfg> candidates(int n_src, int n_tgt, int n_cu) {
  std::vector<LaunchCfg> out;
  const long tiles = ((long)n_src + 255) / 256;
  for (int T : {1, 2, 4}) {
    const long blocks = ((long)n_tgt + 256L * T - 1) / (256L * T);
    if (T > 1 && blocks < 8) continue;
    for (int k = 1; k <= 256; k *= 2) {
      if (k > 1 && k > tiles) break;
      if (blocks * k > 256L * (n_cu > 0 ? n_cu : 256)) break;
      out.push_back(launch_shape(T, k, n_src));
    }
  }
  return out;
}

}  // namespace

extern "C" {

// Autotune the launch shape of a device-resident P3D problem; the winner is
// recorded in the shape cache for the problem's size bucket.  Returns 0 and
// the winning shape, or -1.
int cvtx_hip_autotune_p3d(const cvtx_hip_p3d_problem* p, int reps, int* best_T, int* best_split,
                          float* best_ms) {
  Runtime& rt = runtime();
  std::lock_guard<std::recursive_mutex> lk(rt.mu);
  if (!p || reps < 1) return -1;
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);
  P3DJob j;
  j.src_pv = (const float4*)p->src_pv;
  j.src_a = (const float4*)p->src_a;
  j.n_src = p->n_src;
  j.tgt_pv = (const float4*)p->tgt_pv;
  j.tgt_a = (const float4*)p->tgt_a;
  j.n_tgt = p->n_tgt;
  j.out_vel = p->vel;
  j.out_dvort = p->dvort;
  j.out_visc = p->visc;
  j.reg = p->regularisation;
  j.mask = (p->vel ? OUT_VEL : 0) | (p->dvort ? OUT_DVORT : 0) | (p->visc ? OUT_VISC : 0);
  j.acc = p->accumulate;
  j.sigma = p->sigma;
  j.nu = p->nu;
  j.gtab = d->gtab;
  if (!p3d_supported(j.reg, j.mask) || j.n_tgt <= 0 || j.n_src < 0) return -1;
  // time the shapes on the target order the real calls will use
  if (rt.order_targets && p3d_wants_order(j.reg, j.n_src, j.n_tgt)) {
    j.tgt_perm = order_points(rt, *d, j.tgt_pv, j.n_tgt, SLOT_PERM, d->stream);
    if (!j.tgt_perm) return -1;
  }
  float best = -1.f;
  LaunchCfg win;
  for (const LaunchCfg& c : candidates(j.n_src, j.n_tgt, d->n_cu)) {
    P3DJob jj = j;
    const size_t wb = p3d_workspace_bytes(jj, c);
    if (wb) {
      jj.partial = (double*)device_buffer(rt, *d, SLOT_WORK, wb);
      if (!jj.partial) return -1;
    }
    const float ms = time_shape(d->stream, reps, [&]() { return launch_p3d(jj, c, d->stream); });
    if (ms > 0.f && (best < 0.f || ms < best)) {
      best = ms;
      win = c;
    }
  }
  if (best < 0.f) return -1;
  cache()[Key{ordered_kind(KIND_P3D, j.reg, j.n_src, j.n_tgt), j.reg, j.mask, j.acc,
              size_bucket(j.n_src), size_bucket(j.n_tgt)}] = Entry{win.T, win.n_split, best};
  if (best_T) *best_T = win.T;
  if (best_split) *best_split = win.n_split;
  if (best_ms) *best_ms = best;
  return 0;
}

int cvtx_hip_tune_save(const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  for (auto& kv : cache()) {
    const Key& k = kv.first;
    std::fprintf(f, "%d %d %d %d %d %d %d %d %.6f %d\n", std::get<0>(k), std::get<1>(k),
                 std::get<2>(k), std::get<3>(k), std::get<4>(k), std::get<5>(k), kv.second.T,
                 kv.second.n_split, kv.second.ms, kv.second.wg);
  }
  std::fclose(f);
  return 0;
}

int cvtx_hip_tune_load(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) return -1;
  int kind, reg, mask, acc, bs, bt, T, ns, wg;
  float ms;
  int n = 0;
  char line[256];
  // one entry per line; the workgroup size column may be absent (256)
  while (std::fgets(line, sizeof(line), f)) {
    wg = 256;
    const int got = std::sscanf(line, "%d %d %d %d %d %d %d %d %f %d", &kind, &reg, &mask, &acc,
                                &bs, &bt, &T, &ns, &ms, &wg);
    if (got >= 9 && (T == 1 || T == 2 || T == 4) && ns >= 1 && (wg == 128 || wg == 256)) {
      cache()[Key{kind, reg, mask, acc, bs, bt}] = Entry{T, ns, ms, wg};
      ++n;
    }
  }
  std::fclose(f);
  return n;
}

void cvtx_hip_tune_clear(void) { cache().clear(); }

int cvtx_hip_gaussian_table(float* out, int capacity) {
  if (out == nullptr || capacity < 8 * GTAB_N) return -1;
  build_gaussian_table(out);
  return 8 * GTAB_N;
}

}  // extern "C"
