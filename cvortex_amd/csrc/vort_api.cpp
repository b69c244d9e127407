// vort_api.cpp -- C ABI of the vorticity field (cvtx_P3D_M2M_vort) and the
// Pedrizzetti relaxation (cvtx_P3D_pedrizzetti_relaxation): host-array entry
// points with the reference's signatures, their int-returning seam, and the
// device-resident forms.  Kernels: vort.hip.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include <cvortex/libcvtx.h>
#include <cvortex_hip.h>

#include "api_common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace cvtx;

namespace {

using Lock = std::lock_guard<std::recursive_mutex>;

// Vorticity of sources (spv, sa) at targets tgt, into out[3 nt], on stream s.
// Large problems are Morton-sorted so that the kernel skips source tiles
// outside the cut-off box; `same` marks sources == targets (one sort).
bool vort_core(Runtime& rt, Device& d, const float4* spv, const float4* sa, int ns,
               const float4* tgt, int nt, float* out, int reg, int acc, float sigma, bool same,
               hipStream_t s, const LaunchCfg* forced = nullptr, bool cutoff = true) {
  VortJob j;
  j.src_pv = spv;
  j.src_a = sa;
  j.n_src = ns;
  j.tgt = tgt;
  j.n_tgt = nt;
  j.out = out;
  j.reg = reg;
  j.acc = acc;
  j.sigma = sigma;
  j.cutoff = cutoff;
  const bool sorted = cutoff && rt.order_targets && vort_wants_sort(ns, nt);
  if (sorted) {
    j.src_perm = order_points(rt, d, spv, ns, SLOT_PERM2, s);
    if (!j.src_perm) return false;
    j.tgt_perm = same ? j.src_perm : order_points(rt, d, tgt, nt, SLOT_PERM, s);
    if (!j.tgt_perm) return false;
  }
  const LaunchCfg c = forced ? *forced : vort_default_cfg(ns, nt, d.n_cu, sorted);
  const size_t wb = vort_workspace_bytes(j, c, sorted);
  void* ws = wb ? device_buffer(rt, d, SLOT_WORK, wb) : nullptr;
  if (wb && !ws) return false;
  return check_hip(rt, launch_vort(j, c, ws, wb, s), "vorticity kernel launch");
}

int regularisation_of(Runtime& rt, const cvtx_VortFunc* vf) {
  if (vf == nullptr) {
    set_error(rt, "null regularisation");
    return -1;
  }
  const int reg = reg_from_name(vf->cl_kernel_name_ext);
  if (reg < 0)
    set_error(rt, "regularisation name is not one of singular, winckelmans, planetary, gaussian");
  return reg;
}

}  // namespace

int cvtx::vort_host(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nt, float* out,
                    const cvtx_VortFunc* vf, float sigma, bool cutoff) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  const int reg = regularisation_of(rt, vf);
  if (reg < 0) return -1;
  if (bad_counts(rt, ns, nt, src, mes)) return -1;
  if (nt == 0) return 0;
  if (out == nullptr) {
    set_error(rt, "null result array");
    return -1;
  }
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);
  const size_t sb = sizeof(float4) * (size_t)(ns > 0 ? ns : 1);
  const size_t tb = sizeof(float4) * (size_t)nt;
  const size_t ob = sizeof(float) * 3 * (size_t)nt;
  float4* h_spv = (float4*)host_buffer(rt, *d, SLOT_SRC0, sb);
  float4* h_sa = (float4*)host_buffer(rt, *d, SLOT_SRC1, sb);
  float4* h_t = (float4*)host_buffer(rt, *d, SLOT_TGT0, tb);
  float* h_out = (float*)host_buffer(rt, *d, SLOT_OUT0, ob);
  float4* d_spv = (float4*)device_buffer(rt, *d, SLOT_SRC0, sb);
  float4* d_sa = (float4*)device_buffer(rt, *d, SLOT_SRC1, sb);
  float4* d_t = (float4*)device_buffer(rt, *d, SLOT_TGT0, tb);
  float* d_out = (float*)device_buffer(rt, *d, SLOT_OUT0, ob);
  if (!h_spv || !h_sa || !h_t || !h_out || !d_spv || !d_sa || !d_t || !d_out) return -1;
  gather_p3d(src, ns, h_spv, h_sa);
  gather_pts3(mes, nt, h_t);
  if (!h2d(rt, *d, d_spv, h_spv, sizeof(float4) * (size_t)ns)) return -1;
  if (!h2d(rt, *d, d_sa, h_sa, sizeof(float4) * (size_t)ns)) return -1;
  if (!h2d(rt, *d, d_t, h_t, tb)) return -1;
  if (!vort_core(rt, *d, d_spv, d_sa, ns, d_t, nt, d_out, reg, rt.acc_mode, sigma, false,
                 d->stream, nullptr, cutoff))
    return -1;
  if (!d2h(rt, *d, h_out, d_out, ob)) return -1;
  if (!check_hip(rt, hipStreamSynchronize(d->stream), "vorticity kernels")) return -1;
  std::memcpy(out, h_out, ob);
  return 0;
}

namespace {

int relax_host(cvtx_P3D** p, int n, float fdt, const cvtx_VortFunc* vf, float sigma) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  const int reg = regularisation_of(rt, vf);
  if (reg < 0) return -1;
  if (bad_counts(rt, n, 0, p, nullptr)) return -1;
  if (n == 0) return 0;
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, d->stream);
  const size_t pb = sizeof(float4) * (size_t)n;
  const size_t ob = sizeof(float) * 3 * (size_t)n;
  float4* h_pv = (float4*)host_buffer(rt, *d, SLOT_SRC0, pb);
  float4* h_a = (float4*)host_buffer(rt, *d, SLOT_SRC1, pb);
  float4* d_pv = (float4*)device_buffer(rt, *d, SLOT_SRC0, pb);
  float4* d_a = (float4*)device_buffer(rt, *d, SLOT_SRC1, pb);
  float* d_om = (float*)device_buffer(rt, *d, SLOT_OUT0, ob);
  if (!h_pv || !h_a || !d_pv || !d_a || !d_om) return -1;
  gather_p3d(p, n, h_pv, h_a);
  if (!h2d(rt, *d, d_pv, h_pv, pb) || !h2d(rt, *d, d_a, h_a, pb)) return -1;
  if (!vort_core(rt, *d, d_pv, d_a, n, d_pv, n, d_om, reg, rt.acc_mode, sigma, true, d->stream))
    return -1;
  if (!check_hip(rt, launch_relax(d_a, d_om, n, fdt, d->stream), "relaxation kernel")) return -1;
  if (!d2h(rt, *d, h_a, d_a, pb)) return -1;
  if (!check_hip(rt, hipStreamSynchronize(d->stream), "relaxation")) return -1;
  parallel_for(n, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      p[i]->vorticity.x[0] = h_a[i].x;
      p[i]->vorticity.x[1] = h_a[i].y;
      p[i]->vorticity.x[2] = h_a[i].z;
    }
  });
  return 0;
}

}  // namespace

extern "C" {

// ================================ seam ================================

int cvtx_hip_P3D_M2M_vort(const cvtx_P3D** src, int ns, const bsv_V3f* mes, int nm, bsv_V3f* out,
                          const cvtx_VortFunc* k, float sigma) {
  return vort_host(src, ns, mes, nm, (float*)out, k, sigma, true);
}

int cvtx_hip_P3D_pedrizzetti_relaxation(cvtx_P3D** particles, int n, float fdt,
                                        const cvtx_VortFunc* k, float sigma) {
  return relax_host(particles, n, fdt, k, sigma);
}

// ========================= device-resident API =========================

static int vort_dev(const cvtx_hip_vort_problem* p, const LaunchCfg* forced, void* stream) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (!p || p->n_src < 0 || p->n_tgt < 0 || (p->n_src > 0 && (!p->src_pv || !p->src_a)) ||
      (p->n_tgt > 0 && (!p->tgt || !p->vort)) || p->regularisation < REG_SINGULAR ||
      p->regularisation > REG_GAUSSIAN ||
      !valid_acc(p->accumulate)) {
    set_error(rt, "invalid vorticity problem");
    return -1;
  }
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, pick_stream(d, stream));
  const bool same = p->src_pv == p->tgt && p->n_src == p->n_tgt;
  return vort_core(rt, *d, (const float4*)p->src_pv, (const float4*)p->src_a, p->n_src,
                   (const float4*)p->tgt, p->n_tgt, p->vort, p->regularisation, p->accumulate,
                   p->sigma, same, pick_stream(d, stream), forced)
             ? 0
             : -1;
}

int cvtx_hip_vort_eval(const cvtx_hip_vort_problem* p, void* stream) {
  return vort_dev(p, nullptr, stream);
}

int cvtx_hip_vort_eval_shape(const cvtx_hip_vort_problem* p, int T, int n_split, void* stream) {
  LaunchCfg c;
  if (!p || !make_launch_shape(T, n_split, p->n_src, &c)) {
    set_error(runtime(), "invalid launch shape");
    return -1;
  }
  return vort_dev(p, &c, stream);
}

int cvtx_hip_relax_eval(const float* pv, float* a, int n, float fdt, int regularisation,
                        float sigma, int accumulate, void* stream) {
  Runtime& rt = runtime();
  Lock lk(rt.mu);
  if (n < 0 || (n > 0 && (!pv || !a)) || regularisation < REG_SINGULAR ||
      regularisation > REG_GAUSSIAN || !valid_acc(accumulate)) {
    set_error(rt, "invalid relaxation problem");
    return -1;
  }
  if (n == 0) return 0;
  Device* d = active_device(rt);
  if (!d) return -1;
  StreamUse use(*d, pick_stream(d, stream));
  float* om = (float*)device_buffer(rt, *d, SLOT_OUT2, sizeof(float) * 3 * (size_t)n);
  if (!om) return -1;
  const hipStream_t s = pick_stream(d, stream);
  if (!vort_core(rt, *d, (const float4*)pv, (const float4*)a, n, (const float4*)pv, n, om,
                 regularisation, accumulate, sigma, true, s))
    return -1;
  return check_hip(rt, launch_relax((float4*)a, om, n, fdt, s), "relaxation kernel") ? 0 : -1;
}

}  // extern "C"
