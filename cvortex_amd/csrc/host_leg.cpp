// host_leg.cpp -- CPU evaluation of the direct sums (see host_leg.hpp).
//
// Pair terms (d = target - source unless stated; the reference's formulas,
// src/P3D.cpp:51-144, 146-161, 298-322, src/F3D.cpp:34-85,
// src/P2D.cpp:49-69, 167-194):
//   P3D vel    -g(rho) (d x a_s) / r^3, summed, times 1/(4 pi)
//   P3D dvort  [c g/rho^3 - d (3 g/rho^3 - zeta)(d.c)/r^2] with c = a_t x a_s,
//              summed, times 1/(4 pi sigma^3) (signed sigma)
//   P3D visc   (a_s V_t - a_t V_s) eta(rho), summed, times 2 nu / sigma^2
//   P3D vort   a_s zeta(r/sigma), summed, divided by 4 pi sigma^3; the M2M
//              form only takes sources inside the 5 sigma box
//   P2D vel    (d_y, -d_x) a_s g_2D(rho) / r^2, summed, times 1/(2 pi)
//   P2D visc   (a_s A_t - a_t A_s) eta_2D(rho), summed, times 2 nu / sigma^2
//   F3D vel    Biot-Savart of a straight segment; F3D dvort its stretching
// with rho = r / |sigma|.  A pair at exactly equal coordinates contributes
// nothing; a filament pair whose terms are not finite (a point on the
// filament's line) contributes nothing.
#include "host_leg.hpp"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

namespace cvtx {
namespace host {

namespace {

constexpr double kPi = 3.14159265358979323846;
// sources per partial sum: every target's sum is the sum, in order, of the
// FP64 sums of consecutive chunks of this many sources
constexpr long kChunk = 2048;

struct V3 {
  double x, y, z;
};
inline V3 v3(const bsv_V3f& a) { return {a.x[0], a.x[1], a.x[2]}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline bool same(const bsv_V3f& a, const bsv_V3f& b) {
  return a.x[0] == b.x[0] && a.x[1] == b.x[1] && a.x[2] == b.x[2];
}
inline bool same(const bsv_V2f& a, const bsv_V2f& b) {
  return a.x[0] == b.x[0] && a.x[1] == b.x[1];
}
inline bool finite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

int env_threads(const char* name) {
  const char* s = std::getenv(name);
  if (!s || !*s) return 0;
  const long v = std::strtol(s, nullptr, 10);
  return v > 0 ? (int)(v < 64 ? v : 64) : 0;
}

// body(k) for k in [0, n), handed out one index at a time to nth threads
template <class F>
void run_parallel(long n, int nth, F body) {
  if (n <= 0) return;
  if (nth > n) nth = (int)n;
  if (nth <= 1) {
    for (long k = 0; k < n; ++k) body(k);
    return;
  }
  std::atomic<long> next{0};
  auto worker = [&]() {
    for (long k = next.fetch_add(1); k < n; k = next.fetch_add(1)) body(k);
  };
  std::vector<std::thread> pool;
  pool.reserve(nth - 1);
  for (int t = 1; t < nth; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
}

// For each target i in [0, nt): acc[K] = sum over sources j of pair(j, i, acc)
// (FP64, in chunks of kChunk sources), then store(i, acc).
template <int K, class Pair, class Store>
void direct_sum(long ns, long nt, Pair pair, Store store) {
  if (nt <= 0) return;
  const long nc = ns > 0 ? (ns + kChunk - 1) / kChunk : 0;
  auto chunk_sum = [&](long i, long c, double* part) {
    double a[K] = {};
    const long j1 = (c + 1) * kChunk < ns ? (c + 1) * kChunk : ns;
    for (long j = c * kChunk; j < j1; ++j) pair(j, i, a);
    for (int k = 0; k < K; ++k) part[k] = a[k];
  };
  const double work = (double)ns * (double)nt;
  const int nth = work < 32768.0 ? 1 : threads();
  if (nt >= 8L * nth || nc <= 1) {
    // by target: blocks of targets, each target's chunks in order
    const long blk = nt >= 64L * nth ? 16 : 1;
    run_parallel((nt + blk - 1) / blk, nth, [&](long b) {
      const long i1 = (b + 1) * blk < nt ? (b + 1) * blk : nt;
      for (long i = b * blk; i < i1; ++i) {
        double acc[K] = {}, part[K];
        for (long c = 0; c < nc; ++c) {
          chunk_sum(i, c, part);
          for (int k = 0; k < K; ++k) acc[k] += part[k];
        }
        store(i, acc);
      }
    });
    return;
  }
  // few targets, many sources: every (target, chunk) partial sum in
  // parallel, then each target's partials added in chunk order
  std::vector<double> parts((size_t)nt * nc * K);
  run_parallel(nt * nc, nth, [&](long q) { chunk_sum(q / nc, q % nc, &parts[(size_t)q * K]); });
  for (long i = 0; i < nt; ++i) {
    double acc[K] = {};
    for (long c = 0; c < nc; ++c)
      for (int k = 0; k < K; ++k) acc[k] += parts[((size_t)i * nc + c) * K + k];
    store(i, acc);
  }
}

bool need(const void* fn, const char* what, std::string* err) {
  if (fn) return true;
  if (err) *err = std::string("regularisation has no ") + what + " function";
  return false;
}

inline void put3(bsv_V3f* out, long i, const double* a, double s) {
  out[i].x[0] = (float)(a[0] * s);
  out[i].x[1] = (float)(a[1] * s);
  out[i].x[2] = (float)(a[2] * s);
}

// filament velocity term (without the 1/(4 pi)); false if not finite
inline bool fil_vel(const cvtx_F3D& f, V3 x, V3* v) {
  const V3 a = v3(f.start), b = v3(f.end);
  const V3 r1 = x - a, r2 = x - b, r0 = b - a;
  const V3 c = cross(r1, r2);
  const double c2 = dot(c, c);
  const double n1 = std::sqrt(dot(r1, r1)), n2 = std::sqrt(dot(r2, r2));
  const double k = f.strength * (dot(r0, r1) / n1 - dot(r0, r2) / n2) / c2;
  *v = c * k;
  return std::isfinite(k) && finite(*v);
}

}  // namespace

int threads() {
  int n = env_threads("CVTX_NUM_THREADS");
  if (!n) n = env_threads("OMP_NUM_THREADS");
  if (!n) {
    const unsigned hw = std::thread::hardware_concurrency();
    n = hw == 0 ? 1 : (int)(hw < 64 ? hw : 64);
  }
  return n;
}

// ------------------------------------ P3D ------------------------------------

int p3d_vel(const cvtx_P3D* const* src, long ns, const bsv_V3f* mes, long nm, bsv_V3f* out,
            const cvtx_VortFunc* vf, float sigma, std::string* err) {
  if (!need((const void*)vf->g_3D, "g_3D", err)) return -1;
  float (*g3)(float) = vf->g_3D;
  const double rinv = 1.0 / std::fabs((double)sigma);
  direct_sum<3>(
      ns, nm,
      [&](long j, long i, double* a) {
        const cvtx_P3D& s = *src[j];
        if (same(s.coord, mes[i])) return;
        const V3 d = v3(mes[i]) - v3(s.coord);
        const double r2 = dot(d, d), r = std::sqrt(r2);
        const double k = -(double)g3((float)(r * rinv)) / (r2 * r);
        const V3 t = cross(d, v3(s.vorticity)) * k;
        a[0] += t.x;
        a[1] += t.y;
        a[2] += t.z;
      },
      [&](long i, const double* a) { put3(out, i, a, 1.0 / (4.0 * kPi)); });
  return 0;
}

int p3d_dvort(const cvtx_P3D* const* src, long ns, const cvtx_P3D* const* ind, long ni,
              bsv_V3f* out, const cvtx_VortFunc* vf, float sigma, std::string* err) {
  if (!need((const void*)vf->combined_3D, "combined_3D", err)) return -1;
  void (*comb)(float, float*, float*) = vf->combined_3D;
  const double sg = sigma;
  direct_sum<3>(
      ns, ni,
      [&](long j, long i, double* a) {
        const cvtx_P3D& s = *src[j];
        const cvtx_P3D& t = *ind[i];
        if (same(s.coord, t.coord)) return;
        const V3 d = v3(t.coord) - v3(s.coord);
        const double r2 = dot(d, d);
        const float rho = (float)std::fabs(std::sqrt(r2) / sg);
        float g, z;
        comb(rho, &g, &z);
        const double q3 = (double)rho * rho * rho;
        const V3 c = cross(v3(t.vorticity), v3(s.vorticity));
        const double k1 = g / q3;
        const double k2 = -(3.0 * k1 - z) * dot(d, c) / r2;
        a[0] += c.x * k1 + d.x * k2;
        a[1] += c.y * k1 + d.y * k2;
        a[2] += c.z * k1 + d.z * k2;
      },
      [&](long i, const double* a) { put3(out, i, a, 1.0 / (4.0 * kPi * sg * sg * sg)); });
  return 0;
}

int p3d_visc(const cvtx_P3D* const* src, long ns, const cvtx_P3D* const* ind, long ni,
             bsv_V3f* out, const cvtx_VortFunc* vf, float sigma, float nu, std::string* err) {
  if (!need((const void*)vf->eta_3D, "eta_3D", err)) return -1;
  float (*eta)(float) = vf->eta_3D;
  const double sg = sigma;
  direct_sum<3>(
      ns, ni,
      [&](long j, long i, double* a) {
        const cvtx_P3D& s = *src[j];
        const cvtx_P3D& t = *ind[i];
        if (same(s.coord, t.coord)) return;
        const V3 d = v3(s.coord) - v3(t.coord);
        const double e = eta((float)std::fabs(std::sqrt(dot(d, d)) / sg));
        const V3 x = (v3(s.vorticity) * (double)t.volume - v3(t.vorticity) * (double)s.volume) * e;
        a[0] += x.x;
        a[1] += x.y;
        a[2] += x.z;
      },
      [&](long i, const double* a) { put3(out, i, a, 2.0 * nu / (sg * sg)); });
  return 0;
}

int p3d_vort(const cvtx_P3D* const* src, long ns, const bsv_V3f* mes, long nm, bsv_V3f* out,
             const cvtx_VortFunc* vf, float sigma, bool box_cutoff, std::string* err) {
  if (!need((const void*)vf->zeta_3D, "zeta_3D", err)) return -1;
  float (*zeta)(float) = vf->zeta_3D;
  const float cut = 5.f * sigma;
  const double sg = sigma;
  direct_sum<3>(
      ns, nm,
      [&](long j, long i, double* a) {
        const cvtx_P3D& s = *src[j];
        if (box_cutoff) {
          // the box test on the FP32 differences, as the reference makes it
          const float dx = s.coord.x[0] - mes[i].x[0], dy = s.coord.x[1] - mes[i].x[1],
                      dz = s.coord.x[2] - mes[i].x[2];
          if (!(std::fabs(dx) < cut && std::fabs(dy) < cut && std::fabs(dz) < cut)) return;
        }
        const V3 d = v3(s.coord) - v3(mes[i]);
        const double z = zeta((float)(std::sqrt(dot(d, d)) / sg));
        a[0] += s.vorticity.x[0] * z;
        a[1] += s.vorticity.x[1] * z;
        a[2] += s.vorticity.x[2] * z;
      },
      [&](long i, const double* a) { put3(out, i, a, 1.0 / (4.0 * kPi * sg * sg * sg)); });
  return 0;
}

int p3d_relax(cvtx_P3D* const* p, long n, float fdt, const cvtx_VortFunc* vf, float sigma,
              std::string* err) {
  if (n <= 0) return 0;
  std::vector<bsv_V3f> pts((size_t)n), om((size_t)n);
  for (long i = 0; i < n; ++i) pts[i] = p[i]->coord;
  if (p3d_vort(p, n, pts.data(), n, om.data(), vf, sigma, true, err) != 0) return -1;
  // a_new = (1 - fdt) a + fdt |a| / |w| w, or zero where w = 0
  // (reference src/P3D.cpp:669-707)
  for (long i = 0; i < n; ++i) {
    const V3 a = v3(p[i]->vorticity), w = v3(om[i]);
    const double nw = std::sqrt(dot(w, w));
    V3 r = {0.0, 0.0, 0.0};
    if (nw != 0.0) r = a * (1.0 - (double)fdt) + w * ((double)fdt * std::sqrt(dot(a, a)) / nw);
    p[i]->vorticity.x[0] = (float)r.x;
    p[i]->vorticity.x[1] = (float)r.y;
    p[i]->vorticity.x[2] = (float)r.z;
  }
  return 0;
}

// ------------------------------------ F3D ------------------------------------

void f3d_vel(const cvtx_F3D* const* f, long nf, const bsv_V3f* mes, long nm, bsv_V3f* out) {
  direct_sum<3>(
      nf, nm,
      [&](long j, long i, double* a) {
        V3 v;
        if (!fil_vel(*f[j], v3(mes[i]), &v)) return;
        a[0] += v.x;
        a[1] += v.y;
        a[2] += v.z;
      },
      [&](long i, const double* a) { put3(out, i, a, 1.0 / (4.0 * kPi)); });
}

void f3d_dvort(const cvtx_F3D* const* f, long nf, const cvtx_P3D* const* ind, long ni,
               bsv_V3f* out) {
  // per pair: B al + A x al with
  //   A = -r0 (r0.r1/|r1| - r0.r2/|r2|) / |r1 x r0|^2,
  //   B = 3 |r0 x r1| (1/|r1| - 1/|r2|) / |r0|           (times Gamma)
  direct_sum<3>(
      nf, ni,
      [&](long j, long i, double* acc) {
        const cvtx_F3D& fl = *f[j];
        const cvtx_P3D& t = *ind[i];
        const V3 x = v3(t.coord), al = v3(t.vorticity);
        const V3 r1 = x - v3(fl.start), r2 = x - v3(fl.end), r0 = v3(fl.end) - v3(fl.start);
        const V3 c = cross(r0, r1);
        const double c2 = dot(c, c);
        const double n1 = std::sqrt(dot(r1, r1)), n2 = std::sqrt(dot(r2, r2));
        const double G = fl.strength;
        const double kA = -G * (dot(r0, r1) / n1 - dot(r0, r2) / n2) / c2;
        const double B = 3.0 * G * std::sqrt(c2) * (1.0 / n1 - 1.0 / n2) / std::sqrt(dot(r0, r0));
        const V3 A = r0 * kA;
        const V3 v = al * B + cross(A, al);
        if (!std::isfinite(kA) || !std::isfinite(B) || !finite(v)) return;
        acc[0] += v.x;
        acc[1] += v.y;
        acc[2] += v.z;
      },
      [&](long i, const double* a) { put3(out, i, a, 1.0 / (4.0 * kPi)); });
}

void f3d_inf(const cvtx_F3D* const* f, long nf, const bsv_V3f* mes, const bsv_V3f* dir, long nm,
             float* out) {
  if (nf <= 0 || nm <= 0) return;
  const double s = 1.0 / (4.0 * kPi);
  const int nth = (double)nf * (double)nm < 32768.0 ? 1 : threads();
  run_parallel(nm, nth, [&](long i) {
    const V3 x = v3(mes[i]), u = v3(dir[i]);
    for (long j = 0; j < nf; ++j) {
      V3 v;
      out[(size_t)i * nf + j] = fil_vel(*f[j], x, &v) ? (float)(dot(v, u) * s) : 0.f;
    }
  });
}

// ------------------------------------ P2D ------------------------------------

int p2d_vel(const cvtx_P2D* const* src, long ns, const bsv_V2f* mes, long nm, bsv_V2f* out,
            const cvtx_VortFunc* vf, float sigma, std::string* err) {
  if (!need((const void*)vf->g_2D, "g_2D", err)) return -1;
  float (*g2)(float) = vf->g_2D;
  const double rinv = 1.0 / std::fabs((double)sigma);
  direct_sum<2>(
      ns, nm,
      [&](long j, long i, double* a) {
        const cvtx_P2D& s = *src[j];
        if (same(s.coord, mes[i])) return;
        const double dx = (double)mes[i].x[0] - s.coord.x[0];
        const double dy = (double)mes[i].x[1] - s.coord.x[1];
        const double r2 = dx * dx + dy * dy;
        const double k = s.vorticity * (double)g2((float)(std::sqrt(r2) * rinv)) / r2;
        a[0] += dy * k;
        a[1] -= dx * k;
      },
      [&](long i, const double* a) {
        out[i].x[0] = (float)(a[0] / (2.0 * kPi));
        out[i].x[1] = (float)(a[1] / (2.0 * kPi));
      });
  return 0;
}

int p2d_visc(const cvtx_P2D* const* src, long ns, const cvtx_P2D* const* ind, long ni,
             float* out, const cvtx_VortFunc* vf, float sigma, float nu, std::string* err) {
  if (!need((const void*)vf->eta_2D, "eta_2D", err)) return -1;
  float (*eta)(float) = vf->eta_2D;
  const double sg = sigma;
  direct_sum<1>(
      ns, ni,
      [&](long j, long i, double* a) {
        const cvtx_P2D& s = *src[j];
        const cvtx_P2D& t = *ind[i];
        if (same(s.coord, t.coord)) return;
        const double dx = (double)s.coord.x[0] - t.coord.x[0];
        const double dy = (double)s.coord.x[1] - t.coord.x[1];
        const double e = eta((float)std::fabs(std::sqrt(dx * dx + dy * dy) / sg));
        a[0] += ((double)s.vorticity * t.area - (double)t.vorticity * s.area) * e;
      },
      [&](long i, const double* a) { out[i] = (float)(a[0] * 2.0 * nu / (sg * sg)); });
  return 0;
}

}  // namespace host
}  // namespace cvtx
