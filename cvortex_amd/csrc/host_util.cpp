// host_util.cpp -- host-side helpers: the Gaussian near-field table and the
// default launch shape.
#include <algorithm>
#include <cmath>

#include "kernels.hpp"

namespace cvtx {

namespace {

constexpr double kSqrt2OverPi = 0.79788456080286535588;

// sqrt(2/pi) * sum_k (-rho^2/2)^k / (k! (2k + shift))
double reduced_series(double rho, int shift) {
  const double x = -0.5 * rho * rho;
  double term = 1.0, sum = 0.0;
  for (int k = 0; k < 60; ++k) {
    const double add = term / (2.0 * k + shift);
    sum += add;
    if (std::fabs(add) < 1e-19 * std::fabs(sum)) break;
    term *= x / (k + 1.0);
  }
  return kSqrt2OverPi * sum;
}

// Exact reduced Gaussian kernels in FP64:
//   G = g/rho^3, g = erf(rho/sqrt2) - rho sqrt(2/pi) exp(-rho^2/2)
//   H = (3 G - zeta)/rho^2, zeta = sqrt(2/pi) exp(-rho^2/2)
// (G = sqrt(2/pi) int_0^1 s^2 exp(-rho^2 s^2/2) ds, H the same with s^4.)
void gaussian_GH(double rho, double* G, double* H) {
  if (rho < 1.0) {
    *G = reduced_series(rho, 3);
    *H = reduced_series(rho, 5);
    return;
  }
  const double z = kSqrt2OverPi * std::exp(-0.5 * rho * rho);
  const double g = std::erf(rho / std::sqrt(2.0)) - rho * z;
  *G = g / (rho * rho * rho);
  *H = (3.0 * *G - z) / (rho * rho);
}

// Solve the 4x4 Vandermonde system for the cubic through (u_j, f_j).
void cubic_through(const double* u, const double* f, double* c) {
  double m[4][5];
  for (int r = 0; r < 4; ++r) {
    m[r][0] = 1.0;
    m[r][1] = u[r];
    m[r][2] = u[r] * u[r];
    m[r][3] = u[r] * u[r] * u[r];
    m[r][4] = f[r];
  }
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r = col + 1; r < 4; ++r)
      if (std::fabs(m[r][col]) > std::fabs(m[piv][col])) piv = r;
    for (int k = 0; k < 5; ++k) std::swap(m[col][k], m[piv][k]);
    for (int r = 0; r < 4; ++r) {
      if (r == col) continue;
      const double fct = m[r][col] / m[col][col];
      for (int k = col; k < 5; ++k) m[r][k] -= fct * m[col][k];
    }
  }
  for (int r = 0; r < 4; ++r) c[r] = m[r][4] / m[r][r];
}

}  // namespace

void build_gaussian_table(float* out) {
  // Chebyshev nodes on [0,1] make each cubic a near-minimax fit of its piece.
  double u[4];
  for (int j = 0; j < 4; ++j) u[j] = 0.5 - 0.5 * std::cos((2.0 * j + 1.0) * M_PI / 8.0);
  for (int i = 0; i < GTAB_N; ++i) {
    double fg[4], fh[4], cg[4], ch[4];
    for (int j = 0; j < 4; ++j) {
      const double rho = (i + u[j]) / (double)GTAB_PER_UNIT;
      gaussian_GH(rho, &fg[j], &fh[j]);
      fh[j] /= 3.0;  // the kernels use H/3 (the factor 3 is in the epilogue scale)
    }
    cubic_through(u, fg, cg);
    cubic_through(u, fh, ch);
    for (int c = 0; c < 4; ++c) {
      out[8 * i + c] = (float)cg[c];
      out[8 * i + 4 + c] = (float)ch[c];
    }
  }
}

LaunchCfg default_cfg(int n_src, int n_tgt, int n_cu) {
  const long long cus = n_cu > 0 ? n_cu : 256;
  const long long want = 4 * cus;  // workgroups to fill the chip
  auto blocks = [&](int T) { return ((long long)n_tgt + 256LL * T - 1) / (256LL * T); };
  int T = 1;
  if (blocks(4) >= 2 * want) T = 4;
  else if (blocks(2) >= want / 2) T = 2;
  return partitioned_cfg(T, n_src, n_tgt, n_cu);
}

LaunchCfg partitioned_cfg(int T, int n_src, int n_tgt, int n_cu) {
  LaunchCfg c;
  c.wg = 256;
  c.T = T;
  const long long cus = n_cu > 0 ? n_cu : 256;
  const long long nb = ((long long)n_tgt + 256LL * T - 1) / (256LL * T);
  const long long tiles = ((long long)n_src + 255) / 256;
  if (nb <= 0 || tiles <= 0) return c;  // nothing to split (no targets or sources)
  // Source partitions, for two reasons:
  //   * to fill the chip (4 workgroups per CU) when there are few target
  //     blocks, down to one tile per partition;
  //   * for an even finish: with uniform work per pair, a launch of ~16
  //     workgroups per CU still ends unevenly across the CUs, and ~64 per CU
  //     does better.  Measured: 16384 filaments x 2^20 targets, vel
  //     16.54 -> 16.12 ms (T=1, 4 partitions), vel+dvort 24.43 -> 23.61 ms
  //     (T=2, 4); 2^22 P2D sheet 3577 -> 3501 ms (T=4, 4).  These
  //     partitions keep at least 16 tiles each.
  const long long fill = nb >= 4 * cus ? 1 : (4 * cus + nb - 1) / nb;
  const long long even = nb >= 64 * cus ? 1 : (64 * cus + nb - 1) / nb;
  long long k = std::max(fill, std::min(even, std::max(1LL, tiles / 16)));
  k = std::min(k, std::max(1LL, tiles));
  if (k <= 1) {
    c.n_split = 1;
    c.split_len = 0;
    return c;
  }
  const long long tiles_per = (tiles + k - 1) / k;
  c.split_len = (int)(tiles_per * 256);
  c.n_split = (int)(((long long)n_src + c.split_len - 1) / c.split_len);
  if (c.n_split <= 1) {
    c.n_split = 1;
    c.split_len = 0;
  }
  return c;
}

}  // namespace cvtx
