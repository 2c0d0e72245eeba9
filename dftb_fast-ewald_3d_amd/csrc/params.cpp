// params.cpp -- host-side parameter chain of the MI355X engine.
//
// The real-space cutoff, splitting parameter and Fourier cutoff follow the reference's fast
// path chain (src/dftbp/extlibs/fftewald/fft/ewald_fft_3d_parameters.f90:52-142); the SPME
// mesh and B-spline order replace its NFFT mode count and window width.  The short-range
// gamma closed form and its cutoff search follow src/dftbp/dftb/shortgammafuncs.F90.
#include <cmath>
#include <cstring>

#include "../../include/fftewald_amd.h"
#include "engine.hpp"

namespace ewamd {

static void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// fe/common/ewald_geometry.f90:65-93 (volume from |a1 . (a2 x a3)|, dual basis rows)
int cell_metrics(const double* lat9, Cell* c) {
  std::memset(c, 0, sizeof(*c));
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) c->lat[i][d] = lat9[3 * i + d];
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(lat9[i])) return EW_ERR_NOT_FINITE;
  double x[3];
  cross3(c->lat[1], c->lat[2], x);
  c->volume = std::fabs(c->lat[0][0] * x[0] + c->lat[0][1] * x[1] + c->lat[0][2] * x[2]);
  if (!(c->volume > 1.0e-18)) return EW_ERR_CELL;
  for (int d = 0; d < 3; ++d) c->rec[0][d] = x[d] / c->volume;
  cross3(c->lat[2], c->lat[0], x);
  for (int d = 0; d < 3; ++d) c->rec[1][d] = x[d] / c->volume;
  cross3(c->lat[0], c->lat[1], x);
  for (int d = 0; d < 3; ++d) c->rec[2][d] = x[d] / c->volume;
  for (int i = 0; i < 3; ++i) {
    c->latlen[i] = std::sqrt(c->lat[i][0] * c->lat[i][0] + c->lat[i][1] * c->lat[i][1] +
                             c->lat[i][2] * c->lat[i][2]);
    c->reclen[i] = std::sqrt(c->rec[i][0] * c->rec[i][0] + c->rec[i][1] * c->rec[i][1] +
                             c->rec[i][2] * c->rec[i][2]);
  }
  return EW_OK;
}

// fe/common/ewald_lambert_w.f90:27-56 -- principal branch by Halley's iteration.
static double lambert_w0(double x) {
  if (x <= 0.0) return 0.0;
  double w = std::log1p(x);
  for (int it = 0; it < 60; ++it) {
    const double ew = std::exp(w);
    const double f = w * ew - x;
    const double wp = w + 1.0;
    const double step = f / (ew * wp - 0.5 * (w + 2.0) * f / wp);
    w -= step;
    if (std::fabs(step) <= 1.0e-15 * (1.0 + std::fabs(w))) break;
  }
  return w;
}

// fe/common/ewald_truncation.f90:53-78 -- alpha keeping the rms real-space tail at `budget`.
static double alpha_from_real_budget(double r_cut, double q2, int n, double vol, double budget) {
  const double x = (2.0 / budget) * q2 * std::sqrt(r_cut / (double(n) * vol));
  double w = x > 0.0 ? lambert_w0(x) : -std::log(budget);
  if (w <= 0.0) w = -std::log(budget);
  return std::sqrt(w) / r_cut;
}

// fe/common/ewald_truncation.f90:84-104 -- Fourier cutoff for the rms reciprocal tail.
static double kcut_from_fourier_budget(double q2, int n, double edge, double alpha,
                                       double budget) {
  const double x = (4.0 / (3.0 * edge * edge)) *
                   std::pow(2.0 / (double(n) * alpha * kPi), 2.0 / 3.0) *
                   std::pow(2.0 * q2 / budget, 4.0 / 3.0);
  double w = lambert_w0(x);
  if (w <= 0.0) w = -std::log(budget);
  return std::sqrt(3.0) * alpha * std::sqrt(w);
}

// src/dftbp/dftb/coulomb.F90:2097-2179 (getOptimalAlphaEwald) with diffRecReal 2686-2748:
// bracket and bisect the balance of the 4th/5th reciprocal shell against the 2nd/3rd real
// shell; nSearchIter = 30 (src/dftbp/common/accuracy.F90:114).
static int dftb_optimal_alpha(const Cell& c, double tol, double* alpha_out) {
  double minG = 1e308, minR = 1e308;
  for (int i = 0; i < 3; ++i) {
    minG = std::fmin(minG, 2.0 * kPi * c.reclen[i]);
    minR = std::fmin(minR, c.latlen[i]);
  }
  auto diff = [&](double a) {
    auto gt = [&](double g) { return 4.0 * kPi * (std::exp(-0.25 * g * g / (a * a)) / (c.volume * g * g)); };
    auto rt = [&](double r) { return std::erfc(a * r) / r; };
    return (gt(4.0 * minG) - gt(5.0 * minG)) - (rt(2.0 * minR) - rt(3.0 * minR));
  };
  const double a0 = 1.0e-8, big = 1.79769313486231570e308;
  double a = a0, d = diff(a);
  while (d < -tol && a <= big) { a *= 2.0; d = diff(a); }
  if (a > big || a == a0) return EW_ERR_ARGUMENT;
  double lo = 0.5 * a;
  while (d < tol && a <= big) { a *= 2.0; d = diff(a); }
  if (a > big) return EW_ERR_ARGUMENT;
  double hi = a;
  a = 0.5 * (lo + hi);
  d = diff(a);
  int it = 0;
  while (std::fabs(d) > tol && it <= 30) {
    if (d < 0.0) lo = a; else hi = a;
    a = 0.5 * (lo + hi);
    d = diff(a);
    ++it;
  }
  if (it > 30) return EW_ERR_ARGUMENT;
  *alpha_out = a;
  return EW_OK;
}

int smooth235(int n) {
  int m = n < 1 ? 1 : n;
  for (;; ++m) {
    int r = m;
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    while (r % 5 == 0) r /= 5;
    if (r == 1) return m;
  }
}

int choose_parameters(const Cell& c, int n, double q2sum, const double* opt_tol,
                      double opt_alpha, double opt_r_cut, double opt_k_cut, const int* opt_grid,
                      int opt_order,
                      int alpha_rule, Params* p, std::string* err) {
  double tol = (opt_tol && *opt_tol > 0.0) ? *opt_tol : 1.0e-10;
  if (!(tol > 0.0 && tol < 1.0)) {
    if (err) *err = "tol must lie in (0,1)";
    return EW_ERR_TOLERANCE;
  }
  if (n <= 0) {
    if (err) *err = "n must be positive";
    return EW_ERR_ARGUMENT;
  }
  if (!std::isfinite(opt_k_cut) || !std::isfinite(opt_alpha) || !std::isfinite(opt_r_cut)) {
    if (err) *err = "alpha, r_cut and k_cut overrides must be finite";
    return EW_ERR_ARGUMENT;
  }
  if (!(q2sum > 0.0)) q2sum = double(n);  // unit charges when none are known
  p->tol = tol;
  const double spacing = std::cbrt(c.volume / double(n));
  // 1. real-space cutoff (fe/fft/ewald_fft_3d_parameters.f90:86-111)
  double alpha_dftb = 0.0;
  if (alpha_rule == 1 && !(opt_alpha > 0.0)) {
    int st = dftb_optimal_alpha(c, tol, &alpha_dftb);
    if (st) {
      if (err) *err = "optimal-alpha search failed (getOptimalAlphaEwald)";
      return st;
    }
  }
  const double alpha_pinned = opt_alpha > 0.0 ? opt_alpha : alpha_dftb;
  if (opt_r_cut > 0.0) {
    p->r_cut = opt_r_cut;
  } else if (alpha_pinned > 0.0) {
    p->r_cut = std::sqrt(-std::log(tol)) / alpha_pinned;
  } else {
    p->r_cut = 3.0 * spacing;
    const double mx = std::fmax(c.reclen[0], std::fmax(c.reclen[1], c.reclen[2]));
    const double limit = 0.999 / (3.0 * mx);
    if (limit >= 1.5 * spacing) p->r_cut = std::fmin(p->r_cut, limit);
  }
  // 2. splitting parameter (fe/fft/ewald_fft_3d_parameters.f90:114-121)
  p->alpha = alpha_pinned > 0.0
                 ? alpha_pinned
                 : alpha_from_real_budget(p->r_cut, q2sum, n, c.volume, 0.2 * tol);
  if (!(p->alpha > 0.0) || !(p->r_cut > 0.0)) {
    if (err) *err = "alpha and r_cut must be positive";
    return EW_ERR_ARGUMENT;
  }
  // 3. Fourier cutoff (fe/fft/ewald_fft_3d_parameters.f90:127-130), or the caller's k_cut_in
  p->k_cut = opt_k_cut > 0.0 ? opt_k_cut
                             : kcut_from_fourier_budget(q2sum, n, std::cbrt(c.volume), p->alpha, tol);
  // 4. B-spline order and mesh.  Calibrated against the direct sum: the reciprocal error of
  //    order 10 with a mesh of 2*sigma*m_max points per axis (m_max = k_cut |a_i| / 2 pi)
  //    stays below tol for sigma = 1.25 / 1.5 / 2 at tol >= 1e-7 / >= 1e-9 / below.
  p->order = opt_order > 0 ? opt_order : (tol >= 1.0e-7 ? 8 : 10);
  if (p->order < 2 || p->order > kMaxOrder) {
    if (err) *err = "B-spline order must lie in 2..16";
    return EW_ERR_ARGUMENT;
  }
  const double sigma = tol >= 1.0e-7 ? 1.25 : (tol >= 1.0e-9 ? 1.5 : 2.0);
  for (int a = 0; a < 3; ++a) {
    if (opt_grid && opt_grid[0] > 0) {
      p->K[a] = opt_grid[a];
    } else {
      const double mmax = p->k_cut * c.latlen[a] / (2.0 * kPi);
      int want = int(std::ceil(2.0 * sigma * mmax));
      if (want < 2 * p->order) want = 2 * p->order;
      p->K[a] = smooth235(want);
    }
    if (p->K[a] < p->order || p->K[a] > 4096 || smooth235(p->K[a]) != p->K[a]) {
      if (err) *err = "grid sizes must be 2^a 3^b 5^c, >= the B-spline order and <= 4096";
      return EW_ERR_ARGUMENT;
    }
  }
  p->r_list = p->r_cut;
  return EW_OK;
}

// src/dftbp/dftb/shortgammafuncs.F90:267-296 (gammaSubExprn_)
static double gamma_sub(double r, double t1, double t2) {
  const double d = t1 * t1 - t2 * t2;
  const double t2_4 = t2 * t2 * t2 * t2;
  return std::exp(-t1 * r) * ((0.5 * t2_4 * t1 / (d * d)) -
                              (t2_4 * t2 * t2 - 3.0 * t2_4 * t1 * t1) / (r * d * d * d));
}

// src/dftbp/dftb/shortgammafuncs.F90:97-148, constants from src/dftbp/common/accuracy.F90
double exp_gamma(double r, double ua, double ub) {
  const double tol_same_dist = 1.0e-5, min_hub_diff = 0.3125e-5;
  const double ta = 3.2 * ua, tb = 3.2 * ub;
  if (r < tol_same_dist) {
    if (std::fabs(ua - ub) < min_hub_diff) return -0.5 * (ua + ub);
    const double s = ta + tb;
    return -0.5 * ((ta * tb) / s + (ta * tb) * (ta * tb) / (s * s * s));
  }
  if (std::fabs(ua - ub) < min_hub_diff) {
    const double tm = 0.5 * (ta + tb);
    return std::exp(-tm * r) * (1.0 / r + 0.6875 * tm + 0.1875 * r * tm * tm +
                                0.02083333333333333333 * r * r * tm * tm * tm);
  }
  return gamma_sub(r, ta, tb) + gamma_sub(r, tb, ta);
}

// src/dftbp/dftb/shortgammafuncs.F90:27-93 (minShortGamma = tolShortGamma = 1e-10)
double exp_gamma_cutoff(double u2, double u1) {
  const double cut_value = 1.0e-10, tol_short = 1.0e-10;
  double rab = 1.0;
  while (exp_gamma(rab, u2, u1) > cut_value) rab *= 2.0;
  double min_cut = rab + 0.1, max_cut = 0.5 * rab - 0.1;
  const double max_g = exp_gamma(max_cut, u2, u1), min_g = exp_gamma(min_cut, u2, u1);
  double lower_g = exp_gamma(min_cut, u2, u1);
  double cut = max_cut + 0.1;
  double g = exp_gamma(cut, u2, u1);
  while ((g - lower_g) > tol_short) {
    max_cut = 0.5 * (cut + min_cut);
    if ((max_g >= min_g) == (cut_value >= exp_gamma(max_cut, u2, u1))) {
      min_cut = max_cut;
      lower_g = exp_gamma(min_cut, u2, u1);
    } else {
      cut = max_cut;
      g = exp_gamma(cut, u2, u1);
    }
  }
  return min_cut;
}

// Cardinal B-spline M_p(k+1), k = 0..p-2, by the order recursion, and
// |b(m)|^2 = 1 / |sum_k M_p(k+1) exp(2 pi i m k / K)|^2  (Essmann et al. 1995, eq. 4.4).
void bspline_moduli(int K, int p, double* bsq) {
  long double c[kMaxOrder + 2] = {0.0L};
  c[0] = 0.0L;  // M_2(0)
  c[1] = 1.0L;  // M_2(1)
  for (int k = 2; k < p; ++k) {
    long double nx[kMaxOrder + 2] = {0.0L};
    for (int j = 0; j <= k; ++j) {
      const long double a = j < k ? (long double)j * c[j] : 0.0L;
      const long double b = j > 0 ? (long double)(k + 1 - j) * c[j - 1] : 0.0L;
      nx[j] = (a + b) / k;
    }
    for (int j = 0; j <= k; ++j) c[j] = nx[j];
  }
  // c[j] = M_p(j), j = 0..p-1
  for (int m = 0; m < K; ++m) {
    long double re = 0.0L, im = 0.0L;
    for (int k = 0; k <= p - 2; ++k) {
      const long double ang = 2.0L * 3.14159265358979323846264338327950288L * m * k / K;
      re += c[k + 1] * cosl(ang);
      im += c[k + 1] * sinl(ang);
    }
    const long double d = re * re + im * im;
    bsq[m] = d > 1.0e-14L ? double(1.0L / d) : 0.0;
  }
}

FastDiv make_fastdiv(int d) {
  int s = 0;
  while ((1LL << s) < d) ++s;
  const unsigned long long m = ((1ULL << 32) * ((1ULL << s) - (unsigned long long)d)) / d + 1;
  return FastDiv{unsigned(m), s};
}

// Radix passes of a length n = 2^a 3^b 5^c in the order of `variant`:
//   0: 8s, then a 4 or a 2, then 3s, 5s      1: 4s, then a 2, then 3s, 5s
//   2: the leftover 2 or 4 first, then 8s, 3s, 5s      3: 3s and 5s first, then 8s, 4, 2
//   4: 16s, then an 8, 4 or 2, then 3s, 5s
// Every order is a valid Stockham plan (the twiddle tables follow the same order).
int fft_plan(int n, int* radices, int variant) {
  int a = 0, b = 0, c = 0, r = n;
  while (r > 0 && r % 2 == 0) { ++a; r /= 2; }
  while (r > 0 && r % 3 == 0) { ++b; r /= 3; }
  while (r > 0 && r % 5 == 0) { ++c; r /= 5; }
  if (r != 1) return -1;
  std::vector<int> pw;  // powers of two
  if (variant == 4) {
    for (int i = 0; i < a / 4; ++i) pw.push_back(16);
    if (a % 4) pw.push_back(1 << (a % 4));
  } else if (variant == 1) {
    for (int i = 0; i < a / 2; ++i) pw.push_back(4);
    if (a % 2) pw.push_back(2);
  } else {
    if (variant == 2 && a % 3) pw.push_back(a % 3 == 1 ? 2 : 4);
    for (int i = 0; i < a / 3; ++i) pw.push_back(8);
    if (variant != 2 && a % 3) pw.push_back(a % 3 == 1 ? 2 : 4);
  }
  std::vector<int> odd;
  for (int i = 0; i < b; ++i) odd.push_back(3);
  for (int i = 0; i < c; ++i) odd.push_back(5);
  std::vector<int> all;
  if (variant == 3) {
    all = odd;
    all.insert(all.end(), pw.begin(), pw.end());
  } else {
    all = pw;
    all.insert(all.end(), odd.begin(), odd.end());
  }
  if ((int)all.size() > kMaxRadixPasses) return -1;
  for (size_t i = 0; i < all.size(); ++i) radices[i] = all[i];
  return (int)all.size();
}

}  // namespace ewamd
