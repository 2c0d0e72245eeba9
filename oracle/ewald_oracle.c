/*
 * ewald_oracle.c -- CPU restatement of the reference's periodic Coulomb path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline of the benchmark.  Nothing in the product links or loads it.
 *
 * Each function names the reference file:line it follows.  "fe/" abbreviates
 * src/dftbp/extlibs/fftewald/ of the reference repository.  Plain C99 + OpenMP.
 */
#include "ewald_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846264338327950288
#define SQRT_PI 1.77245385090551602729816748334114518

/* fe/common/ewald_validation.f90:13-20 */
static const double MIN_SEPARATION_SQ = 1.0e-24;
static const double MIN_WAVENUMBER_SQ = 1.0e-24;
static const double NEUTRALITY_TOL = 1.0e-8;

/* ------------------------------------------------------------------------ */
/* geometry, validation, self term                                           */
/* ------------------------------------------------------------------------ */

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* fe/common/ewald_geometry.f90:65-93 */
int orc_cell_metrics(const double *lat9, orc_cell *c) {
  double x[3];
  memset(c, 0, sizeof(*c));
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) c->lat[i][d] = lat9[3 * i + d];
  cross3(c->lat[1], c->lat[2], x);
  c->volume = fabs(c->lat[0][0] * x[0] + c->lat[0][1] * x[1] + c->lat[0][2] * x[2]);
  if (c->volume <= 1.0e-18) return ORC_ERR_VOLUME;
  cross3(c->lat[1], c->lat[2], x);
  for (int d = 0; d < 3; ++d) c->rec[0][d] = x[d] / c->volume;
  cross3(c->lat[2], c->lat[0], x);
  for (int d = 0; d < 3; ++d) c->rec[1][d] = x[d] / c->volume;
  cross3(c->lat[0], c->lat[1], x);
  for (int d = 0; d < 3; ++d) c->rec[2][d] = x[d] / c->volume;
  for (int i = 0; i < 3; ++i) {
    c->latlen[i] = sqrt(c->lat[i][0] * c->lat[i][0] + c->lat[i][1] * c->lat[i][1] +
                        c->lat[i][2] * c->lat[i][2]);
    c->reclen[i] = sqrt(c->rec[i][0] * c->rec[i][0] + c->rec[i][1] * c->rec[i][1] +
                        c->rec[i][2] * c->rec[i][2]);
  }
  return ORC_OK;
}

/* fe/common/ewald_validation.f90:26-48 */
int orc_check_configuration(int n, const double *q) {
  if (n <= 0) return ORC_ERR_N;
  double net = 0.0, absq = 0.0;
  for (int i = 0; i < n; ++i) {
    net += q[i];
    absq += fabs(q[i]);
  }
  if (fabs(net) > NEUTRALITY_TOL * absq) return ORC_ERR_NOT_NEUTRAL;
  return ORC_OK;
}

/* fe/common/ewald_self.f90:12-19 */
double orc_self_potential(double q, double alpha) { return -2.0 * alpha / SQRT_PI * q; }

/* ------------------------------------------------------------------------ */
/* truncation, Lambert W, alpha search                                       */
/* ------------------------------------------------------------------------ */

/* fe/common/ewald_truncation.f90:36-45 */
void orc_direct_cutoffs(double alpha, double tol, double *r_cut, double *k_cut) {
  *r_cut = sqrt(-log(tol)) / alpha;
  *k_cut = 2.0 * alpha * sqrt(-log(tol));
}

/* fe/common/ewald_lambert_w.f90:27-56: Halley iteration from log(1+x). */
double orc_lambert_w0(double x) {
  if (x <= 0.0) return 0.0;
  double w = log(1.0 + x);
  for (int it = 1; it <= 60; ++it) {
    double ew = exp(w);
    double res = w * ew - x;
    double wp1 = w + 1.0;
    double corr = res / (ew * wp1 - 0.5 * (w + 2.0) * res / wp1);
    w = w - corr;
    if (fabs(corr) <= 1.0e-15 * (1.0 + fabs(w))) break;
  }
  return w;
}

/* fe/common/ewald_truncation.f90:53-78 */
double orc_splitting_from_real_budget(double r_cut, double q2sum, int n, double pseudo_volume,
                                      double budget) {
  double arg = (2.0 / budget) * q2sum * sqrt(r_cut / ((double)n * pseudo_volume));
  double w;
  if (arg > 0.0)
    w = orc_lambert_w0(arg);
  else
    w = -log(budget);
  if (w <= 0.0) w = -log(budget);
  return sqrt(w) / r_cut;
}

/* fe/common/ewald_truncation.f90:84-104 */
double orc_cutoff_from_fourier_budget(double q2sum, int n, double box_edge, double alpha,
                                      double budget) {
  double arg = (4.0 / (3.0 * box_edge * box_edge)) *
               pow(2.0 / ((double)n * alpha * PI), 2.0 / 3.0) *
               pow(2.0 * q2sum / budget, 4.0 / 3.0);
  double w = orc_lambert_w0(arg);
  if (w <= 0.0) w = -log(budget);
  return sqrt(3.0) * alpha * sqrt(w);
}

/* fe/common/ewald_alpha_search.f90:174-180 (balance_condition) */
static double balance_condition(double a, double minG, double minR, double vol) {
  double g4 = 4.0 * minG, g5 = 5.0 * minG;
  double f4 = (4.0 * PI / vol) * exp(-(g4 * g4) / (4.0 * a * a)) / (g4 * g4);
  double f5 = (4.0 * PI / vol) * exp(-(g5 * g5) / (4.0 * a * a)) / (g5 * g5);
  /* real_shell_term, fe/common/ewald_alpha_search.f90:60-75 */
  double r2 = erfc(a * (2.0 * minR)) / (2.0 * minR);
  double r3 = erfc(a * (3.0 * minR)) / (3.0 * minR);
  return (f4 - f5) - (r2 - r3);
}

/* src/dftbp/dftb/coulomb.F90:2686-2748 (diffRecReal with gTerm and rTerm) */
static double diff_rec_real(double a, double minG, double minR, double vol) {
  double g4 = 4.0 * minG, g5 = 5.0 * minG;
  double t4 = 4.0 * PI * (exp(-0.25 * g4 * g4 / (a * a)) / (vol * g4 * g4));
  double t5 = 4.0 * PI * (exp(-0.25 * g5 * g5 / (a * a)) / (vol * g5 * g5));
  double r2 = erfc(a * (2.0 * minR)) / (2.0 * minR);
  double r3 = erfc(a * (3.0 * minR)) / (3.0 * minR);
  return (t4 - t5) - (r2 - r3);
}

/* Bracket-and-bisect shared by fe/common/ewald_alpha_search.f90:83-181 and
 * coulomb.F90:2097-2179; they differ in the condition and the bisection cap. */
static int alpha_search(double (*cond)(double, double, double, double), double minG,
                        double minR, double vol, double tol, int max_bisection,
                        double *alpha_out) {
  const double alpha_start = 1.0e-8;
  int status = 0;
  double alpha = alpha_start, left = 0.0, right;
  double imb = cond(alpha, minG, minR, vol);
  while (imb < -tol && alpha <= 1.79769313486231570e308) {
    alpha = 2.0 * alpha;
    imb = cond(alpha, minG, minR, vol);
  }
  if (alpha > 1.79769313486231570e308)
    status = 1;
  else if (alpha == alpha_start)
    status = 2;
  if (status == 0) {
    left = 0.5 * alpha;
    while (imb < tol && alpha <= 1.79769313486231570e308) {
      alpha = 2.0 * alpha;
      imb = cond(alpha, minG, minR, vol);
    }
    if (alpha > 1.79769313486231570e308) status = 3;
  }
  if (status == 0) {
    right = alpha;
    alpha = (left + right) / 2.0;
    int it = 0;
    imb = cond(alpha, minG, minR, vol);
    while (fabs(imb) > tol && it <= max_bisection) {
      if (imb < 0.0)
        left = alpha;
      else
        right = alpha;
      alpha = (left + right) / 2.0;
      imb = cond(alpha, minG, minR, vol);
      it += 1;
    }
    if (it > max_bisection) status = 4;
  }
  *alpha_out = alpha;
  return status;
}

/* fe/common/ewald_alpha_search.f90:83-181 (maxBisection = 100, line 54) */
int orc_balanced_alpha(double minG, double minR, double volume, double tol, double *alpha) {
  return alpha_search(balance_condition, minG, minR, volume, tol, 100, alpha);
}

/* fe/direct/ewald_direct_3d.f90:319-347 */
double orc_balanced_splitting(const orc_cell *c, double tol) {
  double minR = fmin(c->latlen[0], fmin(c->latlen[1], c->latlen[2]));
  double minG = 2.0 * PI * fmin(c->reclen[0], fmin(c->reclen[1], c->reclen[2]));
  double alpha;
  int st = orc_balanced_alpha(minG, minR, c->volume, tol, &alpha);
  if (st != 0) alpha = exp(-0.310104 * log(c->volume) + 0.786382) / 2.0;
  return alpha;
}

/* src/dftbp/dftb/coulomb.F90:2097-2179, nSearchIter = 30 (common/accuracy.F90:114).
 * DFTB+ reciprocal vectors carry the factor 2 pi, hence minG = 2 pi min |b_i|. */
int orc_optimal_alpha_dftb(const double *lat9, double tol, double *alpha) {
  orc_cell c;
  int st = orc_cell_metrics(lat9, &c);
  if (st) return st;
  double minG = 1e308, minR = 1e308;
  for (int i = 0; i < 3; ++i) {
    double g2 = 0.0, r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      double g = 2.0 * PI * c.rec[i][d];
      g2 += g * g;
      r2 += c.lat[i][d] * c.lat[i][d];
    }
    if (g2 < minG) minG = g2;
    if (r2 < minR) minR = r2;
  }
  minG = sqrt(minG);
  minR = sqrt(minR);
  st = alpha_search(diff_rec_real, minG, minR, c.volume, tol, 30, alpha);
  return st ? ORC_ERR_ALPHA_SEARCH : ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* real space                                                                */
/* ------------------------------------------------------------------------ */

/* fe/direct/ewald_direct_3d.f90:192-309 */
int orc_real_space_images(int n, const double *pos, const double *q, const orc_cell *c,
                          double alpha, double r_cut, double *pot, double *force) {
  const double rc2 = r_cut * r_cut;
  int nimg[3];
  for (int d = 0; d < 3; ++d) nimg[d] = (int)ceil(r_cut * c->reclen[d]) + 1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    double acc = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int j = 0; j < n; ++j) {
      for (int nx = -nimg[0]; nx <= nimg[0]; ++nx)
        for (int ny = -nimg[1]; ny <= nimg[1]; ++ny)
          for (int nz = -nimg[2]; nz <= nimg[2]; ++nz) {
            if (i == j && nx == 0 && ny == 0 && nz == 0) continue;
            double dx = pos[3 * i + 0] - pos[3 * j + 0] + nx * c->lat[0][0] + ny * c->lat[1][0] +
                        nz * c->lat[2][0];
            double dy = pos[3 * i + 1] - pos[3 * j + 1] + nx * c->lat[0][1] + ny * c->lat[1][1] +
                        nz * c->lat[2][1];
            double dz = pos[3 * i + 2] - pos[3 * j + 2] + nx * c->lat[0][2] + ny * c->lat[1][2] +
                        nz * c->lat[2][2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < MIN_SEPARATION_SQ) {
              bad = 1;
              continue;
            }
            if (r2 > rc2) continue;
            double r = sqrt(r2);
            acc += q[j] * erfc(alpha * r) / r;
            if (force) {
              double f = q[i] * q[j] *
                         (2.0 * alpha / SQRT_PI * exp(-alpha * alpha * r2) + erfc(alpha * r) / r) /
                         r2;
              fx += f * dx;
              fy += f * dy;
              fz += f * dz;
            }
          }
    }
    pot[i] = acc;
    if (force) {
      force[3 * i + 0] = fx;
      force[3 * i + 1] = fy;
      force[3 * i + 2] = fz;
    }
  }
  return bad ? ORC_ERR_COINCIDENT : ORC_OK;
}

/* fe/common/ewald_cell_list.f90:56-73 */
static int box_count(double reclen, double r_cut) { return (int)floor(1.0 / (r_cut * reclen)); }
/* fe/common/ewald_cell_list.f90:78-91 */
static int box_index(double frac, int nb) {
  int b = (int)(frac * nb);
  return b < nb - 1 ? b : nb - 1;
}

/* fe/fft/ewald_fft_3d_real.f90:32-187 with fe/common/ewald_cell_list.f90:96-141 */
int orc_real_space_cell_list(int n, const double *pos, const double *q, const orc_cell *c,
                             double alpha, double r_cut, double *pot, double *force) {
  int nb[3];
  for (int d = 0; d < 3; ++d) nb[d] = box_count(c->reclen[d], r_cut);
  if (nb[0] < 3 || nb[1] < 3 || nb[2] < 3)
    return orc_real_space_images(n, pos, q, c, alpha, r_cut, pot, force);
  const double rc2 = r_cut * r_cut;
  double *fr = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  int *next = (int *)malloc(sizeof(int) * (size_t)n);
  size_t nbox = (size_t)nb[0] * nb[1] * nb[2];
  int *head = (int *)malloc(sizeof(int) * nbox);
  if (!fr || !next || !head) {
    free(fr);
    free(next);
    free(head);
    return ORC_ERR_ALLOC;
  }
  for (size_t b = 0; b < nbox; ++b) head[b] = -1;
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      double s = c->rec[d][0] * pos[3 * i] + c->rec[d][1] * pos[3 * i + 1] +
                 c->rec[d][2] * pos[3 * i + 2];
      fr[3 * i + d] = s - floor(s);
    }
    int ix = box_index(fr[3 * i], nb[0]), iy = box_index(fr[3 * i + 1], nb[1]),
        iz = box_index(fr[3 * i + 2], nb[2]);
    size_t b = ((size_t)iz * nb[1] + iy) * nb[0] + ix;
    next[i] = head[b];
    head[b] = i;
  }
  int bad = 0;
#pragma omp parallel for schedule(guided) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    int ix = box_index(fr[3 * i], nb[0]), iy = box_index(fr[3 * i + 1], nb[1]),
        iz = box_index(fr[3 * i + 2], nb[2]);
    double acc = 0.0, f3[3] = {0.0, 0.0, 0.0};
    for (int dz = -1; dz <= 1; ++dz) {
      int jz = ((iz + dz) % nb[2] + nb[2]) % nb[2];
      for (int dy = -1; dy <= 1; ++dy) {
        int jy = ((iy + dy) % nb[1] + nb[1]) % nb[1];
        for (int dx = -1; dx <= 1; ++dx) {
          int jx = ((ix + dx) % nb[0] + nb[0]) % nb[0];
          for (int j = head[((size_t)jz * nb[1] + jy) * nb[0] + jx]; j >= 0; j = next[j]) {
            if (j == i) continue;
            double ds[3];
            for (int d = 0; d < 3; ++d) {
              ds[d] = fr[3 * i + d] - fr[3 * j + d];
              ds[d] = ds[d] - round(ds[d]);
            }
            double sv[3];
            for (int d = 0; d < 3; ++d)
              sv[d] = ds[0] * c->lat[0][d] + ds[1] * c->lat[1][d] + ds[2] * c->lat[2][d];
            double r2 = sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2];
            if (r2 < MIN_SEPARATION_SQ) {
              bad = 1;
              continue;
            }
            if (r2 <= rc2) {
              double r = sqrt(r2);
              double ce = erfc(alpha * r);
              acc += q[j] * ce / r;
              if (force) {
                double f =
                    q[i] * q[j] * (2.0 * alpha / SQRT_PI * exp(-alpha * alpha * r2) + ce / r) / r2;
                for (int d = 0; d < 3; ++d) f3[d] += f * sv[d];
              }
            }
          }
        }
      }
    }
    pot[i] = acc;
    if (force)
      for (int d = 0; d < 3; ++d) force[3 * i + d] = f3[d];
  }
  free(fr);
  free(next);
  free(head);
  return bad ? ORC_ERR_COINCIDENT : ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* direct Ewald (the parity reference)                                       */
/* ------------------------------------------------------------------------ */

/* fe/direct/ewald_direct_3d.f90:37-182 */
int orc_direct_ewald(int n, const double *pos, const double *q, const double *lat9, double tol,
                     double alpha_in, double r_cut_in, double k_cut_in, int real_space,
                     int phase_tables, double *pot, double *force, double *alpha_used) {
  int st = orc_check_configuration(n, q);
  if (st) return st;
  orc_cell c;
  st = orc_cell_metrics(lat9, &c);
  if (st) return st;
  if (tol <= 0.0) tol = 1.0e-12; /* defaultTolerance, fe/direct/ewald_direct_3d.f90:24 */
  if (tol >= 1.0) return ORC_ERR_TOL;
  double alpha = alpha_in > 0.0 ? alpha_in : orc_balanced_splitting(&c, tol);
  if (alpha <= 0.0) return ORC_ERR_PARAM;
  double r_cut, k_cut;
  orc_direct_cutoffs(alpha, tol, &r_cut, &k_cut);
  if (r_cut_in > 0.0) r_cut = r_cut_in;
  if (k_cut_in > 0.0) k_cut = k_cut_in;
  if (alpha_used) *alpha_used = alpha;

  st = real_space ? orc_real_space_cell_list(n, pos, q, &c, alpha, r_cut, pot, force)
                  : orc_real_space_images(n, pos, q, &c, alpha, r_cut, pot, force);
  if (st) return st;

  /* Fourier branch: full mode box +-nModes with a k_cut sphere. */
  int nm[3];
  for (int d = 0; d < 3; ++d) nm[d] = (int)ceil(k_cut * c.latlen[d] / (2.0 * PI)) + 1;
  const double kc2 = k_cut * k_cut;
  /* enumerate the retained modes in the reference's loop order (mx, my, mz) */
  size_t cap = (size_t)(2 * nm[0] + 1) * (2 * nm[1] + 1) * (2 * nm[2] + 1);
  int *modes = (int *)malloc(sizeof(int) * 3 * cap);
  double *kvec = (double *)malloc(sizeof(double) * 4 * cap);
  if (!modes || !kvec) {
    free(modes);
    free(kvec);
    return ORC_ERR_ALLOC;
  }
  size_t nmodes = 0;
  for (int mx = -nm[0]; mx <= nm[0]; ++mx)
    for (int my = -nm[1]; my <= nm[1]; ++my)
      for (int mz = -nm[2]; mz <= nm[2]; ++mz) {
        double k[3];
        for (int d = 0; d < 3; ++d)
          k[d] = (mx * c.rec[0][d] + my * c.rec[1][d] + mz * c.rec[2][d]) * 2.0 * PI;
        double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        if (k2 < MIN_WAVENUMBER_SQ) continue;
        if (k2 > kc2) continue;
        modes[3 * nmodes] = mx;
        modes[3 * nmodes + 1] = my;
        modes[3 * nmodes + 2] = mz;
        kvec[4 * nmodes] = k[0];
        kvec[4 * nmodes + 1] = k[1];
        kvec[4 * nmodes + 2] = k[2];
        kvec[4 * nmodes + 3] = 4.0 * PI / c.volume * exp(-k2 / (4.0 * alpha * alpha)) / k2;
        ++nmodes;
      }
  double complex *S = (double complex *)malloc(sizeof(double complex) * (nmodes ? nmodes : 1));
  double complex *ph = NULL;
  int w[3] = {2 * nm[0] + 1, 2 * nm[1] + 1, 2 * nm[2] + 1};
  if (phase_tables) {
    ph = (double complex *)malloc(sizeof(double complex) * (size_t)n * (w[0] + w[1] + w[2]));
    if (!ph) {
      free(modes);
      free(kvec);
      free(S);
      return ORC_ERR_ALLOC;
    }
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
      double complex *p = ph + (size_t)i * (w[0] + w[1] + w[2]);
      for (int d = 0; d < 3; ++d) {
        double s = c.rec[d][0] * pos[3 * i] + c.rec[d][1] * pos[3 * i + 1] +
                   c.rec[d][2] * pos[3 * i + 2];
        for (int m = -nm[d]; m <= nm[d]; ++m) {
          double a = 2.0 * PI * (double)m * s;
          p[m + nm[d]] = cos(a) + I * sin(a);
        }
        p += w[d];
      }
    }
  }
#define PHASE(i, md)                                                                             \
  (ph ? (ph[(size_t)(i) * (w[0] + w[1] + w[2]) + modes[3 * (md)] + nm[0]] *                      \
         ph[(size_t)(i) * (w[0] + w[1] + w[2]) + w[0] + modes[3 * (md) + 1] + nm[1]] *           \
         ph[(size_t)(i) * (w[0] + w[1] + w[2]) + w[0] + w[1] + modes[3 * (md) + 2] + nm[2]])     \
      : (cos(kvec[4 * (md)] * pos[3 * (i)] + kvec[4 * (md) + 1] * pos[3 * (i) + 1] +            \
             kvec[4 * (md) + 2] * pos[3 * (i) + 2]) +                                            \
         I * sin(kvec[4 * (md)] * pos[3 * (i)] + kvec[4 * (md) + 1] * pos[3 * (i) + 1] +        \
                 kvec[4 * (md) + 2] * pos[3 * (i) + 2])))
#pragma omp parallel for schedule(static)
  for (size_t md = 0; md < nmodes; ++md) {
    double complex s = 0.0;
    for (int j = 0; j < n; ++j) s += q[j] * PHASE(j, md);
    S[md] = s;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double acc = 0.0, f3[3] = {0.0, 0.0, 0.0};
    for (size_t md = 0; md < nmodes; ++md) {
      double complex loc = S[md] * conj(PHASE(i, md));
      double wgt = kvec[4 * md + 3];
      acc += wgt * creal(loc);
      if (force) {
        double ff = q[i] * wgt * cimag(loc);
        for (int d = 0; d < 3; ++d) f3[d] -= ff * kvec[4 * md + d];
      }
    }
    pot[i] += acc;
    if (force)
      for (int d = 0; d < 3; ++d) force[3 * i + d] += f3[d];
  }
#undef PHASE
  for (int i = 0; i < n; ++i) pot[i] += orc_self_potential(q[i], alpha);
  free(modes);
  free(kvec);
  free(S);
  free(ph);
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* mixed-radix complex FFT (Stockham autosort)                               */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, nr;
  int radix[32];
  double complex *tw; /* tw[m] = exp(sign 2 pi i m / n) */
} fft_plan1;

static int fft_plan_init(fft_plan1 *p, int n, int sign) {
  p->n = n;
  p->nr = 0;
  int r = n;
  while (r % 4 == 0 && r > 4) {
    p->radix[p->nr++] = 4;
    r /= 4;
  }
  static const int primes[] = {2, 3, 4, 5, 7};
  for (int k = 0; k < 5 && r > 1; ++k)
    while (r % primes[k] == 0) {
      p->radix[p->nr++] = primes[k];
      r /= primes[k];
    }
  for (int f = 11; r > 1; f += 2)
    while (r % f == 0) {
      p->radix[p->nr++] = f;
      r /= f;
    }
  p->tw = (double complex *)malloc(sizeof(double complex) * (size_t)n);
  if (!p->tw) return ORC_ERR_ALLOC;
  for (int m = 0; m < n; ++m) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / n;
    p->tw[m] = (double)cosl(a) + I * (double)(sign * sinl(a));
  }
  return ORC_OK;
}

/* One 1D transform of x (length n, contiguous); work has room for n values. */
static void fft_exec1(const fft_plan1 *p, double complex *x, double complex *work) {
  const int n = p->n;
  double complex *src = x, *dst = work;
  int Ns = 1;
  double complex v[64], y[64];
  for (int s = 0; s < p->nr; ++s) {
    const int R = p->radix[s];
    const int nr = n / R;
    for (int j = 0; j < nr; ++j) {
      int k = j % Ns;
      int tstep = k * (n / (Ns * R)); /* twiddle exponent per t, in units of 1/n */
      for (int t = 0; t < R; ++t) v[t] = src[j + t * nr] * p->tw[(size_t)(t * tstep) % n];
      for (int u = 0; u < R; ++u) {
        double complex a = 0.0;
        for (int t = 0; t < R; ++t) a += v[t] * p->tw[(size_t)((t * u) % R) * nr];
        y[u] = a;
      }
      int base = (j / Ns) * Ns * R + k;
      for (int t = 0; t < R; ++t) dst[base + t * Ns] = y[t];
    }
    double complex *tmp = src;
    src = dst;
    dst = tmp;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(double complex) * (size_t)n);
}

/* Unnormalised 3D DFT, first index fastest, sign +1 = exp(+2 pi i ...). */
int orc_fft3d(int n1, int n2, int n3, double *data, int sign) {
  double complex *g = (double complex *)data;
  int dims[3] = {n1, n2, n3};
  size_t strides[3] = {1, (size_t)n1, (size_t)n1 * n2};
  size_t total = (size_t)n1 * n2 * n3;
  for (int ax = 0; ax < 3; ++ax) {
    int n = dims[ax];
    if (n == 1) continue;
    fft_plan1 plan;
    if (fft_plan_init(&plan, n, sign >= 0 ? 1 : -1)) return ORC_ERR_ALLOC;
    size_t nlines = total / n, st = strides[ax];
    int err = 0;
#pragma omp parallel reduction(| : err)
    {
      double complex *line = (double complex *)malloc(sizeof(double complex) * (size_t)n);
      double complex *work = (double complex *)malloc(sizeof(double complex) * (size_t)n);
      if (!line || !work) err = 1;
#pragma omp for schedule(static)
      for (size_t L = 0; L < nlines; ++L) {
        if (err) continue;
        /* line L -> base offset: split L into (index below axis, index above axis) */
        size_t lo = L % st, hi = L / st;
        size_t base = hi * st * n + lo;
        for (int e = 0; e < n; ++e) line[e] = g[base + (size_t)e * st];
        fft_exec1(&plan, line, work);
        for (int e = 0; e < n; ++e) g[base + (size_t)e * st] = line[e];
      }
      free(line);
      free(work);
    }
    free(plan.tw);
    if (err) return ORC_ERR_ALLOC;
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* the reference fast path: parameter chain + Gaussian NFFT (CPU baseline)   */
/* ------------------------------------------------------------------------ */

/* fe/nfft/fft_backend.f90:73-91 */
int orc_next_fast_length(int n) {
  int m = n > 2 ? n : 2;
  if (m % 2) m += 1;
  for (;;) {
    int r = m;
    while (r % 2 == 0) r /= 2;
    while (r % 3 == 0) r /= 3;
    while (r % 5 == 0) r /= 5;
    while (r % 7 == 0) r /= 7;
    if (r == 1) return m;
    m += 2;
  }
}

/* fe/nfft/nfft.f90:99-109 (windowErrorRate = 2 pi / 3, line 43) */
int orc_window_cutoff_from_budget(double budget) {
  int m = (int)ceil(log(4.0 / budget) / (2.0 * PI / 3.0));
  return m > 2 ? m : 2;
}

/* fe/fft/ewald_fft_3d_parameters.f90:144-168 */
static int modes_along_axis(double k_cut, double latlen) {
  int hi = (int)ceil(k_cut * latlen / (2.0 * PI)) + 1;
  int m = orc_next_fast_length(2 * hi);
  return m > 4 ? m : 4;
}

/* fe/fft/ewald_fft_3d_parameters.f90:52-142 */
int orc_choose_parameters(const orc_cell *c, double q2sum, int n, double tol, double alpha_in,
                          double r_cut_in, double k_cut_in, int nmodes_in, int window_in,
                          orc_fft_params *p) {
  double spacing = pow(c->volume / (double)n, 1.0 / 3.0);
  if (r_cut_in > 0.0)
    p->r_cut = r_cut_in;
  else if (alpha_in > 0.0)
    p->r_cut = sqrt(-log(tol)) / alpha_in;
  else {
    p->r_cut = 3.0 * spacing;
    double mx = fmax(c->reclen[0], fmax(c->reclen[1], c->reclen[2]));
    double limit = 0.999 / (3.0 * mx);
    if (limit >= 1.5 * spacing) p->r_cut = fmin(p->r_cut, limit);
  }
  if (alpha_in > 0.0)
    p->alpha = alpha_in;
  else
    p->alpha = orc_splitting_from_real_budget(p->r_cut, q2sum, n, c->volume, 0.2 * tol);
  if (p->alpha <= 0.0) return ORC_ERR_PARAM;
  p->k_cut = orc_cutoff_from_fourier_budget(q2sum, n, pow(c->volume, 1.0 / 3.0), p->alpha, tol);
  if (k_cut_in > 0.0) p->k_cut = k_cut_in;
  p->window_cutoff = orc_window_cutoff_from_budget(tol);
  if (window_in > 0) p->window_cutoff = window_in;
  for (int d = 0; d < 3; ++d) p->nmodes[d] = modes_along_axis(p->k_cut, c->latlen[d]);
  if (nmodes_in > 0) p->nmodes[0] = p->nmodes[1] = p->nmodes[2] = nmodes_in;
  return ORC_OK;
}

/* fe/fft/ewald_fft_3d_parameters.f90:177-229 */
double orc_grid_memory_gigabytes(const orc_cell *c, double q2sum, int n, double tol,
                                 double alpha_in, double r_cut_in, double k_cut_in, int nmodes_in,
                                 int window_in) {
  orc_fft_params p;
  if (orc_choose_parameters(c, q2sum, n, tol, alpha_in, r_cut_in, k_cut_in, nmodes_in, window_in,
                            &p))
    return -1.0;
  double e = (double)(2 * p.nmodes[0]) * (double)(2 * p.nmodes[1]) * (double)(2 * p.nmodes[2]) +
             2.0 * (double)p.nmodes[0] * (double)p.nmodes[1] * (double)p.nmodes[2];
  return e * 16.0 / (1024.0 * 1024.0 * 1024.0);
}

/* fe/nfft/nfft.f90:61-75 */
static int mode_of_bin(int a, int M) { return a >= M / 2 ? a - M : a; }
/* fe/nfft/nfft.f90:82-92, oversampling 2 */
static double window_shape(int m) { return (4.0 / 3.0) * ((double)m / PI); }
/* fe/nfft/nfft.f90:122-159 */
static void stencil_weights(double t, int n, int m, double b, int *lowest, double *w) {
  double wrapped = t - floor(t);
  double scaled = wrapped * (double)n;
  int lo = (int)floor(scaled);
  double frac = scaled - (double)lo;
  if (lo >= n) lo -= n;
  double pref = 1.0 / sqrt(PI * b);
  for (int o = -m; o <= m; ++o) w[o + m] = pref * exp(-(frac - o) * (frac - o) / b);
  *lowest = lo;
}
/* fe/nfft/nfft.f90:163-189 */
static void deconvolution_table(int M, int n, double b, int *idx, double *fac) {
  for (int a = 0; a < M; ++a) {
    int k = mode_of_bin(a, M);
    idx[a] = ((k % n) + n) % n;
    fac[a] = exp(b * (PI * k / n) * (PI * k / n));
  }
}

/* fe/nfft/nfft.f90:196-305 */
int orc_nfft_adjoint_3d(int n, const double *t1, const double *t2, const double *t3,
                        const double *q, int M1, int M2, int M3, int m, double *shat) {
  int n1 = 2 * M1, n2 = 2 * M2, n3 = 2 * M3;
  double b = window_shape(m);
  size_t ng = (size_t)n1 * n2 * n3;
  double complex *fine = (double complex *)calloc(ng, sizeof(double complex));
  int *idx = (int *)malloc(sizeof(int) * (size_t)(M1 + M2 + M3));
  double *fac = (double *)malloc(sizeof(double) * (size_t)(M1 + M2 + M3));
  if (!fine || !idx || !fac) {
    free(fine);
    free(idx);
    free(fac);
    return ORC_ERR_ALLOC;
  }
  double *re = (double *)fine;
#pragma omp parallel
  {
    double *w1 = (double *)malloc(sizeof(double) * (2 * m + 1) * 3);
    double *w2 = w1 + 2 * m + 1, *w3 = w2 + 2 * m + 1;
#pragma omp for schedule(guided)
    for (int j = 0; j < n; ++j) {
      if (q[j] == 0.0) continue;
      int c1, c2, c3;
      stencil_weights(t1[j], n1, m, b, &c1, w1);
      stencil_weights(t2[j], n2, m, b, &c2, w2);
      stencil_weights(t3[j], n3, m, b, &c3, w3);
      for (int o3 = -m; o3 <= m; ++o3) {
        int i3 = ((c3 + o3) % n3 + n3) % n3;
        for (int o2 = -m; o2 <= m; ++o2) {
          int i2 = ((c2 + o2) % n2 + n2) % n2;
          double cw = q[j] * w2[o2 + m] * w3[o3 + m];
          for (int o1 = -m; o1 <= m; ++o1) {
            int i1 = ((c1 + o1) % n1 + n1) % n1;
            size_t g = ((size_t)i3 * n2 + i2) * n1 + i1;
#pragma omp atomic update
            re[2 * g] += cw * w1[o1 + m];
          }
        }
      }
    }
    free(w1);
  }
  int st = orc_fft3d(n1, n2, n3, (double *)fine, 1);
  if (!st) {
    int *i1 = idx, *i2 = idx + M1, *i3 = idx + M1 + M2;
    double *f1 = fac, *f2 = fac + M1, *f3 = fac + M1 + M2;
    deconvolution_table(M1, n1, b, i1, f1);
    deconvolution_table(M2, n2, b, i2, f2);
    deconvolution_table(M3, n3, b, i3, f3);
    double complex *S = (double complex *)shat;
    for (int a3 = 0; a3 < M3; ++a3)
      for (int a2 = 0; a2 < M2; ++a2)
        for (int a1 = 0; a1 < M1; ++a1)
          S[((size_t)a3 * M2 + a2) * M1 + a1] =
              fine[((size_t)i3[a3] * n2 + i2[a2]) * n1 + i1[a1]] * f1[a1] * f2[a2] * f3[a3];
  }
  free(fine);
  free(idx);
  free(fac);
  return st;
}

/* fe/nfft/nfft.f90:313-404 */
int orc_nfft_forward_3d(int n, const double *t1, const double *t2, const double *t3, int M1,
                        int M2, int M3, int m, const double *fhat, double *f) {
  int n1 = 2 * M1, n2 = 2 * M2, n3 = 2 * M3;
  double b = window_shape(m);
  size_t ng = (size_t)n1 * n2 * n3;
  double complex *fine = (double complex *)calloc(ng, sizeof(double complex));
  int *idx = (int *)malloc(sizeof(int) * (size_t)(M1 + M2 + M3));
  double *fac = (double *)malloc(sizeof(double) * (size_t)(M1 + M2 + M3));
  if (!fine || !idx || !fac) {
    free(fine);
    free(idx);
    free(fac);
    return ORC_ERR_ALLOC;
  }
  int *i1 = idx, *i2 = idx + M1, *i3 = idx + M1 + M2;
  double *f1 = fac, *f2 = fac + M1, *f3 = fac + M1 + M2;
  deconvolution_table(M1, n1, b, i1, f1);
  deconvolution_table(M2, n2, b, i2, f2);
  deconvolution_table(M3, n3, b, i3, f3);
  const double complex *F = (const double complex *)fhat;
  for (int a3 = 0; a3 < M3; ++a3)
    for (int a2 = 0; a2 < M2; ++a2)
      for (int a1 = 0; a1 < M1; ++a1)
        fine[((size_t)i3[a3] * n2 + i2[a2]) * n1 + i1[a1]] =
            F[((size_t)a3 * M2 + a2) * M1 + a1] * f1[a1] * f2[a2] * f3[a3];
  int st = orc_fft3d(n1, n2, n3, (double *)fine, -1);
  if (!st) {
    double complex *out = (double complex *)f;
#pragma omp parallel
    {
      double *w1 = (double *)malloc(sizeof(double) * (2 * m + 1) * 3);
      double *w2 = w1 + 2 * m + 1, *w3 = w2 + 2 * m + 1;
#pragma omp for schedule(guided)
      for (int j = 0; j < n; ++j) {
        int c1, c2, c3;
        stencil_weights(t1[j], n1, m, b, &c1, w1);
        stencil_weights(t2[j], n2, m, b, &c2, w2);
        stencil_weights(t3[j], n3, m, b, &c3, w3);
        double complex acc = 0.0;
        for (int o3 = -m; o3 <= m; ++o3) {
          int j3 = ((c3 + o3) % n3 + n3) % n3;
          for (int o2 = -m; o2 <= m; ++o2) {
            int j2 = ((c2 + o2) % n2 + n2) % n2;
            double ow = w2[o2 + m] * w3[o3 + m];
            for (int o1 = -m; o1 <= m; ++o1) {
              int j1 = ((c1 + o1) % n1 + n1) % n1;
              acc += w1[o1 + m] * ow * fine[((size_t)j3 * n2 + j2) * n1 + j1];
            }
          }
        }
        out[j] = acc;
      }
      free(w1);
    }
  }
  free(fine);
  free(idx);
  free(fac);
  return st;
}

/* fe/fft/ewald_fft_3d_fourier.f90:28-153 */
static int long_range(int n, const double *pos, const double *q, const orc_cell *c,
                      double alpha, const int *M, int m, double *pot, double *force) {
  size_t nm = (size_t)M[0] * M[1] * M[2];
  double *t = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double complex *coef = (double complex *)malloc(sizeof(double complex) * nm);
  double complex *sites = (double complex *)malloc(sizeof(double complex) * (size_t)n);
  if (!t || !coef || !sites) {
    free(t);
    free(coef);
    free(sites);
    return ORC_ERR_ALLOC;
  }
  double *t1 = t, *t2 = t + n, *t3 = t + 2 * n;
  for (int i = 0; i < n; ++i) {
    t1[i] = c->rec[0][0] * pos[3 * i] + c->rec[0][1] * pos[3 * i + 1] + c->rec[0][2] * pos[3 * i + 2];
    t2[i] = c->rec[1][0] * pos[3 * i] + c->rec[1][1] * pos[3 * i + 1] + c->rec[1][2] * pos[3 * i + 2];
    t3[i] = c->rec[2][0] * pos[3 * i] + c->rec[2][1] * pos[3 * i + 1] + c->rec[2][2] * pos[3 * i + 2];
  }
  int st = orc_nfft_adjoint_3d(n, t1, t2, t3, q, M[0], M[1], M[2], m, (double *)coef);
  if (st) goto done;
  double pref = 4.0 * PI / c->volume, i4a2 = 1.0 / (4.0 * alpha * alpha);
  for (int a3 = 0; a3 < M[2]; ++a3) {
    int k3 = mode_of_bin(a3, M[2]);
    for (int a2 = 0; a2 < M[1]; ++a2) {
      int k2 = mode_of_bin(a2, M[1]);
      for (int a1 = 0; a1 < M[0]; ++a1) {
        int k1 = mode_of_bin(a1, M[0]);
        size_t g = ((size_t)a3 * M[1] + a2) * M[0] + a1;
        if (k1 == 0 && k2 == 0 && k3 == 0) {
          coef[g] = 0.0;
          continue;
        }
        double kv[3];
        for (int d = 0; d < 3; ++d)
          kv[d] = 2.0 * PI * (k1 * c->rec[0][d] + k2 * c->rec[1][d] + k3 * c->rec[2][d]);
        double k2s = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        if (k2s < MIN_WAVENUMBER_SQ) {
          coef[g] = 0.0;
          continue;
        }
        coef[g] *= pref * exp(-k2s * i4a2) / k2s;
      }
    }
  }
  st = orc_nfft_forward_3d(n, t1, t2, t3, M[0], M[1], M[2], m, (const double *)coef,
                           (double *)sites);
  if (st) goto done;
  for (int i = 0; i < n; ++i) pot[i] = creal(sites[i]);
  if (force) {
    double complex *der = (double complex *)malloc(sizeof(double complex) * nm);
    if (!der) {
      st = ORC_ERR_ALLOC;
      goto done;
    }
    for (int comp = 0; comp < 3; ++comp) {
      for (int a3 = 0; a3 < M[2]; ++a3) {
        int k3 = mode_of_bin(a3, M[2]);
        for (int a2 = 0; a2 < M[1]; ++a2) {
          int k2 = mode_of_bin(a2, M[1]);
          for (int a1 = 0; a1 < M[0]; ++a1) {
            int k1 = mode_of_bin(a1, M[0]);
            size_t g = ((size_t)a3 * M[1] + a2) * M[0] + a1;
            double kc = 2.0 * PI * (k1 * c->rec[0][comp] + k2 * c->rec[1][comp] +
                                    k3 * c->rec[2][comp]);
            der[g] = (I * kc) * coef[g];
          }
        }
      }
      st = orc_nfft_forward_3d(n, t1, t2, t3, M[0], M[1], M[2], m, (const double *)der,
                               (double *)sites);
      if (st) break;
      for (int i = 0; i < n; ++i) force[3 * i + comp] = q[i] * creal(sites[i]);
    }
    free(der);
  }
done:
  free(t);
  free(coef);
  free(sites);
  return st;
}

/* fe/fft/ewald_fft_3d.f90:36-134 (defaultTolerance 1e-10, line 23) */
int orc_fft_ewald(int n, const double *pos, const double *q, const double *lat9, double tol,
                  double *pot, double *force, orc_fft_params *params_used) {
  int st = orc_check_configuration(n, q);
  if (st) return st;
  orc_cell c;
  st = orc_cell_metrics(lat9, &c);
  if (st) return st;
  if (tol <= 0.0) tol = 1.0e-10;
  if (tol >= 1.0) return ORC_ERR_TOL;
  double q2 = 0.0;
  for (int i = 0; i < n; ++i) q2 += q[i] * q[i];
  orc_fft_params p;
  st = orc_choose_parameters(&c, q2, n, tol, 0, 0, 0, 0, 0, &p);
  if (st) return st;
  if (params_used) *params_used = p;
  double *sr = (double *)malloc(sizeof(double) * (size_t)n);
  double *lr = (double *)malloc(sizeof(double) * (size_t)n);
  double *sf = force ? (double *)malloc(sizeof(double) * 3 * (size_t)n) : NULL;
  double *lf = force ? (double *)malloc(sizeof(double) * 3 * (size_t)n) : NULL;
  st = orc_real_space_cell_list(n, pos, q, &c, p.alpha, p.r_cut, sr, sf);
  if (!st) st = long_range(n, pos, q, &c, p.alpha, p.nmodes, p.window_cutoff, lr, lf);
  if (!st) {
    for (int i = 0; i < n; ++i) pot[i] = sr[i] + lr[i] + orc_self_potential(q[i], p.alpha);
    if (force)
      for (size_t k = 0; k < 3 * (size_t)n; ++k) force[k] = sf[k] + lf[k];
  }
  free(sr);
  free(lr);
  free(sf);
  free(lf);
  return st;
}

/* ------------------------------------------------------------------------ */
/* smooth particle-mesh Ewald (same method and conventions as the HIP path)  */
/* ------------------------------------------------------------------------ */

/* Cardinal B-spline M_p by the order recursion; theta[o] = M_p(w + p - 1 - o),
 * the weight at grid point floor(u) - (p - 1) + o, and dtheta its u-derivative. */
void orc_bspline_weights(double u, int p, double *theta, double *dtheta) {
  double w = u - floor(u);
  double c[64], cp[64];
  c[0] = w;
  c[1] = 1.0 - w;
  for (int k = 2; k < p; ++k) {
    /* c holds M_k(w + j), j = 0..k-1; build M_{k+1}(w + j), j = 0..k */
    if (k == p - 1)
      for (int j = 0; j < k; ++j) cp[j] = c[j];
    double nxt[64];
    for (int j = 0; j <= k; ++j) {
      double a = j < k ? (w + j) * c[j] : 0.0;
      double b = j > 0 ? (k + 1 - w - j) * c[j - 1] : 0.0;
      nxt[j] = (a + b) / k;
    }
    for (int j = 0; j <= k; ++j) c[j] = nxt[j];
  }
  if (p == 2) {
    cp[0] = 1.0; /* M_1 = indicator of [0,1) */
  }
  for (int o = 0; o < p; ++o) {
    int j = p - 1 - o;
    theta[o] = c[j];
    if (dtheta) {
      double a = j < p - 1 ? cp[j] : 0.0;
      double b = j > 0 ? cp[j - 1] : 0.0;
      dtheta[o] = a - b;
    }
  }
}

int orc_bspline_moduli(int K, int p, double *bsq) {
  double th[64];
  orc_bspline_weights(0.0, p, th, NULL); /* th[o] = M_p(p - 1 - o) */
  double mk[64];                          /* mk[k] = M_p(k + 1), k = 0..p-2 */
  for (int k = 0; k <= p - 2; ++k) mk[k] = th[p - 2 - k];
  for (int m = 0; m < K; ++m) {
    long double re = 0.0L, im = 0.0L;
    for (int k = 0; k <= p - 2; ++k) {
      long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m * k / K;
      re += mk[k] * cosl(a);
      im += mk[k] * sinl(a);
    }
    long double d = re * re + im * im;
    bsq[m] = d > 1.0e-14L ? (double)(1.0L / d) : 0.0;
  }
  return ORC_OK;
}

static int signed_mode(int k, int K) { return k <= (K - 1) / 2 ? k : k - K; }
static int is_nyquist(int k, int K) { return (K % 2 == 0) && (k == K / 2); }

int orc_spme_reciprocal(int n, const double *pos, const double *q, const double *lat9,
                        const int *K, int p, double alpha, double *pot_rec, double *force_rec) {
  orc_cell c;
  int st = orc_cell_metrics(lat9, &c);
  if (st) return st;
  if (p < 2 || p > 32 || K[0] < p || K[1] < p || K[2] < p || alpha <= 0.0) return ORC_ERR_PARAM;
  const int K1 = K[0], K2 = K[1], K3 = K[2];
  size_t ng = (size_t)K1 * K2 * K3;
  double complex *grid = (double complex *)calloc(ng, sizeof(double complex));
  double *th = (double *)malloc(sizeof(double) * (size_t)n * 6 * p);
  int *k0 = (int *)malloc(sizeof(int) * 3 * (size_t)n);
  double *bsq = (double *)malloc(sizeof(double) * (size_t)(K1 + K2 + K3));
  if (!grid || !th || !k0 || !bsq) {
    free(grid);
    free(th);
    free(k0);
    free(bsq);
    return ORC_ERR_ALLOC;
  }
  /* weights: th[(i*3 + d)*2p + o] value, + p derivative */
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      double s = c.rec[d][0] * pos[3 * i] + c.rec[d][1] * pos[3 * i + 1] +
                 c.rec[d][2] * pos[3 * i + 2];
      s -= floor(s);
      double u = s * K[d];
      int fl = (int)floor(u);
      if (fl >= K[d]) {
        fl -= K[d];
        u -= K[d];
      }
      double *t = th + ((size_t)i * 3 + d) * 2 * p;
      orc_bspline_weights(u, p, t, t + p);
      k0[3 * i + d] = ((fl - (p - 1)) % K[d] + K[d]) % K[d];
    }
  /* spread (serial: deterministic order) */
  for (int i = 0; i < n; ++i) {
    const double *w1 = th + ((size_t)i * 3 + 0) * 2 * p;
    const double *w2 = th + ((size_t)i * 3 + 1) * 2 * p;
    const double *w3 = th + ((size_t)i * 3 + 2) * 2 * p;
    for (int o3 = 0; o3 < p; ++o3) {
      int g3 = (k0[3 * i + 2] + o3) % K3;
      for (int o2 = 0; o2 < p; ++o2) {
        int g2 = (k0[3 * i + 1] + o2) % K2;
        double cw = q[i] * w2[o2] * w3[o3];
        for (int o1 = 0; o1 < p; ++o1) {
          int g1 = (k0[3 * i] + o1) % K1;
          grid[((size_t)g3 * K2 + g2) * K1 + g1] += cw * w1[o1];
        }
      }
    }
  }
  st = orc_fft3d(K1, K2, K3, (double *)grid, -1);
  if (st) goto out;
  orc_bspline_moduli(K1, p, bsq);
  orc_bspline_moduli(K2, p, bsq + K1);
  orc_bspline_moduli(K3, p, bsq + K1 + K2);
  {
    const double pref = 1.0 / (PI * c.volume), a2 = PI * PI / (alpha * alpha);
#pragma omp parallel for schedule(static)
    for (int k3 = 0; k3 < K3; ++k3)
      for (int k2 = 0; k2 < K2; ++k2)
        for (int k1 = 0; k1 < K1; ++k1) {
          size_t g = ((size_t)k3 * K2 + k2) * K1 + k1;
          if ((k1 == 0 && k2 == 0 && k3 == 0) || is_nyquist(k1, K1) || is_nyquist(k2, K2) ||
              is_nyquist(k3, K3)) {
            grid[g] = 0.0;
            continue;
          }
          int m1 = signed_mode(k1, K1), m2 = signed_mode(k2, K2), m3 = signed_mode(k3, K3);
          double mv[3];
          for (int d = 0; d < 3; ++d)
            mv[d] = m1 * c.rec[0][d] + m2 * c.rec[1][d] + m3 * c.rec[2][d];
          double msq = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
          double wgt = pref * exp(-a2 * msq) / msq * bsq[k1] * bsq[K1 + k2] * bsq[K1 + K2 + k3];
          grid[g] *= wgt;
        }
  }
  st = orc_fft3d(K1, K2, K3, (double *)grid, 1);
  if (st) goto out;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double *w1 = th + ((size_t)i * 3 + 0) * 2 * p;
    const double *w2 = th + ((size_t)i * 3 + 1) * 2 * p;
    const double *w3 = th + ((size_t)i * 3 + 2) * 2 * p;
    double acc = 0.0, du[3] = {0.0, 0.0, 0.0};
    for (int o3 = 0; o3 < p; ++o3) {
      int g3 = (k0[3 * i + 2] + o3) % K3;
      for (int o2 = 0; o2 < p; ++o2) {
        int g2 = (k0[3 * i + 1] + o2) % K2;
        for (int o1 = 0; o1 < p; ++o1) {
          int g1 = (k0[3 * i] + o1) % K1;
          double v = creal(grid[((size_t)g3 * K2 + g2) * K1 + g1]);
          acc += w1[o1] * w2[o2] * w3[o3] * v;
          if (force_rec) {
            du[0] += w1[p + o1] * w2[o2] * w3[o3] * v;
            du[1] += w1[o1] * w2[p + o2] * w3[o3] * v;
            du[2] += w1[o1] * w2[o2] * w3[p + o3] * v;
          }
        }
      }
    }
    pot_rec[i] = acc;
    if (force_rec)
      for (int d = 0; d < 3; ++d)
        force_rec[3 * i + d] =
            -q[i] * (du[0] * K1 * c.rec[0][d] + du[1] * K2 * c.rec[1][d] + du[2] * K3 * c.rec[2][d]);
  }
out:
  free(grid);
  free(th);
  free(k0);
  free(bsq);
  return st;
}

int orc_spme_ewald(int n, const double *pos, const double *q, const double *lat9, const int *K,
                   int p, double alpha, double r_cut, double *pot, double *force) {
  orc_cell c;
  int st = orc_cell_metrics(lat9, &c);
  if (st) return st;
  double *rec = (double *)malloc(sizeof(double) * (size_t)n);
  double *frec = force ? (double *)malloc(sizeof(double) * 3 * (size_t)n) : NULL;
  st = orc_real_space_cell_list(n, pos, q, &c, alpha, r_cut, pot, force);
  if (!st) st = orc_spme_reciprocal(n, pos, q, lat9, K, p, alpha, rec, frec);
  if (!st) {
    for (int i = 0; i < n; ++i) pot[i] += rec[i] + orc_self_potential(q[i], alpha);
    if (force)
      for (size_t k = 0; k < 3 * (size_t)n; ++k) force[k] += frec[k];
  }
  free(rec);
  free(frec);
  return st;
}

/* ------------------------------------------------------------------------ */
/* DFTB short-range gamma                                                    */
/* ------------------------------------------------------------------------ */

/* src/dftbp/common/accuracy.F90:69, 86, 94, 106, 110 */
static const double TOL_SAME_DIST = 1.0e-5;
static const double MIN_HUB_DIFF = 0.3125 * 1.0e-5;
static const double TOL_SHORT_GAMMA = 1.0e-10;

/* src/dftbp/dftb/shortgammafuncs.F90:267-296 */
static double gamma_sub(double r, double t1, double t2) {
  double d = t1 * t1 - t2 * t2;
  return exp(-t1 * r) *
         ((0.5 * t2 * t2 * t2 * t2 * t1 / (d * d)) -
          (t2 * t2 * t2 * t2 * t2 * t2 - 3.0 * t2 * t2 * t2 * t2 * t1 * t1) / (r * d * d * d));
}

/* src/dftbp/dftb/shortgammafuncs.F90:97-148 */
double orc_exp_gamma(double r, double ua, double ub) {
  double ta = 3.2 * ua, tb = 3.2 * ub;
  if (r < TOL_SAME_DIST) {
    if (fabs(ua - ub) < MIN_HUB_DIFF) return -0.5 * (ua + ub);
    return -0.5 * ((ta * tb) / (ta + tb) + (ta * tb) * (ta * tb) / ((ta + tb) * (ta + tb) * (ta + tb)));
  }
  if (fabs(ua - ub) < MIN_HUB_DIFF) {
    double tm = 0.5 * (ta + tb);
    return exp(-tm * r) * (1.0 / r + 0.6875 * tm + 0.1875 * r * tm * tm +
                           0.02083333333333333333 * r * r * tm * tm * tm);
  }
  return gamma_sub(r, ta, tb) + gamma_sub(r, tb, ta);
}

/* src/dftbp/dftb/shortgammafuncs.F90:27-93 */
double orc_exp_gamma_cutoff(double u2, double u1, double min_value) {
  double cv = min_value > 0.0 ? min_value : 1.0e-10;
  double rab = 1.0;
  while (orc_exp_gamma(rab, u2, u1) > cv) rab = 2.0 * rab;
  double minc = rab + 0.1, maxc = 0.5 * rab - 0.1;
  double maxg = orc_exp_gamma(maxc, u2, u1), ming = orc_exp_gamma(minc, u2, u1);
  double lowg = orc_exp_gamma(minc, u2, u1);
  double cut = maxc + 0.1;
  double g = orc_exp_gamma(cut, u2, u1);
  while ((g - lowg) > TOL_SHORT_GAMMA) {
    maxc = 0.5 * (cut + minc);
    if ((maxg >= ming) == (cv >= orc_exp_gamma(maxc, u2, u1))) {
      minc = maxc;
      lowg = orc_exp_gamma(minc, u2, u1);
    } else {
      cut = maxc;
      g = orc_exp_gamma(cut, u2, u1);
    }
  }
  return minc;
}

/* src/dftbp/dftb/shortgamma.F90:753-934 restated as a full image sum: the half neighbour
 * list of src/dftbp/dftb/periodic.F90:565-640 with the symmetric scatter of buildShifts_
 * (918-923) visits every (j, T) once per atom, the on-site term once. */
int orc_short_gamma_shift(int n, const double *pos, const double *q, const double *lat9,
                          const int *species, const double *hubbard_u, int n_species,
                          double *shift) {
  orc_cell c;
  int st = orc_cell_metrics(lat9, &c);
  if (st) return st;
  double *cut = (double *)malloc(sizeof(double) * (size_t)n_species * n_species);
  if (!cut) return ORC_ERR_ALLOC;
  double rmax = 0.0;
  for (int a = 0; a < n_species; ++a)
    for (int b = a; b < n_species; ++b) {
      double v = orc_exp_gamma_cutoff(hubbard_u[b], hubbard_u[a], 0.0);
      cut[a * n_species + b] = cut[b * n_species + a] = v;
      if (v > rmax) rmax = v;
    }
  int nimg[3];
  for (int d = 0; d < 3; ++d) nimg[d] = (int)ceil(rmax * c.reclen[d]) + 1;
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; ++i) {
    double ui = hubbard_u[species[i]];
    double acc = ui * q[i];
    for (int j = 0; j < n; ++j) {
      double uj = hubbard_u[species[j]];
      double rc = cut[species[i] * n_species + species[j]];
      double rc2 = rc * rc + TOL_SAME_DIST * TOL_SAME_DIST;
      for (int nx = -nimg[0]; nx <= nimg[0]; ++nx)
        for (int ny = -nimg[1]; ny <= nimg[1]; ++ny)
          for (int nz = -nimg[2]; nz <= nimg[2]; ++nz) {
            if (i == j && nx == 0 && ny == 0 && nz == 0) continue;
            double dv[3];
            for (int d = 0; d < 3; ++d)
              dv[d] = pos[3 * j + d] - pos[3 * i + d] + nx * c.lat[0][d] + ny * c.lat[1][d] +
                      nz * c.lat[2][d];
            double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            if (r2 > rc2) continue;
            acc -= orc_exp_gamma(sqrt(r2), uj, ui) * q[j];
          }
    }
    shift[i] = acc;
  }
  free(cut);
  return ORC_OK;
}
