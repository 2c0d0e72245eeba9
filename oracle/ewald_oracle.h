/*
 * ewald_oracle.h -- CPU restatement of the reference's periodic Coulomb path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it.  The product library never links, loads or calls it.
 *
 * Conventions (atomic units, as in the reference):
 *   positions  [n][3]  Cartesian, row i = atom i   (== DFTB+ coords(3,N))
 *   lattice    [3][3]  row i = lattice vector a_i  (== DFTB+ latVecs(3,3))
 *   potential  pot[i] = phi(x_i), force[i][3] = -dU/dx_i (physical force)
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root, src/dftbp/extlibs/fftewald abbreviated "fe/").
 */
#ifndef EWALD_ORACLE_H
#define EWALD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference aborts with `error stop`; the restatement
 * returns a code instead so that the tests can probe each guard. */
enum {
  ORC_OK = 0,
  ORC_ERR_N = 1,            /* nParticle <= 0                                  */
  ORC_ERR_NOT_NEUTRAL = 2,  /* |sum q| > 1e-8 sum |q|  (fe/common/ewald_validation.f90:41-46) */
  ORC_ERR_VOLUME = 3,       /* cell volume <= 1e-18    (fe/common/ewald_geometry.f90:78) */
  ORC_ERR_TOL = 4,          /* tol outside (0,1)       (fe/direct/ewald_direct_3d.f90:103-104) */
  ORC_ERR_COINCIDENT = 5,   /* r^2 < 1e-24             (fe/direct/ewald_direct_3d.f90:274-275) */
  ORC_ERR_PARAM = 6,        /* non-positive alpha / cutoff / grid */
  ORC_ERR_ALLOC = 7,
  ORC_ERR_ALPHA_SEARCH = 8  /* getOptimalAlphaEwald failure (coulomb.F90:2172-2177) */
};

typedef struct {
  double lat[3][3];   /* rows a_i */
  double rec[3][3];   /* rows b_i, b_i . a_j = delta_ij, no 2 pi */
  double volume;
  double latlen[3];   /* |a_i| */
  double reclen[3];   /* |b_i| */
} orc_cell;

/* fe/common/ewald_geometry.f90:65-93 */
int orc_cell_metrics(const double *lat9, orc_cell *cell);
/* fe/common/ewald_validation.f90:26-48 (neutrality part; shapes are implicit in C) */
int orc_check_configuration(int n, const double *q);
/* fe/common/ewald_self.f90:12-19 */
double orc_self_potential(double q, double alpha);
/* fe/common/ewald_truncation.f90:36-45 */
void orc_direct_cutoffs(double alpha, double tol, double *r_cut, double *k_cut);
/* fe/common/ewald_lambert_w.f90:27-56 */
double orc_lambert_w0(double x);
/* fe/common/ewald_truncation.f90:53-78 */
double orc_splitting_from_real_budget(double r_cut, double q2sum, int n, double pseudo_volume,
                                      double budget);
/* fe/common/ewald_truncation.f90:84-104 */
double orc_cutoff_from_fourier_budget(double q2sum, int n, double box_edge, double alpha,
                                      double budget);
/* fe/common/ewald_alpha_search.f90:83-181; returns the search status (0 = converged) */
int orc_balanced_alpha(double minG, double minR, double volume, double tol, double *alpha);
/* fe/direct/ewald_direct_3d.f90:319-347 (with the volume-fit fallback) */
double orc_balanced_splitting(const orc_cell *cell, double tol);
/* src/dftbp/dftb/coulomb.F90:2097-2179 with diffRecReal/gTerm/rTerm 2686-2748 and
 * nSearchIter = 30 (src/dftbp/common/accuracy.F90:114).  lat9 rows a_i. */
int orc_optimal_alpha_dftb(const double *lat9, double tol, double *alpha);

/* fe/direct/ewald_direct_3d.f90:192-309: explicit image sum, overwrites pot/force. */
int orc_real_space_images(int n, const double *pos, const double *q, const orc_cell *cell,
                          double alpha, double r_cut, double *pot, double *force);
/* fe/fft/ewald_fft_3d_real.f90:32-187 + fe/common/ewald_cell_list.f90:56-141:
 * linked-cell path when >= 3 boxes per axis, image sum otherwise. */
int orc_real_space_cell_list(int n, const double *pos, const double *q, const orc_cell *cell,
                             double alpha, double r_cut, double *pot, double *force);

/* fe/direct/ewald_direct_3d.f90:37-182.  Pass alpha_in/r_cut_in/k_cut_in <= 0 for "absent".
 * force may be NULL.  real_space selects the real-space summation: 0 = image sum exactly as
 * the reference, 1 = linked-cell (same terms, O(N) pair search) for large cells.
 * phase_tables = 1 evaluates exp(i k.x) from per-axis phase powers instead of one cos/sin
 * per (mode, atom); the terms are identical, round-off differs at the 1e-15 level. */
int orc_direct_ewald(int n, const double *pos, const double *q, const double *lat9, double tol,
                     double alpha_in, double r_cut_in, double k_cut_in, int real_space,
                     int phase_tables, double *pot, double *force, double *alpha_used);

/* ---------- the reference fast path (CPU baseline) ---------- */
/* fe/nfft/fft_backend.f90:73-91 */
int orc_next_fast_length(int n);
/* fe/nfft/nfft.f90:99-109 */
int orc_window_cutoff_from_budget(double budget);
typedef struct {
  double alpha, r_cut, k_cut;
  int window_cutoff;
  int nmodes[3];
} orc_fft_params;
/* fe/fft/ewald_fft_3d_parameters.f90:52-168.  Overrides: <= 0 means absent. */
int orc_choose_parameters(const orc_cell *cell, double q2sum, int n, double tol, double alpha_in,
                          double r_cut_in, double k_cut_in, int nmodes_in, int window_in,
                          orc_fft_params *out);
/* fe/fft/ewald_fft_3d_parameters.f90:177-229, in gigabytes (1024^3 bytes) */
double orc_grid_memory_gigabytes(const orc_cell *cell, double q2sum, int n, double tol,
                                 double alpha_in, double r_cut_in, double k_cut_in, int nmodes_in,
                                 int window_in);
/* Unnormalised in-place 3D complex DFT of a column-major (first index fastest) grid,
 * sign +1 -> exp(+2 pi i ...), the convention of fe/nfft/fft_backend.f90:6-15, 43-58.
 * data holds interleaved (re, im) doubles.  Own mixed-radix (2,3,4,5,7 + generic) code. */
int orc_fft3d(int n1, int n2, int n3, double *data, int sign);
/* fe/nfft/nfft.f90:196-305 / 313-404.  Shat/fhat: M1*M2*M3 complex, first index fastest. */
int orc_nfft_adjoint_3d(int n, const double *t1, const double *t2, const double *t3,
                        const double *q, int M1, int M2, int M3, int m, double *shat);
int orc_nfft_forward_3d(int n, const double *t1, const double *t2, const double *t3, int M1,
                        int M2, int M3, int m, const double *fhat, double *f);
/* fe/fft/ewald_fft_3d.f90:36-134 (real + Fourier + self).  force may be NULL. */
int orc_fft_ewald(int n, const double *pos, const double *q, const double *lat9, double tol,
                  double *pot, double *force, orc_fft_params *params_used);

/* ---------- smooth particle-mesh Ewald (the method the GPU path runs) ---------- */
/* Essmann et al., J. Chem. Phys. 103, 8577 (1995): cardinal B-splines of order p on a
 * K1 x K2 x K3 grid; influence B(m) exp(-pi^2 m^2/alpha^2)/(pi V m^2), m = 0 and the
 * Nyquist planes of even K excluded.  pot_rec gets the reciprocal potential only.
 * force_rec may be NULL. */
int orc_spme_reciprocal(int n, const double *pos, const double *q, const double *lat9,
                        const int *K, int order, double alpha, double *pot_rec,
                        double *force_rec);
/* Full SPME Ewald potential: real space (cell list / image sum) + reciprocal + self. */
int orc_spme_ewald(int n, const double *pos, const double *q, const double *lat9, const int *K,
                   int order, double alpha, double r_cut, double *pot, double *force);
/* |b(m)|^2 moduli of one axis (length K), as used by orc_spme_reciprocal. */
int orc_bspline_moduli(int K, int order, double *bsq);
/* B-spline weights theta[o] = M_p(u - (floor(u) - p + 1 + o)), o = 0..p-1. */
void orc_bspline_weights(double u, int order, double *theta, double *dtheta);

/* ---------- DFTB short-range gamma ---------- */
/* src/dftbp/dftb/shortgammafuncs.F90:97-148 (+ gammaSubExprn_ 267-296) */
double orc_exp_gamma(double r, double ua, double ub);
/* src/dftbp/dftb/shortgammafuncs.F90:27-93 with minShortGamma = tolShortGamma = 1e-10 */
double orc_exp_gamma_cutoff(double u2, double u1, double min_value);
/* shortgamma.F90:753-934 for one U per atom (atom-resolved SCC):
 *   shift_i = U_i q_i - sum_{(j,T) != (i,0), r <= cut(U_i,U_j)} expGamma(r, U_j, U_i) q_j
 * species[i] in 0..n_species-1 indexes hubbard_u; cutoffs are taken per species pair as
 * expGammaCutoff(U_hi, U_lo) with species order as in setupShortGammaCutoffs_. */
int orc_short_gamma_shift(int n, const double *pos, const double *q, const double *lat9,
                          const int *species, const double *hubbard_u, int n_species,
                          double *shift);

#ifdef __cplusplus
}
#endif
#endif
