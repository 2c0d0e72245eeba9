// engine.hpp -- internal declarations of the MI355X periodic Coulomb engine.
//
// Layout of the resident state of one structure (HBM):
//   atoms are renumbered ("sorted order") by the SPME tile that owns their stencil origin, so
//   that spreading, gathering and the pair sum all walk memory in spatial order;
//   perm[k]        sorted -> caller index
//   k0[k]          int4: first grid point of the B-spline stencil on each axis (+ pad)
//   theta[k]       [3][p] B-spline weights (FP64 or FP32)
//   frac[k]        double3 wrapped fractional coordinates (forces, pair build)
//   tile_start[t]  CSR offsets of atoms per tile
//   row_ptr/col/coef  the real-space pair list: coef = erfc(a r)/r - gamma_short(r), summed over
//                  images within r_list, one entry per (i, j) partner
//   grid / cgrid   K1*K2*K3 real mesh and K3*K2*(K1/2+1) half-complex spectrum
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace ewamd {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kSqrtPi = 1.77245385090551602729816748334114518;

// fe/common/ewald_validation.f90:13-20
constexpr double kMinSeparationSq = 1.0e-24;
constexpr double kNeutralityTol = 1.0e-8;
// Largest charge magnitude the entry points accept: the spreading accumulates in fixed point,
// exactly for contributions below 2^15 (FP32 meshes) and sums below 1.3e8 per mesh point;
// atomic and point charges are many orders of magnitude smaller.
constexpr double kMaxCharge = 1.0e4;

constexpr int kMaxOrder = 16;
constexpr int kMaxRadixPasses = 12;
constexpr int kFftBlock = 256;      // threads per FFT workgroup
constexpr int kFftMaxElems = 2048;  // complex elements per FFT workgroup (LDS: 32 KiB in FP64)
constexpr int kMaxSrcTiles = 4;     // source tiles per axis in the spreading table

struct Cell {
  double lat[3][3];  // rows a_i
  double rec[3][3];  // rows b_i, b_i . a_j = delta_ij (no 2 pi)
  double volume;
  double latlen[3], reclen[3];
};

// fe/common/ewald_geometry.f90:65-93
int cell_metrics(const double* lat9, Cell* c);

struct Params {
  double tol, alpha, r_cut, k_cut, r_list;
  int K[3];
  int order;
  int precision;
};

// Parameter chain (host).  Returns an EW_* status and fills p.
int choose_parameters(const Cell& c, int n, double q2sum, const double* opt_tol,
                      double opt_alpha, double opt_r_cut, double opt_k_cut, const int* opt_grid,
                      int opt_order, int alpha_rule, Params* p, std::string* err);

// src/dftbp/dftb/shortgammafuncs.F90:97-148, 27-93
double exp_gamma(double r, double ua, double ub);
double exp_gamma_cutoff(double u2, double u1);

// smallest 2^a 3^b 5^c >= n
int smooth235(int n);

// |b(m)|^2 B-spline moduli of one axis
void bspline_moduli(int K, int order, double* bsq);

// radix plan for a length (radices 2, 3, 4, 5, 8) in the pass order of `variant` (0..3)
constexpr int kRadixVariants = 5;
int fft_plan(int n, int* radices, int variant = 0);

// Division of a non-negative 31-bit integer by a fixed divisor d >= 1 as a multiply-high and a
// shift: q = (umulhi(x, m) + x) >> s, with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1
// (Granlund and Montgomery, PLDI 1994, round-up variant).
struct FastDiv {
  unsigned m;
  int s;
};
FastDiv make_fastdiv(int d);

struct FftPlan {
  int n;
  int npass;
  int radix[kMaxRadixPasses];
  int lines;   // lines per workgroup
  int pitch;   // LDS pitch per line (complex elements)
  FastDiv nd;  // / n
  FastDiv hd;  // / (n / 2 + 1)
  FastDiv ld;  // / lines
  FastDiv nbd[kMaxRadixPasses];  // / (n / radix) of each pass
  FastDiv nsd[kMaxRadixPasses];  // / (product of the earlier radices) of each pass
};

}  // namespace ewamd
