// kernels.hip -- gfx950 device kernels of the periodic Coulomb engine.
//
// Per SCC iteration (charges change, geometry resident):
//   k_gather_charges     dq (caller order) -> q (sorted order)
//   k_spread_region      B-spline spreading into privatised block copies (FP64 LDS accumulation)
//   k_spread_assemble    block copies -> mesh, fixed order per point
//   k_fft_xy_r2c         x-y plane of the mesh -> half spectrum (or k_fft_x_r2c + k_fft_lines<0>)
//   k_fft_lines<2>       z pass fused with the influence multiply and the inverse z pass
//   k_fft_xy_c2r         half spectrum -> convolved mesh plane (or k_fft_lines<1> + k_fft_x_c2r)
//   k_pair_spmv_stream   real-space pair sum over the resident coefficients (auxiliary stream)
//   k_gather_tile        B-spline gather + real space + on-site + self -> shift, energy partials
//   k_energy_final       fixed-order reduction of the energy partials
// Several vectors / structures: the kg_* kernels (same bodies, a block prefix per kind) and
// k_pair_spmm.  Per geometry: the ordering of order.hip, k_bspline (weights), k_pair_tiles
// (linked tiles), k_pair_dense,
// k_compress_cols, k_half_pairs.  Stress and forces: k_recip_virial, k_force_recip, and the
// real-space parts over the pair list (k_pair_real_terms) or the tiles (k_real_virial,
// k_force_real).  External point charges: k_ext_real (with the spreading, FFTs
// and gather above run on the sources and the resident atoms).
//
// The generic FFTs are Stockham autosort transforms of one or more lines staged in LDS: every
// pass loads its butterflies into registers, synchronises, and stores them back in place; the
// register line transforms of wfft.hip replace them for the axis lengths they cover.
#include <hip/hip_runtime.h>

#include "dft.hpp"
#include "kernels.hpp"

namespace ewamd {

// One in-place Stockham pass over `nl` lines of length n (LDS pitch P).
// tw[m] = exp(-2 pi i m / n); the inverse uses its conjugate.
// LDS slot of element i of a line: one pad slot after every eight elements, so that the
// stride-radix stores of the first pass and the strided column reads fall in distinct banks.
__device__ __forceinline__ int line_slot(int i) { return i + (i >> 3); }


// Coordinates (x, y, z) of the flat index i0, i0 + S, i0 + 2S, ... of an E1 x E2 x E3 box (x
// fastest), without a division per step: S is split once into box coordinates and each step
// adds them with carries (x + sx < 2 E1 and y + sy + 1 < 2 E2, so one subtraction each).
struct BoxWalk {
  int x, y, z, sx, sy, sz, E1, E2;
  __device__ __forceinline__ BoxWalk(int i0, int stride, int e1, int e2) : E1(e1), E2(e2) {
    const int r = i0 / e1;
    x = i0 - r * e1;
    z = r / e2;
    y = r - z * e2;
    const int rs = stride / e1;
    sx = stride - rs * e1;
    sz = rs / e2;
    sy = rs - sz * e2;
  }
  __device__ __forceinline__ void next() {
    x += sx;
    int c = x >= E1;
    x -= c ? E1 : 0;
    y += sy + c;
    c = y >= E2;
    y -= c ? E2 : 0;
    z += sz + c;
  }
};

// tw: this pass's table, exp(-2 pi i t k / (Ns RAD)) at [(t - 1) * Ns + k].  BLOCK threads
// cover nl lines of n elements with at most EPT elements per thread (nl * n <= BLOCK * EPT).
template <int RAD, typename R, int BLOCK, int EPT>
__device__ __forceinline__ void stockham_pass(cx<R>* s, int n, int nl, int P, int Ns,
                                              const cx<R>* tw, R dir, FastDiv nbd,
                                              FastDiv nsd) {
  constexpr int MAXB = (EPT + RAD - 1) / RAD;
  const int nb = n / RAD;
  const int total = nl * nb;
  cx<R> v[MAXB][RAD];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = threadIdx.x + b * BLOCK;
    if (idx < total) {
      const int line = fdiv(idx, nbd), j = idx - line * nb;
      const int k = j - fdiv(j, nsd) * Ns;
      const cx<R>* src = s + line * P;
#pragma unroll
      for (int t = 0; t < RAD; ++t) {
        cx<R> a = src[line_slot(j + t * nb)];
        if (t > 0) {
          cx<R> w = tw[(t - 1) * Ns + k];
          if (dir > R(0)) w.y = -w.y;  // table holds exp(-i..); the inverse conjugates it
          a = cmul(a, w);
        }
        v[b][t] = a;
      }
      Dft<RAD, R>::run(v[b], dir);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = threadIdx.x + b * BLOCK;
    if (idx < total) {
      const int line = fdiv(idx, nbd), j = idx - line * nb;
      const int jq = fdiv(j, nsd), k = j - jq * Ns;
      cx<R>* dst = s + line * P;
      const int d0 = jq * Ns * RAD + k;
#pragma unroll
      for (int t = 0; t < RAD; ++t) dst[line_slot(d0 + t * Ns)] = v[b][t];
    }
  }
  __syncthreads();
}

// Copies the length-n twiddle table into LDS (after the line data, at s + nl * pitch).
template <typename R>
__device__ __forceinline__ const cx<R>* stage_twiddles(cx<R>* s, const FftPlan& pl, int nl,
                                                       const cx<R>* __restrict__ tw) {
  cx<R>* t = s + (size_t)nl * pl.pitch;
  for (int i = threadIdx.x; i < pl.n; i += blockDim.x) t[i] = tw[i];
  return t;
}

template <typename R, int BLOCK = kFftBlock, int EPT = kFftMaxElems / kFftBlock>
__device__ void fft_lds(cx<R>* s, const FftPlan& pl, int nl, const cx<R>* tw, R dir) {
  int Ns = 1;
  const cx<R>* t = tw;
  for (int i = 0; i < pl.npass; ++i) {
    const int r = pl.radix[i];
    if (r == 16)
      stockham_pass<16, R, BLOCK, EPT>(s, pl.n, nl, pl.pitch, Ns, t, dir, pl.nbd[i],
                                           pl.nsd[i]);
    else if (r == 8)
      stockham_pass<8, R, BLOCK, EPT>(s, pl.n, nl, pl.pitch, Ns, t, dir, pl.nbd[i],
                                          pl.nsd[i]);
    else if (r == 4)
      stockham_pass<4, R, BLOCK, EPT>(s, pl.n, nl, pl.pitch, Ns, t, dir, pl.nbd[i],
                                          pl.nsd[i]);
    else if (r == 2)
      stockham_pass<2, R, BLOCK, EPT>(s, pl.n, nl, pl.pitch, Ns, t, dir, pl.nbd[i],
                                          pl.nsd[i]);
    else if (r == 3)
      stockham_pass<3, R, BLOCK, EPT>(s, pl.n, nl, pl.pitch, Ns, t, dir, pl.nbd[i],
                                          pl.nsd[i]);
    else
      stockham_pass<5, R, BLOCK, EPT>(s, pl.n, nl, pl.pitch, Ns, t, dir, pl.nbd[i],
                                          pl.nsd[i]);
    t += (r - 1) * Ns;
    Ns *= r;
  }
}

// ----------------------------------------------------------------------------------------
// FFT kernels
// ----------------------------------------------------------------------------------------

// ----------------------------------------------------------------------------------------
// real-space pair sum rows (shared by the pair-sum kernel and the chain kernels that co-run it)
// ----------------------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Partner index of a stored pair: absolute (int) or a signed 16-bit offset from the row
// (cyclic in n), which halves the index stream when rows are in spatial order.
__device__ __forceinline__ int partner(int row, int c, int) { return c; }
__device__ __forceinline__ int partner(int row, short c, int n) {
  int j = row + int(c);
  j += j < 0 ? n : 0;
  j -= j >= n ? n : 0;
  return j;
}

// Real-space pair sum out[i] = sum_j coef_ij q_j over rows r0 + wave, r0 + wave + nw, ...
// below r1: one 64-lane wave per row.  A sweep covers 512 entries: lane l takes entries
// l + 64u (u < 8), so that every load instruction of the wave reads consecutive entries and
// consecutive partners, and all eight loads of a lane are issued before the first is used.
// While a wave gathers one row's partner charges it already loads the next row's bounds,
// offsets and coefficients, so a row costs one memory latency instead of three.
template <typename CT, typename IT>
__device__ __forceinline__ void pair_stream_rows(int n, int r0, int r1, int wave, int nw,
                                                 const long long* __restrict__ row_ptr,
                                                 const IT* __restrict__ col,
                                                 const CT* __restrict__ coef,
                                                 const double* __restrict__ q,
                                                 double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  int row = __builtin_amdgcn_readfirstlane(r0 + wave);
  if (row >= r1) return;  // uniform across the wave
  long long a = row_ptr[row], b = row_ptr[row + 1];
  IT c[8];
  CT w[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const long long k = a + lane + 64 * u;
    c[u] = k < b ? col[k] : IT(0);
    w[u] = k < b ? coef[k] : CT(0);
  }
  while (true) {
    const int nrow = row + nw;
    const bool more = nrow < r1;
    // this row's partners' charges
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = a + lane + 64 * u < b ? q[partner(row, c[u], n)] : 0.0;
    // the next row's bounds and first sweep, in flight during this row's sums
    long long na = 0, nb = 0;
    IT nc[8];
    CT nwt[8];
    if (more) {
      na = row_ptr[nrow];
      nb = row_ptr[nrow + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const long long k = na + lane + 64 * u;
      nc[u] = k < nb ? col[k] : IT(0);
      nwt[u] = k < nb ? coef[k] : CT(0);
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += double(w[u]) * x[u];
    // rows longer than one sweep (dense image sums, long short-range lists)
    for (long long g = a + lane + 512; g < b; g += 512) {
      IT cc[8];
      CT ww[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const long long k = g + 64 * u;
        cc[u] = k < b ? col[k] : IT(0);
        ww[u] = k < b ? coef[k] : CT(0);
      }
      double qv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) qv[u] = g + 64 * u < b ? q[partner(row, cc[u], n)] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += double(ww[u]) * qv[u];
    }
    acc = wave_sum(acc);
    if (lane == 0) out[row] = acc;
    if (!more) break;
    row = nrow;
    a = na;
    b = nb;
    for (int u = 0; u < 8; ++u) {
      c[u] = nc[u];
      w[u] = nwt[u];
    }
  }
}


// Blocks past co.first of a chain kernel launch run rows [co.r0, co.r1) of the pair sum
// (returns true for them; the chain blocks return false and carry on).
template <typename CT>
__device__ __forceinline__ bool co_pairs(const PairCo& co) {
  if ((int)blockIdx.x < co.first) return false;
  const int wpb = blockDim.x >> 6;
  const int wave = ((int)blockIdx.x - co.first) * wpb + (int)(threadIdx.x >> 6);
  const int nw = co.nblk * wpb;
  if (co.col16)
    pair_stream_rows<CT, short>(co.n, co.r0, co.r1, wave, nw, co.row_ptr,
                                static_cast<const short*>(co.col),
                                static_cast<const CT*>(co.coef), co.q, co.out);
  else
    pair_stream_rows<CT, int>(co.n, co.r0, co.r1, wave, nw, co.row_ptr,
                              static_cast<const int*>(co.col), static_cast<const CT*>(co.coef),
                              co.q, co.out);
  return true;
}

// Rows of K1 reals (row = z*K2 + y); workgroup takes pl.lines complex lines = 2*lines rows.
template <typename R>
__device__ __forceinline__ void k_fft_x_r2c_body(const R* __restrict__ grid,
                                                         cx<R>* __restrict__ cg, int nrows,
                                                         int Kh, const FftPlan& pl,
                                                         const cx<R>* __restrict__ tw, int bid) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int K1 = pl.n, nl = pl.lines, P = pl.pitch;
  const int row0 = bid * nl * 2;
  const cx<R>* tws = stage_twiddles(s, pl, nl, tw);
  {
    // all of a thread's loads are issued before any LDS store (one memory latency, not EPT)
    constexpr int EPT = kFftMaxElems / kFftBlock;
    cx<R> z[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int idx = threadIdx.x + k * kFftBlock;
      const int l = fdiv(idx, pl.nd), e = idx - l * K1;
      const int r0 = row0 + 2 * l, r1 = r0 + 1;
      const bool in = idx < nl * K1;
      z[k].x = in && r0 < nrows ? grid[(size_t)r0 * K1 + e] : R(0);
      z[k].y = in && r1 < nrows ? grid[(size_t)r1 * K1 + e] : R(0);
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int idx = threadIdx.x + k * kFftBlock;
      if (idx < nl * K1) {
        const int l = fdiv(idx, pl.nd), e = idx - l * K1;
        s[l * P + line_slot(e)] = z[k];
      }
    }
  }
  __syncthreads();
  fft_lds<R>(s, pl, nl, tws, R(-1));
  for (int idx = threadIdx.x; idx < nl * Kh; idx += kFftBlock) {
    const int l = fdiv(idx, pl.hd), k = idx - l * Kh;
    const int r0 = row0 + 2 * l, r1 = r0 + 1;
    const cx<R> z = s[l * P + line_slot(k)];
    const cx<R> zm = s[l * P + line_slot(k == 0 ? 0 : K1 - k)];
    if (r0 < nrows) cg[(size_t)r0 * Kh + k] = {R(0.5) * (z.x + zm.x), R(0.5) * (z.y - zm.y)};
    if (r1 < nrows) cg[(size_t)r1 * Kh + k] = {R(0.5) * (z.y + zm.y), R(0.5) * (zm.x - z.x)};
  }
}

template <typename R, bool CO>
__global__ __launch_bounds__(kFftBlock) void k_fft_x_r2c(const R* __restrict__ grid,
                                                         cx<R>* __restrict__ cg, int nrows,
                                                         int Kh, FftPlan pl,
                                                         const cx<R>* __restrict__ tw,
                                                         PairCo co) {
  if constexpr (CO) {
    if (co_pairs<R>(co)) return;
  }
  k_fft_x_r2c_body<R>(grid, cg, nrows, Kh, pl, tw, blockIdx.x);
}

template <typename R>
__device__ __forceinline__ void k_fft_x_c2r_body(const cx<R>* __restrict__ cg,
                                                         R* __restrict__ grid, int nrows, int Kh,
                                                         const FftPlan& pl,
                                                         const cx<R>* __restrict__ tw, int bid) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int K1 = pl.n, nl = pl.lines, P = pl.pitch;
  const int row0 = bid * nl * 2;
  const cx<R>* tws = stage_twiddles(s, pl, nl, tw);
  {
    constexpr int EPT = kFftMaxElems / kFftBlock;
    cx<R> a[EPT], b[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int idx = threadIdx.x + j * kFftBlock;
      const int l = fdiv(idx, pl.nd), k = idx - l * K1;
      const int r0 = row0 + 2 * l, r1 = r0 + 1;
      const bool in = idx < nl * K1;
      const int kk = k < Kh ? k : K1 - k;
      a[j] = {R(0), R(0)};
      b[j] = {R(0), R(0)};
      if (in && r0 < nrows) a[j] = cg[(size_t)r0 * Kh + kk];
      if (in && r1 < nrows) b[j] = cg[(size_t)r1 * Kh + kk];
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int idx = threadIdx.x + j * kFftBlock;
      if (idx < nl * K1) {
        const int l = fdiv(idx, pl.nd), k = idx - l * K1;
        cx<R> aa = a[j], bb = b[j];
        if (k >= Kh) {
          aa.y = -aa.y;
          bb.y = -bb.y;
        }
        s[l * P + line_slot(k)] = {aa.x - bb.y, aa.y + bb.x};
      }
    }
  }
  __syncthreads();
  fft_lds<R>(s, pl, nl, tws, R(1));
  for (int idx = threadIdx.x; idx < nl * K1; idx += kFftBlock) {
    const int l = fdiv(idx, pl.nd), e = idx - l * K1;
    const int r0 = row0 + 2 * l, r1 = r0 + 1;
    const cx<R> z = s[l * P + line_slot(e)];
    if (r0 < nrows) grid[(size_t)r0 * K1 + e] = z.x;
    if (r1 < nrows) grid[(size_t)r1 * K1 + e] = z.y;
  }
}

template <typename R, bool CO>
__global__ __launch_bounds__(kFftBlock) void k_fft_x_c2r(const cx<R>* __restrict__ cg,
                                                         R* __restrict__ grid, int nrows, int Kh,
                                                         FftPlan pl,
                                                         const cx<R>* __restrict__ tw,
                                                         PairCo co) {
  if constexpr (CO) {
    if (co_pairs<R>(co)) return;
  }
  k_fft_x_c2r_body<R>(cg, grid, nrows, Kh, pl, tw, blockIdx.x);
}

// Strided complex lines: line L has base (L / A) * B + L % A and element stride `es`.
// MODE 0: forward; 1: inverse; 2: forward, influence multiply, inverse (Z pass).
template <typename R, int MODE>
__device__ __forceinline__ void k_fft_lines_body(cx<R>* __restrict__ cg, int nlines,
                                                         int A, FastDiv ad, long long B,
                                                         long long es,
                                                         const FftPlan& pl,
                                                         const cx<R>* __restrict__ tw,
                                                         InfluenceArgs inf, int bid) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int n = pl.n, nl = pl.lines, P = pl.pitch;
  const int L0 = bid * nl;
  const cx<R>* tws = stage_twiddles(s, pl, nl, tw);
  {
    constexpr int EPT = kFftMaxElems / kFftBlock;
    cx<R> z[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int idx = threadIdx.x + k * kFftBlock;
      const int e = fdiv(idx, pl.ld), l = idx - e * nl;
      const int L = L0 + l;
      z[k] = {R(0), R(0)};
      if (idx < nl * n && L < nlines) {
        const int La = fdiv(L, ad);
        z[k] = cg[(long long)La * B + (L - La * A) + e * es];
      }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int idx = threadIdx.x + k * kFftBlock;
      if (idx < nl * n) {
        const int e = fdiv(idx, pl.ld), l = idx - e * nl;
        s[l * P + line_slot(e)] = z[k];
      }
    }
  }
  __syncthreads();
  fft_lds<R>(s, pl, nl, tws, MODE == 1 ? R(1) : R(-1));
  if (MODE == 2) {
    // influence function on the half spectrum: line L = ky*Kh + kx, element e = kz
    for (int idx = threadIdx.x; idx < nl * n; idx += kFftBlock) {
      const int l = fdiv(idx, pl.nd), kz = idx - l * n;
      const int L = L0 + l;
      if (L >= nlines) continue;
      const int ky = fdiv(L, inf.khd), kx = L - ky * inf.Kh;
      double w = 0.0;
      const bool nyq = (!(inf.K[0] & 1) && kx == inf.K[0] / 2) ||
                       (!(inf.K[1] & 1) && ky == inf.K[1] / 2) ||
                       (!(inf.K[2] & 1) && kz == inf.K[2] / 2);
      if (!nyq && (kx | ky | kz)) {
        const int m1 = kx;
        const int m2 = ky <= (inf.K[1] - 1) / 2 ? ky : ky - inf.K[1];
        const int m3 = kz <= (inf.K[2] - 1) / 2 ? kz : kz - inf.K[2];
        const double vx = m1 * inf.rec[0] + m2 * inf.rec[3] + m3 * inf.rec[6];
        const double vy = m1 * inf.rec[1] + m2 * inf.rec[4] + m3 * inf.rec[7];
        const double vz = m1 * inf.rec[2] + m2 * inf.rec[5] + m3 * inf.rec[8];
        const double msq = vx * vx + vy * vy + vz * vz;
        w = inf.pref * exp(-inf.a2 * msq) / msq * inf.bsq1[kx] * inf.bsq2[ky] * inf.bsq3[kz];
      }
      cx<R>& z = s[l * P + line_slot(kz)];
      z.x = R(z.x * w);
      z.y = R(z.y * w);
    }
    __syncthreads();
    fft_lds<R>(s, pl, nl, tws, R(1));
  }
  for (int idx = threadIdx.x; idx < nl * n; idx += kFftBlock) {
    const int e = fdiv(idx, pl.ld), l = idx - e * nl;
    const int L = L0 + l;
    if (L < nlines) {
      const int La = fdiv(L, ad);
      cg[(long long)La * B + (L - La * A) + e * es] = s[l * P + line_slot(e)];
    }
  }
}

template <typename R, int MODE, bool CO>
__global__ __launch_bounds__(kFftBlock) void k_fft_lines(cx<R>* __restrict__ cg, int nlines,
                                                         int A, FastDiv ad, long long B,
                                                         long long es,
                                                         FftPlan pl,
                                                         const cx<R>* __restrict__ tw,
                                                         InfluenceArgs inf, PairCo co) {
  if constexpr (CO) {
    if (co_pairs<R>(co)) return;
  }
  k_fft_lines_body<R, MODE>(cg, nlines, A, ad, B, es, pl, tw, inf, blockIdx.x);
}

// One x-y plane per workgroup (z = bid): the K1 real rows are packed two per complex line
// (row 2l real, row 2l+1 imaginary) and transformed along x, split into the half spectra of
// the two rows, re-laid out as Kh columns (one per kx) and transformed along y; the plane of the
// half spectrum is written as cg[(z * K1 + ky) * Kh + kx].  LDS: the x layout (line l at
// l * px.pitch) and the y layout (column kx at kx * py.pitch) share one buffer, followed by the
// x and y twiddle tables.
template <typename R>
__device__ __forceinline__ void k_fft_xy_r2c_body(const R* __restrict__ grid,
                                                  cx<R>* __restrict__ cg, const FftPlan& px,
                                                  const FftPlan& py, const cx<R>* __restrict__ twx,
                                                  const cx<R>* __restrict__ twy, int bid) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int K0 = px.n, K1 = py.n, Kh = K0 / 2 + 1, nl = (K1 + 1) / 2;
  const int Px = px.pitch, Py = py.pitch;
  const int span = nl * Px > Kh * Py ? nl * Px : Kh * Py;
  cx<R>* tx = s + span;
  cx<R>* ty = tx + K0;
  for (int i = threadIdx.x; i < K0; i += kPlaneBlock) tx[i] = twx[i];
  for (int i = threadIdx.x; i < K1; i += kPlaneBlock) ty[i] = twy[i];
  const R* plane = grid + (size_t)bid * K1 * K0;
  {
    cx<R> z[kPlaneEpt];
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      const int l = fdiv(idx, px.nd), x = idx - l * K0;
      const bool in = idx < nl * K0;
      z[k].x = in ? plane[(size_t)(2 * l) * K0 + x] : R(0);
      z[k].y = in && 2 * l + 1 < K1 ? plane[(size_t)(2 * l + 1) * K0 + x] : R(0);
    }
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      if (idx < nl * K0) {
        const int l = fdiv(idx, px.nd), x = idx - l * K0;
        s[l * Px + line_slot(x)] = z[k];
      }
    }
  }
  __syncthreads();
  fft_lds<R, kPlaneBlock, kPlaneEpt>(s, px, nl, tx, R(-1));
  {
    // split the packed lines: row r, mode kx  ->  column kx, element r
    cx<R> o[kPlaneEpt];
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      if (idx < K1 * Kh) {
        const int r = fdiv(idx, px.hd), kx = idx - r * Kh;
        const cx<R>* line = s + (r >> 1) * Px;
        const cx<R> z = line[line_slot(kx)];
        const cx<R> zm = line[line_slot(kx == 0 ? 0 : K0 - kx)];
        o[k] = (r & 1) ? cx<R>{R(0.5) * (z.y + zm.y), R(0.5) * (zm.x - z.x)}
                       : cx<R>{R(0.5) * (z.x + zm.x), R(0.5) * (z.y - zm.y)};
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      if (idx < K1 * Kh) {
        const int r = fdiv(idx, px.hd), kx = idx - r * Kh;
        s[kx * Py + line_slot(r)] = o[k];
      }
    }
  }
  __syncthreads();
  fft_lds<R, kPlaneBlock, kPlaneEpt>(s, py, Kh, ty, R(-1));
  cx<R>* out = cg + (size_t)bid * K1 * Kh;
  for (int idx = threadIdx.x; idx < K1 * Kh; idx += kPlaneBlock) {
    const int ky = fdiv(idx, px.hd), kx = idx - ky * Kh;
    out[idx] = s[kx * Py + line_slot(ky)];
  }
}

// Inverse of k_fft_xy_r2c_body: plane z = bid of the half spectrum -> K1 real rows.
template <typename R>
__device__ __forceinline__ void k_fft_xy_c2r_body(const cx<R>* __restrict__ cg,
                                                  R* __restrict__ grid, const FftPlan& px, const FftPlan& py,
                                                  const cx<R>* __restrict__ twx,
                                                  const cx<R>* __restrict__ twy, int bid) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int K0 = px.n, K1 = py.n, Kh = K0 / 2 + 1, nl = (K1 + 1) / 2;
  const int Px = px.pitch, Py = py.pitch;
  const int span = nl * Px > Kh * Py ? nl * Px : Kh * Py;
  cx<R>* tx = s + span;
  cx<R>* ty = tx + K0;
  for (int i = threadIdx.x; i < K0; i += kPlaneBlock) tx[i] = twx[i];
  for (int i = threadIdx.x; i < K1; i += kPlaneBlock) ty[i] = twy[i];
  const cx<R>* in = cg + (size_t)bid * K1 * Kh;
  {
    cx<R> z[kPlaneEpt];
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      z[k] = idx < K1 * Kh ? in[idx] : cx<R>{R(0), R(0)};
    }
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      if (idx < K1 * Kh) {
        const int ky = fdiv(idx, px.hd), kx = idx - ky * Kh;
        s[kx * Py + line_slot(ky)] = z[k];
      }
    }
  }
  __syncthreads();
  fft_lds<R, kPlaneBlock, kPlaneEpt>(s, py, Kh, ty, R(1));
  {
    // pack rows 2l and 2l+1 of each mode into complex line l (Hermitian extension for kx >= Kh)
    cx<R> o[kPlaneEpt];
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      if (idx < nl * K0) {
        const int l = fdiv(idx, px.nd), x = idx - l * K0;
        const int kk = x < Kh ? x : K0 - x;
        cx<R> a = s[kk * Py + line_slot(2 * l)];
        cx<R> b = 2 * l + 1 < K1 ? s[kk * Py + line_slot(2 * l + 1)] : cx<R>{R(0), R(0)};
        if (x >= Kh) {
          a.y = -a.y;
          b.y = -b.y;
        }
        o[k] = {a.x - b.y, a.y + b.x};
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPlaneEpt; ++k) {
      const int idx = threadIdx.x + k * kPlaneBlock;
      if (idx < nl * K0) {
        const int l = fdiv(idx, px.nd), x = idx - l * K0;
        s[l * Px + line_slot(x)] = o[k];
      }
    }
  }
  __syncthreads();
  fft_lds<R, kPlaneBlock, kPlaneEpt>(s, px, nl, tx, R(1));
  R* plane = grid + (size_t)bid * K1 * K0;
  for (int idx = threadIdx.x; idx < K1 * K0; idx += kPlaneBlock) {
    const int r = fdiv(idx, px.nd), x = idx - r * K0;
    const cx<R> z = s[(r >> 1) * Px + line_slot(x)];
    plane[idx] = (r & 1) ? z.y : z.x;
  }
}

template <typename R, bool CO>
__global__ __launch_bounds__(kPlaneBlock) void k_fft_xy_r2c(const R* __restrict__ grid,
                                                            cx<R>* __restrict__ cg, FftPlan px,
                                                            FftPlan py,
                                                            const cx<R>* __restrict__ twx,
                                                            const cx<R>* __restrict__ twy,
                                                            PairCo co) {
  if constexpr (CO) {
    if (co_pairs<R>(co)) return;
  }
  k_fft_xy_r2c_body<R>(grid, cg, px, py, twx, twy, blockIdx.x);
}

template <typename R, bool CO>
__global__ __launch_bounds__(kPlaneBlock) void k_fft_xy_c2r(const cx<R>* __restrict__ cg,
                                                            R* __restrict__ grid, FftPlan px,
                                                            FftPlan py,
                                                            const cx<R>* __restrict__ twx,
                                                            const cx<R>* __restrict__ twy,
                                                            PairCo co) {
  if constexpr (CO) {
    if (co_pairs<R>(co)) return;
  }
  k_fft_xy_c2r_body<R>(cg, grid, px, py, twx, twy, blockIdx.x);
}

// ----------------------------------------------------------------------------------------
// charge spreading
// ----------------------------------------------------------------------------------------

// round(x) as an int64, for |x| < 2^51: x + 1.5 * 2^52 keeps the exponent of 2^52, so the
// rounding to an integer happens in that addition and the bit patterns differ by round(x).
__device__ __forceinline__ long long fix_round(double x) {
  constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
  return __double_as_longlong(x + kMagic) - __double_as_longlong(kMagic);
}

// Fixed-point accumulator of a spreading workgroup's mesh points.  Integer adds are exact and
// commute, so a point's value does not depend on the order in which the lanes add its
// contributions: the spreading is bitwise reproducible.  FP32 mesh: one int64 at 2^-36
// (contributions below 2^15 in magnitude, sums below 1.3e8).  FP64 mesh: an int32 at 2^-4, added
// only when not zero (most contributions are below 1/32), and an int64 remainder at 2^-55
// (contributions and sums below 1.3e8; the remainder of each contribution, |r| <= 2^-5, is exact
// by Sterbenz's lemma and rounded at 2^-56).
template <typename R>
struct OwnedAcc {
  static constexpr bool F32 = sizeof(R) == 4;
  unsigned long long* lo;  // FP32: the value; FP64: the remainder
  int* hi;                 // FP64: the leading part
  __device__ __forceinline__ OwnedAcc(unsigned long long* l, int* h) : lo(l), hi(h) {}
  // the remainders, then (FP64) the leading parts of npts points from base
  __device__ __forceinline__ OwnedAcc(unsigned char* base, int npts)
      : lo(reinterpret_cast<unsigned long long*>(base)),
        hi(reinterpret_cast<int*>(base + sizeof(unsigned long long) * (size_t)npts)) {}
  __device__ __forceinline__ void clear(int i) {
    lo[i] = 0ull;
    if constexpr (!F32) hi[i] = 0;
  }
  __device__ __forceinline__ void add(int i, double v) {
    if constexpr (F32) {
      atomicAdd(lo + i, (unsigned long long)fix_round(v * 0x1p36));
    } else {
      const long long h = fix_round(v * 0x1p4);
      const double r = v - (double)h * 0x1p-4;
      if (h != 0) atomicAdd(hi + i, (int)h);
      atomicAdd(lo + i, (unsigned long long)fix_round(r * 0x1p55));
    }
  }
  __device__ __forceinline__ double value(int i) const {
    const double l = (double)(long long)lo[i];
    if constexpr (F32) return l * 0x1p-36;
    else return (double)hi[i] * 0x1p-4 + l * 0x1p-55;
  }
};


// One workgroup per output tile; atoms are listed by the tile of their stencil origin, and the
// workgroup visits the source tiles whose stencils reach into it.  Each contribution lands in
// exactly one output tile, so the mesh is written once, without global atomics or a clear; the
// tile accumulates in fixed point (OwnedAcc), so the result does not depend on the order of the
// adds.  Work items are (candidate atom, z offset) pairs, so every thread gets a similar share.
template <typename R>
__device__ __forceinline__ void k_spread_checked_body(const int4* __restrict__ k0,
                                                const R* __restrict__ theta,
                                                const double* __restrict__ q,
                                                const int* __restrict__ tile_start,
                                                SpreadArgs g, R* __restrict__ grid, int bid) {
  constexpr int kMaxVisit = kMaxSrcTiles * kMaxSrcTiles * kMaxSrcTiles;
  __shared__ unsigned long long acc_lo[kSpreadTileMax];
  __shared__ int acc_hi[sizeof(R) == 4 ? 1 : kSpreadTileMax];
  OwnedAcc<R> acc(acc_lo, acc_hi);
  __shared__ int src_count[kMaxVisit];
  __shared__ int src_first[kMaxVisit + 1];
  __shared__ int src_begin[kMaxVisit];
  const int p = g.order;
  int t = bid;
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1];
  const int t3 = t / g.nt[1];
  const int lo1 = t1 * g.T[0], lo2 = t2 * g.T[1], lo3 = t3 * g.T[2];
  const int w1 = min(g.T[0], g.K[0] - lo1), w2 = min(g.T[1], g.K[1] - lo2),
            w3 = min(g.T[2], g.K[2] - lo3);
  const int npts = w1 * w2 * w3;
  for (int i = threadIdx.x; i < npts; i += blockDim.x) acc.clear(i);
  const int* e1 = g.src + ((size_t)0 * g.ntmax + t1) * (kMaxSrcTiles + 1);
  const int* e2 = g.src + ((size_t)1 * g.ntmax + t2) * (kMaxSrcTiles + 1);
  const int* e3 = g.src + ((size_t)2 * g.ntmax + t3) * (kMaxSrcTiles + 1);
  const int n1 = e1[0], n2 = e2[0], n3 = e3[0];
  const int ns = n1 * n2 * n3;
  for (int c = threadIdx.x; c < ns; c += blockDim.x) {
    const int c1 = c % n1, c2 = (c / n1) % n2, c3 = c / (n1 * n2);
    const int st = (e3[1 + c3] * g.nt[1] + e2[1 + c2]) * g.nt[0] + e1[1 + c1];
    const int b = tile_start[st];
    src_begin[c] = b;
    src_count[c] = tile_start[st + 1] - b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int c = 0; c < ns; ++c) {
      src_first[c] = run;
      run += src_count[c];
    }
    src_first[ns] = run;
  }
  __syncthreads();
  const int nitems = src_first[ns] * p;
  for (int it = threadIdx.x; it < nitems; it += blockDim.x) {
    const int v = it / p, o3 = it - v * p;
    int c = 0;
    while (src_first[c + 1] <= v) ++c;
    const int a = src_begin[c] + (v - src_first[c]);
    const int4 kk = k0[a];
    int z = kk.z - lo3 + o3;  // position relative to the tile, modulo K3
    if (z < 0) z += g.K[2];
    if (z >= g.K[2]) z -= g.K[2];
    if (z >= w3) continue;
    int r1 = kk.x - lo1, r2 = kk.y - lo2;
    if (r1 < 0) r1 += g.K[0];
    if (r2 < 0) r2 += g.K[1];
    const R* th = theta + (size_t)a * 3 * p;
    const double c3 = q[a] * double(th[2 * p + o3]);
    for (int o2 = 0; o2 < p; ++o2) {
      int y = r2 + o2;
      if (y >= g.K[1]) y -= g.K[1];
      if (y >= w2) continue;
      const double c23 = c3 * double(th[p + o2]);
      const int row = (z * w2 + y) * w1;
      for (int o1 = 0; o1 < p; ++o1) {
        int x = r1 + o1;
        if (x >= g.K[0]) x -= g.K[0];
        if (x >= w1) continue;
        acc.add(row + x, c23 * double(th[o1]));
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < npts; i += blockDim.x) {
    const int x = i % w1, y = (i / w1) % w2, z = i / (w1 * w2);
    grid[((size_t)(lo3 + z) * g.K[1] + (lo2 + y)) * g.K[0] + lo1 + x] = R(acc.value(i));
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_spread_checked(const int4* __restrict__ k0,
                                                const R* __restrict__ theta,
                                                const double* __restrict__ q,
                                                const int* __restrict__ tile_start,
                                                SpreadArgs g, R* __restrict__ grid) {
  k_spread_checked_body<R>(k0, theta, q, tile_start, g, grid, blockIdx.x);
}

// Fast spreading for meshes with K >= T + p - 1 on every axis, where the stencil of an atom
// meets a tile in one contiguous interval per axis.  Per output tile the candidate atoms are
// staged in LDS in chunks -- only those that reach the tile, in a fixed order, with their
// weights (q folded into the z weights) and their valid interval per axis.  Each thread owns a
// 4-point x segment of the tile and accumulates it in registers over the staged atoms (every
// LDS read is a broadcast), so there are no atomics and the sum order is fixed.
template <typename R, int P>
__device__ __forceinline__ void k_spread_body(const int4* __restrict__ k0,
                                                           const R* __restrict__ theta,
                                                           const double* __restrict__ q,
                                                           const int* __restrict__ tile_start,
                                                           SpreadArgs g, R* __restrict__ grid, int bid) {
  constexpr int kMaxVisit = kMaxSrcTiles * kMaxSrcTiles * kMaxSrcTiles;
  constexpr int kWaves = kSpreadThreads / 64;
  extern __shared__ unsigned char smem_raw[];
  const int p = P > 0 ? P : g.order;
  const int cap = g.slots;
  // per-row bitmasks of the staged atoms that touch the row (rows = (z, y) of the tile)
  unsigned long long* rmask = reinterpret_cast<unsigned long long*>(smem_raw);  // [rows][4]
  int4* iv = reinterpret_cast<int4*>(rmask + kSpreadRows * 4);  // [cap]: a | n<<5 | pos<<10
  R* w1s = reinterpret_cast<R*>(iv + cap);                       // [cap][p]
  R* w2s = w1s + (size_t)cap * p;                                // [cap][p]
  R* w3s = w2s + (size_t)cap * p;                                // [cap][p], times q
  __shared__ int src_count[kMaxVisit];
  __shared__ int src_first[kMaxVisit + 1];
  __shared__ int src_begin[kMaxVisit];
  __shared__ int wave_hits[kWaves];
  int t = bid;
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1];
  const int t3 = t / g.nt[1];
  const int lo[3] = {t1 * g.T[0], t2 * g.T[1], t3 * g.T[2]};
  const int w[3] = {min(g.T[0], g.K[0] - lo[0]), min(g.T[1], g.K[1] - lo[1]),
                    min(g.T[2], g.K[2] - lo[2])};
  // this thread's mesh point
  const int mx = threadIdx.x % g.T[0], my = (threadIdx.x / g.T[0]) % g.T[1],
            mz = threadIdx.x / (g.T[0] * g.T[1]);
  const bool owner = mx < w[0] && my < w[1] && mz < w[2];
  R acc = R(0);

  const int* e1 = g.src + ((size_t)0 * g.ntmax + t1) * (kMaxSrcTiles + 1);
  const int* e2 = g.src + ((size_t)1 * g.ntmax + t2) * (kMaxSrcTiles + 1);
  const int* e3 = g.src + ((size_t)2 * g.ntmax + t3) * (kMaxSrcTiles + 1);
  const int n1 = e1[0], n2 = e2[0], n3 = e3[0];
  const int ns = n1 * n2 * n3;
  for (int c = threadIdx.x; c < ns; c += blockDim.x) {
    const int c1 = c % n1, c2 = (c / n1) % n2, c3 = c / (n1 * n2);
    const int st = (e3[1 + c3] * g.nt[1] + e2[1 + c2]) * g.nt[0] + e1[1 + c1];
    const int b = tile_start[st];
    src_begin[c] = b;
    src_count[c] = tile_start[st + 1] - b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int c = 0; c < ns; ++c) {
      src_first[c] = run;
      run += src_count[c];
    }
    src_first[ns] = run;
  }
  __syncthreads();
  const int nvisit = src_first[ns];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int chunk = 0; chunk < nvisit; chunk += cap) {
    for (int i = threadIdx.x; i < kSpreadRows * 4; i += blockDim.x) rmask[i] = 0ULL;
    // stage the candidates of this chunk that reach the tile, in candidate order
    const int v = chunk + (int)threadIdx.x;
    bool hit = false;
    int pk[3] = {0, 0, 0}, a = 0;
    if ((int)threadIdx.x < cap && v < nvisit) {
      int c = 0;
      while (src_first[c + 1] <= v) ++c;
      a = src_begin[c] + (v - src_first[c]);
      const int4 kk = k0[a];
      const int kv[3] = {kk.x, kk.y, kk.z};
      hit = true;
      for (int ax = 0; ax < 3; ++ax) {
        int r = kv[ax] - lo[ax];
        if (r < 0) r += g.K[ax];
        int ia, ib, pos;
        if (r < w[ax]) {  // stencil starts inside the tile
          ia = 0;
          ib = min(p, w[ax] - r);
          pos = r;
        } else {          // it reaches the tile only after wrapping past K
          ia = g.K[ax] - r;
          ib = min(p, g.K[ax] - r + w[ax]);
          pos = 0;
        }
        if (ia >= ib) hit = false;
        pk[ax] = ia | ((ib - ia) << 5) | (pos << 10);
      }
    }
    const unsigned long long hm = __ballot(hit);
    if (lane == 0) wave_hits[wv] = __popcll(hm);
    __syncthreads();
    int sl = __popcll(hm & ((1ULL << lane) - 1ULL));
    int nslot = 0;
    for (int k = 0; k < kWaves; ++k) {
      if (k < wv) sl += wave_hits[k];
      nslot += wave_hits[k];
    }
    if (hit) {
      const R* th = theta + (size_t)a * 3 * p;
      const R qa = R(q[a]);
      for (int o = 0; o < p; ++o) {
        w1s[sl * p + o] = th[o];
        w2s[sl * p + o] = th[p + o];
        w3s[sl * p + o] = qa * th[2 * p + o];
      }
      iv[sl] = make_int4(pk[0], pk[1], pk[2], 0);
      // mark the rows this atom touches
      const int z0 = pk[2] >> 10, nz = (pk[2] >> 5) & 31, y0 = pk[1] >> 10, ny = (pk[1] >> 5) & 31;
      const unsigned long long bit = 1ULL << (sl & 63);
      for (int dz = 0; dz < nz; ++dz)
        for (int dy = 0; dy < ny; ++dy)
          atomicOr(&rmask[((z0 + dz) * w[1] + y0 + dy) * 4 + (sl >> 6)], bit);
    }
    __syncthreads();
    if (owner) {
      // only the staged atoms that touch this point's row, in slot order
      const int row = mz * w[1] + my;
      const int nwords = (nslot + 63) >> 6;
      for (int wd = 0; wd < nwords; ++wd) {
        unsigned long long bits = rmask[row * 4 + wd];
        while (bits) {
          const int sl2 = wd * 64 + __ffsll((long long)bits) - 1;
          bits &= bits - 1ULL;
          const int4 hd = iv[sl2];
          const int dx = mx - (hd.x >> 10);
          if ((unsigned)dx < (unsigned)((hd.x >> 5) & 31)) {
            const int dz = mz - (hd.z >> 10), dy = my - (hd.y >> 10);
            acc += w3s[sl2 * p + (hd.z & 31) + dz] * w2s[sl2 * p + (hd.y & 31) + dy] *
                   w1s[sl2 * p + (hd.x & 31) + dx];
          }
        }
      }
    }
    __syncthreads();
  }
  if (owner) grid[((size_t)(lo[2] + mz) * g.K[1] + (lo[1] + my)) * g.K[0] + lo[0] + mx] = acc;
}

template <typename R, int P>
__global__ __launch_bounds__(kSpreadThreads) void k_spread(const int4* __restrict__ k0,
                                                           const R* __restrict__ theta,
                                                           const double* __restrict__ q,
                                                           const int* __restrict__ tile_start,
                                                           SpreadArgs g, R* __restrict__ grid) {
  k_spread_body<R, P>(k0, theta, q, tile_start, g, grid, blockIdx.x);
}

// Privatised spreading in two steps.
// Step 1 (one workgroup per spread block of whole sort tiles): the atoms of the block -- those
// whose stencil origin lies in it -- add their p^3 weights into a private LDS copy of the block
// plus its (p-1)-point upper halo; the copy is written whole to a resident scratch buffer.
// Step 2 (one thread per mesh point): each point is the fixed-order sum of the at most 2^3
// block copies that cover it.  No global atomics and no mesh clear.
template <typename R, int P>
__device__ __forceinline__ void k_spread_region_body(const int4* __restrict__ k0,
                                                       const R* __restrict__ theta,
                                                       const double* __restrict__ q,
                                                       const int* __restrict__ tile_start,
                                                       RegionArgs g, R* __restrict__ regions, int bid,
                                                       const int* __restrict__ qperm = nullptr,
                                                       R* __restrict__ grid = nullptr) {
  // the private copy accumulates in FP64 in both precision modes (LDS FP64 adds are native
  // and fast; the FP32 ones are not), and is stored in the mesh precision R
  extern __shared__ unsigned char smem_raw[];
  double* reg = reinterpret_cast<double*>(smem_raw);
  const int p = P > 0 ? P : g.order;
  const int RX = g.S[0] + p - 1, RY = g.S[1] + p - 1, RZ = g.S[2] + p - 1;
  const int nreg = RX * RY * RZ;
  for (int i = threadIdx.x; i < nreg; i += blockDim.x) reg[i] = 0.0;
  int b = bid;
  const int b1 = b % g.nb[0];
  b /= g.nb[0];
  const int b2 = b % g.nb[1], b3 = b / g.nb[1];
  const int lo1 = b1 * g.S[0], lo2 = b2 * g.S[1], lo3 = b3 * g.S[2];
  __syncthreads();
  for (int s3 = 0; s3 < g.f[2]; ++s3) {
    const int u3 = b3 * g.f[2] + s3;
    if (u3 >= g.nt[2]) break;
    for (int s2 = 0; s2 < g.f[1]; ++s2) {
      const int u2 = b2 * g.f[1] + s2;
      if (u2 >= g.nt[1]) break;
      // sort tiles u1 .. u1 + f[0] - 1 of one (u2, u3) row are contiguous in atom order
      const int u1a = b1 * g.f[0], u1b = min(u1a + g.f[0], g.nt[0]);
      const int ta = (u3 * g.nt[1] + u2) * g.nt[0];
      const int a0 = tile_start[ta + u1a], a1 = tile_start[ta + u1b];
      const int nitems = (a1 - a0) * p;  // one (atom, z plane) per item
      if (P > 0) {
        // kSpreadBatch items per thread at a time: all their loads (stencil origin, weights,
        // charge) are issued before the first LDS atomic, so a batch costs one memory latency
        constexpr int PP = P > 0 ? P : 1;
        constexpr int NB = kSpreadBatch;
        for (int base = threadIdx.x; base < nitems; base += NB * blockDim.x) {
          int4 kk[NB];
          double c3[NB];
          R w1[NB][PP], w2[NB][PP];
          int o3v[NB];
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            const int it = base + u * (int)blockDim.x;
            const bool ok = it < nitems;
            const int a = ok ? a0 + it / PP : a0, o3 = ok ? it % PP : 0;
            o3v[u] = ok ? o3 : -1;
            kk[u] = k0[a];
            const R* th = theta + (size_t)a * 3 * PP;
#pragma unroll
            for (int o = 0; o < PP; ++o) {
              w1[u][o] = th[o];
              w2[u][o] = th[PP + o];
            }
            c3[u] = (qperm ? q[qperm[a]] : q[a]) * double(th[2 * PP + o3]);
          }
#pragma unroll
          for (int u = 0; u < NB; ++u) {
            if (o3v[u] < 0) continue;
            double* plane = reg + ((size_t)(kk[u].z - lo3 + o3v[u]) * RY + (kk[u].y - lo2)) * RX +
                            (kk[u].x - lo1);
#pragma unroll
            for (int o2 = 0; o2 < PP; ++o2) {
              const double c = c3[u] * double(w2[u][o2]);
#pragma unroll
              for (int o1 = 0; o1 < PP; ++o1)
                atomicAdd(plane + o2 * RX + o1, c * double(w1[u][o1]));
            }
          }
        }
        continue;
      }
      for (int it = threadIdx.x; it < nitems; it += blockDim.x) {
        const int a = a0 + it / p, o3 = it % p;
        const int4 kk = k0[a];
        const R* th = theta + (size_t)a * 3 * p;
        // sorted charges, or the caller's through the sort permutation
        const double c3 = (qperm ? q[qperm[a]] : q[a]) * double(th[2 * p + o3]);
        double* plane = reg + ((size_t)(kk.z - lo3 + o3) * RY + (kk.y - lo2)) * RX + (kk.x - lo1);
        if (P > 0) {
          double w1[P > 0 ? P : 1], w2[P > 0 ? P : 1];
#pragma unroll
          for (int o = 0; o < (P > 0 ? P : 1); ++o) {
            w1[o] = double(th[o]);
            w2[o] = double(th[p + o]);
          }
#pragma unroll
          for (int o2 = 0; o2 < (P > 0 ? P : 1); ++o2) {
            const double c = c3 * w2[o2];
#pragma unroll
            for (int o1 = 0; o1 < (P > 0 ? P : 1); ++o1)
              atomicAdd(plane + o2 * RX + o1, c * w1[o1]);
          }
        } else {
          for (int o2 = 0; o2 < p; ++o2) {
            const double c = c3 * double(th[p + o2]);
            for (int o1 = 0; o1 < p; ++o1) atomicAdd(plane + o2 * RX + o1, c * double(th[o1]));
          }
        }
      }
    }
  }
  __syncthreads();
  if (g.direct == 1) {
    // the copy (block + halo, wrapped onto the periodic mesh) is added into the mesh, which was
    // cleared before the launch; untouched halo points are skipped
    for (int i = threadIdx.x; i < nreg; i += blockDim.x) {
      const double v = reg[i];
      if (v == 0.0) continue;
      const int ix = i % RX, t = i / RX;
      const int iy = t % RY, iz = t / RY;
      int gx = lo1 + ix, gy = lo2 + iy, gz = lo3 + iz;
      while (gx >= g.K[0]) gx -= g.K[0];
      while (gy >= g.K[1]) gy -= g.K[1];
      while (gz >= g.K[2]) gz -= g.K[2];
      unsafeAtomicAdd(grid + ((size_t)gz * g.K[1] + gy) * g.K[0] + gx, R(v));
    }
    return;
  }
  R* out = regions + (size_t)bid * g.region_stride;
  for (int i = threadIdx.x; i < nreg; i += blockDim.x) out[i] = R(reg[i]);
}

template <typename R, int P, bool CO>
__global__ __launch_bounds__(256) void k_spread_region(const int4* __restrict__ k0,
                                                       const R* __restrict__ theta,
                                                       const double* __restrict__ q,
                                                       const int* __restrict__ tile_start,
                                                       RegionArgs g, R* __restrict__ regions,
                                                       const int* __restrict__ qperm,
                                                       R* __restrict__ grid, PairCo co) {
  if constexpr (CO) {
    if (co_pairs<R>(co)) return;
  }
  k_spread_region_body<R, P>(k0, theta, q, tile_start, g, regions, blockIdx.x, qperm, grid);
}

// Owner-computes spreading (RegionArgs::direct == 2): one workgroup per block of whole sort
// tiles owns the block's mesh points.  The atoms whose stencils can reach the block are those of
// the sort tiles holding the cells lo-(p-1) .. lo+w-1 of each axis (cyclically); in each tile
// row they form one or two contiguous runs of the sorted atoms.  Pass 1 walks those runs as one
// flattened index and keeps, for each atom whose stencil reaches the block, its in-block z
// planes as (atom, plane) items in an LDS list; pass 2 adds each item's in-block part of its
// plane into a fixed-point LDS copy of the block alone (no halo; OwnedAcc).  The copy is then
// stored into the
// mesh: every mesh point is written once, with no copies, no assembly, no mesh clear and no
// device atomics.  Atoms near a block face are read by the workgroups of both blocks.
template <typename R, int P>
__device__ __forceinline__ void k_spread_owned_body(const int4* __restrict__ k0,
                                                    const R* __restrict__ theta,
                                                    const double* __restrict__ q,
                                                    const int* __restrict__ tile_start,
                                                    RegionArgs g, R* __restrict__ grid, int bid,
                                                    const int* __restrict__ qperm) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ int run_a0[kOwnedMaxRuns];       // first sorted atom of run r
  __shared__ int run_off[kOwnedMaxRuns + 1];  // atoms in the runs before r
  __shared__ int n_items;
  const int p = P > 0 ? P : g.order;
  const int S0 = g.S[0], S1 = g.S[1], S2 = g.S[2];
  const int PX = owned_row_pitch(g.S), PZ = owned_plane_pitch(g.S);
  const int nreg = PZ * S2;
  OwnedAcc<R> reg(smem_raw, nreg);
  int* items = reinterpret_cast<int*>(smem_raw + owned_point_bytes(sizeof(R) == 4) * (size_t)nreg);
  for (int i = threadIdx.x; i < nreg; i += blockDim.x) reg.clear(i);
  int blk[3];
  {
    int b = bid;
    blk[0] = b % g.nb[0];
    b /= g.nb[0];
    blk[1] = b % g.nb[1];
    blk[2] = b / g.nb[1];
  }
  // per axis: the block's first point and width, and the cyclic run of sort tiles (first,
  // count) that holds the stencil origins lo-(p-1) .. lo+w-1
  int lo[3], w[3], tf[3], tc[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int T = g.S[a] / g.f[a];
    lo[a] = blk[a] * g.S[a];
    w[a] = min(g.S[a], g.K[a] - lo[a]);
    int c0 = lo[a] - (p - 1);
    if (c0 < 0) c0 += g.K[a];
    tf[a] = c0 / T;
    int cnt = (lo[a] + w[a] - 1) / T - tf[a];
    if (cnt < 0) cnt += g.nt[a];
    tc[a] = min(cnt + 1, g.nt[a]);
  }
  // the runs: tile row (iz, iy), x tiles tf[0] .. tf[0]+tc[0]-1, split where they wrap
  const int xe = tf[0] + tc[0];
  const int pieces = xe > g.nt[0] ? 2 : 1;
  const int nruns = tc[1] * tc[2] * pieces;  // the host keeps this within kOwnedMaxRuns
  for (int r = threadIdx.x; r < nruns; r += blockDim.x) {
    const int row = r / pieces, piece = r - row * pieces;
    const int iz = row / tc[1], iy = row - iz * tc[1];
    const int u3 = tf[2] + iz < g.nt[2] ? tf[2] + iz : tf[2] + iz - g.nt[2];
    const int u2 = tf[1] + iy < g.nt[1] ? tf[1] + iy : tf[1] + iy - g.nt[1];
    const int ta = (u3 * g.nt[1] + u2) * g.nt[0];
    const int ua = piece == 0 ? tf[0] : 0;
    const int ub = piece == 0 ? min(xe, g.nt[0]) : xe - g.nt[0];
    const int a0 = tile_start[ta + ua];
    run_a0[r] = a0;
    run_off[r + 1] = tile_start[ta + ub] - a0;
  }
  if (threadIdx.x == 0) {
    run_off[0] = 0;
    n_items = 0;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int r = 0; r < nruns; ++r) run_off[r + 1] += run_off[r];
  __syncthreads();
  const int natoms = run_off[nruns];
  const int chunk = owned_items(g) / p;  // atoms per pass: each keeps at most p items
  for (int j0 = 0; j0 < natoms; j0 += chunk) {
    const int j1 = min(natoms, j0 + chunk);
    // pass 1
    for (int j = j0 + (int)threadIdx.x; j < j1; j += blockDim.x) {
      int lr = 0, hr = nruns - 1;  // the last run that starts at or before j
      while (lr < hr) {
        const int mid = (lr + hr + 1) >> 1;
        if (run_off[mid] <= j) lr = mid;
        else hr = mid - 1;
      }
      const int a = run_a0[lr] + (j - run_off[lr]);
      const int4 kk = k0[a];
      int r[3] = {kk.x - lo[0], kk.y - lo[1], kk.z - lo[2]};
      bool reach = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        // the representative in [-(p-1), K-(p-1)) of the origin relative to the block
        if (r[d] >= g.K[d] - (p - 1)) r[d] -= g.K[d];
        else if (r[d] < -(p - 1)) r[d] += g.K[d];
        reach = reach && r[d] < w[d];
      }
      if (!reach) continue;
      const int za = max(0, -r[2]), zb = min(p, w[2] - r[2]);
      const int pos = atomicAdd(&n_items, zb - za);
      for (int o3 = za; o3 < zb; ++o3) items[pos + o3 - za] = (a << 4) | o3;
    }
    __syncthreads();
    const int m = n_items;
    // pass 2
    if (P > 0) {
      constexpr int PP = P > 0 ? P : 1;
      // FP32 weights: two items in flight per thread, so that eight waves per SIMD fit
      constexpr int NB = sizeof(R) == 4 ? 2 : kSpreadBatch;
      for (int base = threadIdx.x; base < m; base += NB * blockDim.x) {
        int4 kk[NB];
        double c3[NB];
        R w1[NB][PP], w2[NB][PP];
        int o3v[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const int it = base + u * (int)blockDim.x;
          const int e = items[it < m ? it : base];
          const int a = e >> 4, o3 = e & 15;
          o3v[u] = it < m ? o3 : -1;
          kk[u] = k0[a];
          const R* th = theta + (size_t)a * 3 * PP;
#pragma unroll
          for (int o = 0; o < PP; ++o) {
            w1[u][o] = th[o];
            w2[u][o] = th[PP + o];
          }
          c3[u] = (qperm ? q[qperm[a]] : q[a]) * double(th[2 * PP + o3]);
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          if (o3v[u] < 0) continue;
          int r[3] = {kk[u].x - lo[0], kk[u].y - lo[1], kk[u].z - lo[2]};
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            if (r[d] >= g.K[d] - (PP - 1)) r[d] -= g.K[d];
            else if (r[d] < -(PP - 1)) r[d] += g.K[d];
          }
          const int z = r[2] + o3v[u];
          bool vx[PP];
#pragma unroll
          for (int o1 = 0; o1 < PP; ++o1) vx[o1] = (unsigned)(r[0] + o1) < (unsigned)w[0];
#pragma unroll
          for (int o2 = 0; o2 < PP; ++o2) {
            const int y = r[1] + o2;
            if ((unsigned)y >= (unsigned)w[1]) continue;
            const double c = c3[u] * double(w2[u][o2]);
            const int row = z * PZ + y * PX + r[0];
#pragma unroll
            for (int o1 = 0; o1 < PP; ++o1)
              if (vx[o1]) reg.add(row + o1, c * double(w1[u][o1]));
          }
        }
      }
    } else {
      for (int it = threadIdx.x; it < m; it += blockDim.x) {
        const int e = items[it];
        const int a = e >> 4, o3 = e & 15;
        const int4 kk = k0[a];
        const R* th = theta + (size_t)a * 3 * p;
        int r[3] = {kk.x - lo[0], kk.y - lo[1], kk.z - lo[2]};
        for (int d = 0; d < 3; ++d) {
          if (r[d] >= g.K[d] - (p - 1)) r[d] -= g.K[d];
          else if (r[d] < -(p - 1)) r[d] += g.K[d];
        }
        const int z = r[2] + o3;
        const double c3 = (qperm ? q[qperm[a]] : q[a]) * double(th[2 * p + o3]);
        for (int o2 = 0; o2 < p; ++o2) {
          const int y = r[1] + o2;
          if ((unsigned)y >= (unsigned)w[1]) continue;
          const double c = c3 * double(th[p + o2]);
          for (int o1 = 0; o1 < p; ++o1) {
            const int x = r[0] + o1;
            if ((unsigned)x < (unsigned)w[0]) reg.add(z * PZ + y * PX + x, c * double(th[o1]));
          }
        }
      }
    }
    __syncthreads();  // every item of this pass is added before the list is refilled
    if (threadIdx.x == 0) n_items = 0;
    __syncthreads();
  }
  BoxWalk wk((int)threadIdx.x, blockDim.x, S0, S1);
  for (int i = threadIdx.x; i < S0 * S1 * S2; i += blockDim.x) {
    if (wk.x < w[0] && wk.y < w[1] && wk.z < w[2])
      grid[((size_t)(lo[2] + wk.z) * g.K[1] + lo[1] + wk.y) * g.K[0] + lo[0] + wk.x] =
          R(reg.value(wk.z * PZ + wk.y * PX + wk.x));
    wk.next();
  }
}

// FP32 weights leave room for eight waves per SIMD (at most 64 VGPRs): with item lists sized
// to the block, eight 256-thread blocks then share a CU.
template <typename R, int P, bool CO>
__global__ __launch_bounds__(256, (sizeof(R) == 4 && !CO) ? 8 : 1) void k_spread_owned(
    const int4* __restrict__ k0,
                                                      const R* __restrict__ theta,
                                                      const double* __restrict__ q,
                                                      const int* __restrict__ tile_start,
                                                      RegionArgs g, R* __restrict__ grid,
                                                      const int* __restrict__ qperm, PairCo co) {
  if constexpr (CO) {
    if (co_pairs<R>(co)) return;
  }
  k_spread_owned_body<R, P>(k0, theta, q, tile_start, g, grid, blockIdx.x, qperm);
}

// Clears n elements (the mesh before a direct spreading): four per thread, grid-stride.
template <typename R>
__global__ __launch_bounds__(256) void k_clear(R* __restrict__ p, long long n) {
  const long long stride = 4LL * blockDim.x * gridDim.x;
  for (long long i = 4LL * (blockIdx.x * (long long)blockDim.x + threadIdx.x); i < n; i += stride) {
    if (i + 3 < n) {
      p[i] = R(0);
      p[i + 1] = R(0);
      p[i + 2] = R(0);
      p[i + 3] = R(0);
    } else {
      for (long long j = i; j < n; ++j) p[j] = R(0);
    }
  }
}

// Assembly of the privatised copies, one workgroup per block of the mesh: a point at offset l
// (per axis) inside its block receives its own copy at l and, for small l, the halo of the
// copies of the preceding blocks at l + (widths of the blocks in between); they are summed in a
// fixed order.  Everything but the offsets and the loads is uniform across the workgroup.
template <typename R>
__device__ __forceinline__ void k_spread_assemble_body(const R* __restrict__ regions,
                                                         RegionArgs g, R* __restrict__ grid, int bid) {
  const int p = g.order;
  const int RX = g.S[0] + p - 1, RY = g.S[1] + p - 1;
  const int part = bid % kAssembleSplit;
  int b[3];
  {
    int t = bid / kAssembleSplit;
    b[0] = t % g.nb[0];
    t /= g.nb[0];
    b[1] = t % g.nb[1];
    b[2] = t / g.nb[1];
  }
  // per axis: candidate blocks d = 0..3 back (wrapping), the offset shift of each and its width
  int cb[3][4], sh[3][4], lim[3][4], nc[3], w[3], lo[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = b[a] * g.S[a];
    w[a] = min(g.S[a], g.K[a] - lo[a]);
    nc[a] = 0;
    int shift = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      if (d >= g.nb[a] && d > 0) break;
      int t = b[a] - d;
      if (t < 0) t += g.nb[a];
      const int wt = min(g.S[a], g.K[a] - t * g.S[a]);
      if (d > 0) shift += wt;
      // offsets l + shift inside copy t (width wt, halo p-1) are valid for l < wt + p - 1 - shift
      const int l_max = wt + p - 1 - shift;
      if (l_max <= 0) break;
      cb[a][d] = t;
      sh[a][d] = shift;
      lim[a][d] = l_max;
      nc[a] = d + 1;
    }
  }
  // every mesh point of the block, x fastest; this workgroup takes part `part` of them.  The
  // candidate loops have block-uniform bounds and the loads are predicated, so a thread keeps
  // the loads of kAssembleUnroll points in flight at once; per point the copies are summed in
  // the same (d3, d2, d1) order whatever the split.
  const int wxy = w[0] * w[1];
  const int npts = wxy * w[2];
  const int chunk = (npts + kAssembleSplit - 1) / kAssembleSplit;
  const int p0 = part * chunk, p1 = min(npts, p0 + chunk);
  BoxWalk wk(p0 + (int)threadIdx.x, blockDim.x, w[0], w[1]);
  for (int base = p0 + (int)threadIdx.x; base < p1; base += kAssembleUnroll * blockDim.x) {
    int lx[kAssembleUnroll], ly[kAssembleUnroll], lz[kAssembleUnroll];
    bool ok[kAssembleUnroll];
    R sum[kAssembleUnroll];
#pragma unroll
    for (int u = 0; u < kAssembleUnroll; ++u) {
      ok[u] = base + u * (int)blockDim.x < p1;
      lx[u] = wk.x;
      ly[u] = wk.y;
      lz[u] = wk.z;
      wk.next();
      sum[u] = R(0);
    }
    for (int d3 = 0; d3 < nc[2]; ++d3) {
      for (int d2 = 0; d2 < nc[1]; ++d2) {
        const size_t base_copy = (size_t)(cb[2][d3] * g.nb[1] + cb[1][d2]) * g.nb[0];
        for (int d1 = 0; d1 < nc[0]; ++d1) {
          const R* __restrict__ copy = regions + (base_copy + cb[0][d1]) * g.region_stride;
          R v[kAssembleUnroll];
#pragma unroll
          for (int u = 0; u < kAssembleUnroll; ++u) {
            const bool in = ok[u] && lz[u] < lim[2][d3] && ly[u] < lim[1][d2] &&
                            lx[u] < lim[0][d1];
            // one copy holds (S + p - 1)^3 points: 32-bit offsets
            const int off = ((lz[u] + sh[2][d3]) * RY + (ly[u] + sh[1][d2])) * RX + lx[u] +
                            sh[0][d1];
            v[u] = in ? copy[off] : R(0);
          }
#pragma unroll
          for (int u = 0; u < kAssembleUnroll; ++u) sum[u] += v[u];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kAssembleUnroll; ++u)
      if (ok[u])
        grid[(size_t)((lo[2] + lz[u]) * g.K[1] + (lo[1] + ly[u])) * g.K[0] + lo[0] + lx[u]] =
            sum[u];
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_spread_assemble(const R* __restrict__ regions,
                                                         RegionArgs g, R* __restrict__ grid) {
  k_spread_assemble_body<R>(regions, g, grid, blockIdx.x);
}

// ----------------------------------------------------------------------------------------
// per-atom kernels
// ----------------------------------------------------------------------------------------

__device__ __forceinline__ void k_gather_charges_body(int n, const int* __restrict__ perm,
                                 const double* __restrict__ dq, double* __restrict__ qs,
                                 double* __restrict__ zero, int bid) {
  const int k = bid * blockDim.x + threadIdx.x;
  if (k < n) {
    qs[k] = dq[perm[k]];
    if (zero) zero[k] = 0.0;
  }
}

__global__ void k_gather_charges(int n, const int* __restrict__ perm,
                                 const double* __restrict__ dq, double* __restrict__ qs,
                                 double* __restrict__ zero) {
  k_gather_charges_body(n, perm, dq, qs, zero, blockIdx.x);
}


// Real-space pair sum: out[i] = sum_j coef_ij q_j, one 64-lane wave per atom (row).  A sweep
// covers 512 entries: lane l takes entries l + 64u (u < 8), so that every load instruction of
// the wave reads consecutive entries and consecutive partners, and all eight loads of a lane
// are issued before the first is used.
template <typename CT, typename IT>
__device__ __forceinline__ void k_pair_spmv_body(int n, const long long* __restrict__ row_ptr,
                                                   const IT* __restrict__ col,
                                                   const CT* __restrict__ coef,
                                                   const double* __restrict__ q,
                                                   double* __restrict__ out, int bid) {
  const int lane = threadIdx.x & 63;
  // wave-uniform row: its bounds come through the scalar cache
  const int row = __builtin_amdgcn_readfirstlane(bid * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (row >= n) return;
  const long long a = row_ptr[row], b = row_ptr[row + 1];
  double acc = 0.0;
  for (long long g = a + lane; g < b; g += 512) {
    IT c[8];
    CT w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const long long k = g + 64 * u;
      c[u] = k < b ? col[k] : IT(0);
      w[u] = k < b ? coef[k] : CT(0);
    }
    double qv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) qv[u] = g + 64 * u < b ? q[partner(row, c[u], n)] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += double(w[u]) * qv[u];
  }
  acc = wave_sum(acc);
  if (lane == 0) out[row] = acc;
}

// Streaming form of the same sum for large row lists: a bounded grid of waves walks the rows
// (wave w takes rows w, w + W, ...), and while it gathers the charges of one row it already
// loads the bounds, partner offsets and coefficients of its next row, so that a row costs one
// memory latency instead of three.  Products are summed in k_pair_spmv_body's order (sweeps
// of 512 entries, lane l taking entries l + 64u), so both give the same result bit for bit.
// The bounded grid also leaves room on every CU for the kernels that run beside it.
template <typename CT, typename IT>
__global__ __launch_bounds__(256) void k_pair_spmv_stream(int n,
                                                          const long long* __restrict__ row_ptr,
                                                          const IT* __restrict__ col,
                                                          const CT* __restrict__ coef,
                                                          const double* __restrict__ q,
                                                          double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  int row = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (row >= n) return;  // uniform across the wave
  long long a = row_ptr[row], b = row_ptr[row + 1];
  IT c[8];
  CT w[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const long long k = a + lane + 64 * u;
    c[u] = k < b ? col[k] : IT(0);
    w[u] = k < b ? coef[k] : CT(0);
  }
  while (true) {
    const int nrow = row + nw;
    const bool more = nrow < n;
    // this row's partners' charges
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = a + lane + 64 * u < b ? q[partner(row, c[u], n)] : 0.0;
    // the next row's bounds and first sweep, in flight during this row's sums
    long long na = 0, nb = 0;
    IT nc[8];
    CT nwt[8];
    if (more) {
      na = row_ptr[nrow];
      nb = row_ptr[nrow + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const long long k = na + lane + 64 * u;
      nc[u] = k < nb ? col[k] : IT(0);
      nwt[u] = k < nb ? coef[k] : CT(0);
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += double(w[u]) * x[u];
    // rows longer than one sweep (dense image sums, long short-range lists)
    for (long long g = a + lane + 512; g < b; g += 512) {
      IT cc[8];
      CT ww[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const long long k = g + 64 * u;
        cc[u] = k < b ? col[k] : IT(0);
        ww[u] = k < b ? coef[k] : CT(0);
      }
      double qv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) qv[u] = g + 64 * u < b ? q[partner(row, cc[u], n)] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += double(ww[u]) * qv[u];
    }
    acc = wave_sum(acc);
    if (lane == 0) out[row] = acc;
    if (!more) break;
    row = nrow;
    a = na;
    b = nb;
    for (int u = 0; u < 8; ++u) {
      c[u] = nc[u];
      w[u] = nwt[u];
    }
  }
}

template <typename CT, typename IT>
__global__ __launch_bounds__(256) void k_pair_spmv(int n, const long long* __restrict__ row_ptr,
                                                   const IT* __restrict__ col,
                                                   const CT* __restrict__ coef,
                                                   const double* __restrict__ q,
                                                   double* __restrict__ out) {
  k_pair_spmv_body<CT, IT>(n, row_ptr, col, coef, q, out, blockIdx.x);
}

#ifndef EW_PAIR_MULTI_BLOCKS
#define EW_PAIR_MULTI_BLOCKS 256  // 1 workgroup (4 waves) per CU on 256 CUs
#endif

// Charges of NV vectors interleaved per atom (row j holds the NV charges of atom j, padded to
// an even count) so that one gather of a partner fetches all of its charges.
template <int NV>
__global__ __launch_bounds__(256) void k_interleave_charges(int n, const double* __restrict__ q,
                                                            double* __restrict__ qil) {
  constexpr int S = spmm_stride(NV);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int v = 0; v < S; ++v) qil[(size_t)i * S + v] = v < NV ? q[(size_t)v * n + i] : 0.0;
}

// Multi-vector pair sum: a bounded grid of waves walks the rows as k_pair_spmv_stream does
// (the next row's bounds, offsets and coefficients in flight while this row's charges are
// gathered); the coefficients and offsets of a row are loaded once for the NV vectors and each
// partner's NV charges come with one gather of its interleaved row.  Per vector the products
// are summed in k_pair_spmv_body's order, so every vector's result is the single-vector sum
// bit for bit.
template <typename CT, typename IT, int NV>
__global__ __launch_bounds__(256) void k_pair_spmm(PairMulti pm) {
  constexpr int S = spmm_stride(NV);
  constexpr int H = S == 1 ? 1 : S / 2;  // 16-byte chunks per partner
  const int lane = threadIdx.x & 63;
  const int n = pm.n;
  const int nw = gridDim.x * (blockDim.x >> 6);
  int row =
      __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (row >= n) return;  // uniform across the wave
  const long long* __restrict__ row_ptr = pm.row_ptr;
  const IT* __restrict__ col = static_cast<const IT*>(pm.col);
  const CT* __restrict__ coef = static_cast<const CT*>(pm.coef);
  const double* __restrict__ qil = pm.qil;
  auto gather = [&](int j, bool ok, double* x) {
    if constexpr (S == 1) {
      x[0] = ok ? qil[j] : 0.0;
    } else {
      const double2* __restrict__ src = reinterpret_cast<const double2*>(qil + (size_t)j * S);
#pragma unroll
      for (int hh = 0; hh < H; ++hh) {
        const double2 t = ok ? src[hh] : make_double2(0.0, 0.0);
        x[2 * hh] = t.x;
        x[2 * hh + 1] = t.y;
      }
    }
  };
  long long a = row_ptr[row], b = row_ptr[row + 1];
  IT c[8];
  CT w[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const long long k = a + lane + 64 * u;
    c[u] = k < b ? col[k] : IT(0);
    w[u] = k < b ? coef[k] : CT(0);
  }
  while (true) {
    const int nrow = row + nw;
    const bool more = nrow < n;
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    double x[8][S];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool ok = a + lane + 64 * u < b;
      gather(ok ? partner(row, c[u], n) : 0, ok, x[u]);
    }
    long long na = 0, nb = 0;
    IT nc[8];
    CT nwt[8];
    if (more) {
      na = row_ptr[nrow];
      nb = row_ptr[nrow + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const long long k = na + lane + 64 * u;
      nc[u] = k < nb ? col[k] : IT(0);
      nwt[u] = k < nb ? coef[k] : CT(0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += double(w[u]) * x[u][v];
    for (long long g = a + lane + 512; g < b; g += 512) {
      IT cc[8];
      CT ww[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const long long k = g + 64 * u;
        cc[u] = k < b ? col[k] : IT(0);
        ww[u] = k < b ? coef[k] : CT(0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const bool ok = g + 64 * u < b;
        gather(ok ? partner(row, cc[u], n) : 0, ok, x[u]);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] += double(ww[u]) * x[u][v];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double t = wave_sum(acc[v]);
      if (lane == 0) pm.out[(size_t)v * n + row] = t;
    }
    if (!more) break;
    row = nrow;
    a = na;
    b = nb;
    for (int u = 0; u < 8; ++u) {
      c[u] = nc[u];
      w[u] = nwt[u];
    }
  }
}

template <typename CT, typename IT>
static hipError_t launch_spmm_nv(const PairMulti& pm, hipStream_t st) {
  const dim3 grid(pm.beside ? std::min((pm.n + 3) / 4, EW_PAIR_MULTI_BLOCKS) : (pm.n + 3) / 4),
      block(256), igrid((pm.n + 255) / 256);
  switch (pm.nvec) {
#define EW_NV(V)                                                                             \
  case V:                                                                                    \
    hipLaunchKernelGGL((k_interleave_charges<V>), igrid, block, 0, st, pm.n, pm.q, pm.qil);  \
    hipLaunchKernelGGL((k_pair_spmm<CT, IT, V>), grid, block, 0, st, pm);                    \
    break;
    EW_NV(1) EW_NV(2) EW_NV(3) EW_NV(4) EW_NV(5) EW_NV(6) EW_NV(7) EW_NV(8)
#undef EW_NV
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pair_spmm(const PairMulti& pm, bool fp32, hipStream_t st) {
  if (pm.col16)
    return fp32 ? launch_spmm_nv<float, short>(pm, st) : launch_spmm_nv<double, short>(pm, st);
  return fp32 ? launch_spmm_nv<float, int>(pm, st) : launch_spmm_nv<double, int>(pm, st);
}

// ---- symmetric pair storage ----

// Partner slot of tile B in tile A's ascending partner list [k0, k1), or -1.
__device__ __forceinline__ int half_find(const int* __restrict__ ptile, int k0, int k1, int B) {
  while (k0 < k1) {
    const int m = (k0 + k1) >> 1;
    const int v = ptile[m];
    if (v == B) return m;
    if (v < B)
      k0 = m + 1;
    else
      k1 = m;
  }
  return -1;
}

// One thread per atom i (tile A): its entries j of the full list with (tile(j), j) after
// (A, i), per chunk of A.  COUNT: rc[base + chunk * nA + i_local] = count; else: written at
// rc[...] onwards, in the order of the full row.
template <bool COUNT, typename CT>
__global__ void k_half_pairs(int n, const long long* __restrict__ row_ptr,
                             const int* __restrict__ col, const CT* __restrict__ coef,
                             const int* __restrict__ tile_of, const int* __restrict__ tile_start,
                             HalfPairs hp, int* __restrict__ rc, unsigned short* __restrict__ out,
                             CT* __restrict__ coef_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int A = tile_of[i];
  const int a0 = tile_start[A], nA = tile_start[A + 1] - a0, il = i - a0;
  const int c0 = hp.tile_chunk[A], nch = hp.tile_chunk[A + 1] - c0;
  const int k0 = hp.chunk_tile[c0], k1 = hp.chunk_tile[c0 + nch];
  const int base = hp.rc_base[A] + il;
  int cnt[kHalfMaxChunks];
#pragma unroll
  for (int c = 0; c < kHalfMaxChunks; ++c) cnt[c] = 0;
  for (long long k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
    const int j = col[k];
    const int B = tile_of[j];
    if (B < A || (B == A && j <= i)) continue;
    const int slot = half_find(hp.ptile, k0, k1, B);
    if (slot < 0) continue;  // cannot happen: every partner tile is in A's list
    const int rel = hp.ptile_chunk[slot];
    if (COUNT) {
      for (int c = 0; c < kHalfMaxChunks; ++c)
        if (c == rel) ++cnt[c];
    } else {
      int pos = rc[base + rel * nA];
      for (int c = 0; c < kHalfMaxChunks; ++c)
        if (c == rel) pos += cnt[c]++;
      out[pos] = (unsigned short)((il << 8) | (hp.ptile_off[slot] + j - tile_start[B]));
      coef_out[pos] = coef[k];
    }
  }
  if (COUNT)
    for (int c = 0; c < nch; ++c) {
      int v = 0;
#pragma unroll
      for (int cc = 0; cc < kHalfMaxChunks; ++cc)
        if (cc == c) v = cnt[cc];
      rc[base + c * nA] = v;
    }
}

// out += sum over the stored pairs of tile A (= bid): c_ij q_j into i and c_ij q_i into j, in LDS
// per chunk, then one global atomic add per atom and chunk.
template <typename CT>
__device__ __forceinline__ void k_pair_half_body(const int* __restrict__ tile_start,
                                                 const HalfPairs& hp,
                                                 const double* __restrict__ q,
                                                 double* __restrict__ out, int bid) {
  __shared__ double qa[256], acc_a[256], qc[256], acc_c[256];
  const int A = bid;
  const int a0 = tile_start[A], nA = tile_start[A + 1] - a0;
  const int tid = threadIdx.x;
  if (tid < nA) {
    qa[tid] = q[a0 + tid];
    acc_a[tid] = 0.0;
  }
  const unsigned short* __restrict__ ent = hp.entry;
  const CT* __restrict__ cf = static_cast<const CT*>(hp.coef);
  for (int c = hp.tile_chunk[A]; c < hp.tile_chunk[A + 1]; ++c) {
    const int k0 = hp.chunk_tile[c], k1 = hp.chunk_tile[c + 1];
    const int na = hp.chunk_atoms[c];
    int gj = -1;  // this thread's staged atom (global index)
    if (tid < na) {
      int k = k0;
      while (k + 1 < k1 && hp.ptile_off[k + 1] <= tid) ++k;
      const int B = hp.ptile[k], t0 = tile_start[B], r = tid - hp.ptile_off[k];
      if (r < tile_start[B + 1] - t0) gj = t0 + r;  // inside the partner tile
      qc[tid] = gj >= 0 ? q[gj] : 0.0;
      acc_c[tid] = 0.0;
    }
    __syncthreads();
    const int e1 = hp.chunk_entry[c + 1];
    for (int e = hp.chunk_entry[c] + tid; e < e1; e += blockDim.x) {
      const unsigned v = ent[e];
      const int i = v >> 8, j = v & 255;
      const double w = double(cf[e]);
      atomicAdd(&acc_a[i], w * qc[j]);
      atomicAdd(&acc_c[j], w * qa[i]);
    }
    __syncthreads();
    if (gj >= 0) unsafeAtomicAdd(out + gj, acc_c[tid]);
    __syncthreads();
  }
  if (tid < nA) unsafeAtomicAdd(out + a0 + tid, acc_a[tid]);
}

template <typename CT>
__global__ __launch_bounds__(256) void k_pair_half(const int* __restrict__ tile_start,
                                                   HalfPairs hp, const double* __restrict__ q,
                                                   double* __restrict__ out) {
  k_pair_half_body<CT>(tile_start, hp, q, out, blockIdx.x);
}

// Rewrites absolute partner indices as signed 16-bit offsets from the row; flags rows whose
// partners lie further than +-32767 atoms away (then the 32-bit list is kept).
__global__ __launch_bounds__(256) void k_compress_cols(int n, const long long* __restrict__ row_ptr,
                                                       const int* __restrict__ col,
                                                       short* __restrict__ off,
                                                       int* __restrict__ too_far) {
  // one wave per row, lanes along the row
  const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (i >= n) return;
  const int lane = threadIdx.x & 63;
  for (long long k = row_ptr[i] + lane; k < row_ptr[i + 1]; k += 64) {
    int d = col[k] - i;
    if (d >= (n + 1) / 2) d -= n;
    if (d < -(n / 2)) d += n;
    if (d < -32768 || d > 32767) {
      *too_far = 1;
      d = 0;
    }
    off[k] = short(d);
  }
}

// B-spline gather of the convolved mesh, plus the real-space sum, the on-site short-gamma
// term and the Ewald self term (fe/common/ewald_self.f90:17), written in caller order.
template <typename R>
__device__ __forceinline__ void k_gather_combine_body(
    int n, const int4* __restrict__ k0, const R* __restrict__ theta,
    const R* __restrict__ grid, const double* __restrict__ q, const double* __restrict__ phi_real,
    const double* __restrict__ onsite, const int* __restrict__ perm, GatherArgs g,
    double* __restrict__ shift, double* __restrict__ epart, int bid) {
  __shared__ double red[256 / 64];
  const int k = bid * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (k < n) {
    const int p = g.order;
    const int4 kk = k0[k];
    const R* th = theta + (size_t)k * 3 * p;
    int i1[kMaxOrder];
#pragma unroll
    for (int o = 0; o < kMaxOrder; ++o)
      if (o < p) {
        int x = kk.x + o;
        i1[o] = x >= g.K[0] ? x - g.K[0] : x;
      }
    R acc = R(0), comp = R(0);
    for (int o3 = 0; o3 < p; ++o3) {
      int z = kk.z + o3;
      if (z >= g.K[2]) z -= g.K[2];
      for (int o2 = 0; o2 < p; ++o2) {
        int y = kk.y + o2;
        if (y >= g.K[1]) y -= g.K[1];
        const R* row = grid + ((size_t)z * g.K[1] + y) * g.K[0];
        R part = R(0);
#pragma unroll
        for (int o = 0; o < kMaxOrder; ++o)
          if (o < p) part += th[o] * row[i1[o]];
        // compensated accumulation of the 36 row partials (Kahan)
        const R yv = part * th[p + o2] * th[2 * p + o3] - comp;
        const R tv = acc + yv;
        comp = (tv - acc) - yv;
        acc = tv;
      }
    }
    const double qk = q[k];
    double phi = phi_real[k] + double(acc) + (onsite ? onsite[k] : 0.0) * qk + g.self_coef * qk;
    if (g.qsum) phi += g.bg * g.qsum[0];
    shift[perm[k]] = phi;
    e = 0.5 * phi * qk;
  }
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    epart[bid] = s;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_gather_combine(
    int n, const int4* __restrict__ k0, const R* __restrict__ theta,
    const R* __restrict__ grid, const double* __restrict__ q, const double* __restrict__ phi_real,
    const double* __restrict__ onsite, const int* __restrict__ perm, GatherArgs g,
    double* __restrict__ shift, double* __restrict__ epart) {
  k_gather_combine_body<R>(n, k0, theta, grid, q, phi_real, onsite, perm, g, shift, epart, blockIdx.x);
}

// B-spline gather per mesh tile: the tile's atoms (their stencil origins lie in it) read the
// tile plus a (p-1)-point halo, staged once in LDS; four lanes share an atom's p^2 rows and are
// reduced by shuffles.  Then the pair sum, on-site and self terms are added and the shift is
// written in caller order; one energy partial per tile.
// MO (mesh only): the mesh potential of each atom goes to shift[a] in sorted order (no other
// term, no energy); k_shift_combine adds the rest once the pair sum is done.
template <typename R, int P, bool MO = false>
__device__ __forceinline__ void k_gather_tile_body(
    const int4* __restrict__ k0, const R* __restrict__ theta, const R* __restrict__ grid,
    const double* __restrict__ q, const double* __restrict__ phi_real,
    const double* __restrict__ onsite, const int* __restrict__ perm,
    const int* __restrict__ tile_start, GatherArgs g, double* __restrict__ shift,
    double* __restrict__ epart, int bid) {
  extern __shared__ unsigned char smem_raw[];
  R* ths = reinterpret_cast<R*>(smem_raw);  // [64][3p] weights of the 64 atoms of one pass
  R* reg = ths + 64 * 3 * g.order;          // staged mesh region
  __shared__ double red[4];
  const int p = P > 0 ? P : g.order;
  int t = bid;
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1];
  const int t3 = t / g.nt[1];
  const int lo1 = t1 * g.T[0], lo2 = t2 * g.T[1], lo3 = t3 * g.T[2];
  const int E1 = min(g.T[0], g.K[0] - lo1) + p - 1, E2 = min(g.T[1], g.K[1] - lo2) + p - 1,
            E3 = min(g.T[2], g.K[2] - lo3) + p - 1;
  const int nreg = E1 * E2 * E3;
  BoxWalk wk(threadIdx.x, 256, E1, E2);
  for (int i0 = 0; i0 < nreg; i0 += 16 * 256) {
    // sixteen loads in flight per thread before the LDS stores
    R v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int i = i0 + threadIdx.x + k * 256;
      v[k] = R(0);
      if (i < nreg) {
        int gx = lo1 + wk.x, gy = lo2 + wk.y, gz = lo3 + wk.z;
        while (gx >= g.K[0]) gx -= g.K[0];
        while (gy >= g.K[1]) gy -= g.K[1];
        while (gz >= g.K[2]) gz -= g.K[2];
        v[k] = grid[(size_t)(gz * g.K[1] + gy) * g.K[0] + gx];
      }
      wk.next();
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int i = i0 + threadIdx.x + k * 256;
      if (i < nreg) reg[i] = v[k];
    }
  }
  __syncthreads();
  const int a0 = tile_start[bid], a1 = tile_start[bid + 1];
  const int sub = threadIdx.x & 3, grp = threadIdx.x >> 2;
  const int rows = p * p;
  double e = 0.0;
  for (int base = a0; base < a1; base += 64) {  // uniform trip count across the workgroup
    const int a = base + grp;
    const bool live = a < a1;
    const int nb = min(64, a1 - base) * 3 * p;
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) ths[i] = theta[(size_t)base * 3 * p + i];
    __syncthreads();
    double part = 0.0;
    if (live) {
      const int4 kk = k0[a];
      int r1 = kk.x - lo1, r2 = kk.y - lo2, r3 = kk.z - lo3;
      if (r1 < 0) r1 += g.K[0];
      if (r2 < 0) r2 += g.K[1];
      if (r3 < 0) r3 += g.K[2];
      const R* th = ths + grp * 3 * p;
      R acc = R(0), comp = R(0);
      R w1r[P > 0 ? P : 1];
      if (P > 0) {
#pragma unroll
        for (int o1 = 0; o1 < (P > 0 ? P : 1); ++o1) w1r[o1] = th[o1];
      }
      for (int r = sub; r < rows; r += 4) {
        const int o3 = r / p, o2 = r - o3 * p;
        const R* row = reg + ((r3 + o3) * E2 + (r2 + o2)) * E1 + r1;
        R v = R(0);
        if (P > 0) {
          R rv[P > 0 ? P : 1];
#pragma unroll
          for (int o1 = 0; o1 < (P > 0 ? P : 1); ++o1) rv[o1] = row[o1];
#pragma unroll
          for (int o1 = 0; o1 < (P > 0 ? P : 1); ++o1) v += w1r[o1] * rv[o1];
        } else {
          for (int o1 = 0; o1 < p; ++o1) v += th[o1] * row[o1];
        }
        // compensated (Kahan) accumulation of the row partials
        const R yv = v * th[p + o2] * th[2 * p + o3] - comp;
        const R tv = acc + yv;
        comp = (tv - acc) - yv;
        acc = tv;
      }
      part = double(acc) - double(comp);
    }
    part += __shfl_xor(part, 1, 64);
    part += __shfl_xor(part, 2, 64);
    if (MO) {
      if (live && sub == 0) shift[a] = part;
      continue;
    }
    if (live && sub == 0) {
      const double qa = q[a];
      const double phi =
          phi_real[a] + part + (onsite ? onsite[a] : 0.0) * qa + g.self_coef * qa +
          (g.qsum ? g.bg * g.qsum[0] : 0.0);
      shift[perm[a]] = phi;
      e += 0.5 * phi * qa;
    }
  }
  if (MO) return;
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) epart[bid] = red[0] + red[1] + red[2] + red[3];
}

template <typename R, int P, bool MO>
__global__ __launch_bounds__(256) void k_gather_tile(
    const int4* __restrict__ k0, const R* __restrict__ theta, const R* __restrict__ grid,
    const double* __restrict__ q, const double* __restrict__ phi_real,
    const double* __restrict__ onsite, const int* __restrict__ perm,
    const int* __restrict__ tile_start, GatherArgs g, double* __restrict__ shift,
    double* __restrict__ epart) {
  k_gather_tile_body<R, P, MO>(k0, theta, grid, q, phi_real, onsite, perm, tile_start, g, shift,
                               epart, blockIdx.x);
}

// The shift of each atom from its mesh potential (k_gather_tile<.., true>, sorted order) and
// the real-space pair sum, plus the on-site, self and background terms, written in caller order;
// one energy partial 1/2 sum shift q per workgroup, in a fixed order.
__global__ __launch_bounds__(256) void k_shift_combine(int n, const double* __restrict__ mesh,
                                                       const double* __restrict__ phi_real,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ onsite,
                                                       const int* __restrict__ perm, GatherArgs g,
                                                       double* __restrict__ shift,
                                                       double* __restrict__ epart) {
  __shared__ double red[4];
  const int k = blockIdx.x * 256 + threadIdx.x;
  double e = 0.0;
  if (k < n) {
    const double qk = q[k];
    const double phi = phi_real[k] + mesh[k] + (onsite ? onsite[k] : 0.0) * qk +
                       g.self_coef * qk + (g.qsum ? g.bg * g.qsum[0] : 0.0);
    shift[perm[k]] = phi;
    e = 0.5 * phi * qk;
  }
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) epart[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__device__ __forceinline__ void k_energy_final_body(int nparts, const double* __restrict__ epart,
                               double* __restrict__ energy, int bid) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += epart[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *energy = red[0];
}

__global__ void k_energy_final(int nparts, const double* __restrict__ epart,
                               double* __restrict__ energy) {
  k_energy_final_body(nparts, epart, energy, blockIdx.x);
}

// ----------------------------------------------------------------------------------------
// geometry-time kernels
// ----------------------------------------------------------------------------------------

// B-spline weights theta[k][axis][o] = M_p(u - k0 - o) by the order recursion.
template <typename R>
__global__ void k_bspline(int n, const double* __restrict__ frac, BsplineArgs g,
                          R* __restrict__ theta) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int p = g.order;
  for (int a = 0; a < 3; ++a) {
    const double u = frac[3 * k + a] * g.K[a];
    const double w = u - floor(u);
    double c[kMaxOrder + 1];
    c[0] = w;
    c[1] = 1.0 - w;
    for (int kk = 2; kk < p; ++kk) {
      // c[j] = M_kk(w + j) -> M_{kk+1}(w + j), j = kk..0 (descending keeps c[j-1] intact)
      c[kk] = (kk + 1 - w - kk) * c[kk - 1] / kk;
      for (int j = kk - 1; j >= 1; --j) c[j] = ((w + j) * c[j] + (kk + 1 - w - j) * c[j - 1]) / kk;
      c[0] = w * c[0] / kk;
    }
    R* out = theta + ((size_t)k * 3 + a) * p;
    for (int o = 0; o < p; ++o) out[o] = R(c[p - 1 - o]);
  }
}

// Pair coefficient of one image: erfc(a r)/r within r_cut minus the short-range gamma within
// its species-pair cutoff (src/dftbp/dftb/shortgamma.F90:918-923 sign convention).
__device__ __forceinline__ double pair_term(double r2, const PairArgs& g, int si, int sj,
                                            bool* inside) {
  double c = 0.0;
  bool in = false;
  if (r2 <= g.rc2) {
    const double r = sqrt(r2);
    c += erfc(g.alpha * r) / r;
    in = true;
  }
  if (g.ns > 0) {
    const int sp = si * g.ns + sj;
    if (r2 <= g.gcut2[sp]) {
      const double r = sqrt(r2);
      const double ua = g.hub[sj], ub = g.hub[si];
      const double ta = 3.2 * ua, tb = 3.2 * ub;
      double gm;
      if (fabs(ua - ub) < 0.3125e-5) {
        const double tm = 0.5 * (ta + tb);
        gm = exp(-tm * r) * (1.0 / r + 0.6875 * tm + 0.1875 * r * tm * tm +
                             0.02083333333333333333 * r * r * tm * tm * tm);
      } else {
        auto sub = [r](double t1, double t2) {
          const double d = t1 * t1 - t2 * t2;
          const double t24 = t2 * t2 * t2 * t2;
          return exp(-t1 * r) * ((0.5 * t24 * t1 / (d * d)) -
                                 (t24 * t2 * t2 - 3.0 * t24 * t1 * t1) / (r * d * d * d));
        };
        gm = sub(ta, tb) + sub(tb, ta);
      }
      c -= gm;
      in = true;
    }
  }
  *inside = in;
  return c;
}

// Linked tiles: every partner tile within +-D along each axis, minimum image by fractional
// rounding (fe/fft/ewald_fft_3d_real.f90:150-157).  COUNT: row lengths; else fill.  One wave
// per atom: the lanes take 64 consecutive candidates of a partner tile at a time (coalesced
// coordinate loads), and a ballot compacts the pairs inside r_list, so a row lists its partners
// in candidate order -- partner tiles in (z, y, x) order, atoms ascending within a tile.
// Stores partner j of row i at entry k: the index itself (out.col), or its signed 16-bit offset
// from the row taken modulo n into [-n/2, n/2) (out.off; out.too_far is set and 0 stored when
// it does not fit), the layout the pair sums read.
__device__ __forceinline__ void store_partner(const PartnerOut& out, long long k, int i, int j) {
  if (out.col) {
    out.col[k] = j;
    return;
  }
  int d = j - i;
  if (d >= (out.n + 1) / 2) d -= out.n;
  if (d < -(out.n / 2)) d += out.n;
  if (d < -32768 || d > 32767) {
    *out.too_far = 1;
    d = 0;
  }
  out.off[k] = short(d);
}

template <bool COUNT, typename CT>
__global__ __launch_bounds__(256) void k_pair_tiles(int n, const double* __restrict__ frac,
                                                    const int* __restrict__ sp,
                                                    const int* __restrict__ tile_of,
                                                    const int* __restrict__ tile_start,
                                                    PairArgs g, long long* __restrict__ row_ptr,
                                                    PartnerOut po,
                                                    CT* __restrict__ coef,
                                                    int* __restrict__ err) {
  const int lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (i >= n) return;  // uniform across the wave
  const double fi0 = frac[3 * i], fi1 = frac[3 * i + 1], fi2 = frac[3 * i + 2];
  const int si = sp ? sp[i] : 0;
  int t = tile_of[i];
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1], t3 = t / g.nt[1];
  long long pos = COUNT ? 0 : row_ptr[i];
  const long long end = COUNT ? 0 : row_ptr[i + 1];
  long long cnt = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int b1 = g.full[0] ? 0 : t1 - g.D[0], n1 = g.full[0] ? g.nt[0] : 2 * g.D[0] + 1;
  const int b2 = g.full[1] ? 0 : t2 - g.D[1], n2 = g.full[1] ? g.nt[1] : 2 * g.D[1] + 1;
  const int b3 = g.full[2] ? 0 : t3 - g.D[2], n3 = g.full[2] ? g.nt[2] : 2 * g.D[2] + 1;
  for (int c3 = 0; c3 < n3; ++c3) {
    const int u3 = ((b3 + c3) % g.nt[2] + g.nt[2]) % g.nt[2];
    for (int c2 = 0; c2 < n2; ++c2) {
      const int u2 = ((b2 + c2) % g.nt[1] + g.nt[1]) % g.nt[1];
      for (int c1 = 0; c1 < n1; ++c1) {
        const int u1 = ((b1 + c1) % g.nt[0] + g.nt[0]) % g.nt[0];
        const int tt = (u3 * g.nt[1] + u2) * g.nt[0] + u1;
        const int j1 = tile_start[tt + 1];
        for (int j0 = tile_start[tt]; j0 < j1; j0 += 64) {
          const int j = j0 + lane;
          bool in = false;
          double c = 0.0;
          if (j < j1 && j != i) {
            double s0 = fi0 - frac[3 * j], s1 = fi1 - frac[3 * j + 1], s2 = fi2 - frac[3 * j + 2];
            s0 -= rint(s0);
            s1 -= rint(s1);
            s2 -= rint(s2);
            const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
            const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
            const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
            const double r2 = x * x + y * y + z * z;
            if (r2 <= g.rl2) {
              if (r2 < kMinSeparationSq)
                *err = 1;
              else
                c = pair_term(r2, g, si, sp ? sp[j] : 0, &in);
            }
          }
          const unsigned long long mask = __ballot(in);
          if (COUNT) {
            cnt += __popcll(mask);
          } else {
            const long long k = pos + __popcll(mask & below);
            if (in && k < end) {  // the fill never runs past the counted row length
              store_partner(po, k, i, j);
              coef[k] = CT(c);
            }
            pos += __popcll(mask);
          }
        }
      }
    }
  }
  if (COUNT && lane == 0) row_ptr[i] = cnt;
}

// The same list, one workgroup per sort tile: the partner candidates every atom of the tile
// shares (the atoms of its partner tiles, in the same order) are staged in LDS a chunk at a
// time, and each wave scans them for one atom of the tile after another, so the coordinates
// are read from HBM once per tile instead of once per atom.  Rows come out exactly as from
// k_pair_tiles.  Needs at most kPairMaxCandTiles partner tiles and kPairMaxTileAtoms atoms in
// a tile.
constexpr int kPairStage = 1024;        // candidates staged per chunk
constexpr int kPairMaxCandTiles = 1024;
constexpr int kPairMaxTileAtoms = 1024;

template <bool COUNT, typename CT>
__global__ __launch_bounds__(256) void k_pair_tiles_lds(const double* __restrict__ frac,
                                                        const int* __restrict__ sp,
                                                        const int* __restrict__ tile_start,
                                                        PairArgs g, long long* __restrict__ row_ptr,
                                                        PartnerOut po,
                                                        CT* __restrict__ coef,
                                                        int* __restrict__ err) {
  __shared__ double cxs[kPairStage], cys[kPairStage], czs[kPairStage];
  __shared__ int cjs[kPairStage], css[kPairStage];
  __shared__ int ctile[kPairMaxCandTiles], cstart[kPairMaxCandTiles + 1];
  __shared__ long long apos[kPairMaxTileAtoms];
  // fill: per wave, the candidates inside r_list queue up (in candidate order) so that the
  // pair terms are evaluated 64 at a time on full waves
  __shared__ int qj[COUNT ? 1 : 4][128], qs[COUNT ? 1 : 4][128];
  __shared__ double qr[COUNT ? 1 : 4][128];
  const int A = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  int t = A;
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1], t3 = t / g.nt[1];
  const int b1 = g.full[0] ? 0 : t1 - g.D[0], n1 = g.full[0] ? g.nt[0] : 2 * g.D[0] + 1;
  const int b2 = g.full[1] ? 0 : t2 - g.D[1], n2 = g.full[1] ? g.nt[1] : 2 * g.D[1] + 1;
  const int b3 = g.full[2] ? 0 : t3 - g.D[2], n3 = g.full[2] ? g.nt[2] : 2 * g.D[2] + 1;
  const int ntc = n1 * n2 * n3;  // <= kPairMaxCandTiles (checked by the launcher)
  // partner tiles in (z, y, x) order, with the running count of their atoms
  for (int c = threadIdx.x; c < ntc; c += blockDim.x) {
    const int c1 = c % n1, c2 = (c / n1) % n2, c3 = c / (n1 * n2);
    const int u1 = ((b1 + c1) % g.nt[0] + g.nt[0]) % g.nt[0];
    const int u2 = ((b2 + c2) % g.nt[1] + g.nt[1]) % g.nt[1];
    const int u3 = ((b3 + c3) % g.nt[2] + g.nt[2]) % g.nt[2];
    ctile[c] = (u3 * g.nt[1] + u2) * g.nt[0] + u1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int c = 0; c < ntc; ++c) {
      cstart[c] = run;
      run += tile_start[ctile[c] + 1] - tile_start[ctile[c]];
    }
    cstart[ntc] = run;
  }
  const int ia0 = tile_start[A], na = tile_start[A + 1] - ia0;  // na <= kPairMaxTileAtoms
  if (na == 0) return;  // uniform: an empty tile has no rows
  for (int a = threadIdx.x; a < na; a += blockDim.x) apos[a] = COUNT ? 0 : row_ptr[ia0 + a];
  __syncthreads();
  const int total = cstart[ntc];
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int k0 = 0; k0 < total; k0 += kPairStage) {
    const int m = min(kPairStage, total - k0);
    __syncthreads();  // the previous chunk is no longer read
    for (int idx = threadIdx.x; idx < m; idx += blockDim.x) {
      const int k = k0 + idx;
      int lo = 0, hi = ntc - 1;  // the partner tile holding candidate k
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cstart[mid] <= k) lo = mid;
        else hi = mid - 1;
      }
      const int j = tile_start[ctile[lo]] + (k - cstart[lo]);
      cxs[idx] = frac[3 * j];
      cys[idx] = frac[3 * j + 1];
      czs[idx] = frac[3 * j + 2];
      cjs[idx] = j;
      css[idx] = sp ? sp[j] : 0;
    }
    __syncthreads();
    for (int a = wv; a < na; a += nwv) {
      const int i = ia0 + a;
      const double fi0 = frac[3 * i], fi1 = frac[3 * i + 1], fi2 = frac[3 * i + 2];
      const int si = sp ? sp[i] : 0;
      long long pos = apos[a];
      const long long end = COUNT ? 0 : row_ptr[i + 1];
      if constexpr (!COUNT) {
        // the queue of this wave: qn entries (j, r^2, species of j), flushed 64 at a time
        int* wj = qj[wv];
        int* ws = qs[wv];
        double* wr = qr[wv];
        int qn = 0;
        auto flush = [&](int cnt) {
          bool in = false;
          double v = 0.0;
          int j = -1;
          if (lane < cnt) {
            j = wj[lane];
            v = pair_term(wr[lane], g, si, ws[lane], &in);
          }
          const unsigned long long mask = __ballot(in);
          const long long k = pos + __popcll(mask & below);
          if (in && k < end) {  // the fill never runs past the counted row length
            store_partner(po, k, i, j);
            coef[k] = CT(v);
          }
          pos += __popcll(mask);
        };
        for (int c0 = 0; c0 < m; c0 += 64) {
          const int c = c0 + lane;
          bool near = false;
          double r2 = 0.0;
          int j = -1;
          if (c < m) {
            j = cjs[c];
            if (j != i) {
              double s0 = fi0 - cxs[c], s1 = fi1 - cys[c], s2 = fi2 - czs[c];
              s0 -= rint(s0);
              s1 -= rint(s1);
              s2 -= rint(s2);
              const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
              const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
              const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
              r2 = x * x + y * y + z * z;
              if (r2 <= g.rl2) {
                if (r2 < kMinSeparationSq)
                  *err = 1;
                else
                  near = true;
              }
            }
          }
          const unsigned long long mask = __ballot(near);
          if (near) {
            const int slot = qn + __popcll(mask & below);
            wj[slot] = j;
            wr[slot] = r2;
            ws[slot] = css[c];
          }
          qn += __popcll(mask);
          if (qn >= 64) {
            flush(64);
            // move the overflow (fewer than 64 entries) to the front
            if (lane < qn - 64) {
              wj[lane] = wj[64 + lane];
              wr[lane] = wr[64 + lane];
              ws[lane] = ws[64 + lane];
            }
            qn -= 64;
          }
        }
        if (qn > 0) flush(qn);
        if (lane == 0) apos[a] = pos;
        continue;
      }
      for (int c0 = 0; c0 < m; c0 += 64) {
        const int c = c0 + lane;
        bool in = false;
        double v = 0.0;
        int j = -1;
        if (c < m) {
          j = cjs[c];
          if (j != i) {
            double s0 = fi0 - cxs[c], s1 = fi1 - cys[c], s2 = fi2 - czs[c];
            s0 -= rint(s0);
            s1 -= rint(s1);
            s2 -= rint(s2);
            const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
            const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
            const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
            const double r2 = x * x + y * y + z * z;
            if (r2 <= g.rl2) {
              if (r2 < kMinSeparationSq)
                *err = 1;
              else
                v = pair_term(r2, g, si, css[c], &in);
            }
          }
        }
        const unsigned long long mask = __ballot(in);
        if (!COUNT) {
          const long long k = pos + __popcll(mask & below);
          if (in && k < end) {  // the fill never runs past the counted row length
            store_partner(po, k, i, j);
            coef[k] = CT(v);
          }
        }
        pos += __popcll(mask);
      }
      if (lane == 0) apos[a] = pos;
    }
  }
  if (COUNT) {
    __syncthreads();
    for (int a = threadIdx.x; a < na; a += blockDim.x) row_ptr[ia0 + a] = apos[a];
  }
}

// Small cells: every (i, j) pair summed over all images within r_list
// (fe/direct/ewald_direct_3d.f90:237-300).  Row i holds all n partners.
template <typename CT>
__global__ void k_pair_dense(int n, const double* __restrict__ frac, const int* __restrict__ sp,
                             PairArgs g, int* __restrict__ col, CT* __restrict__ coef,
                             int* __restrict__ err) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int i = int(idx / n), j = int(idx - (long long)i * n);
  double s0 = frac[3 * i] - frac[3 * j], s1 = frac[3 * i + 1] - frac[3 * j + 1],
         s2 = frac[3 * i + 2] - frac[3 * j + 2];
  const int si = sp ? sp[i] : 0, sj = sp ? sp[j] : 0;
  double c = 0.0;
  for (int a = -g.D[0]; a <= g.D[0]; ++a)
    for (int b = -g.D[1]; b <= g.D[1]; ++b)
      for (int e = -g.D[2]; e <= g.D[2]; ++e) {
        if (i == j && a == 0 && b == 0 && e == 0) continue;
        const double f0 = s0 + a, f1 = s1 + b, f2 = s2 + e;
        const double x = f0 * g.lat[0] + f1 * g.lat[3] + f2 * g.lat[6];
        const double y = f0 * g.lat[1] + f1 * g.lat[4] + f2 * g.lat[7];
        const double z = f0 * g.lat[2] + f1 * g.lat[5] + f2 * g.lat[8];
        const double r2 = x * x + y * y + z * z;
        if (r2 > g.rl2) continue;
        if (r2 < kMinSeparationSq) {
          *err = 1;
          continue;
        }
        bool in;
        c += pair_term(r2, g, si, sj, &in);
      }
  col[idx] = j;
  coef[idx] = CT(c);
}

// ----------------------------------------------------------------------------------------
// forces (once per geometry step; fe/fft/ewald_fft_3d.f90:128-132)
// ----------------------------------------------------------------------------------------

// Reciprocal force from the convolved mesh: F_i = -q_i sum_k grad theta_i(k) Phi(k), with
// grad theta = sum_a dtheta/du_a K_a b_a.  Weights and derivatives are rebuilt from frac.
template <typename R>
__global__ void k_force_recip(int n, const double* __restrict__ frac, const int4* __restrict__ k0,
                              const R* __restrict__ grid, const double* __restrict__ q,
                              const int* __restrict__ perm, ForceArgs g,
                              double* __restrict__ force) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int p = g.order;
  double th[3][kMaxOrder], dth[3][kMaxOrder];
  for (int a = 0; a < 3; ++a) {
    const double u = frac[3 * k + a] * g.K[a];
    const double w = u - floor(u);
    double c[kMaxOrder + 1], cp[kMaxOrder + 1];
    c[0] = w;
    c[1] = 1.0 - w;
    cp[0] = 1.0;
    cp[1] = 0.0;
    for (int kk = 2; kk < p; ++kk) {
      if (kk == p - 1)
        for (int j = 0; j < kk; ++j) cp[j] = c[j];
      c[kk] = (1.0 - w) * c[kk - 1] / kk;
      for (int j = kk - 1; j >= 1; --j) c[j] = ((w + j) * c[j] + (kk + 1 - w - j) * c[j - 1]) / kk;
      c[0] = w * c[0] / kk;
    }
    for (int o = 0; o < p; ++o) {
      const int j = p - 1 - o;
      th[a][o] = c[j];
      dth[a][o] = (j < p - 1 ? cp[j] : 0.0) - (j > 0 ? cp[j - 1] : 0.0);
    }
  }
  const int4 kk = k0[k];
  double d1 = 0.0, d2 = 0.0, d3 = 0.0;
  for (int o3 = 0; o3 < p; ++o3) {
    int z = kk.z + o3;
    if (z >= g.K[2]) z -= g.K[2];
    for (int o2 = 0; o2 < p; ++o2) {
      int y = kk.y + o2;
      if (y >= g.K[1]) y -= g.K[1];
      const R* row = grid + ((size_t)z * g.K[1] + y) * g.K[0];
      double v0 = 0.0, v1 = 0.0;
      for (int o1 = 0; o1 < p; ++o1) {
        int x = kk.x + o1;
        if (x >= g.K[0]) x -= g.K[0];
        const double v = double(row[x]);
        v0 += th[0][o1] * v;
        v1 += dth[0][o1] * v;
      }
      d1 += v1 * th[1][o2] * th[2][o3];
      d2 += v0 * dth[1][o2] * th[2][o3];
      d3 += v0 * th[1][o2] * dth[2][o3];
    }
  }
  const double qk = q[k];
  const double s1 = d1 * g.K[0], s2 = d2 * g.K[1], s3 = d3 * g.K[2];
  const int i = perm[k];
  for (int d = 0; d < 3; ++d)
    force[3 * i + d] = -qk * (s1 * g.rec[d] + s2 * g.rec[3 + d] + s3 * g.rec[6 + d]);
}

// Real-space Ewald force, added to force[]: sum_j q_i q_j (2a/sqrt(pi) exp(-a^2 r^2) +
// erfc(a r)/r) / r^2 (x_i - x_j), over linked tiles (minimum image) or, in DENSE mode, over
// every partner and image (fe/fft/ewald_fft_3d_real.f90:166-172, direct_3d.f90:284-297).
template <bool DENSE>
__global__ void k_force_real(int n, const double* __restrict__ frac, const int* __restrict__ tile_of,
                             const int* __restrict__ tile_start, const double* __restrict__ q,
                             const int* __restrict__ perm, PairArgs g, double* __restrict__ force) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double fi0 = frac[3 * i], fi1 = frac[3 * i + 1], fi2 = frac[3 * i + 2];
  const double two_a_sqrtpi = 2.0 * g.alpha / kSqrtPi;
  double f[3] = {0.0, 0.0, 0.0};
  auto add = [&](double s0, double s1, double s2, int j) {
    const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
    const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
    const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
    const double r2 = x * x + y * y + z * z;
    if (r2 > g.rc2 || r2 < kMinSeparationSq) return;
    const double r = sqrt(r2);
    const double c = q[j] * (two_a_sqrtpi * exp(-g.alpha * g.alpha * r2) + erfc(g.alpha * r) / r) / r2;
    f[0] += c * x;
    f[1] += c * y;
    f[2] += c * z;
  };
  if (DENSE) {
    for (int j = 0; j < n; ++j) {
      const double s0 = fi0 - frac[3 * j], s1 = fi1 - frac[3 * j + 1], s2 = fi2 - frac[3 * j + 2];
      for (int a = -g.D[0]; a <= g.D[0]; ++a)
        for (int b = -g.D[1]; b <= g.D[1]; ++b)
          for (int e = -g.D[2]; e <= g.D[2]; ++e) {
            if (i == j && a == 0 && b == 0 && e == 0) continue;
            add(s0 + a, s1 + b, s2 + e, j);
          }
    }
  } else {
    int t = tile_of[i];
    const int t1 = t % g.nt[0];
    t /= g.nt[0];
    const int t2 = t % g.nt[1], t3 = t / g.nt[1];
    const int b1 = g.full[0] ? 0 : t1 - g.D[0], n1 = g.full[0] ? g.nt[0] : 2 * g.D[0] + 1;
    const int b2 = g.full[1] ? 0 : t2 - g.D[1], n2 = g.full[1] ? g.nt[1] : 2 * g.D[1] + 1;
    const int b3 = g.full[2] ? 0 : t3 - g.D[2], n3 = g.full[2] ? g.nt[2] : 2 * g.D[2] + 1;
    for (int c3 = 0; c3 < n3; ++c3) {
      const int u3 = ((b3 + c3) % g.nt[2] + g.nt[2]) % g.nt[2];
      for (int c2 = 0; c2 < n2; ++c2) {
        const int u2 = ((b2 + c2) % g.nt[1] + g.nt[1]) % g.nt[1];
        for (int c1 = 0; c1 < n1; ++c1) {
          const int u1 = ((b1 + c1) % g.nt[0] + g.nt[0]) % g.nt[0];
          const int tt = (u3 * g.nt[1] + u2) * g.nt[0] + u1;
          for (int j = tile_start[tt]; j < tile_start[tt + 1]; ++j) {
            if (j == i) continue;
            double s0 = fi0 - frac[3 * j], s1 = fi1 - frac[3 * j + 1], s2 = fi2 - frac[3 * j + 2];
            s0 -= rint(s0);
            s1 -= rint(s1);
            s2 -= rint(s2);
            add(s0, s1, s2, j);
          }
        }
      }
    }
  }
  const double qi = q[i];
  const int o = perm[i];
  for (int d = 0; d < 3; ++d) force[3 * o + d] += qi * f[d];
}

// Total charge in a fixed order: one workgroup, a strided partial per thread, a fixed tree.
__device__ __forceinline__ void k_charge_sum_body(int n, const double* __restrict__ q,
                                                  double* __restrict__ qsum, int) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += q[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) qsum[0] = red[0];
}

__global__ __launch_bounds__(256) void k_charge_sum(int n, const double* __restrict__ q,
                                                    double* __restrict__ qsum) {
  k_charge_sum_body(n, q, qsum, blockIdx.x);
}

// ----------------------------------------------------------------------------------------
// grouped kernels: one launch covers the blocks of many structures
// ----------------------------------------------------------------------------------------

// Evaluation slot and local block index of this workgroup (binary search over the block
// prefix; the prefix is uniform per workgroup, so the loads go through the scalar cache).
// Work item `item` of a grouped launch: its evaluation slot, and in `local` its block index
// within that evaluation's structure (binary search over the per-kind block prefix).
__device__ __forceinline__ int group_slot(const GroupMap& m, int item, int& local) {
  int lo = 0, hi = m.count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (m.prefix[mid] <= item)
      lo = mid;
    else
      hi = mid;
  }
  local = item - m.prefix[lo];
  return m.slot[lo];
}

// The work items of a grouped kernel.  LOOP (persistent grid): workgroup w takes items
// w, w + gridDim.x, ..., a barrier between consecutive items (their bodies reuse the
// workgroup's LDS); otherwise one workgroup per item, compiled without the loop.
#define EW_GROUP_RUN(LOOP, m, item)                                     \
  if constexpr (LOOP) {                                                 \
    for (int gi_ = blockIdx.x; gi_ < (m).total; gi_ += gridDim.x) {     \
      item(gi_);                                                        \
      __syncthreads();                                                  \
    }                                                                   \
  } else {                                                              \
    item((int)blockIdx.x);                                              \
  }

template <bool LOOP>
__global__ __launch_bounds__(256) void kg_gather_charges(const StructDesc* __restrict__ D,
                                                         const IoSlot* __restrict__ io,
                                                         GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const IoSlot& o = io[group_slot(m, gi, b)];
    const StructDesc& d = D[o.sid];
    k_gather_charges_body(d.n, d.perm, o.dq, d.qs, d.half ? d.phi_real : nullptr, b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, int P, bool LOOP>
__global__ __launch_bounds__(256) void kg_spread_region(const StructDesc* __restrict__ D,
                                                        const IoSlot* __restrict__ io,
                                                        GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    k_spread_region_body<R, P>(d.k0, static_cast<const R*>(d.theta), d.qs, d.tile_start, d.region,
                               static_cast<R*>(d.region.scratch), b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, int P, bool LOOP>
__global__ __launch_bounds__(256) void kg_spread_owned(const StructDesc* __restrict__ D,
                                                       const IoSlot* __restrict__ io,
                                                       GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    k_spread_owned_body<R, P>(d.k0, static_cast<const R*>(d.theta), d.qs, d.tile_start, d.region,
                              static_cast<R*>(d.grid), b, nullptr);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, bool LOOP>
__global__ __launch_bounds__(256) void kg_spread_assemble(const StructDesc* __restrict__ D,
                                                          const IoSlot* __restrict__ io,
                                                          GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    k_spread_assemble_body<R>(static_cast<const R*>(d.region.scratch), d.region,
                              static_cast<R*>(d.grid), b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, int P, bool LOOP>
__global__ __launch_bounds__(kSpreadThreads) void kg_spread(const StructDesc* __restrict__ D,
                                                            const IoSlot* __restrict__ io,
                                                            GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    k_spread_body<R, P>(d.k0, static_cast<const R*>(d.theta), d.qs, d.tile_start, d.spread,
                        static_cast<R*>(d.grid), b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, bool LOOP>
__global__ __launch_bounds__(256) void kg_spread_checked(const StructDesc* __restrict__ D,
                                                         const IoSlot* __restrict__ io,
                                                         GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    k_spread_checked_body<R>(d.k0, static_cast<const R*>(d.theta), d.qs, d.tile_start, d.spread,
                             static_cast<R*>(d.grid), b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, bool INVERSE, bool LOOP>
__global__ __launch_bounds__(kFftBlock) void kg_fft_x(const StructDesc* __restrict__ D,
                                                      const IoSlot* __restrict__ io,
                                                      GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    const int nrows = d.K[1] * d.K[2], Kh = d.K[0] / 2 + 1;
    if (INVERSE)
      k_fft_x_c2r_body<R>(static_cast<const cx<R>*>(d.cgrid), static_cast<R*>(d.grid), nrows, Kh,
                          d.px, static_cast<const cx<R>*>(d.tw[0]), b);
    else
      k_fft_x_r2c_body<R>(static_cast<const R*>(d.grid), static_cast<cx<R>*>(d.cgrid), nrows, Kh,
                          d.px, static_cast<const cx<R>*>(d.tw[0]), b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

// AXIS 1: y lines (MODE 0 forward, 1 inverse); AXIS 2: z lines with the influence (MODE 2).
template <typename R, int AXIS, int MODE, bool LOOP>
__global__ __launch_bounds__(kFftBlock) void kg_fft_lines(const StructDesc* __restrict__ D,
                                                          const IoSlot* __restrict__ io,
                                                          GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    const int Kh = d.K[0] / 2 + 1;
    cx<R>* cg = static_cast<cx<R>*>(d.cgrid);
    if (AXIS == 1)
      k_fft_lines_body<R, MODE>(cg, Kh * d.K[2], Kh, d.yd, (long long)d.K[1] * Kh, (long long)Kh,
                                d.py, static_cast<const cx<R>*>(d.tw[1]), d.inf, b);
    else
      k_fft_lines_body<R, MODE>(cg, Kh * d.K[1], Kh * d.K[1], d.zd, 0LL, (long long)Kh * d.K[1],
                                d.pz, static_cast<const cx<R>*>(d.tw[2]), d.inf, b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, bool INVERSE, bool LOOP>
__global__ __launch_bounds__(kPlaneBlock) void kg_fft_xy(const StructDesc* __restrict__ D,
                                                         const IoSlot* __restrict__ io,
                                                         GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    if (INVERSE)
      k_fft_xy_c2r_body<R>(static_cast<const cx<R>*>(d.cgrid), static_cast<R*>(d.grid), d.px,
                           d.py, static_cast<const cx<R>*>(d.tw[0]),
                           static_cast<const cx<R>*>(d.tw[1]), b);
    else
      k_fft_xy_r2c_body<R>(static_cast<const R*>(d.grid), static_cast<cx<R>*>(d.cgrid), d.px,
                           d.py, static_cast<const cx<R>*>(d.tw[0]),
                           static_cast<const cx<R>*>(d.tw[1]), b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename CT, bool LOOP>
__global__ __launch_bounds__(256) void kg_pair_spmv(const StructDesc* __restrict__ D,
                                                    const IoSlot* __restrict__ io, GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    if (d.half)
      k_pair_half_body<CT>(d.tile_start, d.hp, d.qs, d.phi_real, b);
    else if (d.col16)
      k_pair_spmv_body<CT, short>(d.n, d.row_ptr, static_cast<const short*>(d.col),
                                  static_cast<const CT*>(d.coef), d.qs, d.phi_real, b);
    else
      k_pair_spmv_body<CT, int>(d.n, d.row_ptr, static_cast<const int*>(d.col),
                                static_cast<const CT*>(d.coef), d.qs, d.phi_real, b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, int P, bool LOOP>
__global__ __launch_bounds__(256) void kg_gather_tile(const StructDesc* __restrict__ D,
                                                      const IoSlot* __restrict__ io,
                                                      GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const IoSlot& o = io[group_slot(m, gi, b)];
    const StructDesc& d = D[o.sid];
    k_gather_tile_body<R, P>(d.k0, static_cast<const R*>(d.theta), static_cast<const R*>(d.grid),
                             d.qs, d.phi_real, d.onsite, d.perm, d.tile_start, d.ga, o.shift,
                             d.epart, b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, bool LOOP>
__global__ __launch_bounds__(256) void kg_gather_combine(const StructDesc* __restrict__ D,
                                                         const IoSlot* __restrict__ io,
                                                         GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const IoSlot& o = io[group_slot(m, gi, b)];
    const StructDesc& d = D[o.sid];
    k_gather_combine_body<R>(d.n, d.k0, static_cast<const R*>(d.theta),
                             static_cast<const R*>(d.grid), d.qs, d.phi_real, d.onsite, d.perm,
                             d.ga, o.shift, d.epart, b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <bool LOOP>
__global__ __launch_bounds__(256) void kg_charge_sum(const StructDesc* __restrict__ D,
                                                     const IoSlot* __restrict__ io,
                                                     GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const StructDesc& d = D[io[group_slot(m, gi, b)].sid];
    k_charge_sum_body(d.n, d.qs, const_cast<double*>(d.ga.qsum), b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <bool LOOP>
__global__ __launch_bounds__(256) void kg_energy_final(const StructDesc* __restrict__ D,
                                                       const IoSlot* __restrict__ io,
                                                       GroupMap m) {
  auto item = [&](int gi) {
    int b;
    const IoSlot& o = io[group_slot(m, gi, b)];
    const StructDesc& d = D[o.sid];
    k_energy_final_body(d.nparts, d.epart, o.energy, b);
  };
  EW_GROUP_RUN(LOOP, m, item)
}

template <typename R, bool LOOP>
static hipError_t launch_group_t(const GroupLaunch& G, hipStream_t st) {
  const StructDesc* D = G.desc;
  const IoSlot* io = G.io;
  auto run = [&](int kind) { return G.blocks[kind] > 0; };
  // persistent: at most kGroupWgPerCu workgroups per CU (of 256 threads; the 1024-thread plane
  // kernels one per CU), each walking its share of the kind's work items
  auto grid = [&](int kind) {
    const int cap = kind == kGkPlane ? kNumCus : kGroupWgPerCu * kNumCus;
    return dim3(LOOP ? std::min(G.blocks[kind], cap) : G.blocks[kind]);
  };
  if (run(kGkCharges))
    hipLaunchKernelGGL(kg_gather_charges<LOOP>, grid(kGkCharges), dim3(256), 0, st, D, io,
                       G.map[kGkCharges]);
  if (run(kGkChargeSum))
    hipLaunchKernelGGL(kg_charge_sum<LOOP>, grid(kGkChargeSum), dim3(256), 0, st, D, io,
                       G.map[kGkChargeSum]);
  const bool fork = G.aux && (run(kGkPairs) || G.has_pm);
  if (fork) {
    hipError_t e = hipEventRecord(G.fork_ev, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(G.aux, G.fork_ev, 0);
    if (e != hipSuccess) return e;
    if (G.has_pm) {
      e = launch_pair_spmm(G.pm, sizeof(R) == sizeof(float), G.aux);
      if (e != hipSuccess) return e;
    } else {
      hipLaunchKernelGGL((kg_pair_spmv<R, LOOP>), grid(kGkPairs), dim3(256), 0, G.aux, D, io,
                         G.map[kGkPairs]);
    }
    e = hipEventRecord(G.join_ev, G.aux);
    if (e != hipSuccess) return e;
  }
#define EW_ORDER_SWITCH(MACRO) \
  switch (G.order) {           \
    case 4: MACRO(4); break;   \
    case 5: MACRO(5); break;   \
    case 6: MACRO(6); break;   \
    case 7: MACRO(7); break;   \
    case 8: MACRO(8); break;   \
    case 10: MACRO(10); break; \
    default: MACRO(0); break;  \
  }
  if (run(kGkRegion)) {
    if (G.smem[kGkRegion] > 65536) {
      const hipError_t e = prepare_region_kernels();
      if (e != hipSuccess) return e;
    }
#define EW_G_REGION(PP)                                                                \
  hipLaunchKernelGGL((kg_spread_region<R, PP, LOOP>), grid(kGkRegion), dim3(256),             \
                     G.smem[kGkRegion], st, D, io, G.map[kGkRegion])
    EW_ORDER_SWITCH(EW_G_REGION)
#undef EW_G_REGION
    hipLaunchKernelGGL((kg_spread_assemble<R, LOOP>), grid(kGkAssemble), dim3(256), 0, st, D, io,
                       G.map[kGkAssemble]);
  }
  if (run(kGkOwned)) {
    if (G.smem[kGkOwned] > 65536) {
      const hipError_t e = prepare_region_kernels();
      if (e != hipSuccess) return e;
    }
#define EW_G_OWNED(PP)                                                                    \
  hipLaunchKernelGGL((kg_spread_owned<R, PP, LOOP>), grid(kGkOwned), dim3(256),               \
                     G.smem[kGkOwned], st, D, io, G.map[kGkOwned])
    EW_ORDER_SWITCH(EW_G_OWNED)
#undef EW_G_OWNED
  }
  if (run(kGkSlotted)) {
#define EW_G_SLOTTED(PP)                                                                  \
  hipLaunchKernelGGL((kg_spread<R, PP, LOOP>), grid(kGkSlotted), dim3(kSpreadThreads),         \
                     G.smem[kGkSlotted], st, D, io, G.map[kGkSlotted])
    EW_ORDER_SWITCH(EW_G_SLOTTED)
#undef EW_G_SLOTTED
  }
  if (run(kGkChecked))
    hipLaunchKernelGGL((kg_spread_checked<R, LOOP>), grid(kGkChecked), dim3(256), 0, st, D, io,
                       G.map[kGkChecked]);
  for (const WaveGroup& w : G.wave) {
    const hipError_t e = launch_wave_fft_batch<R>(w.l, w.io, w.n, st);
    if (e != hipSuccess) return e;
  }
  if (run(kGkFftZ)) {
    if (run(kGkPlane))
      hipLaunchKernelGGL((kg_fft_xy<R, false, LOOP>), grid(kGkPlane), dim3(kPlaneBlock),
                         G.smem[kGkPlane], st, D, io, G.map[kGkPlane]);
    if (run(kGkFftX)) {
      hipLaunchKernelGGL((kg_fft_x<R, false, LOOP>), grid(kGkFftX), dim3(kFftBlock), G.smem[kGkFftX],
                         st, D, io, G.map[kGkFftX]);
      hipLaunchKernelGGL((kg_fft_lines<R, 1, 0, LOOP>), grid(kGkFftY), dim3(kFftBlock),
                         G.smem[kGkFftY], st, D, io, G.map[kGkFftY]);
    }
    hipLaunchKernelGGL((kg_fft_lines<R, 2, 2, LOOP>), grid(kGkFftZ), dim3(kFftBlock), G.smem[kGkFftZ],
                       st, D, io, G.map[kGkFftZ]);
    if (run(kGkFftX)) {
      hipLaunchKernelGGL((kg_fft_lines<R, 1, 1, LOOP>), grid(kGkFftY), dim3(kFftBlock),
                         G.smem[kGkFftY], st, D, io, G.map[kGkFftY]);
      hipLaunchKernelGGL((kg_fft_x<R, true, LOOP>), grid(kGkFftX), dim3(kFftBlock), G.smem[kGkFftX],
                         st, D, io, G.map[kGkFftX]);
    }
    if (run(kGkPlane))
      hipLaunchKernelGGL((kg_fft_xy<R, true, LOOP>), grid(kGkPlane), dim3(kPlaneBlock),
                         G.smem[kGkPlane], st, D, io, G.map[kGkPlane]);
  }
  if (fork) {
    const hipError_t e = hipStreamWaitEvent(st, G.join_ev, 0);
    if (e != hipSuccess) return e;
  } else if (G.has_pm) {
    const hipError_t e = launch_pair_spmm(G.pm, sizeof(R) == sizeof(float), st);
    if (e != hipSuccess) return e;
  } else if (run(kGkPairs)) {
    hipLaunchKernelGGL((kg_pair_spmv<R, LOOP>), grid(kGkPairs), dim3(256), 0, st, D, io,
                       G.map[kGkPairs]);
  }
  if (run(kGkGatherTile)) {
#define EW_G_GATHER(PP)                                                                   \
  hipLaunchKernelGGL((kg_gather_tile<R, PP, LOOP>), grid(kGkGatherTile), dim3(256),            \
                     G.smem[kGkGatherTile], st, D, io, G.map[kGkGatherTile])
    EW_ORDER_SWITCH(EW_G_GATHER)
#undef EW_G_GATHER
  }
  if (run(kGkGatherAtom))
    hipLaunchKernelGGL((kg_gather_combine<R, LOOP>), grid(kGkGatherAtom), dim3(256), 0, st, D, io,
                       G.map[kGkGatherAtom]);
  hipLaunchKernelGGL(kg_energy_final<LOOP>, grid(kGkEnergy), dim3(256), 0, st, D, io,
                     G.map[kGkEnergy]);
#undef EW_ORDER_SWITCH
  return hipGetLastError();
}

template <typename R>
hipError_t launch_group(const GroupLaunch& G, hipStream_t st) {
  return G.persistent ? launch_group_t<R, true>(G, st) : launch_group_t<R, false>(G, st);
}

template hipError_t launch_group<double>(const GroupLaunch&, hipStream_t);
template hipError_t launch_group<float>(const GroupLaunch&, hipStream_t);

// ----------------------------------------------------------------------------------------
// virial (stress)
// ----------------------------------------------------------------------------------------

// Sums 6 per-thread components over the workgroup (256 threads) into part[blockIdx.x * 6 ..].
__device__ __forceinline__ void block_sum6(double (&v)[6], double* __restrict__ part) {
  __shared__ double red[6][4];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double w = wave_sum(v[c]);
    if ((threadIdx.x & 63) == 0) red[c][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    part[(size_t)blockIdx.x * 6 + c] = red[c][0] + red[c][1] + red[c][2] + red[c][3];
  }
}

// Reciprocal virial of SPME (Essmann et al. 1995, eq. 2.8): over the full spectrum,
// W_ab = sum_m E_m (delta_ab - 2 (1 + pi^2 m^2 / alpha^2) m_a m_b / m^2), E_m = w(m) |F(Q)(m)|^2 / 2,
// with the influence function w of the evaluation; the half spectrum counts kx > 0 twice.  The
// modes the evaluation drops (m = 0, Nyquist planes of even meshes) are dropped here too.
template <typename R>
__global__ __launch_bounds__(256) void k_recip_virial(const cx<R>* __restrict__ cg,
                                                      InfluenceArgs inf,
                                                      double* __restrict__ part) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)inf.Kh * inf.K[1] * inf.K[2];
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (id < total) {
    const int kx = int(id % inf.Kh);
    const int ky = int((id / inf.Kh) % inf.K[1]);
    const int kz = int(id / ((long long)inf.Kh * inf.K[1]));
    const bool nyq = (!(inf.K[0] & 1) && kx == inf.K[0] / 2) ||
                     (!(inf.K[1] & 1) && ky == inf.K[1] / 2) ||
                     (!(inf.K[2] & 1) && kz == inf.K[2] / 2);
    if (!nyq && (kx | ky | kz)) {
      const int m1 = kx;
      const int m2 = ky <= (inf.K[1] - 1) / 2 ? ky : ky - inf.K[1];
      const int m3 = kz <= (inf.K[2] - 1) / 2 ? kz : kz - inf.K[2];
      const double mx = m1 * inf.rec[0] + m2 * inf.rec[3] + m3 * inf.rec[6];
      const double my = m1 * inf.rec[1] + m2 * inf.rec[4] + m3 * inf.rec[7];
      const double mz = m1 * inf.rec[2] + m2 * inf.rec[5] + m3 * inf.rec[8];
      const double msq = mx * mx + my * my + mz * mz;
      const double w = inf.pref * exp(-inf.a2 * msq) / msq * inf.bsq1[kx] * inf.bsq2[ky] *
                       inf.bsq3[kz];
      const cx<R> z = cg[id];
      const double e = (kx == 0 ? 0.5 : 1.0) * w * (double(z.x) * z.x + double(z.y) * z.y);
      const double f = 2.0 * (1.0 + inf.a2 * msq) / msq;
      v[0] = e * (1.0 - f * mx * mx);
      v[1] = e * (1.0 - f * my * my);
      v[2] = e * (1.0 - f * mz * mz);
      v[3] = -e * f * mx * my;
      v[4] = -e * f * mx * mz;
      v[5] = -e * f * my * mz;
    }
  }
  block_sum6(v, part);
}

// Real-space virial W_ab = 1/2 sum over ordered pairs (i, j + images) within r_cut of
// -q_i q_j phi'(r) x_a x_b / r, phi(r) = erfc(alpha r) / r, x = r_i - r_j (the Ewald real-space
// energy only, as the Coulomb stress of the reference: coulomb.F90:2861-2900).
template <bool DENSE>
__global__ __launch_bounds__(256) void k_real_virial(int n, const double* __restrict__ frac,
                                                     const int* __restrict__ tile_of,
                                                     const int* __restrict__ tile_start,
                                                     const double* __restrict__ q, PairArgs g,
                                                     double* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const double fi0 = frac[3 * i], fi1 = frac[3 * i + 1], fi2 = frac[3 * i + 2];
    const double two_a_sqrtpi = 2.0 * g.alpha / kSqrtPi;
    auto add = [&](double s0, double s1, double s2, int j) {
      const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
      const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
      const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
      const double r2 = x * x + y * y + z * z;
      if (r2 > g.rc2 || r2 < kMinSeparationSq) return;
      const double r = sqrt(r2);
      // -phi'(r) / r
      const double c = q[j] * (two_a_sqrtpi * exp(-g.alpha * g.alpha * r2) + erfc(g.alpha * r) / r) / r2;
      v[0] += c * x * x;
      v[1] += c * y * y;
      v[2] += c * z * z;
      v[3] += c * x * y;
      v[4] += c * x * z;
      v[5] += c * y * z;
    };
    if (DENSE) {
      for (int j = 0; j < n; ++j) {
        const double s0 = fi0 - frac[3 * j], s1 = fi1 - frac[3 * j + 1], s2 = fi2 - frac[3 * j + 2];
        for (int a = -g.D[0]; a <= g.D[0]; ++a)
          for (int b = -g.D[1]; b <= g.D[1]; ++b)
            for (int e = -g.D[2]; e <= g.D[2]; ++e) {
              if (i == j && a == 0 && b == 0 && e == 0) continue;
              add(s0 + a, s1 + b, s2 + e, j);
            }
      }
    } else {
      int t = tile_of[i];
      const int t1 = t % g.nt[0];
      t /= g.nt[0];
      const int t2 = t % g.nt[1], t3 = t / g.nt[1];
      const int b1 = g.full[0] ? 0 : t1 - g.D[0], n1 = g.full[0] ? g.nt[0] : 2 * g.D[0] + 1;
      const int b2 = g.full[1] ? 0 : t2 - g.D[1], n2 = g.full[1] ? g.nt[1] : 2 * g.D[1] + 1;
      const int b3 = g.full[2] ? 0 : t3 - g.D[2], n3 = g.full[2] ? g.nt[2] : 2 * g.D[2] + 1;
      for (int c3 = 0; c3 < n3; ++c3) {
        const int u3 = ((b3 + c3) % g.nt[2] + g.nt[2]) % g.nt[2];
        for (int c2 = 0; c2 < n2; ++c2) {
          const int u2 = ((b2 + c2) % g.nt[1] + g.nt[1]) % g.nt[1];
          for (int c1 = 0; c1 < n1; ++c1) {
            const int u1 = ((b1 + c1) % g.nt[0] + g.nt[0]) % g.nt[0];
            const int tt = (u3 * g.nt[1] + u2) * g.nt[0] + u1;
            for (int j = tile_start[tt]; j < tile_start[tt + 1]; ++j) {
              if (j == i) continue;
              double s0 = fi0 - frac[3 * j], s1 = fi1 - frac[3 * j + 1], s2 = fi2 - frac[3 * j + 2];
              s0 -= rint(s0);
              s1 -= rint(s1);
              s2 -= rint(s2);
              add(s0, s1, s2, j);
            }
          }
        }
      }
    }
    const double h = 0.5 * q[i];
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] *= h;
  }
  block_sum6(v, part);
}

template <typename R>
hipError_t launch_recip_virial(const SpmeLaunch& L, double* partials, int* nparts,
                               hipStream_t st) {
  const long long total = (long long)(L.K[0] / 2 + 1) * L.K[1] * L.K[2];
  const int nb = int((total + 255) / 256);
  hipLaunchKernelGGL(k_recip_virial<R>, dim3(nb), dim3(256), 0, st,
                     static_cast<const cx<R>*>(L.cgrid), L.inf, partials);
  *nparts = nb;
  return hipGetLastError();
}
template hipError_t launch_recip_virial<double>(const SpmeLaunch&, double*, int*, hipStream_t);
template hipError_t launch_recip_virial<float>(const SpmeLaunch&, double*, int*, hipStream_t);

// Real-space force and virial over the resident pair list (linked mode), one wave per row:
// the partners' displacements come from the sorted fractional coordinates (minimum image, the
// list's own convention), and only pairs within r_cut contribute (the list reaches r_list).
// Per pair c = q_j (2 alpha / sqrt(pi) exp(-alpha^2 r^2) + erfc(alpha r) / r) / r^2, the same
// term as k_force_real / k_real_virial.
template <typename IT, bool VIRIAL>
__global__ __launch_bounds__(256) void k_pair_real_terms(int n, const long long* __restrict__ row_ptr,
                                                         const IT* __restrict__ col,
                                                         const double* __restrict__ frac,
                                                         const double* __restrict__ q,
                                                         const int* __restrict__ perm, PairArgs g,
                                                         double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int row = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (row < n) {
    const double fi0 = frac[3 * row], fi1 = frac[3 * row + 1], fi2 = frac[3 * row + 2];
    const double two_a_sqrtpi = 2.0 * g.alpha / kSqrtPi;
    const long long a = row_ptr[row], b = row_ptr[row + 1];
    for (long long k = a + lane; k < b; k += 64) {
      const int j = partner(row, col[k], n);
      double s0 = fi0 - frac[3 * j], s1 = fi1 - frac[3 * j + 1], s2 = fi2 - frac[3 * j + 2];
      s0 -= rint(s0);
      s1 -= rint(s1);
      s2 -= rint(s2);
      const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
      const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
      const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
      const double r2 = x * x + y * y + z * z;
      if (r2 > g.rc2 || r2 < kMinSeparationSq) continue;
      const double r = sqrt(r2);
      const double c =
          q[j] * (two_a_sqrtpi * exp(-g.alpha * g.alpha * r2) + erfc(g.alpha * r) / r) / r2;
      if (VIRIAL) {
        v[0] += c * x * x;
        v[1] += c * y * y;
        v[2] += c * z * z;
        v[3] += c * x * y;
        v[4] += c * x * z;
        v[5] += c * y * z;
      } else {
        v[0] += c * x;
        v[1] += c * y;
        v[2] += c * z;
      }
    }
  }
  if (VIRIAL) {
    const double h = row < n ? 0.5 * q[row] : 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] *= h;
    block_sum6(v, out);
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = wave_sum(v[c]);
    if (row < n && lane == 0) {
      const double qi = q[row];
      const int o = perm[row];
#pragma unroll
      for (int c = 0; c < 3; ++c) out[3 * o + c] += qi * v[c];
    }
  }
}

hipError_t launch_real_virial(int n, const double* frac, const int* tile_of,
                              const int* tile_start, const double* q, const PairArgs& pa,
                              bool dense, double* partials, int* nparts, hipStream_t st,
                              const PairRows& rows) {
  if (rows.row_ptr && !dense) {
    const int nb = real_virial_parts(n, true);
    if (rows.col16)
      hipLaunchKernelGGL((k_pair_real_terms<short, true>), dim3(nb), dim3(256), 0, st, n,
                         rows.row_ptr, static_cast<const short*>(rows.col), frac, q,
                         static_cast<const int*>(nullptr), pa, partials);
    else
      hipLaunchKernelGGL((k_pair_real_terms<int, true>), dim3(nb), dim3(256), 0, st, n,
                         rows.row_ptr, static_cast<const int*>(rows.col), frac, q,
                         static_cast<const int*>(nullptr), pa, partials);
    *nparts = nb;
    return hipGetLastError();
  }
  const int nb = real_virial_parts(n, false);
  if (dense)
    hipLaunchKernelGGL(k_real_virial<true>, dim3(nb), dim3(256), 0, st, n, frac, tile_of,
                       tile_start, q, pa, partials);
  else
    hipLaunchKernelGGL(k_real_virial<false>, dim3(nb), dim3(256), 0, st, n, frac, tile_of,
                       tile_start, q, pa, partials);
  *nparts = nb;
  return hipGetLastError();
}

// ----------------------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------------------

template <typename R>
static size_t fft_smem(const FftPlan& pl) {
  return sizeof(cx<R>) * ((size_t)pl.lines * pl.pitch + pl.n);
}

template <typename R, bool CO>
static hipError_t launch_spme_t(const SpmeLaunch& L, hipStream_t st, hipEvent_t* ev,
                                hipEvent_t ev_region) {
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  if (ev_region && L.region.region_stride <= 0) (void)hipEventRecord(ev_region, st);
  // co-scheduled pair sum: rows split over the spreading launch (a share in percent) and the
  // FFT launches of the full transform (the rest, equally), each launch running its rows on
  // pair blocks appended after its own blocks; waves of pair blocks per launch: kCoWaves
  const bool co_on = CO && L.pair.r1 > 0 && !L.forward_only;
  const int nfft = L.plane ? 3 : 5;
  const bool co_spread = co_on && L.region.region_stride > 0 && L.pair_spread_share > 0;
  int co_next = 0;
  int co_left = nfft + (co_spread ? 1 : 0);
  auto co_for = [&](int chain_blocks, int threads, bool spread_launch, dim3* grid_out) {
    PairCo c = L.pair;
    c.first = 0x7fffffff;
    *grid_out = dim3(chain_blocks);
    if (!co_on || (spread_launch && !co_spread)) return c;
    const int n = L.pair.r1;
    int take;
    if (spread_launch) {
      take = (int)((long long)n * L.pair_spread_share / 100);
    } else {
      take = (n - co_next) / co_left;
    }
    --co_left;
    if (co_left == 0) take = n - co_next;
    c.r0 = co_next;
    c.r1 = co_next + take;
    co_next = c.r1;
    c.first = chain_blocks;
    c.nblk = std::max(1, kCoWaves / (threads / 64));
    if (c.r1 > c.r0) *grid_out = dim3(chain_blocks + c.nblk);
    else c.first = 0x7fffffff;
    return c;
  };
  const int K1 = L.K[0], K2 = L.K[1], K3 = L.K[2], Kh = K1 / 2 + 1;
  R* grid = static_cast<R*>(L.grid);
  cx<R>* cg = static_cast<cx<R>*>(L.cgrid);
  const cx<R>* tw1 = static_cast<const cx<R>*>(L.tw[0]);
  const cx<R>* tw2 = static_cast<const cx<R>*>(L.tw[1]);
  const cx<R>* tw3 = static_cast<const cx<R>*>(L.tw[2]);
  const dim3 sgrid(L.spread.nt[0] * L.spread.nt[1] * L.spread.nt[2]);
  if (L.region.region_stride > 0) {
    const RegionArgs& rg = L.region;
    const int nblk = rg.nb[0] * rg.nb[1] * rg.nb[2];
    const size_t smem = sizeof(double) * (size_t)(rg.S[0] + rg.order - 1) *
                        (rg.S[1] + rg.order - 1) * (rg.S[2] + rg.order - 1);
    const R* th = static_cast<const R*>(L.theta);
    R* scr = static_cast<R*>(rg.scratch);
    if (rg.direct == 1) {
      const long long G = (long long)K1 * K2 * K3;
      const int nb = (int)std::min<long long>(2048, (G / 4 + 255) / 256 + 1);
      hipLaunchKernelGGL(k_clear<R>, dim3(nb), dim3(256), 0, st, grid, G);
    }
    dim3 rgrid;
    const PairCo rco = co_for(nblk, 256, true, &rgrid);
    if (rg.direct == 2) {
      const size_t osmem = owned_smem(rg, sizeof(R) == 4);
#define EW_OWNED(PP)                                                                          \
  hipLaunchKernelGGL((k_spread_owned<R, PP, CO>), rgrid, dim3(256), osmem, st, L.k0, th, L.q,      \
                     L.tile_start, rg, grid, L.qperm, rco)
      switch (rg.order) {
        case 4: EW_OWNED(4); break;
        case 5: EW_OWNED(5); break;
        case 6: EW_OWNED(6); break;
        case 7: EW_OWNED(7); break;
        case 8: EW_OWNED(8); break;
        case 10: EW_OWNED(10); break;
        default: EW_OWNED(0); break;
      }
#undef EW_OWNED
    } else {
#define EW_REGION(PP)                                                                         \
  hipLaunchKernelGGL((k_spread_region<R, PP, CO>), rgrid, dim3(256), smem, st, L.k0, th, L.q,      \
                     L.tile_start, rg, scr, L.qperm, grid, rco)
      switch (rg.order) {
        case 4: EW_REGION(4); break;
        case 5: EW_REGION(5); break;
        case 6: EW_REGION(6); break;
        case 7: EW_REGION(7); break;
        case 8: EW_REGION(8); break;
        case 10: EW_REGION(10); break;
        default: EW_REGION(0); break;
      }
#undef EW_REGION
    }
    if (ev_region) (void)hipEventRecord(ev_region, st);
    if (!rg.direct)
      hipLaunchKernelGGL(k_spread_assemble<R>, dim3(nblk * kAssembleSplit), dim3(256), 0, st,
                         static_cast<const R*>(scr), rg, grid);
  } else if (L.spread.slots > 0) {
    const size_t smem = sizeof(unsigned long long) * kSpreadRows * 4 +
                        sizeof(int4) * (size_t)L.spread.slots +
                        sizeof(R) * 3 * (size_t)L.spread.slots * L.spread.order;
    const R* th = static_cast<const R*>(L.theta);
#define EW_SPREAD(PP)                                                                          \
  hipLaunchKernelGGL((k_spread<R, PP>), sgrid, dim3(kSpreadThreads), smem, st, L.k0, th, L.q,    \
                     L.tile_start, L.spread, grid)
    switch (L.spread.order) {
      case 4: EW_SPREAD(4); break;
      case 5: EW_SPREAD(5); break;
      case 6: EW_SPREAD(6); break;
      case 7: EW_SPREAD(7); break;
      case 8: EW_SPREAD(8); break;
      case 10: EW_SPREAD(10); break;
      case 12: EW_SPREAD(12); break;
      default: EW_SPREAD(0); break;
    }
#undef EW_SPREAD
  } else {
    hipLaunchKernelGGL(k_spread_checked<R>, sgrid, dim3(256), 0, st, L.k0,
                       static_cast<const R*>(L.theta), L.q, L.tile_start, L.spread, grid);
  }
  if (L.spread_done) (void)hipEventRecord(L.spread_done, st);
  mark(0);
  if (L.wave && !L.forward_only && !co_on) return launch_wave_fft<R>(L, st, ev);
  const int nrows = K2 * K3;
  const int gx = (nrows + 2 * L.px.lines - 1) / (2 * L.px.lines);
  const int ny = Kh * K3;
  const size_t psmem = plane_smem(L.px, L.py, sizeof(cx<R>));
  dim3 g1;
  if (L.plane) {
    // x and y in one pass per plane (the y phase event follows immediately)
    const PairCo c1 = co_for(K3, kPlaneBlock, false, &g1);
    hipLaunchKernelGGL((k_fft_xy_r2c<R, CO>), g1, dim3(kPlaneBlock), psmem, st, grid, cg, L.px,
                       L.py, tw1, tw2, c1);
    mark(1);
  } else {
    const PairCo c1 = co_for(gx, kFftBlock, false, &g1);
    hipLaunchKernelGGL((k_fft_x_r2c<R, CO>), g1, dim3(kFftBlock), fft_smem<R>(L.px), st, grid,
                       cg, nrows, Kh, L.px, tw1, c1);
    mark(1);
    const PairCo c2 = co_for((ny + L.py.lines - 1) / L.py.lines, kFftBlock, false, &g1);
    hipLaunchKernelGGL((k_fft_lines<R, 0, CO>), g1, dim3(kFftBlock), fft_smem<R>(L.py), st, cg, ny,
                       Kh, make_fastdiv(Kh), (long long)K2 * Kh, (long long)Kh, L.py, tw2, L.inf,
                       c2);
  }
  mark(2);
  const int nz = Kh * K2;
  if (L.forward_only) {
    hipLaunchKernelGGL((k_fft_lines<R, 0, CO>), dim3((nz + L.pz.lines - 1) / L.pz.lines),
                       dim3(kFftBlock), fft_smem<R>(L.pz), st, cg, nz, Kh * K2,
                       make_fastdiv(Kh * K2), 0LL, (long long)Kh * K2, L.pz, tw3, L.inf,
                       PairCo{});
    return hipGetLastError();
  }
  const PairCo c3 = co_for((nz + L.pz.lines - 1) / L.pz.lines, kFftBlock, false, &g1);
  hipLaunchKernelGGL((k_fft_lines<R, 2, CO>), g1, dim3(kFftBlock), fft_smem<R>(L.pz), st, cg, nz,
                     Kh * K2, make_fastdiv(Kh * K2), 0LL, (long long)Kh * K2, L.pz, tw3, L.inf,
                     c3);
  mark(3);
  if (L.plane) {
    mark(4);
    const PairCo c4 = co_for(K3, kPlaneBlock, false, &g1);
    hipLaunchKernelGGL((k_fft_xy_c2r<R, CO>), g1, dim3(kPlaneBlock), psmem, st, cg, grid, L.px,
                       L.py, tw1, tw2, c4);
  } else {
    const PairCo c4 = co_for((ny + L.py.lines - 1) / L.py.lines, kFftBlock, false, &g1);
    hipLaunchKernelGGL((k_fft_lines<R, 1, CO>), g1, dim3(kFftBlock), fft_smem<R>(L.py), st, cg, ny,
                       Kh, make_fastdiv(Kh), (long long)K2 * Kh, (long long)Kh, L.py, tw2, L.inf,
                       c4);
    mark(4);
    const PairCo c5 = co_for(gx, kFftBlock, false, &g1);
    hipLaunchKernelGGL((k_fft_x_c2r<R, CO>), g1, dim3(kFftBlock), fft_smem<R>(L.px), st, cg,
                       grid, nrows, Kh, L.px, tw1, c5);
  }
  mark(5);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_spme(const SpmeLaunch& L, hipStream_t st, hipEvent_t* ev, hipEvent_t ev_region) {
  return L.pair.r1 > 0 ? launch_spme_t<R, true>(L, st, ev, ev_region)
                       : launch_spme_t<R, false>(L, st, ev, ev_region);
}

hipError_t prepare_plane_kernels() {
  static hipError_t done = [] {
    const int b = (int)kPlaneMaxLds;
    const void* fns[] = {
        reinterpret_cast<const void*>(&k_fft_xy_r2c<double, false>),
        reinterpret_cast<const void*>(&k_fft_xy_r2c<float, false>),
        reinterpret_cast<const void*>(&k_fft_xy_c2r<double, false>),
        reinterpret_cast<const void*>(&k_fft_xy_c2r<float, false>),
        reinterpret_cast<const void*>(&k_fft_xy_r2c<double, true>),
        reinterpret_cast<const void*>(&k_fft_xy_r2c<float, true>),
        reinterpret_cast<const void*>(&k_fft_xy_c2r<double, true>),
        reinterpret_cast<const void*>(&k_fft_xy_c2r<float, true>),
        reinterpret_cast<const void*>(&kg_fft_xy<double, false, false>),
        reinterpret_cast<const void*>(&kg_fft_xy<float, false, false>),
        reinterpret_cast<const void*>(&kg_fft_xy<double, true, false>),
        reinterpret_cast<const void*>(&kg_fft_xy<float, true, false>),
        reinterpret_cast<const void*>(&kg_fft_xy<double, false, true>),
        reinterpret_cast<const void*>(&kg_fft_xy<float, false, true>),
        reinterpret_cast<const void*>(&kg_fft_xy<double, true, true>),
        reinterpret_cast<const void*>(&kg_fft_xy<float, true, true>)};
    for (const void* f : fns) {
      const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, b);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }();
  return done;
}

template <typename R, int P>
static hipError_t region_attr(int bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spread_region<R, P, false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spread_region<R, P, true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spread_owned<R, P, false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes - (int)kOwnedStatic);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spread_owned<R, P, true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes - (int)kOwnedStatic);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&kg_spread_owned<R, P, false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes - (int)kOwnedStatic);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&kg_spread_owned<R, P, true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes - (int)kOwnedStatic);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&kg_spread_region<R, P, false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&kg_spread_region<R, P, true>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template <typename R>
static hipError_t region_attrs(int b) {
  const hipError_t es[] = {region_attr<R, 4>(b), region_attr<R, 5>(b), region_attr<R, 6>(b),
                           region_attr<R, 7>(b), region_attr<R, 8>(b), region_attr<R, 10>(b),
                           region_attr<R, 0>(b)};
  for (hipError_t e : es)
    if (e != hipSuccess) return e;
  return hipSuccess;
}

hipError_t prepare_region_kernels() {
  static hipError_t done = [] {
    const hipError_t e = region_attrs<double>((int)kRegionMaxLds);
    return e != hipSuccess ? e : region_attrs<float>((int)kRegionMaxLds);
  }();
  return done;
}

template hipError_t launch_spme<double>(const SpmeLaunch&, hipStream_t, hipEvent_t*, hipEvent_t);
template hipError_t launch_spme<float>(const SpmeLaunch&, hipStream_t, hipEvent_t*, hipEvent_t);

hipError_t launch_gather_charges(int n, const int* perm, const double* dq, double* qs,
                                 double* zero, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_charges, dim3((n + 255) / 256), dim3(256), 0, st, n, perm, dq, qs,
                     zero);
  return hipGetLastError();
}

hipError_t launch_charge_sum(int n, const double* qs, double* qsum, hipStream_t st) {
  hipLaunchKernelGGL(k_charge_sum, dim3(1), dim3(256), 0, st, n, qs, qsum);
  return hipGetLastError();
}

#ifndef EW_PAIR_STREAM_BLOCKS
#define EW_PAIR_STREAM_BLOCKS 512  // default beside the chain: 2 workgroups (8 waves) per CU
#endif

hipError_t launch_pair_spmv(int n, const long long* row_ptr, const void* col, bool col16,
                            const void* coef, bool fp32, const double* q, double* out,
                            bool beside, hipStream_t st, int beside_blocks) {
  const int rows_blocks = (n + 3) / 4;
  const dim3 block(256);
  const int bounded = beside_blocks > 0 ? beside_blocks : EW_PAIR_STREAM_BLOCKS;
  const dim3 grid(beside ? std::min(rows_blocks, bounded) : rows_blocks);
#define EW_SPMV(CT, IT)                                                                      \
  hipLaunchKernelGGL((k_pair_spmv_stream<CT, IT>), grid, block, 0, st, n, row_ptr,           \
                     static_cast<const IT*>(col), static_cast<const CT*>(coef), q, out)
  if (col16) {
    if (fp32)
      EW_SPMV(float, short);
    else
      EW_SPMV(double, short);
  } else {
    if (fp32)
      EW_SPMV(float, int);
    else
      EW_SPMV(double, int);
  }
#undef EW_SPMV
  return hipGetLastError();
}

hipError_t launch_half_count(int n, const long long* row_ptr, const int* col, const int* tile_of,
                             const int* tile_start, const HalfPairs& hp, int* rc,
                             hipStream_t st) {
  hipLaunchKernelGGL((k_half_pairs<true, double>), dim3((n + 127) / 128), dim3(128), 0, st, n,
                     row_ptr, col, static_cast<const double*>(nullptr), tile_of, tile_start, hp,
                     rc, static_cast<unsigned short*>(nullptr), static_cast<double*>(nullptr));
  return hipGetLastError();
}

hipError_t launch_half_fill(int n, const long long* row_ptr, const int* col, const void* coef,
                            bool fp32, const int* tile_of, const int* tile_start,
                            const HalfPairs& hp, const int* rc, unsigned short* entry,
                            void* coef_out, hipStream_t st) {
  int* rcw = const_cast<int*>(rc);
  if (fp32)
    hipLaunchKernelGGL((k_half_pairs<false, float>), dim3((n + 127) / 128), dim3(128), 0, st, n,
                       row_ptr, col, static_cast<const float*>(coef), tile_of, tile_start, hp,
                       rcw, entry, static_cast<float*>(coef_out));
  else
    hipLaunchKernelGGL((k_half_pairs<false, double>), dim3((n + 127) / 128), dim3(128), 0, st,
                       n, row_ptr, col, static_cast<const double*>(coef), tile_of, tile_start,
                       hp, rcw, entry, static_cast<double*>(coef_out));
  return hipGetLastError();
}

hipError_t launch_pair_half(int ntiles, const int* tile_start, const HalfPairs& hp, bool fp32,
                            const double* q, double* out, hipStream_t st) {
  if (fp32)
    hipLaunchKernelGGL(k_pair_half<float>, dim3(ntiles), dim3(256), 0, st, tile_start, hp, q,
                       out);
  else
    hipLaunchKernelGGL(k_pair_half<double>, dim3(ntiles), dim3(256), 0, st, tile_start, hp, q,
                       out);
  return hipGetLastError();
}

hipError_t launch_compress_cols(int n, const long long* row_ptr, const int* col, short* off,
                                int* too_far, hipStream_t st) {
  hipLaunchKernelGGL(k_compress_cols, dim3((n + 3) / 4), dim3(256), 0, st, n, row_ptr, col,
                     off, too_far);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_gather_combine(int n, const int4* k0, const void* theta, const void* grid,
                                 const double* q, const double* phi_real, const double* onsite,
                                 const int* perm, const GatherArgs& g, double* shift,
                                 double* epart, double* energy, hipStream_t st) {
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_gather_combine<R>, dim3(nb), dim3(256), 0, st, n, k0,
                     static_cast<const R*>(theta), static_cast<const R*>(grid), q, phi_real,
                     onsite, perm, g, shift, epart);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, nb, epart, energy);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_gather_tile(const int4* k0, const void* theta, const void* grid,
                              const double* q, const double* phi_real, const double* onsite,
                              const int* perm, const int* tile_start, const GatherArgs& g,
                              double* shift, double* epart, double* energy, hipStream_t st) {
  const int ntiles = g.nt[0] * g.nt[1] * g.nt[2];
  const size_t smem = gather_tile_smem(g.T, g.order, sizeof(R));
  const R* th = static_cast<const R*>(theta);
  const R* gr = static_cast<const R*>(grid);
#define EW_GATHER(PP)                                                                            \
  hipLaunchKernelGGL((k_gather_tile<R, PP, false>), dim3(ntiles), dim3(256), smem, st, k0, th, gr,\
                     q, phi_real, onsite, perm, tile_start, g, shift, epart)
  switch (g.order) {
    case 4: EW_GATHER(4); break;
    case 5: EW_GATHER(5); break;
    case 6: EW_GATHER(6); break;
    case 7: EW_GATHER(7); break;
    case 8: EW_GATHER(8); break;
    case 10: EW_GATHER(10); break;
    default: EW_GATHER(0); break;
  }
#undef EW_GATHER
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, ntiles, epart, energy);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_gather_mesh(const int4* k0, const void* theta, const void* grid,
                              const int* tile_start, const GatherArgs& g, double* mesh,
                              hipStream_t st) {
  const int ntiles = g.nt[0] * g.nt[1] * g.nt[2];
  const size_t smem = gather_tile_smem(g.T, g.order, sizeof(R));
  const R* th = static_cast<const R*>(theta);
  const R* gr = static_cast<const R*>(grid);
#define EW_GATHER(PP)                                                                            \
  hipLaunchKernelGGL((k_gather_tile<R, PP, true>), dim3(ntiles), dim3(256), smem, st, k0, th, gr, \
                     nullptr, nullptr, nullptr, nullptr, tile_start, g, mesh, nullptr)
  switch (g.order) {
    case 4: EW_GATHER(4); break;
    case 5: EW_GATHER(5); break;
    case 6: EW_GATHER(6); break;
    case 7: EW_GATHER(7); break;
    case 8: EW_GATHER(8); break;
    case 10: EW_GATHER(10); break;
    default: EW_GATHER(0); break;
  }
#undef EW_GATHER
  return hipGetLastError();
}
template hipError_t launch_gather_mesh<double>(const int4*, const void*, const void*, const int*,
                                               const GatherArgs&, double*, hipStream_t);
template hipError_t launch_gather_mesh<float>(const int4*, const void*, const void*, const int*,
                                              const GatherArgs&, double*, hipStream_t);

hipError_t launch_shift_combine(int n, const double* mesh, const double* phi_real,
                                const double* q, const double* onsite, const int* perm,
                                const GatherArgs& g, double* shift, double* epart,
                                double* energy, hipStream_t st) {
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_shift_combine, dim3(nb), dim3(256), 0, st, n, mesh, phi_real, q, onsite,
                     perm, g, shift, epart);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, nb, epart, energy);
  return hipGetLastError();
}

template hipError_t launch_gather_tile<double>(const int4*, const void*, const void*,
                                               const double*, const double*, const double*,
                                               const int*, const int*, const GatherArgs&, double*,
                                               double*, double*, hipStream_t);
template hipError_t launch_gather_tile<float>(const int4*, const void*, const void*,
                                              const double*, const double*, const double*,
                                              const int*, const int*, const GatherArgs&, double*,
                                              double*, double*, hipStream_t);

template hipError_t launch_gather_combine<double>(int, const int4*, const void*, const void*,
                                                  const double*, const double*, const double*,
                                                  const int*, const GatherArgs&, double*, double*,
                                                  double*, hipStream_t);
template hipError_t launch_gather_combine<float>(int, const int4*, const void*, const void*,
                                                 const double*, const double*, const double*,
                                                 const int*, const GatherArgs&, double*, double*,
                                                 double*, hipStream_t);

hipError_t launch_bspline(int n, const double* frac, const BsplineArgs& g, void* theta, bool fp32,
                          hipStream_t st) {
  if (fp32)
    hipLaunchKernelGGL(k_bspline<float>, dim3((n + 127) / 128), dim3(128), 0, st, n, frac, g,
                       static_cast<float*>(theta));
  else
    hipLaunchKernelGGL(k_bspline<double>, dim3((n + 127) / 128), dim3(128), 0, st, n, frac, g,
                       static_cast<double*>(theta));
  return hipGetLastError();
}

// The tile-staged build when its bounds hold, else one wave per atom.
static bool pair_lds_fits(const PairArgs& g, int max_tile_atoms) {
  long long ntc = 1;
  for (int a = 0; a < 3; ++a) ntc *= g.full[a] ? g.nt[a] : 2 * g.D[a] + 1;
  return ntc <= kPairMaxCandTiles && max_tile_atoms <= kPairMaxTileAtoms;
}

hipError_t launch_pair_count(int n, const double* frac, const int* sp, const int* tile_of,
                             const int* tile_start, const PairArgs& g, int max_tile_atoms,
                             long long* counts, int* err, hipStream_t st) {
  const int ntiles = g.nt[0] * g.nt[1] * g.nt[2];
  if (pair_lds_fits(g, max_tile_atoms))
    hipLaunchKernelGGL((k_pair_tiles_lds<true, double>), dim3(ntiles), dim3(256), 0, st, frac,
                       sp, tile_start, g, counts, PartnerOut{}, nullptr, err);
  else
    hipLaunchKernelGGL((k_pair_tiles<true, double>), dim3((n + 3) / 4), dim3(256), 0, st, n,
                       frac, sp, tile_of, tile_start, g, counts, PartnerOut{}, nullptr, err);
  return hipGetLastError();
}

hipError_t launch_pair_fill(int n, const double* frac, const int* sp, const int* tile_of,
                            const int* tile_start, const PairArgs& g, int max_tile_atoms,
                            long long* row_ptr, int* col, short* off, int* too_far, void* coef,
                            bool fp32, int* err, hipStream_t st) {
  const int ntiles = g.nt[0] * g.nt[1] * g.nt[2];
  const PartnerOut po{col, off, too_far, n};
  if (pair_lds_fits(g, max_tile_atoms)) {
    if (fp32)
      hipLaunchKernelGGL((k_pair_tiles_lds<false, float>), dim3(ntiles), dim3(256), 0, st, frac,
                         sp, tile_start, g, row_ptr, po, static_cast<float*>(coef), err);
    else
      hipLaunchKernelGGL((k_pair_tiles_lds<false, double>), dim3(ntiles), dim3(256), 0, st,
                         frac, sp, tile_start, g, row_ptr, po, static_cast<double*>(coef), err);
    return hipGetLastError();
  }
  if (fp32)
    hipLaunchKernelGGL((k_pair_tiles<false, float>), dim3((n + 3) / 4), dim3(256), 0, st, n,
                       frac, sp, tile_of, tile_start, g, row_ptr, po, static_cast<float*>(coef),
                       err);
  else
    hipLaunchKernelGGL((k_pair_tiles<false, double>), dim3((n + 3) / 4), dim3(256), 0, st, n,
                       frac, sp, tile_of, tile_start, g, row_ptr, po, static_cast<double*>(coef),
                       err);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_forces(int n, const double* frac, const int4* k0, const void* grid,
                         const double* q, const int* perm, const int* tile_of,
                         const int* tile_start, const ForceArgs& fa, const PairArgs& pa,
                         bool dense, double* force, hipStream_t st, const PairRows& rows) {
  const dim3 grd((n + 127) / 128), blk(128);
  hipLaunchKernelGGL(k_force_recip<R>, grd, blk, 0, st, n, frac, k0, static_cast<const R*>(grid),
                     q, perm, fa, force);
  if (rows.row_ptr && !dense) {
    const dim3 rgrd((n + 3) / 4), rblk(256);
    if (rows.col16)
      hipLaunchKernelGGL((k_pair_real_terms<short, false>), rgrd, rblk, 0, st, n, rows.row_ptr,
                         static_cast<const short*>(rows.col), frac, q, perm, pa, force);
    else
      hipLaunchKernelGGL((k_pair_real_terms<int, false>), rgrd, rblk, 0, st, n, rows.row_ptr,
                         static_cast<const int*>(rows.col), frac, q, perm, pa, force);
  } else if (dense)
    hipLaunchKernelGGL(k_force_real<true>, grd, blk, 0, st, n, frac, tile_of, tile_start, q, perm,
                       pa, force);
  else
    hipLaunchKernelGGL(k_force_real<false>, grd, blk, 0, st, n, frac, tile_of, tile_start, q,
                       perm, pa, force);
  return hipGetLastError();
}
template hipError_t launch_forces<double>(int, const double*, const int4*, const void*,
                                          const double*, const int*, const int*, const int*,
                                          const ForceArgs&, const PairArgs&, bool, double*,
                                          hipStream_t, const PairRows&);
template hipError_t launch_forces<float>(int, const double*, const int4*, const void*,
                                         const double*, const int*, const int*, const int*,
                                         const ForceArgs&, const PairArgs&, bool, double*,
                                         hipStream_t, const PairRows&);

hipError_t launch_pair_dense(int n, const double* frac, const int* sp, const PairArgs& g,
                             int* col, void* coef, bool fp32, int* err, hipStream_t st) {
  const long long total = (long long)n * n;
  const dim3 grid((unsigned)((total + 255) / 256));
  if (fp32)
    hipLaunchKernelGGL(k_pair_dense<float>, grid, dim3(256), 0, st, n, frac, sp, g, col,
                       static_cast<float*>(coef), err);
  else
    hipLaunchKernelGGL(k_pair_dense<double>, grid, dim3(256), 0, st, n, frac, sp, g, col,
                       static_cast<double*>(coef), err);
  return hipGetLastError();
}

// ----------------------------------------------------------------------------------------
// external point charges
// ----------------------------------------------------------------------------------------

// Calls f(s0, s1, s2, j) for the source images a target at fractional (f0, f1, f2) in sort tile
// `tile` may see, sources lane, lane + lanes, ... of each tile (the lanes of a wave share a
// target): linked mode walks the source tiles around the target's tile with the minimum image;
// dense mode takes every source over explicit image shells.
template <bool DENSE, typename F>
__device__ __forceinline__ void for_source_images(const ExtArgs& g, double f0, double f1,
                                                  double f2, int tile, int lane, int lanes,
                                                  F&& f) {
  if (DENSE) {
    for (int j = lane; j < g.m; j += lanes) {
      const double s0 = f0 - g.frac[3 * j], s1 = f1 - g.frac[3 * j + 1], s2 = f2 - g.frac[3 * j + 2];
      for (int a = -g.D[0]; a <= g.D[0]; ++a)
        for (int b = -g.D[1]; b <= g.D[1]; ++b)
          for (int e = -g.D[2]; e <= g.D[2]; ++e) f(s0 + a, s1 + b, s2 + e, j);
    }
    return;
  }
  int t = tile;
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1], t3 = t / g.nt[1];
  const int b1 = g.full[0] ? 0 : t1 - g.D[0], n1 = g.full[0] ? g.nt[0] : 2 * g.D[0] + 1;
  const int b2 = g.full[1] ? 0 : t2 - g.D[1], n2 = g.full[1] ? g.nt[1] : 2 * g.D[1] + 1;
  const int b3 = g.full[2] ? 0 : t3 - g.D[2], n3 = g.full[2] ? g.nt[2] : 2 * g.D[2] + 1;
  for (int c3 = 0; c3 < n3; ++c3) {
    const int u3 = ((b3 + c3) % g.nt[2] + g.nt[2]) % g.nt[2];
    for (int c2 = 0; c2 < n2; ++c2) {
      const int u2 = ((b2 + c2) % g.nt[1] + g.nt[1]) % g.nt[1];
      for (int c1 = 0; c1 < n1; ++c1) {
        const int u1 = ((b1 + c1) % g.nt[0] + g.nt[0]) % g.nt[0];
        const int tt = (u3 * g.nt[1] + u2) * g.nt[0] + u1;
        for (int j = g.start[tt] + lane; j < g.start[tt + 1]; j += lanes) {
          double s0 = f0 - g.frac[3 * j], s1 = f1 - g.frac[3 * j + 1], s2 = f2 - g.frac[3 * j + 2];
          s0 -= rint(s0);
          s1 -= rint(s1);
          s2 -= rint(s2);
          f(s0, s1, s2, j);
        }
      }
    }
  }
}

// An image closer than tolSameDist (1e-5, src/dftbp/common/accuracy.F90:69) is the source on the
// target; erfArgLimit_ = 10 (coulomb.F90:214) bounds the Gaussian correction of a blurred source.
constexpr double kExtSameDist2 = 1.0e-10;
constexpr double kErfArgLimit = 10.0;

// Real-space potential of the external charges at each resident atom (sorted order), one wave
// or one thread per atom (ext_wave): sum_s q_s [erfc(alpha d) - erfc(d / w_s) (blurred, d < 10 w_s)] / sqrt(d^2 + eps^2)
// over the source images within the cutoff (ewaldReal, src/dftbp/dftb/coulomb.F90:2523-2575:
// erf(d/w) - erf(alpha d) written through erfc).  A source on the atom contributes the self
// term -2 alpha / sqrt(pi) q_s instead (ewald, coulomb.F90:2341-2343).
// WAVE: one wave per target, its lanes along the sources (many sources per tile); else one
// thread per target (few sources per tile, where a wave's lanes would mostly idle).
template <bool WAVE>
__device__ __forceinline__ int ext_target(int& lane) {
  if (WAVE) {
    lane = threadIdx.x & 63;
    return __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  }
  lane = 0;
  return blockIdx.x * blockDim.x + threadIdx.x;
}

template <bool DENSE, bool WAVE>
__global__ __launch_bounds__(256) void k_ext_real(int n, const double* __restrict__ frac,
                                                  const int* __restrict__ tile_of, ExtArgs g,
                                                  double* __restrict__ phi) {
  int lane;
  const int i = ext_target<WAVE>(lane);
  if (i >= n) return;
  double acc = 0.0;
  for_source_images<DENSE>(g, frac[3 * i], frac[3 * i + 1], frac[3 * i + 2], DENSE ? 0 : tile_of[i],
                           lane, WAVE ? 64 : 1, [&](double s0, double s1, double s2, int j) {
    const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
    const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
    const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
    const double r2 = x * x + y * y + z * z;
    if (r2 > g.rc2) return;
    if (r2 < kExtSameDist2) {
      acc += g.self_coef * g.q[j];
      return;
    }
    const double r = sqrt(r2);
    double t = erfc(g.alpha * r);
    if (g.blur) {
      const double w = g.blur[j];
      if (r < kErfArgLimit * w) t -= erfc(r / w);
    }
    acc += g.q[j] * t / (g.eps2 > 0.0 ? sqrt(r2 + g.eps2) : r);
  });
  if (WAVE) acc = wave_sum(acc);
  if (lane == 0) phi[i] = acc;
}

// Real-space gradient with respect to the target position of the same sum (derivEwaldReal,
// src/dftbp/dftb/coulomb.F90:2577-2620): per image v = x_t - y_s + L, |v| = d,
//   q_s v [-(2/sqrt(pi)) alpha d exp(-alpha^2 d^2) - erfc(alpha d)] / d^3
//   + q_s v [(2/sqrt(pi)) f exp(-f^2) + erfc(f)] / d^3,  f = d / w  (blurred, d < 10 w),
// skipping a source on the target.  The Gaussian width is the source's (blur) or the
// target's (tblur: the external charges as targets of the atoms).  out[3 perm[i] + c] =
// scale[i] times the sum (scale null: 1).
template <bool DENSE, bool WAVE>
__global__ __launch_bounds__(256) void k_ext_real_grad(int n, const double* __restrict__ frac,
                                                       const int* __restrict__ tile_of,
                                                       const int* __restrict__ perm,
                                                       const double* __restrict__ scale,
                                                       ExtArgs g, double* __restrict__ out) {
  int lane;
  const int i = ext_target<WAVE>(lane);
  if (i >= n) return;
  const double two_sqrtpi = 2.0 / kSqrtPi;
  const double wt = g.tblur ? g.tblur[i] : 0.0;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for_source_images<DENSE>(g, frac[3 * i], frac[3 * i + 1], frac[3 * i + 2], DENSE ? 0 : tile_of[i],
                           lane, WAVE ? 64 : 1, [&](double s0, double s1, double s2, int j) {
    const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
    const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
    const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
    const double r2 = x * x + y * y + z * z;
    if (r2 > g.rc2 || r2 < kExtSameDist2) return;
    const double r = sqrt(r2);
    const double ar = g.alpha * r;
    double c = -two_sqrtpi * ar * exp(-ar * ar) - erfc(ar);
    const double w = g.tblur ? wt : (g.blur ? g.blur[j] : 0.0);
    if (w > 0.0 && r < kErfArgLimit * w) {
      const double f = r / w;
      c += two_sqrtpi * f * exp(-f * f) + erfc(f);
    }
    c *= g.q[j] / (r2 * r);
    a0 += c * x;
    a1 += c * y;
    a2 += c * z;
  });
  if (WAVE) {
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
  }
  if (lane == 0) {
    const double sc = scale ? scale[i] : 1.0;
    const int o = perm[i];
    out[3 * o] = sc * a0;
    out[3 * o + 1] = sc * a1;
    out[3 * o + 2] = sc * a2;
  }
}

// A wave per target pays off from about 16 sources per sort tile.
static bool ext_wave(const ExtArgs& g) {
  return g.m >= 16 * g.nt[0] * g.nt[1] * g.nt[2];
}

hipError_t launch_ext_real(int n, const double* frac, const int* tile_of, const ExtArgs& g,
                           bool dense, double* phi, hipStream_t st) {
  const dim3 blk(256), wgrd((n + 3) / 4), tgrd((n + 255) / 256);
  if (ext_wave(g)) {
    if (dense)
      hipLaunchKernelGGL((k_ext_real<true, true>), wgrd, blk, 0, st, n, frac, tile_of, g, phi);
    else
      hipLaunchKernelGGL((k_ext_real<false, true>), wgrd, blk, 0, st, n, frac, tile_of, g, phi);
  } else if (dense) {
    hipLaunchKernelGGL((k_ext_real<true, false>), tgrd, blk, 0, st, n, frac, tile_of, g, phi);
  } else {
    hipLaunchKernelGGL((k_ext_real<false, false>), tgrd, blk, 0, st, n, frac, tile_of, g, phi);
  }
  return hipGetLastError();
}

hipError_t launch_ext_real_grad(int n, const double* frac, const int* tile_of, const int* perm,
                                const double* scale, const ExtArgs& g, bool dense, double* out,
                                hipStream_t st) {
  const dim3 blk(256), wgrd((n + 3) / 4), tgrd((n + 255) / 256);
  if (ext_wave(g)) {
    if (dense)
      hipLaunchKernelGGL((k_ext_real_grad<true, true>), wgrd, blk, 0, st, n, frac, tile_of, perm,
                         scale, g, out);
    else
      hipLaunchKernelGGL((k_ext_real_grad<false, true>), wgrd, blk, 0, st, n, frac, tile_of, perm,
                         scale, g, out);
  } else if (dense) {
    hipLaunchKernelGGL((k_ext_real_grad<true, false>), tgrd, blk, 0, st, n, frac, tile_of, perm,
                       scale, g, out);
  } else {
    hipLaunchKernelGGL((k_ext_real_grad<false, false>), tgrd, blk, 0, st, n, frac, tile_of, perm,
                       scale, g, out);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_force_recip(int n, const double* frac, const int4* k0, const void* grid,
                              const double* q, const int* perm, const ForceArgs& fa,
                              double* force, hipStream_t st) {
  hipLaunchKernelGGL(k_force_recip<R>, dim3((n + 127) / 128), dim3(128), 0, st, n, frac, k0,
                     static_cast<const R*>(grid), q, perm, fa, force);
  return hipGetLastError();
}
template hipError_t launch_force_recip<double>(int, const double*, const int4*, const void*,
                                               const double*, const int*, const ForceArgs&,
                                               double*, hipStream_t);
template hipError_t launch_force_recip<float>(int, const double*, const int4*, const void*,
                                              const double*, const int*, const ForceArgs&,
                                              double*, hipStream_t);

}  // namespace ewamd
