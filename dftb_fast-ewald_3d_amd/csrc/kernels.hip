// kernels.hip -- gfx950 device kernels of the periodic Coulomb engine.
//
// Per SCC iteration (charges change, geometry resident):
//   k_gather_charges   dq (caller order) -> q (sorted order)
//   k_spread           B-spline charge spreading, one workgroup per mesh tile, LDS accumulation
//   k_fft_x_r2c        mesh rows -> half spectrum (two real rows per complex transform)
//   k_fft_lines        Y pass; Z pass fused with the influence multiply and the inverse Z pass
//   k_fft_x_c2r        half spectrum -> convolved mesh
//   k_pair_spmv        real-space pair sum over the resident coefficients, one wave per atom
//   k_gather_combine   B-spline gather + real space + on-site + self -> shift, energy partials
//   k_energy_final     fixed-order reduction of the energy partials
// Per geometry: k_bspline (weights), k_pair_count/k_pair_fill (linked tiles), k_pair_dense.
//
// All FFTs are Stockham autosort transforms of one or more lines staged in LDS: every pass
// loads its butterflies into registers, synchronises, and stores them back in place.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace ewamd {

// ----------------------------------------------------------------------------------------
// complex helpers
// ----------------------------------------------------------------------------------------
template <typename R>
struct cx {
  R x, y;
};
template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return {a.x * s, a.y * s}; }
// multiply by s*i, s = +-1
template <typename R>
__device__ __forceinline__ cx<R> cmuli(cx<R> a, R s) { return {-s * a.y, s * a.x}; }

// Small DFTs in registers; dir = -1 forward (exp(-2 pi i tu/R)), +1 inverse.
template <int RAD, typename R>
struct Dft;

template <typename R>
struct Dft<2, R> {
  __device__ static void run(cx<R>* v, R) {
    cx<R> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <typename R>
struct Dft<4, R> {
  __device__ static void run(cx<R>* v, R s) {
    cx<R> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cx<R> t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), s);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};
template <typename R>
struct Dft<3, R> {
  __device__ static void run(cx<R>* v, R s) {
    const R c = R(-0.5), sn = R(0.866025403784438646763723170752936183);
    cx<R> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    cx<R> b = cadd(v[0], cscale(t1, c));
    cx<R> d = cmuli(cscale(t2, sn), s);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(b, d);
    v[2] = csub(b, d);
  }
};
template <typename R>
struct Dft<5, R> {
  __device__ static void run(cx<R>* v, R s) {
    const R c1 = R(0.309016994374947424102293417182819059);
    const R c2 = R(-0.809016994374947424102293417182819059);
    const R s1 = R(0.951056516295153572116439333379382143);
    const R s2 = R(0.587785252292473129168705954639072769);
    cx<R> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    cx<R> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    cx<R> b1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
    cx<R> b2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
    cx<R> d1 = cmuli(cadd(cscale(t3, s1), cscale(t4, s2)), s);
    cx<R> d2 = cmuli(csub(cscale(t3, s2), cscale(t4, s1)), s);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(b1, d1);
    v[4] = csub(b1, d1);
    v[2] = cadd(b2, d2);
    v[3] = csub(b2, d2);
  }
};

// One in-place Stockham pass over `nl` lines of length n (LDS pitch P).
// tw[m] = exp(-2 pi i m / n); the inverse uses its conjugate.
template <int RAD, typename R>
__device__ __forceinline__ void stockham_pass(cx<R>* s, int n, int nl, int P, int Ns,
                                              const cx<R>* __restrict__ tw, R dir) {
  constexpr int EPT = kFftMaxElems / kFftBlock;  // complex elements per thread and pass
  constexpr int MAXB = (EPT + RAD - 1) / RAD;
  const int nb = n / RAD;
  const int total = nl * nb;
  const int tscale = n / (Ns * RAD);
  cx<R> v[MAXB][RAD];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = threadIdx.x + b * kFftBlock;
    if (idx < total) {
      const int line = idx / nb, j = idx - line * nb;
      const int k = j % Ns;
      const cx<R>* src = s + line * P;
#pragma unroll
      for (int t = 0; t < RAD; ++t) {
        cx<R> a = src[j + t * nb];
        if (t > 0) {
          cx<R> w = tw[t * k * tscale];
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
    const int idx = threadIdx.x + b * kFftBlock;
    if (idx < total) {
      const int line = idx / nb, j = idx - line * nb;
      const int k = j % Ns;
      cx<R>* dst = s + line * P + (j / Ns) * Ns * RAD + k;
#pragma unroll
      for (int t = 0; t < RAD; ++t) dst[t * Ns] = v[b][t];
    }
  }
  __syncthreads();
}

template <typename R>
__device__ void fft_lds(cx<R>* s, const FftPlan& pl, int nl, const cx<R>* __restrict__ tw,
                        R dir) {
  int Ns = 1;
  for (int i = 0; i < pl.npass; ++i) {
    const int r = pl.radix[i];
    if (r == 4)
      stockham_pass<4, R>(s, pl.n, nl, pl.pitch, Ns, tw, dir);
    else if (r == 2)
      stockham_pass<2, R>(s, pl.n, nl, pl.pitch, Ns, tw, dir);
    else if (r == 3)
      stockham_pass<3, R>(s, pl.n, nl, pl.pitch, Ns, tw, dir);
    else
      stockham_pass<5, R>(s, pl.n, nl, pl.pitch, Ns, tw, dir);
    Ns *= r;
  }
}

// ----------------------------------------------------------------------------------------
// FFT kernels
// ----------------------------------------------------------------------------------------

// Rows of K1 reals (row = z*K2 + y); workgroup takes pl.lines complex lines = 2*lines rows.
template <typename R>
__global__ __launch_bounds__(kFftBlock) void k_fft_x_r2c(const R* __restrict__ grid,
                                                         cx<R>* __restrict__ cg, int nrows,
                                                         int Kh, FftPlan pl,
                                                         const cx<R>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int K1 = pl.n, nl = pl.lines, P = pl.pitch;
  const int row0 = blockIdx.x * nl * 2;
  for (int idx = threadIdx.x; idx < nl * K1; idx += kFftBlock) {
    const int l = idx / K1, e = idx - l * K1;
    const int r0 = row0 + 2 * l, r1 = r0 + 1;
    cx<R> z;
    z.x = r0 < nrows ? grid[(size_t)r0 * K1 + e] : R(0);
    z.y = r1 < nrows ? grid[(size_t)r1 * K1 + e] : R(0);
    s[l * P + e] = z;
  }
  __syncthreads();
  fft_lds<R>(s, pl, nl, tw, R(-1));
  for (int idx = threadIdx.x; idx < nl * Kh; idx += kFftBlock) {
    const int l = idx / Kh, k = idx - l * Kh;
    const int r0 = row0 + 2 * l, r1 = r0 + 1;
    const cx<R> z = s[l * P + k];
    const cx<R> zm = s[l * P + (k == 0 ? 0 : K1 - k)];
    if (r0 < nrows) cg[(size_t)r0 * Kh + k] = {R(0.5) * (z.x + zm.x), R(0.5) * (z.y - zm.y)};
    if (r1 < nrows) cg[(size_t)r1 * Kh + k] = {R(0.5) * (z.y + zm.y), R(0.5) * (zm.x - z.x)};
  }
}

template <typename R>
__global__ __launch_bounds__(kFftBlock) void k_fft_x_c2r(const cx<R>* __restrict__ cg,
                                                         R* __restrict__ grid, int nrows, int Kh,
                                                         FftPlan pl,
                                                         const cx<R>* __restrict__ tw) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int K1 = pl.n, nl = pl.lines, P = pl.pitch;
  const int row0 = blockIdx.x * nl * 2;
  for (int idx = threadIdx.x; idx < nl * K1; idx += kFftBlock) {
    const int l = idx / K1, k = idx - l * K1;
    const int r0 = row0 + 2 * l, r1 = r0 + 1;
    cx<R> a = {R(0), R(0)}, b = {R(0), R(0)};
    const int kk = k < Kh ? k : K1 - k;
    if (r0 < nrows) a = cg[(size_t)r0 * Kh + kk];
    if (r1 < nrows) b = cg[(size_t)r1 * Kh + kk];
    if (k >= Kh) {
      a.y = -a.y;
      b.y = -b.y;
    }
    s[l * P + k] = {a.x - b.y, a.y + b.x};
  }
  __syncthreads();
  fft_lds<R>(s, pl, nl, tw, R(1));
  for (int idx = threadIdx.x; idx < nl * K1; idx += kFftBlock) {
    const int l = idx / K1, e = idx - l * K1;
    const int r0 = row0 + 2 * l, r1 = r0 + 1;
    const cx<R> z = s[l * P + e];
    if (r0 < nrows) grid[(size_t)r0 * K1 + e] = z.x;
    if (r1 < nrows) grid[(size_t)r1 * K1 + e] = z.y;
  }
}

// Strided complex lines: line L has base (L / A) * B + L % A and element stride `es`.
// MODE 0: forward; 1: inverse; 2: forward, influence multiply, inverse (Z pass).
template <typename R, int MODE>
__global__ __launch_bounds__(kFftBlock) void k_fft_lines(cx<R>* __restrict__ cg, int nlines,
                                                         int A, long long B, long long es,
                                                         FftPlan pl,
                                                         const cx<R>* __restrict__ tw,
                                                         InfluenceArgs inf) {
  extern __shared__ unsigned char smem_raw[];
  cx<R>* s = reinterpret_cast<cx<R>*>(smem_raw);
  const int n = pl.n, nl = pl.lines, P = pl.pitch;
  const int L0 = blockIdx.x * nl;
  for (int idx = threadIdx.x; idx < nl * n; idx += kFftBlock) {
    const int e = idx / nl, l = idx - e * nl;
    const int L = L0 + l;
    cx<R> z = {R(0), R(0)};
    if (L < nlines) z = cg[(long long)(L / A) * B + (L % A) + e * es];
    s[l * P + e] = z;
  }
  __syncthreads();
  fft_lds<R>(s, pl, nl, tw, MODE == 1 ? R(1) : R(-1));
  if (MODE == 2) {
    // influence function on the half spectrum: line L = ky*Kh + kx, element e = kz
    for (int idx = threadIdx.x; idx < nl * n; idx += kFftBlock) {
      const int l = idx / n, kz = idx - l * n;
      const int L = L0 + l;
      if (L >= nlines) continue;
      const int ky = L / inf.Kh, kx = L - ky * inf.Kh;
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
      cx<R>& z = s[l * P + kz];
      z.x = R(z.x * w);
      z.y = R(z.y * w);
    }
    __syncthreads();
    fft_lds<R>(s, pl, nl, tw, R(1));
  }
  for (int idx = threadIdx.x; idx < nl * n; idx += kFftBlock) {
    const int e = idx / nl, l = idx - e * nl;
    const int L = L0 + l;
    if (L < nlines) cg[(long long)(L / A) * B + (L % A) + e * es] = s[l * P + e];
  }
}

// ----------------------------------------------------------------------------------------
// charge spreading
// ----------------------------------------------------------------------------------------

// One workgroup per output tile; atoms are listed by the tile of their stencil origin, and the
// workgroup visits the source tiles whose stencils reach into it.  Each contribution lands in
// exactly one output tile, so the mesh is written once, without global atomics or a clear.
template <typename R>
__global__ __launch_bounds__(256) void k_spread(const int4* __restrict__ k0,
                                                const R* __restrict__ theta,
                                                const double* __restrict__ q,
                                                const int* __restrict__ tile_start,
                                                SpreadArgs g, R* __restrict__ grid) {
  __shared__ R acc[kSpreadTileMax];
  __shared__ int src_first[kMaxSrcTiles * kMaxSrcTiles * kMaxSrcTiles + 1];
  __shared__ int src_begin[kMaxSrcTiles * kMaxSrcTiles * kMaxSrcTiles];
  const int p = g.order;
  int t = blockIdx.x;
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1];
  const int t3 = t / g.nt[1];
  const int lo1 = t1 * g.T[0], lo2 = t2 * g.T[1], lo3 = t3 * g.T[2];
  const int w1 = min(g.T[0], g.K[0] - lo1), w2 = min(g.T[1], g.K[1] - lo2),
            w3 = min(g.T[2], g.K[2] - lo3);
  const int npts = w1 * w2 * w3;
  for (int i = threadIdx.x; i < npts; i += blockDim.x) acc[i] = R(0);
  // flatten the (source tile, atom) visits
  const int* e1 = g.src + ((size_t)0 * g.ntmax + t1) * (kMaxSrcTiles + 1);
  const int* e2 = g.src + ((size_t)1 * g.ntmax + t2) * (kMaxSrcTiles + 1);
  const int* e3 = g.src + ((size_t)2 * g.ntmax + t3) * (kMaxSrcTiles + 1);
  const int n1 = e1[0], n2 = e2[0], n3 = e3[0];
  const int ns = n1 * n2 * n3;
  if (threadIdx.x == 0) {
    int run = 0;
    for (int c = 0; c < ns; ++c) {
      const int c1 = c % n1, c2 = (c / n1) % n2, c3 = c / (n1 * n2);
      const int s1 = e1[1 + c1], s2 = e2[1 + c2], s3 = e3[1 + c3];
      const int st = (s3 * g.nt[1] + s2) * g.nt[0] + s1;
      src_first[c] = run;
      src_begin[c] = tile_start[st];
      run += tile_start[st + 1] - tile_start[st];
    }
    src_first[ns] = run;
  }
  __syncthreads();
  const int nvisit = src_first[ns];
  for (int v = threadIdx.x; v < nvisit; v += blockDim.x) {
    int c = 0;
    while (src_first[c + 1] <= v) ++c;
    const int a = src_begin[c] + (v - src_first[c]);
    const int4 kk = k0[a];
    // offsets o whose mesh point (k0 + o) mod K falls inside [lo, lo + w)
    int r1 = kk.x - lo1, r2 = kk.y - lo2, r3 = kk.z - lo3;
    if (r1 < 0) r1 += g.K[0];
    if (r2 < 0) r2 += g.K[1];
    if (r3 < 0) r3 += g.K[2];
    const R* th = theta + (size_t)a * 3 * p;
    const R qa = R(q[a]);
    for (int o3 = 0; o3 < p; ++o3) {
      int z = r3 + o3;
      if (z >= g.K[2]) z -= g.K[2];
      if (z >= w3) continue;
      const R c3 = qa * th[2 * p + o3];
      for (int o2 = 0; o2 < p; ++o2) {
        int y = r2 + o2;
        if (y >= g.K[1]) y -= g.K[1];
        if (y >= w2) continue;
        const R c23 = c3 * th[p + o2];
        R* row = acc + (z * w2 + y) * w1;
        for (int o1 = 0; o1 < p; ++o1) {
          int x = r1 + o1;
          if (x >= g.K[0]) x -= g.K[0];
          if (x >= w1) continue;
          atomicAdd(row + x, c23 * th[o1]);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < npts; i += blockDim.x) {
    const int x = i % w1, y = (i / w1) % w2, z = i / (w1 * w2);
    grid[((size_t)(lo3 + z) * g.K[1] + (lo2 + y)) * g.K[0] + lo1 + x] = acc[i];
  }
}

// ----------------------------------------------------------------------------------------
// per-atom kernels
// ----------------------------------------------------------------------------------------

__global__ void k_gather_charges(int n, const int* __restrict__ perm,
                                 const double* __restrict__ dq, double* __restrict__ qs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) qs[k] = dq[perm[k]];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Real-space pair sum: out[i] = sum_j coef_ij q_j, one 64-lane wave per atom (row).
template <typename CT>
__global__ __launch_bounds__(256) void k_pair_spmv(int n, const long long* __restrict__ row_ptr,
                                                   const int* __restrict__ col,
                                                   const CT* __restrict__ coef,
                                                   const double* __restrict__ q,
                                                   double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= n) return;  // uniform across the wave
  const long long a = row_ptr[row], b = row_ptr[row + 1];
  double acc = 0.0;
  for (long long k = a + lane; k < b; k += 64) acc += double(coef[k]) * q[col[k]];
  acc = wave_sum(acc);
  if (lane == 0) out[row] = acc;
}

// B-spline gather of the convolved mesh, plus the real-space sum, the on-site short-gamma
// term and the Ewald self term (fe/common/ewald_self.f90:17), written in caller order.
template <typename R>
__global__ __launch_bounds__(256) void k_gather_combine(
    int n, const int4* __restrict__ k0, const R* __restrict__ theta,
    const R* __restrict__ grid, const double* __restrict__ q, const double* __restrict__ phi_real,
    const double* __restrict__ onsite, const int* __restrict__ perm, GatherArgs g,
    double* __restrict__ shift, double* __restrict__ epart) {
  __shared__ double red[256 / 64];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
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
    shift[perm[k]] = phi;
    e = 0.5 * phi * qk;
  }
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    epart[blockIdx.x] = s;
  }
}

__global__ void k_energy_final(int nparts, const double* __restrict__ epart,
                               double* __restrict__ energy) {
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
// rounding (fe/fft/ewald_fft_3d_real.f90:150-157).  COUNT: row lengths; else fill.
template <bool COUNT, typename CT>
__global__ void k_pair_tiles(int n, const double* __restrict__ frac, const int* __restrict__ sp,
                             const int* __restrict__ tile_of, const int* __restrict__ tile_start,
                             PairArgs g, long long* __restrict__ row_ptr, int* __restrict__ col,
                             CT* __restrict__ coef, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double fi0 = frac[3 * i], fi1 = frac[3 * i + 1], fi2 = frac[3 * i + 2];
  const int si = sp ? sp[i] : 0;
  int t = tile_of[i];
  const int t1 = t % g.nt[0];
  t /= g.nt[0];
  const int t2 = t % g.nt[1], t3 = t / g.nt[1];
  long long pos = COUNT ? 0 : row_ptr[i];
  const long long end = COUNT ? 0 : row_ptr[i + 1];
  int cnt = 0;
  for (int d3 = -g.D[2]; d3 <= g.D[2]; ++d3) {
    const int u3 = ((t3 + d3) % g.nt[2] + g.nt[2]) % g.nt[2];
    for (int d2 = -g.D[1]; d2 <= g.D[1]; ++d2) {
      const int u2 = ((t2 + d2) % g.nt[1] + g.nt[1]) % g.nt[1];
      for (int d1 = -g.D[0]; d1 <= g.D[0]; ++d1) {
        const int u1 = ((t1 + d1) % g.nt[0] + g.nt[0]) % g.nt[0];
        const int tt = (u3 * g.nt[1] + u2) * g.nt[0] + u1;
        for (int j = tile_start[tt]; j < tile_start[tt + 1]; ++j) {
          if (j == i) continue;
          double s0 = fi0 - frac[3 * j], s1 = fi1 - frac[3 * j + 1], s2 = fi2 - frac[3 * j + 2];
          s0 -= rint(s0);
          s1 -= rint(s1);
          s2 -= rint(s2);
          const double x = s0 * g.lat[0] + s1 * g.lat[3] + s2 * g.lat[6];
          const double y = s0 * g.lat[1] + s1 * g.lat[4] + s2 * g.lat[7];
          const double z = s0 * g.lat[2] + s1 * g.lat[5] + s2 * g.lat[8];
          const double r2 = x * x + y * y + z * z;
          if (r2 > g.rl2) continue;
          if (r2 < kMinSeparationSq) {
            *err = 1;
            continue;
          }
          bool in;
          const double c = pair_term(r2, g, si, sp ? sp[j] : 0, &in);
          if (!in) continue;
          if (COUNT) {
            ++cnt;
          } else if (pos < end) {  // the fill never runs past the counted row length
            col[pos] = j;
            coef[pos] = CT(c);
            ++pos;
          }
        }
      }
    }
  }
  if (COUNT) row_ptr[i] = cnt;
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
// launchers
// ----------------------------------------------------------------------------------------

template <typename R>
static size_t fft_smem(const FftPlan& pl) {
  return sizeof(cx<R>) * (size_t)pl.lines * pl.pitch;
}

template <typename R>
hipError_t launch_spme(const SpmeLaunch& L, hipStream_t st, hipEvent_t* ev) {
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  const int K1 = L.K[0], K2 = L.K[1], K3 = L.K[2], Kh = K1 / 2 + 1;
  R* grid = static_cast<R*>(L.grid);
  cx<R>* cg = static_cast<cx<R>*>(L.cgrid);
  const cx<R>* tw1 = static_cast<const cx<R>*>(L.tw[0]);
  const cx<R>* tw2 = static_cast<const cx<R>*>(L.tw[1]);
  const cx<R>* tw3 = static_cast<const cx<R>*>(L.tw[2]);
  hipLaunchKernelGGL(k_spread<R>, dim3(L.spread.nt[0] * L.spread.nt[1] * L.spread.nt[2]),
                     dim3(256), 0, st, L.k0, static_cast<const R*>(L.theta), L.q, L.tile_start,
                     L.spread, grid);
  mark(0);
  const int nrows = K2 * K3;
  const int gx = (nrows + 2 * L.px.lines - 1) / (2 * L.px.lines);
  hipLaunchKernelGGL(k_fft_x_r2c<R>, dim3(gx), dim3(kFftBlock), fft_smem<R>(L.px), st, grid, cg,
                     nrows, Kh, L.px, tw1);
  mark(1);
  const int ny = Kh * K3;
  hipLaunchKernelGGL((k_fft_lines<R, 0>), dim3((ny + L.py.lines - 1) / L.py.lines),
                     dim3(kFftBlock), fft_smem<R>(L.py), st, cg, ny, Kh, (long long)K2 * Kh,
                     (long long)Kh, L.py, tw2, L.inf);
  mark(2);
  const int nz = Kh * K2;
  hipLaunchKernelGGL((k_fft_lines<R, 2>), dim3((nz + L.pz.lines - 1) / L.pz.lines),
                     dim3(kFftBlock), fft_smem<R>(L.pz), st, cg, nz, Kh * K2, 0LL,
                     (long long)Kh * K2, L.pz, tw3, L.inf);
  mark(3);
  hipLaunchKernelGGL((k_fft_lines<R, 1>), dim3((ny + L.py.lines - 1) / L.py.lines),
                     dim3(kFftBlock), fft_smem<R>(L.py), st, cg, ny, Kh, (long long)K2 * Kh,
                     (long long)Kh, L.py, tw2, L.inf);
  mark(4);
  hipLaunchKernelGGL(k_fft_x_c2r<R>, dim3(gx), dim3(kFftBlock), fft_smem<R>(L.px), st, cg, grid,
                     nrows, Kh, L.px, tw1);
  mark(5);
  return hipGetLastError();
}

template hipError_t launch_spme<double>(const SpmeLaunch&, hipStream_t, hipEvent_t*);
template hipError_t launch_spme<float>(const SpmeLaunch&, hipStream_t, hipEvent_t*);

hipError_t launch_gather_charges(int n, const int* perm, const double* dq, double* qs,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_gather_charges, dim3((n + 255) / 256), dim3(256), 0, st, n, perm, dq, qs);
  return hipGetLastError();
}

hipError_t launch_pair_spmv(int n, const long long* row_ptr, const int* col, const void* coef,
                            bool fp32, const double* q, double* out, hipStream_t st) {
  const dim3 grid((n + 3) / 4), block(256);
  if (fp32)
    hipLaunchKernelGGL(k_pair_spmv<float>, grid, block, 0, st, n, row_ptr, col,
                       static_cast<const float*>(coef), q, out);
  else
    hipLaunchKernelGGL(k_pair_spmv<double>, grid, block, 0, st, n, row_ptr, col,
                       static_cast<const double*>(coef), q, out);
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

hipError_t launch_pair_count(int n, const double* frac, const int* sp, const int* tile_of,
                             const int* tile_start, const PairArgs& g, long long* counts,
                             int* err, hipStream_t st) {
  hipLaunchKernelGGL((k_pair_tiles<true, double>), dim3((n + 127) / 128), dim3(128), 0, st, n,
                     frac, sp, tile_of, tile_start, g, counts, nullptr, nullptr, err);
  return hipGetLastError();
}

hipError_t launch_pair_fill(int n, const double* frac, const int* sp, const int* tile_of,
                            const int* tile_start, const PairArgs& g, long long* row_ptr,
                            int* col, void* coef, bool fp32, int* err, hipStream_t st) {
  if (fp32)
    hipLaunchKernelGGL((k_pair_tiles<false, float>), dim3((n + 127) / 128), dim3(128), 0, st, n,
                       frac, sp, tile_of, tile_start, g, row_ptr, col, static_cast<float*>(coef),
                       err);
  else
    hipLaunchKernelGGL((k_pair_tiles<false, double>), dim3((n + 127) / 128), dim3(128), 0, st,
                       n, frac, sp, tile_of, tile_start, g, row_ptr, col,
                       static_cast<double*>(coef), err);
  return hipGetLastError();
}

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

}  // namespace ewamd
