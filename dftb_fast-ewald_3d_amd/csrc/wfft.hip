// wfft.hip -- register line transforms of the SPME mesh (gfx950).
//
// An axis of length N = N1 N2 is transformed with N2 lanes per line and LW = 64 / N2 lines per
// wave, the line index fastest across the lanes: along y and z a wave's loads then read LW
// consecutive kx of one mesh row at once.  Lane (l, t) holds the elements t + N2 j (j < N1) of
// line l and runs
//   a DFT of length N1 over j in registers (compile-time twiddles),
//   the twiddles W_N^(t k1) (table of W_N^m),
//   one exchange through the wave's own LDS buffer: lane (l, t') then holds the k1 = t' + N2 i
//   of every t,
//   a DFT of length N2 over t in registers,
// which leaves X[k1 + N1 k2].  The inverse runs the same steps backwards.  Every workgroup is
// one wave: no wave waits for another, so the scheduler overlaps one wave's loads with
// another's arithmetic, and all index arithmetic is compile-time.
//
//   k_wfft_x_r2c   rows of K1 reals, two per complex line, -> half-spectrum rows (x pass)
//   k_wfft_lines   y lines, forward or inverse
//   k_wfft_z       z lines: forward, influence function (table), inverse
//   k_wfft_x_c2r   half-spectrum rows -> rows of K1 reals (inverse x pass)
//   k_influence_table  B(m) exp(-pi^2 m^2 / alpha^2) / (pi V m^2) on the half spectrum, once
//                      per geometry
#include <hip/hip_runtime.h>

#include "dft.hpp"
#include "kernels.hpp"

namespace ewamd {

template <typename R, int N1_, int N2_>
struct WLine {
  static constexpr int N1 = N1_, N2 = N2_, N = N1 * N2;
  static constexpr int LW = 64 / N2;                // lines per wave
  static constexpr int KS = (N1 + N2 - 1) / N2;     // k1 values per lane after the exchange
  static constexpr int XP = 64 + LW;                // LDS pitch of one k1 of the exchange
  static constexpr int kExch = N2 > 1 ? N1 * XP : 1;  // exchange buffer (complex elements)
  static_assert(64 % N2 == 0, "lanes per line must divide the wave");
};

// v[k1] *= W_N^(s t k1): tw holds W_N^m = exp(-2 pi i m / N); s = +1 takes its conjugate.
template <typename R, int N1, int N>
__device__ __forceinline__ void line_twiddles(cx<R>* v, int t, const cx<R>* __restrict__ tw,
                                              R s) {
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) {
    cx<R> w = tw[t * k1];
    if (s > R(0)) w.y = -w.y;
    v[k1] = cmul(v[k1], w);
  }
}

// Elements t + N2 j of line l (v[j]) -> X[k1 + N1 k2] for k1 = t + N2 i (w[i][k2], i < KS).  xs:
// the wave's exchange buffer; every lane of the wave takes part.  Everything is inlined into the
// kernel, so v and w stay in registers.
template <typename R, int N1, int N2>
__device__ __forceinline__ void line_forward(cx<R>* v, cx<R> (*w)[N2], cx<R>* xs,
                                             const cx<R>* __restrict__ tw, int l, int t, R s) {
  using W = WLine<R, N1, N2>;
  DftN<N1, R>::run(v, s);
  if constexpr (N2 > 1) {
    line_twiddles<R, N1, W::N>(v, t, tw, s);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) xs[k1 * W::XP + t * W::LW + l] = v[k1];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < W::KS; ++i) {
      const int k1 = t + N2 * i;
#pragma unroll
      for (int tt = 0; tt < N2; ++tt)
        w[i][tt] = (W::KS * N2 == N1 || k1 < N1) ? xs[k1 * W::XP + tt * W::LW + l]
                                                 : cx<R>{R(0), R(0)};
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < W::KS; ++i) DftN<N2, R>::run(w[i], s);
  } else {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) w[k1][0] = v[k1];
  }
}

// The inverse steps of line_forward: X[k1 + N1 k2] (w) -> elements t + N2 j (v).
template <typename R, int N1, int N2>
__device__ __forceinline__ void line_backward(cx<R> (*w)[N2], cx<R>* v, cx<R>* xs,
                                              const cx<R>* __restrict__ tw, int l, int t, R s) {
  using W = WLine<R, N1, N2>;
  if constexpr (N2 > 1) {
#pragma unroll
    for (int i = 0; i < W::KS; ++i) DftN<N2, R>::run(w[i], s);
#pragma unroll
    for (int i = 0; i < W::KS; ++i) {
      const int k1 = t + N2 * i;
      if (W::KS * N2 == N1 || k1 < N1) {
#pragma unroll
        for (int tt = 0; tt < N2; ++tt) xs[k1 * W::XP + tt * W::LW + l] = w[i][tt];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) v[k1] = xs[k1 * W::XP + t * W::LW + l];
    __syncthreads();
    line_twiddles<R, N1, W::N>(v, t, tw, s);
  } else {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) v[k1] = w[k1][0];
  }
  DftN<N1, R>::run(v, s);
}

// z lines (line L = ky Kh + kx, element kz at L + kz es): forward transform, the influence
// function (same layout as the spectrum), inverse transform, in place.
template <typename R, int N1, int N2, bool BATCH>
__global__ __launch_bounds__(64) void k_wfft_z(cx<R>* __restrict__ cg_a, int nlines, int es,
                                               const cx<R>* __restrict__ tw,
                                               const R* __restrict__ inf_a,
                                               const WaveIo* __restrict__ bt) {
  using W = WLine<R, N1, N2>;
  cx<R>* __restrict__ cg = BATCH ? static_cast<cx<R>*>(bt[blockIdx.y].cgrid) : cg_a;
  const R* __restrict__ inf = BATCH ? static_cast<const R*>(bt[blockIdx.y].inf) : inf_a;
  __shared__ cx<R> xs[W::kExch];
  const int lane = threadIdx.x, l = lane % W::LW, t = lane / W::LW;
  const int L = blockIdx.x * W::LW + l;
  const bool ok = L < nlines;
  const int Lc = ok ? L : nlines - 1;  // lanes past the last line repeat it and store nothing
  cx<R> v[N1];
  R f[W::KS][N2];
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = cg[Lc + (t + N2 * j) * es];
#pragma unroll
  for (int i = 0; i < W::KS; ++i) {
    const int k1 = t + N2 * i;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2)
      f[i][k2] = (W::KS * N2 == N1 || k1 < N1) ? inf[Lc + (k1 + N1 * k2) * es] : R(0);
  }
  cx<R> w[W::KS][N2];
  line_forward<R, N1, N2>(v, w, xs, tw, l, t, R(-1));
#pragma unroll
  for (int i = 0; i < W::KS; ++i)
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) w[i][k2] = cscale(w[i][k2], f[i][k2]);
  line_backward<R, N1, N2>(w, v, xs, tw, l, t, R(1));
  if (ok) {
#pragma unroll
    for (int j = 0; j < N1; ++j) cg[L + (t + N2 * j) * es] = v[j];
  }
}

// y lines (line L = kz Kh + kx, element ky at kz K2 Kh + kx + ky Kh), forward (DIR = -1) or
// inverse (+1), in place.
template <typename R, int N1, int N2, int DIR, bool BATCH>
__global__ __launch_bounds__(64) void k_wfft_lines(cx<R>* __restrict__ cg_a, int nlines, int Kh,
                                                   FastDiv khd, int plane,
                                                   const cx<R>* __restrict__ tw,
                                                   const WaveIo* __restrict__ bt) {
  using W = WLine<R, N1, N2>;
  cx<R>* __restrict__ cg = BATCH ? static_cast<cx<R>*>(bt[blockIdx.y].cgrid) : cg_a;
  __shared__ cx<R> xs[W::kExch];
  const int lane = threadIdx.x, l = lane % W::LW, t = lane / W::LW;
  const int L = blockIdx.x * W::LW + l;
  const bool ok = L < nlines;
  const int Lc = ok ? L : nlines - 1;
  const int kz = fdiv(Lc, khd);
  const int base = kz * plane + (Lc - kz * Kh);
  cx<R> v[N1];
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = cg[base + (t + N2 * j) * Kh];
  cx<R> w[W::KS][N2];
  line_forward<R, N1, N2>(v, w, xs, tw, l, t, R(DIR));
  if (ok) {
#pragma unroll
    for (int i = 0; i < W::KS; ++i) {
      const int k1 = t + N2 * i;
      if (W::KS * N2 == N1 || k1 < N1) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) cg[base + (k1 + N1 * k2) * Kh] = w[i][k2];
      }
    }
  }
}

template <typename R, int N1, int N2>
struct WRow {
  using W = WLine<R, N1, N2>;
  static constexpr int N = W::N, LW = W::LW, Kh = N / 2 + 1;
  static constexpr int LP = LW + 1;  // LDS pitch of one element index across the wave's lines
  static constexpr int a = N * LP, b = W::kExch, c = Kh * (2 * LW + 1);
  static constexpr int kBuf = a > b ? (a > c ? a : c) : (b > c ? b : c);
};

// x pass, forward: rows r0 .. r0 + 2 LW - 1 of K1 = N reals (row 2l real part, row 2l + 1
// imaginary part of line l) -> their half spectra (Kh = N/2 + 1 complex per row).  Rows are
// loaded and stored contiguously through LDS.
template <typename R, int N1, int N2, bool BATCH>
__global__ __launch_bounds__(64) void k_wfft_x_r2c(const R* __restrict__ grid_a,
                                                   cx<R>* __restrict__ cg_a, int nrows,
                                                   const cx<R>* __restrict__ tw,
                                                   const WaveIo* __restrict__ bt) {
  using W = WLine<R, N1, N2>;
  const R* __restrict__ grid = BATCH ? static_cast<const R*>(bt[blockIdx.y].grid) : grid_a;
  cx<R>* __restrict__ cg = BATCH ? static_cast<cx<R>*>(bt[blockIdx.y].cgrid) : cg_a;
  using X = WRow<R, N1, N2>;
  constexpr int N = W::N, LW = W::LW, LP = X::LP, Kh = X::Kh;
  __shared__ cx<R> buf[X::kBuf];
  const int lane = threadIdx.x, l = lane % LW, t = lane / LW;
  const int r0 = blockIdx.x * 2 * LW;
  const int nr = min(2 * LW, nrows - r0);
  R* br = reinterpret_cast<R*>(buf);
  {
    constexpr int E = 2 * LW * N / 64;
    R x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int fl = lane + 64 * e, r = fl / N;
      x[e] = r < nr ? grid[(size_t)r0 * N + fl] : R(0);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int fl = lane + 64 * e, r = fl / N, n = fl - r * N;
      br[2 * (n * LP + (r >> 1)) + (r & 1)] = x[e];
    }
  }
  __syncthreads();
  cx<R> v[N1];
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = buf[(t + N2 * j) * LP + l];
  __syncthreads();
  cx<R> w[W::KS][N2];
  line_forward<R, N1, N2>(v, w, buf, tw, l, t, R(-1));
#pragma unroll
  for (int i = 0; i < W::KS; ++i) {
    const int k1 = t + N2 * i;
    if (W::KS * N2 == N1 || k1 < N1) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) buf[(k1 + N1 * k2) * LP + l] = w[i][k2];
    }
  }
  __syncthreads();
  // split each packed line into its two rows' half spectra, Kh modes per row, x fastest
  cx<R>* out = cg + (size_t)r0 * Kh;
  for (int fl = lane; fl < 2 * LW * Kh; fl += 64) {
    const int r = fl / Kh, kx = fl - r * Kh, ll = r >> 1;
    const cx<R> z = buf[kx * LP + ll];
    const cx<R> zm = buf[(kx == 0 ? 0 : N - kx) * LP + ll];
    if (r < nr)
      out[fl] = (r & 1) ? cx<R>{R(0.5) * (z.y + zm.y), R(0.5) * (zm.x - z.x)}
                        : cx<R>{R(0.5) * (z.x + zm.x), R(0.5) * (z.y - zm.y)};
  }
}

// x pass, inverse: half-spectrum rows r0 .. r0 + 2 LW - 1 -> rows of N reals.  Rows 2l and
// 2l + 1 are extended to full spectra (X[N - k] = conj X[k]) and packed as A + iB into line l.
template <typename R, int N1, int N2, bool BATCH>
__global__ __launch_bounds__(64) void k_wfft_x_c2r(const cx<R>* __restrict__ cg_a,
                                                   R* __restrict__ grid_a, int nrows,
                                                   const cx<R>* __restrict__ tw,
                                                   const WaveIo* __restrict__ bt) {
  using W = WLine<R, N1, N2>;
  const cx<R>* __restrict__ cg = BATCH ? static_cast<const cx<R>*>(bt[blockIdx.y].cgrid) : cg_a;
  R* __restrict__ grid = BATCH ? static_cast<R*>(bt[blockIdx.y].grid) : grid_a;
  using X = WRow<R, N1, N2>;
  constexpr int N = W::N, LW = W::LW, LP = X::LP, Kh = X::Kh, RP = 2 * LW + 1;
  __shared__ cx<R> buf[X::kBuf];
  const int lane = threadIdx.x, l = lane % LW, t = lane / LW;
  const int r0 = blockIdx.x * 2 * LW;
  const int nr = min(2 * LW, nrows - r0);
  const cx<R>* in = cg + (size_t)r0 * Kh;
  {
    constexpr int E = (2 * LW * Kh + 63) / 64;
    cx<R> x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int fl = lane + 64 * e, r = fl / Kh;
      x[e] = fl < 2 * LW * Kh && r < nr ? in[fl] : cx<R>{R(0), R(0)};
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int fl = lane + 64 * e, r = fl / Kh, kx = fl - r * Kh;
      if (fl < 2 * LW * Kh) buf[kx * RP + r] = x[e];
    }
  }
  __syncthreads();
  cx<R> v[N1];
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    const int n = t + N2 * j;
    const int kk = n < Kh ? n : N - n;
    cx<R> a = buf[kk * RP + 2 * l], b = buf[kk * RP + 2 * l + 1];
    if (n >= Kh) {
      a.y = -a.y;
      b.y = -b.y;
    }
    v[j] = {a.x - b.y, a.y + b.x};
  }
  __syncthreads();
  cx<R> w[W::KS][N2];
  line_forward<R, N1, N2>(v, w, buf, tw, l, t, R(1));
#pragma unroll
  for (int i = 0; i < W::KS; ++i) {
    const int k1 = t + N2 * i;
    if (W::KS * N2 == N1 || k1 < N1) {
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) buf[(k1 + N1 * k2) * LP + l] = w[i][k2];
    }
  }
  __syncthreads();
  R* out = grid + (size_t)r0 * N;
  constexpr int E = 2 * LW * N / 64;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int fl = lane + 64 * e, r = fl / N, n = fl - r * N;
    const cx<R> z = buf[n * LP + (r >> 1)];
    if (r < nr) out[fl] = (r & 1) ? z.y : z.x;
  }
}

// ---------------------------------------------------------------------------------------
// x-y planes of a square (N x N) plane per workgroup: the x and y transforms of one z plane in
// one launch, N/2 packed lines (x) and N/2 packed columns (y) of N2 lanes each.
// ---------------------------------------------------------------------------------------
template <typename R, int N_, int N2_>
struct WPlane {
  static constexpr int N = N_, N2 = N2_, N1 = N / N2, NL = N / 2, Kh = N / 2 + 1;
  static constexpr int LW = 64 / N2, NT = NL * N2, NW = NT / 64;
  static constexpr int XP = 64 + LW;
  // line-major layout [m][n] of the packed lines (pitch PA = N + N2: the lanes of a stage read
  // l PA + t, distinct modulo 64 elements), and the spectrum rows [ky][kx] (pitch PO >= Kh with
  // PO = LW modulo 64: the lanes of a stage write t PO + l, distinct)
  static constexpr int PA = N + N2;
  static constexpr int PO = ((Kh - LW + 63) / 64) * 64 + LW;
  static constexpr int a = NL * PA, b = NW * N1 * XP, c = N * PO;
  static constexpr int kElems = a > b ? (a > c ? a : c) : (b > c ? b : c);
  static_assert(NT % 64 == 0 && NT <= 1024, "plane workgroup");
  static_assert(N1 % N2 == 0, "balanced exchange");
};

template <typename R, int N, int N2>
size_t wplane_smem() {
  return sizeof(cx<R>) * (size_t)WPlane<R, N, N2>::kElems;
}

// Forward: plane z = bid of the real mesh (N rows of N reals) -> plane of the half spectrum
// cg[(z N + ky) Kh + kx].  Rows 2m, 2m + 1 are packed as line m; after the x transform the
// spectra of the two rows are separated (Z[k], conj Z[N - k]); the real columns kx = 0 and
// kx = N/2 are packed as column 0 (X[.][0] + i X[.][N/2]) so the y transform runs over N/2
// complex columns, and are separated again on the way out.
template <typename R, int N, int N2, bool BATCH>
__global__ __launch_bounds__(N / 2 * N2) void k_wplane_r2c(const R* __restrict__ grid_a,
                                                           cx<R>* __restrict__ cg_a,
                                                           const cx<R>* __restrict__ tw,
                                                           const WaveIo* __restrict__ bt) {
  using P = WPlane<R, N, N2>;
  const R* __restrict__ grid = BATCH ? static_cast<const R*>(bt[blockIdx.y].grid) : grid_a;
  cx<R>* __restrict__ cg = BATCH ? static_cast<cx<R>*>(bt[blockIdx.y].cgrid) : cg_a;
  constexpr int N1 = P::N1, LW = P::LW, PA = P::PA, PO = P::PO, Kh = P::Kh, KS = N1 / N2;
  extern __shared__ unsigned char smem_raw[];
  cx<R>* sh = reinterpret_cast<cx<R>*>(smem_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, l = lane % LW, t = lane / LW;
  const int m = wv * LW + l;  // this lane's packed line (x stage) and packed column (y stage)
  cx<R>* xs = sh + wv * (N1 * P::XP);
  const R* plane = grid + (size_t)blockIdx.x * N * N;
  // rows -> packed lines [m][n]
  {
    R* br = reinterpret_cast<R*>(sh);
    constexpr int E = N * N / P::NT;
    R x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = plane[threadIdx.x + P::NT * e];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int fl = threadIdx.x + P::NT * e, r = fl / N, n = fl - r * N;
      br[2 * ((r >> 1) * PA + n) + (r & 1)] = x[e];
    }
  }
  __syncthreads();
  cx<R> v[N1];
  cx<R> w[KS][N2];
#pragma unroll
  for (int j = 0; j < N1; ++j) v[j] = sh[m * PA + t + N2 * j];
  __syncthreads();
  line_forward<R, N1, N2>(v, w, xs, tw, l, t, R(-1));
  // Z_m[k], natural order, back into the line-major layout
#pragma unroll
  for (int i = 0; i < KS; ++i)
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) sh[m * PA + t + N2 * i + N1 * k2] = w[i][k2];
  __syncthreads();
  // packed column c = m: X_ky[c] of row ky = 2mm + odd from Z_mm[c] and Z_mm[N - c]
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    const int ky = t + N2 * j, mm = ky >> 1;
    const cx<R> z = sh[mm * PA + m];
    const cx<R> zm = sh[mm * PA + (m == 0 ? N / 2 : N - m)];
    if (m == 0) {
      // X_ky[0] + i X_ky[N/2]: real parts of Z_mm[0], Z_mm[N/2] (even row) or imaginary parts
      v[j] = (ky & 1) ? cx<R>{z.y, zm.y} : cx<R>{z.x, zm.x};
    } else {
      v[j] = (ky & 1) ? cx<R>{R(0.5) * (z.y + zm.y), R(0.5) * (zm.x - z.x)}
                      : cx<R>{R(0.5) * (z.x + zm.x), R(0.5) * (z.y - zm.y)};
    }
  }
  __syncthreads();
  line_forward<R, N1, N2>(v, w, xs, tw, l, t, R(-1));
  // Y_c[ky] -> rows [ky][c]
#pragma unroll
  for (int i = 0; i < KS; ++i)
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) sh[(t + N2 * i + N1 * k2) * PO + m] = w[i][k2];
  __syncthreads();
  cx<R>* out = cg + (size_t)blockIdx.x * N * Kh;
  for (int fl = threadIdx.x; fl < N * Kh; fl += P::NT) {
    const int ky = fl / Kh, kx = fl - ky * Kh;
    cx<R> o;
    if (kx == 0 || kx == N / 2) {
      // column 0 holds F0 + i FN of two Hermitian spectra: F0 = (Y[k] + conj Y[-k]) / 2,
      // FN = (Y[k] - conj Y[-k]) / (2i)
      const cx<R> y = sh[ky * PO], ym = sh[(ky == 0 ? 0 : N - ky) * PO];
      o = kx == 0 ? cx<R>{R(0.5) * (y.x + ym.x), R(0.5) * (y.y - ym.y)}
                  : cx<R>{R(0.5) * (y.y + ym.y), R(0.5) * (ym.x - y.x)};
    } else {
      o = sh[ky * PO + kx];
    }
    out[fl] = o;
  }
}

// Inverse of k_wplane_r2c: plane z = bid of the half spectrum -> N rows of N reals.
template <typename R, int N, int N2, bool BATCH>
__global__ __launch_bounds__(N / 2 * N2) void k_wplane_c2r(const cx<R>* __restrict__ cg_a,
                                                           R* __restrict__ grid_a,
                                                           const cx<R>* __restrict__ tw,
                                                           const WaveIo* __restrict__ bt) {
  using P = WPlane<R, N, N2>;
  const cx<R>* __restrict__ cg = BATCH ? static_cast<const cx<R>*>(bt[blockIdx.y].cgrid) : cg_a;
  R* __restrict__ grid = BATCH ? static_cast<R*>(bt[blockIdx.y].grid) : grid_a;
  constexpr int N1 = P::N1, LW = P::LW, PA = P::PA, PO = P::PO, Kh = P::Kh, KS = N1 / N2;
  extern __shared__ unsigned char smem_raw[];
  cx<R>* sh = reinterpret_cast<cx<R>*>(smem_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, l = lane % LW, t = lane / LW;
  const int m = wv * LW + l;
  cx<R>* xs = sh + wv * (N1 * P::XP);
  const cx<R>* in = cg + (size_t)blockIdx.x * N * Kh;
  {
    constexpr int E = (N * Kh + P::NT - 1) / P::NT;
    cx<R> x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int fl = threadIdx.x + P::NT * e;
      x[e] = fl < N * Kh ? in[fl] : cx<R>{R(0), R(0)};
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int fl = threadIdx.x + P::NT * e, ky = fl / Kh, kx = fl - ky * Kh;
      if (fl < N * Kh) sh[ky * PO + kx] = x[e];
    }
  }
  __syncthreads();
  cx<R> v[N1];
  cx<R> w[KS][N2];
  // packed column c = m (column 0: H[.][0] + i H[.][N/2])
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    const int ky = t + N2 * j;
    const cx<R> h = sh[ky * PO + m];
    if (m == 0) {
      const cx<R> hn = sh[ky * PO + N / 2];
      v[j] = {h.x - hn.y, h.y + hn.x};
    } else {
      v[j] = h;
    }
  }
  __syncthreads();
  line_forward<R, N1, N2>(v, w, xs, tw, l, t, R(1));
  // G_c[y] -> [y][c] (column 0: g0[y] + i gN[y], both real)
#pragma unroll
  for (int i = 0; i < KS; ++i)
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) sh[(t + N2 * i + N1 * k2) * PO + m] = w[i][k2];
  __syncthreads();
  // packed line m: rows 2m (A) and 2m + 1 (B) extended to full spectra, z = A + iB
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    const int n = t + N2 * j;
    const int kk = n < Kh ? n : N - n;
    cx<R> a, b;
    if (kk == 0 || kk == N / 2) {
      const cx<R> pa = sh[(2 * m) * PO], pb = sh[(2 * m + 1) * PO];
      a = {kk == 0 ? pa.x : pa.y, R(0)};
      b = {kk == 0 ? pb.x : pb.y, R(0)};
    } else {
      a = sh[(2 * m) * PO + kk];
      b = sh[(2 * m + 1) * PO + kk];
      if (n >= Kh) {
        a.y = -a.y;
        b.y = -b.y;
      }
    }
    v[j] = {a.x - b.y, a.y + b.x};
  }
  __syncthreads();
  line_forward<R, N1, N2>(v, w, xs, tw, l, t, R(1));
#pragma unroll
  for (int i = 0; i < KS; ++i)
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) sh[m * PA + t + N2 * i + N1 * k2] = w[i][k2];
  __syncthreads();
  R* plane = grid + (size_t)blockIdx.x * N * N;
  constexpr int E = N * N / P::NT;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int fl = threadIdx.x + P::NT * e, r = fl / N, n = fl - r * N;
    const cx<R> z = sh[(r >> 1) * PA + n];
    plane[fl] = (r & 1) ? z.y : z.x;
  }
}

// Influence function of the half spectrum, index (kz K2 + ky) Kh + kx: zero at m = 0 and on the
// Nyquist planes of even axes (whose +-m partners differ in a sheared cell), else
// B(m) exp(-pi^2 |m|^2 / alpha^2) / (pi V |m|^2) with |m| from the dual basis, in FP64.
template <typename R>
__global__ __launch_bounds__(256) void k_influence_table(InfluenceArgs inf, R* __restrict__ tab) {
  const int Kh = inf.Kh, K2 = inf.K[1], K3 = inf.K[2];
  const long long total = (long long)Kh * K2 * K3;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int kx = int(idx % Kh);
    const long long rest = idx / Kh;
    const int ky = int(rest % K2), kz = int(rest / K2);
    double w = 0.0;
    const bool nyq = (!(inf.K[0] & 1) && kx == inf.K[0] / 2) || (!(K2 & 1) && ky == K2 / 2) ||
                     (!(K3 & 1) && kz == K3 / 2);
    if (!nyq && (kx | ky | kz)) {
      const int m1 = kx;
      const int m2 = ky <= (K2 - 1) / 2 ? ky : ky - K2;
      const int m3 = kz <= (K3 - 1) / 2 ? kz : kz - K3;
      const double vx = m1 * inf.rec[0] + m2 * inf.rec[3] + m3 * inf.rec[6];
      const double vy = m1 * inf.rec[1] + m2 * inf.rec[4] + m3 * inf.rec[7];
      const double vz = m1 * inf.rec[2] + m2 * inf.rec[5] + m3 * inf.rec[8];
      const double msq = vx * vx + vy * vy + vz * vz;
      w = inf.pref * exp(-inf.a2 * msq) / msq * inf.bsq1[kx] * inf.bsq2[ky] * inf.bsq3[kz];
    }
    tab[idx] = R(w);
  }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

// (axis length, lanes per line) with a compiled register transform
#define EW_WAVE_SHAPES(X)                                                                     \
  X(32, 4) X(32, 8) X(48, 4) X(64, 4) X(64, 8) X(64, 16) X(96, 4) X(96, 8) X(128, 4)          \
  X(128, 8) X(128, 16) X(256, 8) X(256, 16)

// (plane side, lanes per line) with a compiled register plane transform (square x-y planes)
#define EW_PLANE_SHAPES(X) X(32, 4) X(64, 8) X(96, 4) X(128, 8)

bool wave_plane_shape(int n, int lanes) {
#define EW_CASE(NN, LL) \
  if (n == NN && lanes == LL) return true;
  EW_PLANE_SHAPES(EW_CASE)
#undef EW_CASE
  return false;
}

int wave_plane_default_lanes(int n) {
  for (int l : {8, 4})
    if (wave_plane_shape(n, l)) return l;
  return 0;
}

bool wave_fft_shape(int n, int lanes) {
#define EW_CASE(NN, LL) \
  if (n == NN && lanes == LL) return true;
  EW_WAVE_SHAPES(EW_CASE)
#undef EW_CASE
  return false;
}

int wave_fft_default_lanes(int n, bool fp32) {
  // FP32: 16 lines of 8 bytes (128 bytes of a row) per wave where the length allows; FP64: 8
  // lines of 16 bytes
  const int want = fp32 ? 4 : 8;
  if (wave_fft_shape(n, want)) return want;
  for (int l : {4, 8, 16})
    if (wave_fft_shape(n, l)) return l;
  return 0;
}

template <typename R>
hipError_t launch_influence_table(const InfluenceArgs& inf, void* tab, hipStream_t st) {
  const long long total = (long long)inf.Kh * inf.K[1] * inf.K[2];
  const int nb = (int)std::min<long long>(4096, (total + 255) / 256);
  hipLaunchKernelGGL(k_influence_table<R>, dim3(nb), dim3(256), 0, st, inf, static_cast<R*>(tab));
  return hipGetLastError();
}
template hipError_t launch_influence_table<double>(const InfluenceArgs&, void*, hipStream_t);
template hipError_t launch_influence_table<float>(const InfluenceArgs&, void*, hipStream_t);

namespace {

// A launch of the passes: one structure (bt null: the pointers of L) or a batch of count
// structures of the same mesh and lanes (bt: their meshes, spectra and influence tables, one per
// blockIdx.y).
struct WaveCall {
  const SpmeLaunch* L;
  const WaveIo* bt;
  int count;
};

template <typename R, int N1, int N2>
void wave_x(bool forward, const WaveCall& c, hipStream_t st) {
  const SpmeLaunch& L = *c.L;
  constexpr int LW = 64 / N2;
  const int nrows = L.K[1] * L.K[2];
  const dim3 g((nrows + 2 * LW - 1) / (2 * LW), c.count);
  const cx<R>* tw = static_cast<const cx<R>*>(L.twp[0]);
  if (forward) {
    if (c.bt)
      hipLaunchKernelGGL((k_wfft_x_r2c<R, N1, N2, true>), g, dim3(64), 0, st, nullptr, nullptr,
                         nrows, tw, c.bt);
    else
      hipLaunchKernelGGL((k_wfft_x_r2c<R, N1, N2, false>), g, dim3(64), 0, st,
                         static_cast<const R*>(L.grid), static_cast<cx<R>*>(L.cgrid), nrows, tw,
                         nullptr);
  } else {
    if (c.bt)
      hipLaunchKernelGGL((k_wfft_x_c2r<R, N1, N2, true>), g, dim3(64), 0, st, nullptr, nullptr,
                         nrows, tw, c.bt);
    else
      hipLaunchKernelGGL((k_wfft_x_c2r<R, N1, N2, false>), g, dim3(64), 0, st,
                         static_cast<const cx<R>*>(L.cgrid), static_cast<R*>(L.grid), nrows, tw,
                         nullptr);
  }
}

template <typename R, int N1, int N2>
void wave_y(bool forward, const WaveCall& c, hipStream_t st) {
  const SpmeLaunch& L = *c.L;
  constexpr int LW = 64 / N2;
  const int Kh = L.K[0] / 2 + 1, nlines = Kh * L.K[2];
  const dim3 g((nlines + LW - 1) / LW, c.count);
  const cx<R>* tw = static_cast<const cx<R>*>(L.twp[1]);
  cx<R>* cg = c.bt ? nullptr : static_cast<cx<R>*>(L.cgrid);
  const FastDiv khd = make_fastdiv(Kh);
  if (forward) {
    if (c.bt)
      hipLaunchKernelGGL((k_wfft_lines<R, N1, N2, -1, true>), g, dim3(64), 0, st, cg, nlines, Kh,
                         khd, Kh * L.K[1], tw, c.bt);
    else
      hipLaunchKernelGGL((k_wfft_lines<R, N1, N2, -1, false>), g, dim3(64), 0, st, cg, nlines,
                         Kh, khd, Kh * L.K[1], tw, nullptr);
  } else {
    if (c.bt)
      hipLaunchKernelGGL((k_wfft_lines<R, N1, N2, 1, true>), g, dim3(64), 0, st, cg, nlines, Kh,
                         khd, Kh * L.K[1], tw, c.bt);
    else
      hipLaunchKernelGGL((k_wfft_lines<R, N1, N2, 1, false>), g, dim3(64), 0, st, cg, nlines, Kh,
                         khd, Kh * L.K[1], tw, nullptr);
  }
}

template <typename R, int N1, int N2>
void wave_z(const WaveCall& c, hipStream_t st) {
  const SpmeLaunch& L = *c.L;
  constexpr int LW = 64 / N2;
  const int Kh = L.K[0] / 2 + 1, nlines = Kh * L.K[1];
  const dim3 g((nlines + LW - 1) / LW, c.count);
  const cx<R>* tw = static_cast<const cx<R>*>(L.twp[2]);
  if (c.bt)
    hipLaunchKernelGGL((k_wfft_z<R, N1, N2, true>), g, dim3(64), 0, st, nullptr, nlines, nlines,
                       tw, nullptr, c.bt);
  else
    hipLaunchKernelGGL((k_wfft_z<R, N1, N2, false>), g, dim3(64), 0, st,
                       static_cast<cx<R>*>(L.cgrid), nlines, nlines, tw,
                       static_cast<const R*>(L.inf_tab), nullptr);
}

template <typename R, int N, int N2>
hipError_t wave_plane(bool forward, const WaveCall& c, hipStream_t st) {
  const SpmeLaunch& L = *c.L;
  const size_t smem = wplane_smem<R, N, N2>();
  // more than 64 KiB of dynamic LDS needs the kernel attribute (set once per kernel)
  static const hipError_t attr = [&] {
    const void* ks[] = {reinterpret_cast<const void*>(&k_wplane_r2c<R, N, N2, false>),
                        reinterpret_cast<const void*>(&k_wplane_c2r<R, N, N2, false>),
                        reinterpret_cast<const void*>(&k_wplane_r2c<R, N, N2, true>),
                        reinterpret_cast<const void*>(&k_wplane_c2r<R, N, N2, true>)};
    hipError_t e = hipSuccess;
    for (const void* k : ks)
      if (e == hipSuccess)
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    return e;
  }();
  if (attr != hipSuccess) return attr;
  const dim3 g(L.K[2], c.count);
  const dim3 b(WPlane<R, N, N2>::NT);
  const cx<R>* tw = static_cast<const cx<R>*>(L.twp[0]);
  if (forward) {
    if (c.bt)
      hipLaunchKernelGGL((k_wplane_r2c<R, N, N2, true>), g, b, smem, st, nullptr, nullptr, tw,
                         c.bt);
    else
      hipLaunchKernelGGL((k_wplane_r2c<R, N, N2, false>), g, b, smem, st,
                         static_cast<const R*>(L.grid), static_cast<cx<R>*>(L.cgrid), tw, nullptr);
  } else {
    if (c.bt)
      hipLaunchKernelGGL((k_wplane_c2r<R, N, N2, true>), g, b, smem, st, nullptr, nullptr, tw,
                         c.bt);
    else
      hipLaunchKernelGGL((k_wplane_c2r<R, N, N2, false>), g, b, smem, st,
                         static_cast<const cx<R>*>(L.cgrid), static_cast<R*>(L.grid), tw, nullptr);
  }
  return hipGetLastError();
}

template <typename R>
bool dispatch_plane(bool forward, const WaveCall& c, hipStream_t st, hipError_t* err) {
  const SpmeLaunch& L = *c.L;
#define EW_CASE(NN, LL)                                        \
  if (L.K[0] == NN && L.lanes[0] == LL) {                      \
    *err = wave_plane<R, NN, LL>(forward, c, st);              \
    return true;                                               \
  }
  if (L.K[0] == L.K[1]) {
    EW_PLANE_SHAPES(EW_CASE)
  }
#undef EW_CASE
  return false;
}

template <typename R>
bool dispatch_x(bool forward, const WaveCall& c, hipStream_t st) {
#define EW_CASE(NN, LL)                                        \
  if (c.L->K[0] == NN && c.L->lanes[0] == LL) {                \
    wave_x<R, NN / LL, LL>(forward, c, st);                    \
    return true;                                               \
  }
  EW_WAVE_SHAPES(EW_CASE)
#undef EW_CASE
  return false;
}
template <typename R>
bool dispatch_y(bool forward, const WaveCall& c, hipStream_t st) {
#define EW_CASE(NN, LL)                                        \
  if (c.L->K[1] == NN && c.L->lanes[1] == LL) {                \
    wave_y<R, NN / LL, LL>(forward, c, st);                    \
    return true;                                               \
  }
  EW_WAVE_SHAPES(EW_CASE)
#undef EW_CASE
  return false;
}
template <typename R>
bool dispatch_z(const WaveCall& c, hipStream_t st) {
#define EW_CASE(NN, LL)                                        \
  if (c.L->K[2] == NN && c.L->lanes[2] == LL) {                \
    wave_z<R, NN / LL, LL>(c, st);                             \
    return true;                                               \
  }
  EW_WAVE_SHAPES(EW_CASE)
#undef EW_CASE
  return false;
}

template <typename R>
hipError_t run_wave_fft(const WaveCall& c, hipStream_t st, hipEvent_t* ev) {
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  if (c.L->wave == 2) {
    // x and y of each plane in one launch (no y phase: its events follow at once)
    hipError_t e = hipSuccess;
    if (!dispatch_plane<R>(true, c, st, &e)) return hipErrorInvalidValue;
    if (e != hipSuccess) return e;
    mark(1);
    mark(2);
    if (!dispatch_z<R>(c, st)) return hipErrorInvalidValue;
    mark(3);
    mark(4);
    dispatch_plane<R>(false, c, st, &e);
    mark(5);
    return e != hipSuccess ? e : hipGetLastError();
  }
  if (!dispatch_x<R>(true, c, st)) return hipErrorInvalidValue;
  mark(1);
  if (!dispatch_y<R>(true, c, st)) return hipErrorInvalidValue;
  mark(2);
  if (!dispatch_z<R>(c, st)) return hipErrorInvalidValue;
  mark(3);
  dispatch_y<R>(false, c, st);
  mark(4);
  dispatch_x<R>(false, c, st);
  mark(5);
  return hipGetLastError();
}

}  // namespace

template <typename R>
hipError_t launch_wave_fft(const SpmeLaunch& L, hipStream_t st, hipEvent_t* ev) {
  return run_wave_fft<R>(WaveCall{&L, nullptr, 1}, st, ev);
}
template hipError_t launch_wave_fft<double>(const SpmeLaunch&, hipStream_t, hipEvent_t*);
template hipError_t launch_wave_fft<float>(const SpmeLaunch&, hipStream_t, hipEvent_t*);

template <typename R>
hipError_t launch_wave_fft_batch(const SpmeLaunch& L, const WaveIo* bt, int count,
                                 hipStream_t st) {
  if (count < 1) return hipSuccess;
  return run_wave_fft<R>(WaveCall{&L, bt, count}, st, nullptr);
}
template hipError_t launch_wave_fft_batch<double>(const SpmeLaunch&, const WaveIo*, int,
                                                  hipStream_t);
template hipError_t launch_wave_fft_batch<float>(const SpmeLaunch&, const WaveIo*, int,
                                                 hipStream_t);

}  // namespace ewamd
