// dft.hpp -- complex arithmetic and small DFTs in registers (device code shared by the
// Stockham passes of kernels.hip and the register line transforms of wfft.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

#include "engine.hpp"

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

template <typename R>
struct Dft<8, R> {
  __device__ static void run(cx<R>* v, R s) {
    // two radix-4 halves (even and odd inputs), then the eighth-root twiddles
    cx<R> e[4] = {v[0], v[2], v[4], v[6]};
    cx<R> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, R>::run(e, s);
    Dft<4, R>::run(o, s);
    const R c = R(0.707106781186547524400844362104849039);
    const cx<R> w1 = {c * (o[1].x - s * o[1].y), c * (o[1].y + s * o[1].x)};
    const cx<R> w2 = cmuli(o[2], s);
    const cx<R> w3 = {c * (-o[3].x - s * o[3].y), c * (-o[3].y + s * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  }
};

template <typename R>
struct Dft<16, R> {
  // 4 x 4: DFT-4 over the stride-4 subsequences, twiddles W16^(n2 k1), DFT-4 across them
  __device__ static void run(cx<R>* v, R s) {
    const R c1 = R(0.923879532511286756128183189396788933);  // cos(pi/8)
    const R s1 = R(0.382683432365089771728459984030398866);  // sin(pi/8)
    const R c2 = R(0.707106781186547524400844362104849039);  // cos(pi/4)
    cx<R> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cx<R> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      Dft<4, R>::run(t, s);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
    // W16^m = cos(pi m / 8) + i s sin(pi m / 8) for m = n2 k1
    a[1][1] = cmul(a[1][1], cx<R>{c1, s * s1});
    a[1][2] = cmul(a[1][2], cx<R>{c2, s * c2});
    a[1][3] = cmul(a[1][3], cx<R>{s1, s * c1});
    a[2][1] = cmul(a[2][1], cx<R>{c2, s * c2});
    a[2][2] = cmuli(a[2][2], s);
    a[2][3] = cmul(a[2][3], cx<R>{-c2, s * c2});
    a[3][1] = cmul(a[3][1], cx<R>{s1, s * c1});
    a[3][2] = cmul(a[3][2], cx<R>{-c2, s * c2});
    a[3][3] = cmul(a[3][3], cx<R>{-c1, -s * s1});
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cx<R> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Dft<4, R>::run(t, s);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// ----------------------------------------------------------------------------------------
// compile-time roots of unity and loops
// ----------------------------------------------------------------------------------------
namespace ct {
constexpr double kPiC = 3.14159265358979323846264338327950288;
// Taylor series on |x| <= pi/4 (terms to x^25: far below double rounding there)
constexpr double sin_small(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k <= 12; ++k) {
    term *= -x2 / double((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}
constexpr double cos_small(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k <= 12; ++k) {
    term *= -x2 / double((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}
// cos (want_sin = 0) or sin (1) of 2 pi m / n: the quadrant and the reflection about pi/4 are
// taken in integers, so the series only sees angles in [0, pi/4].
constexpr double root(long m, long n, int want_sin) {
  m %= n;
  if (m < 0) m += n;
  const long q = (4 * m) / n, r4 = 4 * m - q * n;  // angle = (pi/2) (q + r4/n)
  const bool flip = 2 * r4 > n;
  const double th = flip ? (kPiC / 2) * double(n - r4) / double(n)
                         : (kPiC / 2) * double(r4) / double(n);
  double cc = cos_small(th), ss = sin_small(th);
  if (flip) {
    const double t = cc;
    cc = ss;
    ss = t;
  }
  double c = cc, s = ss;
  if (q == 1) {
    c = -ss;
    s = cc;
  } else if (q == 2) {
    c = -cc;
    s = -ss;
  } else if (q == 3) {
    c = ss;
    s = -cc;
  }
  return want_sin ? s : c;
}
}  // namespace ct

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>): the index is a constant
// expression inside f, so tables and twiddles indexed by it fold into immediates.
template <int... I, typename F>
__device__ __forceinline__ void sfor_impl(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

// DFT of any length 2^a 3^b 5^c in registers (natural order in and out): the hand-written
// radices directly, other lengths by one Cooley-Tukey split M = A B with compile-time twiddles:
// n = B na + nb, k = ka + A kb; DFT_A over na for each nb, twiddle W_M^(nb ka), DFT_B over nb.
constexpr bool dft_is_base(int m) {
  return m == 1 || m == 2 || m == 3 || m == 4 || m == 5 || m == 8 || m == 16;
}
constexpr int dft_split(int m) {
  return m % 16 == 0 ? 16 : m % 8 == 0 ? 8 : m % 4 == 0 ? 4 : m % 5 == 0 ? 5 : m % 3 == 0 ? 3 : 2;
}
template <int M, typename R, bool BASE = dft_is_base(M)>
struct DftN;
template <int M, typename R>
struct DftN<M, R, true> {
  __device__ __forceinline__ static void run(cx<R>* v, R s) {
    if constexpr (M > 1) Dft<M, R>::run(v, s);
  }
};
template <int M, typename R>
struct DftN<M, R, false> {
  static constexpr int A = dft_split(M), B = M / A;
  __device__ __forceinline__ static void run(cx<R>* v, R s) {
    cx<R> t[B][A];
    sfor<B>([&](auto nbc) {
      constexpr int NB = decltype(nbc)::value;
      cx<R> u[A];
#pragma unroll
      for (int a = 0; a < A; ++a) u[a] = v[B * a + NB];
      DftN<A, R>::run(u, s);
      sfor<A>([&](auto kac) {
        constexpr int KA = decltype(kac)::value;
        if constexpr ((NB * KA) % M == 0) {
          t[NB][KA] = u[KA];
        } else {
          constexpr R c = R(ct::root(NB * KA, M, 0)), sn = R(ct::root(NB * KA, M, 1));
          t[NB][KA] = cmul(u[KA], cx<R>{c, s * sn});
        }
      });
    });
    sfor<A>([&](auto kac) {
      constexpr int KA = decltype(kac)::value;
      cx<R> u[B];
#pragma unroll
      for (int b = 0; b < B; ++b) u[b] = t[b][KA];
      DftN<B, R>::run(u, s);
#pragma unroll
      for (int b = 0; b < B; ++b) v[KA + A * b] = u[b];
    });
  }
};

__device__ __forceinline__ int fdiv(int x, FastDiv f) {
  return int((__umulhi(unsigned(x), f.m) + unsigned(x)) >> f.s);
}

}  // namespace ewamd
