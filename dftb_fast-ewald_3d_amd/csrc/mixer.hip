// mixer.hip -- SCC charge mixing on the device: DFTB+'s simple and modified Broyden mixers with
// their vectors resident in HBM, so that an SCC cycle keeps the charges on the GPU between the
// shift (ew_shift_device) and the mixer.
//
// Modified Broyden (Johnson, PRB 38, 12807 (1988)) as src/dftbp/mixer/broydenmixer.F90:204-342
// computes it; one mix is three launches on the caller's stream and no host round trip:
//   k_mix_dots    every inner product the step needs, in one sweep over the vectors
//                 (<dF_i|dF>, <dF_i|qDiff> for the stored history, |qDiff|^2, |dF|^2,
//                 <dF|qDiff>), as per-block partial sums;
//   k_mix_solve   one workgroup: the partials summed in a fixed order, the weights, the
//                 matrix a, beta = w_i w_j a_ij + omega0^2 delta_ij, its LU factorisation with
//                 partial pivoting and the transposed solve (LAPACK dgetrf / dgetrs('T'),
//                 broydenmixer.F90:318-320);
//   k_mix_update  the new charges, |dF(m-1)> and |u(m-1)> in one elementwise sweep.
// Every sum runs in a fixed order, so a mix is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/fftewald_amd.h"
#include "handle.hpp"

namespace ewamd {
namespace {

constexpr int kMixGroup = 8;                       // history vectors per dot-product group
constexpr int kMixSlots = 3 + 2 * kMixGroup;       // partial sums per group and block
constexpr int kMixMaxBlocks = 256;                 // blocks along the vectors in k_mix_dots
constexpr int kMixMaxIter = 1024;                  // longest history (LDS of k_mix_solve)
constexpr int kMixMaxDots = 3 + 2 * (kMixMaxIter - 2);

__global__ __launch_bounds__(256) void k_mix_simple(int n, double alpha, double* __restrict__ q,
                                                    const double* __restrict__ qd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) q[e] = q[e] + alpha * qd[e];
}

// First Broyden iteration: store the vectors, simple mix (broydenmixer.F90:272-277).
__global__ __launch_bounds__(256) void k_mix_first(int n, double alpha, double* __restrict__ q,
                                                   const double* __restrict__ qd,
                                                   double* __restrict__ qlast,
                                                   double* __restrict__ qdlast) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double x = q[e], y = qd[e];
  qlast[e] = x;
  qdlast[e] = y;
  q[e] = x + alpha * y;
}

__device__ __forceinline__ double mix_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Partial sums over this block's elements.  Group g = blockIdx.y covers the history vectors
// [g*kMixGroup, ...); group 0 also takes |qDiff|^2, |d|^2 and <d|qDiff> with d = qDiff -
// qDiffLast.  partials[(g * gridDim.x + blockIdx.x) * kMixSlots + slot].
__global__ __launch_bounds__(256) void k_mix_dots(int n, int nhist, const double* __restrict__ qd,
                                                  const double* __restrict__ qdlast,
                                                  const double* __restrict__ dF,
                                                  double* __restrict__ partials) {
  __shared__ double red[4][kMixSlots];
  const int g = blockIdx.y;
  const int i0 = g * kMixGroup;
  const int nk = min(kMixGroup, nhist - i0);  // may be <= 0 for group 0 without history
  double acc[kMixSlots];
#pragma unroll
  for (int s = 0; s < kMixSlots; ++s) acc[s] = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const double x = qd[e];
    const double d = x - qdlast[e];
    if (g == 0) {
      acc[0] += x * x;
      acc[1] += d * d;
      acc[2] += d * x;
    }
#pragma unroll
    for (int k = 0; k < kMixGroup; ++k) {
      if (k < nk) {
        const double f = dF[(size_t)(i0 + k) * n + e];
        acc[3 + 2 * k] += f * d;
        acc[4 + 2 * k] += f * x;
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < kMixSlots; ++s) {
    const double t = mix_wave_sum(acc[s]);
    if (lane == 0) red[wv][s] = t;
  }
  __syncthreads();
  if (threadIdx.x < kMixSlots) {
    const int s = threadIdx.x;
    const double t = ((red[0][s] + red[1][s]) + red[2][s]) + red[3][s];
    partials[((size_t)g * gridDim.x + blockIdx.x) * kMixSlots + s] = t;
  }
}

struct MixSolveArgs {
  int nhist;   // stored history vectors used now (nn - 2)
  int m;       // size of the system (nn - 1)
  int ld;      // leading dimension of aa (max_iter - 1)
  int bx;      // blocks along the vectors in k_mix_dots
  double omega0, min_weight, max_weight, weight_fac;
  const double* partials;
  double* aa;    // [ld][ld] symmetric, persistent
  double* ww;    // [ld] weights, persistent
  double* beta;  // [m][m] scratch, row-major
  int* ipiv;     // [m] scratch
  double* coef;  // [m] out: ww_i * c_i
  double* scal;  // [1] out: 1 / |d|
};

__global__ __launch_bounds__(256) void k_mix_solve(MixSolveArgs A) {
  __shared__ double dots[kMixMaxDots];
  __shared__ double cc[kMixMaxIter];
  __shared__ double wws[kMixMaxIter];
  __shared__ double pv[256];
  __shared__ int pi[256];
  __shared__ double inv_norm_s;
  const int t = threadIdx.x;
  const int ndots = 3 + 2 * A.nhist;
  for (int i = t; i < A.nhist; i += blockDim.x) wws[i] = A.ww[i];
  // the partial sums, block by block in a fixed order
  for (int dI = t; dI < ndots; dI += blockDim.x) {
    int g, s;
    if (dI < 3) {
      g = 0;
      s = dI;
    } else {
      const int i = (dI - 3) >> 1;
      g = i / kMixGroup;
      s = 3 + 2 * (i % kMixGroup) + ((dI - 3) & 1);
    }
    double v = 0.0;
    for (int b = 0; b < A.bx; ++b) v += A.partials[((size_t)g * A.bx + b) * kMixSlots + s];
    dots[dI] = v;
  }
  __syncthreads();
  const int m = A.m, ld = A.ld;
  if (t == 0) {
    // weight of this iteration (broydenmixer.F90:285-293)
    double w = sqrt(dots[0]);
    w = w > A.weight_fac / A.max_weight ? A.weight_fac / w : A.max_weight;
    w = fmax(w, A.min_weight);
    wws[m - 1] = w;
    A.ww[m - 1] = w;
    // 1 / |dF| (296-300)
    const double nrm = fmax(sqrt(dots[1]), 2.220446049250313e-16);
    inv_norm_s = 1.0 / nrm;
    A.scal[0] = inv_norm_s;
    A.aa[(size_t)(m - 1) * ld + (m - 1)] = 1.0;
    cc[m - 1] = w * (inv_norm_s * dots[2]);
  }
  __syncthreads();
  const double inv_norm = inv_norm_s;
  // a and c (304-311)
  for (int i = t; i < A.nhist; i += blockDim.x) {
    const double a = inv_norm * dots[3 + 2 * i];
    A.aa[(size_t)i * ld + (m - 1)] = a;
    A.aa[(size_t)(m - 1) * ld + i] = a;
    cc[i] = wws[i] * dots[4 + 2 * i];
  }
  __syncthreads();
  // beta(r, c) = w_r w_c a(r, c) + omega0^2 delta_rc (312-315)
  double* B = A.beta;
  const double om2 = A.omega0 * A.omega0;
  for (int k = t; k < m * m; k += blockDim.x) {
    const int r = k / m, c = k % m;
    double v = wws[r] * wws[c] * A.aa[(size_t)r * ld + c];
    if (r == c) v += om2;
    B[k] = v;
  }
  __syncthreads();
  // LU with partial pivoting, unblocked (dgetf2): pivot = first row of largest |B(i,k)|
  for (int k = 0; k < m; ++k) {
    double best = -1.0;
    int bi = m;
    for (int i = k + t; i < m; i += blockDim.x) {
      const double v = fabs(B[(size_t)i * m + k]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    pv[t] = best;
    pi[t] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        const double o = pv[t + s];
        const int oi = pi[t + s];
        if (o > pv[t] || (o == pv[t] && oi < pi[t])) {
          pv[t] = o;
          pi[t] = oi;
        }
      }
      __syncthreads();
    }
    const int p = pi[0] < m ? pi[0] : k;  // no comparable entry (NaN column): keep row k
    if (t == 0) A.ipiv[k] = p;
    if (p != k)
      for (int c = t; c < m; c += blockDim.x) {
        const double x = B[(size_t)k * m + c];
        B[(size_t)k * m + c] = B[(size_t)p * m + c];
        B[(size_t)p * m + c] = x;
      }
    __syncthreads();
    const double piv = B[(size_t)k * m + k];
    if (piv != 0.0) {
      const double r = 1.0 / piv;
      for (int i = k + 1 + t; i < m; i += blockDim.x) B[(size_t)i * m + k] *= r;
    }
    __syncthreads();
    const int w = m - k - 1;
    for (int q = t; q < w * w; q += blockDim.x) {
      const int i = k + 1 + q / w, j = k + 1 + q % w;
      B[(size_t)i * m + j] -= B[(size_t)i * m + k] * B[(size_t)k * m + j];
    }
    __syncthreads();
  }
  // beta^T x = c with beta = P L U (dgetrs 'T'): U^T y = c, L^T z = y, x = P z
  if (t == 0) {
    for (int i = 0; i < m; ++i) {
      double v = cc[i];
      for (int k = 0; k < i; ++k) v -= B[(size_t)k * m + i] * cc[k];
      cc[i] = v / B[(size_t)i * m + i];
    }
    for (int i = m - 1; i >= 0; --i) {
      double v = cc[i];
      for (int k = i + 1; k < m; ++k) v -= B[(size_t)k * m + i] * cc[k];
      cc[i] = v;
    }
    for (int k = m - 1; k >= 0; --k) {
      const int p = A.ipiv[k];
      if (p != k) {
        const double x = cc[k];
        cc[k] = cc[p];
        cc[p] = x;
      }
    }
    for (int i = 0; i < m; ++i) A.coef[i] = wws[i] * cc[i];
  }
}

// New charges, |dF(m-1)> and |u(m-1)> (broydenmixer.F90:323-340).
__global__ __launch_bounds__(256) void k_mix_update(int n, int nhist, double alpha,
                                                    const double* __restrict__ coef,
                                                    const double* __restrict__ scal,
                                                    double* __restrict__ q,
                                                    const double* __restrict__ qd,
                                                    double* __restrict__ qlast,
                                                    double* __restrict__ qdlast,
                                                    double* __restrict__ dF,
                                                    double* __restrict__ uu) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double inv = scal[0];
  const double x = q[e], y = qd[e];
  const double dfn = inv * (y - qdlast[e]);
  const double u = alpha * dfn + inv * (x - qlast[e]);
  qlast[e] = x;
  qdlast[e] = y;
  double r = x + alpha * y;
  for (int i = 0; i < nhist; ++i) r = r - coef[i] * uu[(size_t)i * n + e];
  r = r - coef[nhist] * u;
  dF[(size_t)nhist * n + e] = dfn;
  uu[(size_t)nhist * n + e] = u;
  q[e] = r;
}

}  // namespace
}  // namespace ewamd

using namespace ewamd;

struct ew_mixer {
  ew_mixer_options opt{};
  hipStream_t stream = nullptr;
  std::string err = "no error";
  int n = 0;       // elements (0: not reset)
  int iter = 0;    // mixes since the last reset
  int ld = 0;      // history length max_iter - 1
  int bx = 1;      // k_mix_dots blocks along the vectors
  double *qlast = nullptr, *qdlast = nullptr, *dF = nullptr, *uu = nullptr;
  double *aa = nullptr, *ww = nullptr, *beta = nullptr, *coef = nullptr, *scal = nullptr;
  double *partials = nullptr, *hq = nullptr, *hqd = nullptr;
  int* ipiv = nullptr;

  int fail(int code, const std::string& msg) {
    err = msg;
    global_error() = msg;
    return code;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : {(void*)qlast, (void*)qdlast, (void*)dF, (void*)uu, (void*)aa, (void*)ww,
                    (void*)beta, (void*)coef, (void*)scal, (void*)partials, (void*)hq,
                    (void*)hqd, (void*)ipiv})
      if (p) (void)hipFree(p);
    qlast = qdlast = dF = uu = aa = ww = beta = coef = scal = partials = hq = hqd = nullptr;
    ipiv = nullptr;
    n = 0;
  }
};

#define MCHK(m, expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return (m)->fail(EW_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) +   \
                                          " at " #expr);                                     \
  } while (0)

namespace {

template <typename T>
int mix_alloc(ew_mixer* M, T** p, size_t count) {
  if (hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)) !=
      hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return M->fail(EW_ERR_MEMORY, "mixer device allocation failed");
  }
  return EW_OK;
}

int enqueue_mix(ew_mixer* M, double* q, const double* qd, hipStream_t s) {
  const int n = M->n;
  const dim3 blk(256), grd((n + 255) / 256);
  const double alpha = M->opt.mix_param;
  if (M->opt.kind == EW_MIXER_SIMPLE) {
    hipLaunchKernelGGL(k_mix_simple, grd, blk, 0, s, n, alpha, q, qd);
    MCHK(M, hipGetLastError());
    return EW_OK;
  }
  const int nn = M->iter;  // this mix's iteration number (1-based)
  if (nn == 1) {
    hipLaunchKernelGGL(k_mix_first, grd, blk, 0, s, n, alpha, q, qd, M->qlast, M->qdlast);
    MCHK(M, hipGetLastError());
    return EW_OK;
  }
  const int nhist = nn - 2;
  const int ng = std::max(1, (nhist + kMixGroup - 1) / kMixGroup);
  hipLaunchKernelGGL(k_mix_dots, dim3(M->bx, ng), blk, 0, s, n, nhist, qd, M->qdlast, M->dF,
                     M->partials);
  MCHK(M, hipGetLastError());
  MixSolveArgs A{};
  A.nhist = nhist;
  A.m = nn - 1;
  A.ld = M->ld;
  A.bx = M->bx;
  A.omega0 = M->opt.omega0;
  A.min_weight = M->opt.min_weight;
  A.max_weight = M->opt.max_weight;
  A.weight_fac = M->opt.weight_fac;
  A.partials = M->partials;
  A.aa = M->aa;
  A.ww = M->ww;
  A.beta = M->beta;
  A.ipiv = M->ipiv;
  A.coef = M->coef;
  A.scal = M->scal;
  hipLaunchKernelGGL(k_mix_solve, dim3(1), blk, 0, s, A);
  MCHK(M, hipGetLastError());
  hipLaunchKernelGGL(k_mix_update, grd, blk, 0, s, n, nhist, alpha, M->coef, M->scal, q, qd,
                     M->qlast, M->qdlast, M->dF, M->uu);
  MCHK(M, hipGetLastError());
  return EW_OK;
}

int begin_mix(ew_mixer* M) {
  if (M->n <= 0) return M->fail(EW_ERR_ARGUMENT, "mix before ew_mixer_reset");
  if (M->opt.kind == EW_MIXER_BROYDEN && M->iter + 1 > M->opt.max_iter)
    return M->fail(EW_ERR_ARGUMENT, "Broyden mixer: Maximal nr. of steps exceeded");
  M->iter += 1;
  return EW_OK;
}

}  // namespace

extern "C" int ew_mixer_create(const ew_mixer_options* opts, ew_mixer** out) {
  if (!out) return set_global_error(EW_ERR_ARGUMENT, "ew_mixer_create: out is NULL");
  *out = nullptr;
  ew_mixer_options o{};
  if (opts) o = *opts;
  if (o.kind != EW_MIXER_BROYDEN && o.kind != EW_MIXER_SIMPLE)
    return set_global_error(EW_ERR_ARGUMENT, "ew_mixer_create: unknown mixer kind");
  if (o.mix_param == 0.0) o.mix_param = 0.2;
  if (o.max_iter == 0) o.max_iter = 100;
  if (o.omega0 == 0.0) o.omega0 = 0.01;
  if (o.min_weight == 0.0) o.min_weight = 1.0;
  if (o.max_weight == 0.0) o.max_weight = 1.0e5;
  if (o.weight_fac == 0.0) o.weight_fac = 1.0e-2;
  // TBroydenMixer%init asserts (broydenmixer.F90:124-126)
  if (!(o.mix_param > 0.0) || !(o.omega0 > 0.0) || o.max_iter < 1 || o.max_iter > kMixMaxIter)
    return set_global_error(EW_ERR_ARGUMENT,
                            "ew_mixer_create: need mix_param > 0, omega0 > 0, 1 <= max_iter <= " +
                                std::to_string(kMixMaxIter));
  if (hipSetDevice(o.device) != hipSuccess) {
    (void)hipGetLastError();
    return set_global_error(EW_ERR_DEVICE, "ew_mixer_create: no such device");
  }
  ew_mixer* M = new ew_mixer();
  M->opt = o;
  M->ld = std::max(1, o.max_iter - 1);
  // a blocking stream, as the engine's: ordered with work on the null stream
  if (hipStreamCreate(&M->stream) != hipSuccess) {
    (void)hipGetLastError();
    delete M;
    return set_global_error(EW_ERR_DEVICE, "ew_mixer_create: stream creation failed");
  }
  *out = M;
  return EW_OK;
}

extern "C" int ew_mixer_destroy(ew_mixer* M) {
  if (!M) return EW_OK;
  (void)hipSetDevice(M->opt.device);
  M->release();
  if (M->stream) (void)hipStreamDestroy(M->stream);
  delete M;
  return EW_OK;
}

extern "C" const char* ew_mixer_last_error(const ew_mixer* M) {
  return M ? M->err.c_str() : global_error().c_str();
}

// TBroydenMixer%reset (broydenmixer.F90:144-169): new length, history cleared.
extern "C" int ew_mixer_reset(ew_mixer* M, int n_elem) {
  if (!M) return set_global_error(EW_ERR_ARGUMENT, "ew_mixer_reset: mixer is NULL");
  if (n_elem <= 0) return M->fail(EW_ERR_ARGUMENT, "ew_mixer_reset: n_elem must be > 0");
  MCHK(M, hipSetDevice(M->opt.device));
  if (n_elem != M->n) {
    M->release();
    const size_t n = (size_t)n_elem, ld = (size_t)M->ld;
    M->bx = std::min(kMixMaxBlocks, std::max(1, (n_elem + 2047) / 2048));
    const size_t ng = (ld + kMixGroup - 1) / kMixGroup;
    if (mix_alloc(M, &M->qlast, n) || mix_alloc(M, &M->qdlast, n) || mix_alloc(M, &M->hq, n) ||
        mix_alloc(M, &M->hqd, n) || mix_alloc(M, &M->aa, ld * ld) || mix_alloc(M, &M->ww, ld) ||
        mix_alloc(M, &M->beta, ld * ld) || mix_alloc(M, &M->coef, ld) ||
        mix_alloc(M, &M->scal, 1) || mix_alloc(M, &M->ipiv, ld) ||
        mix_alloc(M, &M->partials, std::max<size_t>(ng, 1) * M->bx * kMixSlots)) {
      M->release();
      return EW_ERR_MEMORY;
    }
    if (M->opt.kind == EW_MIXER_BROYDEN &&
        (mix_alloc(M, &M->dF, ld * n) || mix_alloc(M, &M->uu, ld * n))) {
      M->release();
      return EW_ERR_MEMORY;
    }
    M->n = n_elem;
  }
  M->iter = 0;
  MCHK(M, hipMemsetAsync(M->ww, 0, sizeof(double) * M->ld, M->stream));
  MCHK(M, hipMemsetAsync(M->aa, 0, sizeof(double) * M->ld * M->ld, M->stream));
  MCHK(M, hipStreamSynchronize(M->stream));
  return EW_OK;
}

extern "C" int ew_mixer_mix_device(ew_mixer* M, double* d_q_inp_result, const double* d_q_diff,
                                   void* stream) {
  if (!M) return set_global_error(EW_ERR_ARGUMENT, "ew_mixer_mix_device: mixer is NULL");
  if (!d_q_inp_result || !d_q_diff) return M->fail(EW_ERR_ARGUMENT, "ew_mixer_mix_device: NULL");
  int st = begin_mix(M);
  if (st) return st;
  MCHK(M, hipSetDevice(M->opt.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : M->stream;
  st = enqueue_mix(M, d_q_inp_result, d_q_diff, s);
  if (st) M->iter -= 1;
  return st;
}

extern "C" int ew_mixer_mix(ew_mixer* M, double* q_inp_result, const double* q_diff) {
  if (!M) return set_global_error(EW_ERR_ARGUMENT, "ew_mixer_mix: mixer is NULL");
  if (!q_inp_result || !q_diff) return M->fail(EW_ERR_ARGUMENT, "ew_mixer_mix: NULL array");
  for (int i = 0; i < M->n; ++i)
    if (!std::isfinite(q_inp_result[i]) || !std::isfinite(q_diff[i]))
      return M->fail(EW_ERR_NOT_FINITE, "ew_mixer_mix: non-finite charge");
  int st = begin_mix(M);
  if (st) return st;
  MCHK(M, hipSetDevice(M->opt.device));
  hipStream_t s = M->stream;
  const size_t bytes = sizeof(double) * (size_t)M->n;
  MCHK(M, hipMemcpyAsync(M->hq, q_inp_result, bytes, hipMemcpyHostToDevice, s));
  MCHK(M, hipMemcpyAsync(M->hqd, q_diff, bytes, hipMemcpyHostToDevice, s));
  st = enqueue_mix(M, M->hq, M->hqd, s);
  if (st) {
    M->iter -= 1;
    return st;
  }
  MCHK(M, hipMemcpyAsync(q_inp_result, M->hq, bytes, hipMemcpyDeviceToHost, s));
  MCHK(M, hipStreamSynchronize(s));
  return EW_OK;
}

extern "C" int ew_mixer_iteration(const ew_mixer* M, int* iter) {
  if (!M || !iter) return set_global_error(EW_ERR_ARGUMENT, "ew_mixer_iteration: NULL argument");
  *iter = M->iter;
  return EW_OK;
}
