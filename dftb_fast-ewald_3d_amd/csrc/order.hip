// order.hip -- per-geometry ordering of the atoms on the device.
//
// Every atom gets its wrapped fractional coordinates, its B-spline stencil origin per axis
// (floor(s K) - (p - 1), modulo K) and the SPME sort tile of that origin; a stable radix sort by
// tile (hipCUB) gives the order the kernels walk, the per-atom arrays are gathered into it, and
// the tile starts are the exclusive sum of the tile counts.  Within a tile the atoms keep the
// caller's order, as a stable counting sort on the host would leave them.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace ewamd {

__global__ __launch_bounds__(256) void k_frac_tiles(int n, const double* __restrict__ coords,
                                                    OrderArgs g, double* __restrict__ frac_raw,
                                                    int4* __restrict__ k0_raw,
                                                    int* __restrict__ key, int* __restrict__ idx,
                                                    int* __restrict__ counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  int kk[3], tt[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = g.rec[3 * a] * x + g.rec[3 * a + 1] * y + g.rec[3 * a + 2] * z;
    s -= floor(s);
    frac_raw[3 * i + a] = s;
    int fl = int(floor(s * g.K[a]));
    if (fl >= g.K[a]) fl -= g.K[a];
    int k0 = fl - (g.order - 1);
    if (k0 < 0) k0 += g.K[a];
    kk[a] = k0;
    tt[a] = k0 / g.T[a];
  }
  const int t = (tt[2] * g.nt[1] + tt[1]) * g.nt[0] + tt[0];
  k0_raw[i] = make_int4(kk[0], kk[1], kk[2], 0);
  key[i] = t;
  idx[i] = i;
  atomicAdd(counts + t, 1);
}

__global__ __launch_bounds__(256) void k_gather_order(int n, const int* __restrict__ perm,
                                                      const double* __restrict__ frac_raw,
                                                      const int4* __restrict__ k0_raw,
                                                      const int* __restrict__ sp_raw,
                                                      const double* __restrict__ hub,
                                                      double* __restrict__ frac,
                                                      int4* __restrict__ k0,
                                                      int* __restrict__ sp,
                                                      double* __restrict__ onsite) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = perm[k];
  frac[3 * k] = frac_raw[3 * i];
  frac[3 * k + 1] = frac_raw[3 * i + 1];
  frac[3 * k + 2] = frac_raw[3 * i + 2];
  k0[k] = k0_raw[i];
  if (sp_raw) {
    const int si = sp_raw[i];
    sp[k] = si;
    onsite[k] = hub[si];  // -expGamma(0, U, U) = U  (shortgammafuncs.F90:128-131)
  }
}

static int key_bits(int ntiles) {
  int b = 1;
  while ((1LL << b) < (long long)ntiles) ++b;
  return b;
}

size_t order_temp_bytes(int n, int ntiles) {
  size_t a = 0, b = 0, c = 0;
  int* p = nullptr;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, p, p, p, p, n, 0, key_bits(ntiles));
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, p, p, ntiles + 1);
  (void)hipcub::DeviceReduce::Max(nullptr, c, p, p, ntiles);
  return std::max(a, std::max(b, c));
}

hipError_t launch_order(const OrderArgs& g, const OrderBufs& o, hipStream_t st) {
  const int n = g.n, ntiles = g.ntiles;
  hipError_t e = hipMemsetAsync(o.counts, 0, sizeof(int) * (ntiles + 1), st);
  if (e != hipSuccess) return e;
  const dim3 grd((n + 255) / 256), blk(256);
  hipLaunchKernelGGL(k_frac_tiles, grd, blk, 0, st, n, o.coords, g, o.frac_raw, o.k0_raw, o.key_in,
                     o.idx_in, o.counts);
  size_t bytes = o.temp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(o.temp, bytes, o.key_in, o.tile_of, o.idx_in, o.perm, n,
                                         0, key_bits(ntiles), st);
  if (e != hipSuccess) return e;
  bytes = o.temp_bytes;
  e = hipcub::DeviceReduce::Max(o.temp, bytes, o.counts, o.max_tile, ntiles, st);
  if (e != hipSuccess) return e;
  bytes = o.temp_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(o.temp, bytes, o.counts, o.tile_start, ntiles + 1, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather_order, grd, blk, 0, st, n, o.perm, o.frac_raw, o.k0_raw, o.sp_raw,
                     o.hub, o.frac, o.k0, o.species, o.onsite);
  return hipGetLastError();
}

size_t row_scan_temp_bytes(int n) {
  size_t b = 0;
  long long* p = nullptr;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, p, p, n + 1);
  return b;
}

hipError_t launch_row_scan(const long long* counts, long long* row_ptr, int n, void* temp,
                           size_t temp_bytes, hipStream_t st) {
  size_t bytes = temp_bytes;
  return hipcub::DeviceScan::ExclusiveSum(temp, bytes, counts, row_ptr, n + 1, st);
}

}  // namespace ewamd
