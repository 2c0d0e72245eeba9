// engine.hip -- host side of the MI355X periodic Coulomb engine and its C ABI.
//
// ew_set_geometry is the geometry step (TCoulomb%updateLatVecs/updateCoords and
// TShortGamma%updateCoords, src/dftbp/dftb/coulomb.F90:509-620, shortgamma.F90:177-195): it
// orders the atoms by SPME tile, uploads them and builds the resident pair list.  ew_shift is
// one SCC iteration (TCoulomb%updateShifts, coulomb.F90:786-880): gather charges, spread, FFT,
// influence, inverse FFT, pair sum, gather and combine -- nine kernels on the handle's stream.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fftewald_amd.h"
#include "engine.hpp"
#include "handle.hpp"
#include "group.hpp"
#include "kernels.hpp"

using namespace ewamd;

namespace ewamd {

std::string& global_error() {
  thread_local std::string msg = "no error";
  return msg;
}

int set_global_error(int code, const std::string& msg) {
  global_error() = msg;
  return code;
}

}  // namespace ewamd

namespace {

// FFT workgroup shape for lines of length n: want_lines lines, at most as many as fit the LDS
// budget; radix passes in the order of `variant` (fft_plan).
FftPlan make_plan(int n, int want_lines, int variant = 0) {
  FftPlan pl{};
  pl.n = n;
  pl.npass = fft_plan(n, pl.radix, variant);
  int lines = std::max(1, std::min(kFftMaxElems / n, want_lines));
  pl.lines = lines;
  // one pad slot per eight elements (line_slot) and an odd pitch between lines
  pl.pitch = n + n / 8 + 1;
  if (!(pl.pitch & 1)) ++pl.pitch;
  pl.nd = make_fastdiv(n);
  pl.hd = make_fastdiv(n / 2 + 1);
  pl.ld = make_fastdiv(lines);
  int ns = 1;
  for (int i = 0; i < pl.npass; ++i) {
    pl.nbd[i] = make_fastdiv(n / pl.radix[i]);
    pl.nsd[i] = make_fastdiv(ns);
    ns *= pl.radix[i];
  }
  return pl;
}

// Per-axis source tiles of every output tile for the spreading kernel.
bool build_source_table(int K, int T, int nt, int p, int ntmax, int axis, std::vector<int>& tab) {
  for (int t = 0; t < nt; ++t) {
    const int lo = t * T, w = std::min(T, K - lo);
    int* e = &tab[((size_t)axis * ntmax + t) * (kMaxSrcTiles + 1)];
    e[0] = 0;
    for (int s = 0; s < nt; ++s) {
      const int s_lo = s * T, s_hi = std::min(K, (s + 1) * T);
      bool hit = false;
      for (int k0 = s_lo; k0 < s_hi && !hit; ++k0)
        for (int o = 0; o < p && !hit; ++o) {
          int x = (k0 + o) % K - lo;
          if (x < 0) x += K;
          if (x < w) hit = true;
        }
      if (hit) {
        if (e[0] >= kMaxSrcTiles) return false;
        e[1 + e[0]++] = s;
      }
    }
  }
  return true;
}

template <typename R>
void twiddles(int n, std::vector<R>& out, int variant) {
  // Per-pass tables, concatenated in pass order: a pass of radix r after a product Ns of the
  // earlier radices holds exp(-2 pi i t k / (Ns r)) at [(t - 1) * Ns + k], t = 1..r-1,
  // k = 0..Ns-1, so that the lanes of a butterfly read consecutive entries.  The sizes
  // telescope to n - 1 entries; one spare keeps the table n long.
  int radix[kMaxRadixPasses];
  const int np = fft_plan(n, radix, variant);
  out.assign(2 * (size_t)n, R(0));
  size_t o = 0;
  int Ns = 1;
  for (int i = 0; i < np; ++i) {
    const int r = radix[i];
    for (int t = 1; t < r; ++t)
      for (int k = 0; k < Ns; ++k, ++o) {
        const long double a =
            2.0L * 3.14159265358979323846264338327950288L * t * k / ((long double)Ns * r);
        out[2 * o] = R(cosl(a));
        out[2 * o + 1] = R(-sinl(a));
      }
    Ns *= r;
  }
}

// W_n^m = exp(-2 pi i m / n) for m < n, interleaved (re, im).
template <typename R>
void plain_roots(int n, std::vector<R>& out) {
  out.assign(2 * (size_t)n, R(0));
  for (int m = 0; m < n; ++m) {
    const long double a = 2.0L * 3.14159265358979323846264338327950288L * m / (long double)n;
    out[2 * (size_t)m] = R(cosl(a));
    out[2 * (size_t)m + 1] = R(-sinl(a));
  }
}

long long estimate_pairs(int n, const Cell& c, double r_list) {
  const double vol_sphere = 4.0 / 3.0 * kPi * r_list * r_list * r_list;
  double per_atom = double(n) / c.volume * vol_sphere;
  return (long long)std::ceil(double(n) * per_atom * 1.05 + 64.0);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------

extern "C" const char* ew_last_error(const ew_handle* h) {
  return h ? h->err.c_str() : global_error().c_str();
}

// ---- the allocation-free evaluation's scratch (ew_shift_device_noalloc) ----
// Charges, potentials, energy partials, mesh, spectrum and spreading copies of one evaluation,
// carved from one static device array that each device holds until the process ends.  In host
// code `scratch` only names the symbol: its device address comes from hipGetSymbolAddress, once
// per device (with that device current), and only that address is passed to kernels.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static void* g_scratch_dev[kMaxScratchDevices] = {};
static std::mutex g_scratch_mutex;

// Resolves the device address of `scratch` on the current device `device` (once per device).
static int resolve_static_scratch(int device) {
  if (device < 0 || device >= kMaxScratchDevices)
    return set_global_error(EW_ERR_DEVICE, "device ordinal beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_dev[device]) return EW_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return set_global_error(EW_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                               hipGetErrorString(e));
  }
  g_scratch_dev[device] = p;
  return EW_OK;
}

int ewamd::handle_create(const ew_options& o, hipStream_t borrowed, ew_handle** out) {
  *out = nullptr;
  if (o.precision != EW_PRECISION_FP64 && o.precision != EW_PRECISION_MIXED)
    return set_global_error(EW_ERR_ARGUMENT, "unknown precision mode");
  if (o.fft_mode < 0 || o.fft_mode > 2)
    return set_global_error(EW_ERR_ARGUMENT, "fft_mode must be 0, 1 or 2");
  if (o.pair_storage < 0 || o.pair_storage > 1)
    return set_global_error(EW_ERR_ARGUMENT, "pair_storage must be 0 or 1");
  if (o.profile < 0 || o.profile > 2)
    return set_global_error(EW_ERR_ARGUMENT, "profile must be 0, 1 or 2");
  if (o.tune < 0 || o.tune > 3) return set_global_error(EW_ERR_ARGUMENT, "tune must be 0..3");
  if (o.group_grid < 0 || o.group_grid > 1)
    return set_global_error(EW_ERR_ARGUMENT, "group_grid must be 0 or 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return set_global_error(EW_ERR_DEVICE, "no HIP device available");
  }
  if (o.device < 0 || o.device >= ndev)
    return set_global_error(EW_ERR_ARGUMENT, "device ordinal out of range");
  if (hipSetDevice(o.device) != hipSuccess)
    return set_global_error(EW_ERR_DEVICE, "hipSetDevice failed");
  if (const int st = resolve_static_scratch(o.device)) return st;
  ew_handle* h = new ew_handle();
  h->opt = o;
  h->fp32 = o.precision == EW_PRECISION_MIXED;
  if (borrowed) {
    h->stream = borrowed;
    h->own_stream = false;
  } else if (hipStreamCreate(&h->stream) != hipSuccess ||
             hipStreamCreate(&h->aux) != hipSuccess ||
             hipEventCreateWithFlags(&h->fork_ev, kSyncEventFlags) != hipSuccess ||
             hipEventCreateWithFlags(&h->join_ev, kSyncEventFlags) != hipSuccess ||
             hipEventCreateWithFlags(&h->spread_ev, kSyncEventFlags) != hipSuccess) {
    (void)hipGetLastError();
    if (h->spread_ev) (void)hipEventDestroy(h->spread_ev);
    if (h->join_ev) (void)hipEventDestroy(h->join_ev);
    if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
    if (h->aux) (void)hipStreamDestroy(h->aux);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return set_global_error(EW_ERR_DEVICE, "hipStreamCreate failed");
  }
  if (hipEventCreateWithFlags(&h->last_ev, kSyncEventFlags) != hipSuccess) {
    (void)hipGetLastError();
    if (h->spread_ev) (void)hipEventDestroy(h->spread_ev);
    if (h->join_ev) (void)hipEventDestroy(h->join_ev);
    if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
    if (h->aux) (void)hipStreamDestroy(h->aux);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return set_global_error(EW_ERR_DEVICE, "hipEventCreate failed");
  }
  if (o.profile == 1)
    for (auto& e : h->ev)
      if (hipEventCreate(&e) != hipSuccess) {
        if (h->own_stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return set_global_error(EW_ERR_DEVICE, "hipEventCreate failed");
      }
  // timing only: no system-scope fence (cache write-back) at each event, which would slow the
  // kernels it separates
  if (o.profile == 2)
    for (auto& e : h->tev)
      if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
        if (h->own_stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return set_global_error(EW_ERR_DEVICE, "hipEventCreate failed");
      }
  *out = h;
  return EW_OK;
}

extern "C" int ew_create(const ew_options* opts, ew_handle** out) {
  if (!out) return set_global_error(EW_ERR_ARGUMENT, "ew_create: out is NULL");
  *out = nullptr;
  ew_options o{};
  if (opts) o = *opts;
  const int st = handle_create(o, nullptr, out);
  if (st != EW_OK) global_error() = "ew_create: " + global_error();
  return st;
}

extern "C" int ew_destroy(ew_handle* h) {
  if (!h) return EW_OK;
  (void)hipSetDevice(h->opt.device);
  h->release();
  h->trim();
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : h->tev)
    if (e) (void)hipEventDestroy(e);
  if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
  if (h->join_ev) (void)hipEventDestroy(h->join_ev);
  if (h->spread_ev) (void)hipEventDestroy(h->spread_ev);
  if (h->last_ev) (void)hipEventDestroy(h->last_ev);
  if (h->aux) (void)hipStreamDestroy(h->aux);
  if (h->stream && h->own_stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return EW_OK;
}

static int validate_inputs(ew_handle* h, int n, const double* coords, const double* latvecs) {
  if (n <= 0) return h->fail(EW_ERR_ARGUMENT, "nParticle must be positive");
  if (!coords || !latvecs) return h->fail(EW_ERR_ARGUMENT, "coords and latvecs are required");
  for (size_t i = 0; i < 3 * (size_t)n; ++i)
    if (!std::isfinite(coords[i])) return h->fail(EW_ERR_NOT_FINITE, "non-finite coordinate");
  return EW_OK;
}

// Symmetric storage of the linked-tile pair list (each unordered pair once, grouped by the tile
// of its lower atom and by chunks of <= 256 atoms of that tile's partner tiles), built from the
// full row list; on success the row list is freed.  Skipped (h->half stays false) when a tile
// holds more than 255 atoms or a tile has more than kHalfMaxChunks chunks.
static int build_half_pairs(ew_handle* h, const std::vector<int>& ts, hipStream_t s) {
  const int nt0 = h->nt[0], nt1 = h->nt[1], nt2 = h->nt[2];
  const int ntiles = nt0 * nt1 * nt2;
  for (int t = 0; t < ntiles; ++t)
    if (ts[t + 1] - ts[t] > 255) return EW_OK;
  const PairArgs& pa = h->pa;
  std::vector<int> tile_chunk(ntiles + 1, 0), chunk_tile(1, 0), chunk_atoms, ptile, ptile_off,
      ptile_chunk, rc_base(ntiles + 1, 0);
  std::vector<int> cand;
  for (int A = 0; A < ntiles; ++A) {
    const int t1 = A % nt0, t2 = (A / nt0) % nt1, t3 = A / (nt0 * nt1);
    const int nt[3] = {nt0, nt1, nt2}, tc[3] = {t1, t2, t3};
    int b[3], m[3];
    for (int a = 0; a < 3; ++a) {
      b[a] = pa.full[a] ? 0 : tc[a] - pa.D[a];
      m[a] = pa.full[a] ? nt[a] : 2 * pa.D[a] + 1;
    }
    cand.clear();
    for (int c3 = 0; c3 < m[2]; ++c3)
      for (int c2 = 0; c2 < m[1]; ++c2)
        for (int c1 = 0; c1 < m[0]; ++c1) {
          const int u1 = ((b[0] + c1) % nt0 + nt0) % nt0;
          const int u2 = ((b[1] + c2) % nt1 + nt1) % nt1;
          const int u3 = ((b[2] + c3) % nt2 + nt2) % nt2;
          const int B = (u3 * nt1 + u2) * nt0 + u1;
          if (B >= A) cand.push_back(B);
        }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    if (ts[A + 1] == ts[A]) cand.clear();  // an empty tile has no pairs
    int rel = -1, fill = 0;
    for (int B : cand) {
      const int nb = ts[B + 1] - ts[B];
      if (nb == 0) continue;  // empty partner tiles hold no atoms
      if (rel < 0 || fill + nb > 256) {  // open a chunk
        ++rel;
        if (rel >= kHalfMaxChunks) return EW_OK;
        chunk_atoms.push_back(0);
        chunk_tile.push_back(chunk_tile.back());
        fill = 0;
      }
      ptile.push_back(B);
      ptile_off.push_back(fill);
      ptile_chunk.push_back(rel);
      fill += nb;
      chunk_atoms.back() = fill;
      chunk_tile.back() = (int)ptile.size();
    }
    tile_chunk[A + 1] = (int)chunk_atoms.size();
    rc_base[A + 1] = rc_base[A] + (rel + 1) * (ts[A + 1] - ts[A]);
  }
  const int nchunks = (int)chunk_atoms.size();
  // invariants the kernels rely on: each chunk stages exactly the atoms of its partner tiles
  // (at most 256), and every partner tile belongs to a chunk of its tile
  for (int c = 0; c < nchunks; ++c) {
    int atoms = 0;
    for (int k = chunk_tile[c]; k < chunk_tile[c + 1]; ++k) {
      if (ptile_off[k] != atoms || ptile_chunk[k] < 0) return EW_OK;
      atoms += ts[ptile[k] + 1] - ts[ptile[k]];
    }
    if (atoms != chunk_atoms[c] || atoms > 256) return EW_OK;
  }
  int *d_tc, *d_ct, *d_ce, *d_ca, *d_pt, *d_po, *d_pc, *d_rb, *d_rc;
  CHK(h->alloc(&d_tc, ntiles + 1));
  CHK(h->alloc(&d_ct, nchunks + 1));
  CHK(h->alloc(&d_ce, nchunks + 1));
  CHK(h->alloc(&d_ca, std::max(1, nchunks)));
  CHK(h->alloc(&d_pt, std::max<size_t>(1, ptile.size())));
  CHK(h->alloc(&d_po, std::max<size_t>(1, ptile.size())));
  CHK(h->alloc(&d_pc, std::max<size_t>(1, ptile.size())));
  CHK(h->alloc(&d_rb, ntiles + 1));
  CHK(h->alloc(&d_rc, std::max(1, rc_base[ntiles])));
  auto up = [&](int* d, const std::vector<int>& v) {
    return v.empty() ? hipSuccess
                     : hipMemcpyAsync(d, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, s);
  };
  HIPCHK(h, up(d_tc, tile_chunk));
  HIPCHK(h, up(d_ct, chunk_tile));
  HIPCHK(h, up(d_ca, chunk_atoms));
  HIPCHK(h, up(d_pt, ptile));
  HIPCHK(h, up(d_po, ptile_off));
  HIPCHK(h, up(d_pc, ptile_chunk));
  HIPCHK(h, up(d_rb, rc_base));
  HalfPairs hp{};
  hp.tile_chunk = d_tc;
  hp.chunk_tile = d_ct;
  hp.chunk_entry = d_ce;
  hp.chunk_atoms = d_ca;
  hp.ptile = d_pt;
  hp.ptile_off = d_po;
  hp.ptile_chunk = d_pc;
  hp.rc_base = d_rb;
  HIPCHK(h, launch_half_count(h->n, h->row_ptr, h->col, h->tile_of, h->tile_start, hp, d_rc, s));
  std::vector<int> rc(rc_base[ntiles]);
  if (!rc.empty())
    HIPCHK(h, hipMemcpyAsync(rc.data(), d_rc, sizeof(int) * rc.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  // starting positions: chunk by chunk, rows in order within a chunk
  std::vector<int> chunk_entry(nchunks + 1, 0);
  long long run = 0;
  for (int A = 0; A < ntiles; ++A) {
    const int nA = ts[A + 1] - ts[A];
    for (int c = tile_chunk[A]; c < tile_chunk[A + 1]; ++c) {
      const int rel = c - tile_chunk[A];
      chunk_entry[c] = (int)run;
      for (int i = 0; i < nA; ++i) {
        int& v = rc[rc_base[A] + rel * nA + i];
        const int cnt = v;
        v = (int)run;
        run += cnt;
      }
    }
  }
  if (run >= (1LL << 31)) return EW_OK;  // stays on the row list
  chunk_entry[nchunks] = (int)run;
  HIPCHK(h, up(d_ce, chunk_entry));
  HIPCHK(h, up(d_rc, rc));
  unsigned short* ent;
  unsigned char* cf;
  CHK(h->alloc(&ent, (size_t)run + 8));
  CHK(h->alloc(&cf, ((size_t)run + 8) * (h->fp32 ? sizeof(float) : sizeof(double))));
  HIPCHK(h, launch_half_fill(h->n, h->row_ptr, h->col, h->coef, h->fp32, h->tile_of,
                             h->tile_start, hp, d_rc, ent, cf, s));
  HIPCHK(h, hipStreamSynchronize(s));
  hp.entry = ent;
  hp.coef = cf;
  h->free_buf(d_rc);
  h->free_buf(h->col);
  h->free_buf(h->coef);
  h->col = nullptr;
  h->coef = nullptr;
  h->hp = hp;
  h->half = true;
  h->half_entries = run;
  return EW_OK;
}

// Points in the order the kernels see them: fractional coordinates wrapped into [0, 1), the
// B-spline stencil origin per axis (floor(s K) - (p - 1), modulo K), the SPME tile of that
// origin, and a stable counting sort by tile.  Needs the handle's mesh and tiles.
struct TileOrder {
  std::vector<int> perm;      // sorted position -> caller index
  std::vector<int> start;     // [ntiles + 1] first sorted position of each tile
  std::vector<double> frac;   // sorted [k][3]
  std::vector<int4> k0;       // sorted stencil origins
  std::vector<int> tile;      // sorted tile indices
  int max_tile = 0;           // most points in one tile
};

static void order_by_tile(const ew_handle* h, int n, const double* coords, TileOrder& o) {
  const Cell& c = h->cell;
  const int* K = h->prm.K;
  const int p = h->prm.order;
  const int ntiles = h->nt[0] * h->nt[1] * h->nt[2];
  std::vector<double> fr(3 * (size_t)n);
  std::vector<int> k0o(3 * (size_t)n), tile(n);
  o.start.assign(ntiles + 1, 0);
  for (int i = 0; i < n; ++i) {
    int tt[3];
    for (int a = 0; a < 3; ++a) {
      double s = c.rec[a][0] * coords[3 * i] + c.rec[a][1] * coords[3 * i + 1] +
                 c.rec[a][2] * coords[3 * i + 2];
      s -= std::floor(s);
      fr[3 * (size_t)i + a] = s;
      int fl = int(std::floor(s * K[a]));
      if (fl >= K[a]) fl -= K[a];
      int k0 = fl - (p - 1);
      if (k0 < 0) k0 += K[a];
      k0o[3 * (size_t)i + a] = k0;
      tt[a] = k0 / h->T[a];
    }
    tile[i] = (tt[2] * h->nt[1] + tt[1]) * h->nt[0] + tt[0];
    o.start[tile[i] + 1]++;
  }
  o.max_tile = 0;
  for (int t = 0; t < ntiles; ++t) o.max_tile = std::max(o.max_tile, o.start[t + 1]);
  for (int t = 0; t < ntiles; ++t) o.start[t + 1] += o.start[t];
  o.perm.assign(n, 0);
  std::vector<int> fill(o.start.begin(), o.start.end() - 1);
  for (int i = 0; i < n; ++i) o.perm[fill[tile[i]]++] = i;
  o.frac.assign(3 * (size_t)n, 0.0);
  o.k0.assign(n, make_int4(0, 0, 0, 0));
  o.tile.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    const int i = o.perm[k];
    for (int a = 0; a < 3; ++a) o.frac[3 * (size_t)k + a] = fr[3 * (size_t)i + a];
    o.k0[k] = make_int4(k0o[3 * (size_t)i], k0o[3 * (size_t)i + 1], k0o[3 * (size_t)i + 2], 0);
    o.tile[k] = tile[i];
  }
}

// ------------------------------------------------------------------------------------------
// Launch shapes (ew_tuning): fixed rules, their application, the tuning cache and the tuner
// ------------------------------------------------------------------------------------------

static int enqueue_eval(ew_handle* h, const double* d_dq, double* d_shift, double* d_energy,
                        hipStream_t s);

// Privatised spreading blocks of f[a] sort tiles per axis, if the mesh holds them (a copy may
// cover no point twice: K >= S + p - 1) and their FP64 LDS copy fits max_bytes.
static bool region_shape(const ew_handle* h, const int* f, size_t max_bytes, RegionArgs* out) {
  const int* K = h->prm.K;
  const int p = h->prm.order;
  RegionArgs rg{};
  for (int a = 0; a < 3; ++a) {
    if (f[a] < 1) return false;
    rg.K[a] = K[a];
    rg.f[a] = f[a];
    rg.S[a] = f[a] * h->T[a];
    rg.nt[a] = h->nt[a];
    rg.nb[a] = (h->nt[a] + f[a] - 1) / f[a];
    if (rg.S[a] > K[a] || K[a] < rg.S[a] + p - 1) return false;
  }
  rg.order = p;
  rg.region_stride = (long long)(rg.S[0] + p - 1) * (rg.S[1] + p - 1) * (rg.S[2] + p - 1);
  if (sizeof(double) * (size_t)rg.region_stride > max_bytes) return false;
  if (out) *out = rg;
  return true;
}

// Sort tiles per axis whose atoms an owned block of rg's shape may have to read: a block starts
// on a tile boundary, so the stencil origins lo-(p-1) .. lo+S-1 it gathers lie in its own f
// tiles and the ceil((p-1)/T) tiles before them, one more when they wrap through a partial
// last tile (K not a multiple of T).
static void owned_tile_span(const ew_handle* h, const RegionArgs& rg, int* tcm) {
  for (int a = 0; a < 3; ++a) {
    const int T = h->T[a];
    const int partial = h->prm.K[a] % T ? 1 : 0;
    tcm[a] = std::min(h->nt[a], rg.f[a] + (h->prm.order - 1 + T - 1) / T + partial);
  }
}

// Entries of an owned block's item list: p planes for every atom the block may read (the tiles
// of owned_tile_span at the geometry's fullest tile), so one pass takes them all; at most
// kOwnedItems (a block with more atoms takes several passes).
static int owned_item_count(const ew_handle* h, const RegionArgs& rg) {
  int tcm[3];
  owned_tile_span(h, rg, tcm);
  const long long atoms = (long long)std::max(1, h->max_tile_atoms) * tcm[0] * tcm[1] * tcm[2];
  const long long want = (atoms * h->prm.order + 63) / 64 * 64;
  return (int)std::max<long long>(h->prm.order, std::min<long long>(kOwnedItems, want));
}

// Owned blocks can take a block shape if the block copy and the item list fit the LDS, the tile
// runs fit the kernel's tables and an item (atom << 4 | plane) fits 31 bits.
static bool owned_fits(const ew_handle* h, const RegionArgs& rg) {
  int tcm[3];
  owned_tile_span(h, rg, tcm);
  RegionArgs sized = rg;
  sized.items = owned_item_count(h, rg);
  return owned_smem(sized, h->fp32) + kOwnedStatic <= kRegionMaxLds &&
         2 * tcm[1] * tcm[2] <= kOwnedMaxRuns && h->prm.order <= 15 && h->n < (1 << 27);
}

// The spreading's write-out by the fixed rule: owned blocks (on the 98,304-atom cell at 128^3,
// 0.211 against 0.224 ms per FP64 evaluation with block copies plus assembly, and 0.160
// against 0.166 ms mixed; direct write-out of the copies with device atomics measured between
// the two); the autotuner chooses per structure.
constexpr int kDirectRule = 2;
// The pair sum beside the chain on the auxiliary stream by the fixed rule (co-scheduling it in
// the chain's launches measured slower on the 98,304-atom cell: 0.226 against 0.189 ms mixed;
// faster on the 12,288-atom water box: 0.096 against 0.097 ms with 20 % of the rows in the
// spreading launch); the autotuner chooses per structure.
constexpr int kCoschedRule = 0;
constexpr int kCoSpreadShareRule = 0;

// The fixed rules: the largest whole-sort-tile spreading blocks whose FP64 copy fits 60 KiB;
// fused x-y plane transforms in mixed precision on planes that occupy a 1024-thread workgroup
// (an FP64 plane needs up to 133 KiB of LDS, one workgroup per CU, and small planes leave
// most of the workgroup idle); x passes of as many rows as fit, y and z passes of 16 lines;
// radix 16 first (two radix-4 stages in registers per pass: on the 98,304-atom cell at 128^3
// 0.164 against 0.172 ms per mixed evaluation, 0.222 against 0.226 ms in FP64).
static ew_tuning rule_tuning(const ew_handle* h) {
  ew_tuning t{};
  const int* K = h->prm.K;
  long long best_pts = 0;
  RegionArgs best_rg{};
  for (int f3 = 1; f3 <= 4; ++f3)
    for (int f2 = 1; f2 <= 4; ++f2)
      for (int f1 = 1; f1 <= 2; ++f1) {
        const int f[3] = {f1, f2, f3};
        RegionArgs rg;
        if (!region_shape(h, f, kRegionMaxBytes, &rg)) continue;
        const long long pts = (long long)rg.S[0] * rg.S[1] * rg.S[2];
        if (pts > best_pts) {
          best_pts = pts;
          best_rg = rg;
          for (int a = 0; a < 3; ++a) t.spread_block[a] = f[a];
        }
      }
  t.spread_direct = 0;
  if (best_pts > 0) {
    if (kDirectRule == 2 && owned_fits(h, best_rg)) t.spread_direct = 2;
    else if (kDirectRule == 1 && h->own_stream) t.spread_direct = 1;
  }
  if (t.spread_direct == 2) {
    // owned blocks keep no halo in LDS: blocks of 4096 mesh points balance the atoms read twice
    // near block faces against workgroups per CU (98,304-atom cell at 128^3, spreading alone:
    // 36 us with 16 x 16 x 16 against 38-44 us for 1024-2048 points and 48-53 us for 8192)
    static const int pref[][3] = {{1, 2, 2}, {2, 2, 1}, {1, 4, 1}, {1, 1, 4}};
    for (const auto& f : pref) {
      RegionArgs rg;
      if (region_shape(h, f, kRegionMaxLds, &rg) && owned_fits(h, rg)) {
        for (int a = 0; a < 3; ++a) t.spread_block[a] = f[a];
        break;
      }
    }
  }
  const bool big_plane = (long long)(K[0] / 2 + 1) * K[1] >= kPlaneMinPoints;
  t.fft_planes = (h->opt.fft_mode == 2 || (h->opt.fft_mode == 0 && h->fp32 && big_plane)) ? 1 : 0;
  t.fft_lines[0] = (K[1] * K[2] + 1) / 2;
  t.fft_lines[1] = 16;
  t.fft_lines[2] = 16;
  for (int a = 0; a < 3; ++a) t.fft_radix[a] = 4;
  t.pair_cosched = kCoschedRule;
  t.pair_spread_share = kCoSpreadShareRule;
  // register transforms: x-y plane kernels on square planes of a compiled side with a compiled z
  // length, else line passes where every axis length has a compiled shape
  t.fft_wave = 1;
  for (int a = 0; a < 3; ++a) {
    t.fft_lanes[a] = wave_fft_default_lanes(K[a], h->fp32);
    if (!t.fft_lanes[a]) t.fft_wave = 0;
  }
  if (K[0] == K[1] && wave_plane_default_lanes(K[0]) && wave_fft_default_lanes(K[2], h->fp32)) {
    t.fft_wave = 2;
    t.fft_lanes[0] = t.fft_lanes[1] = wave_plane_default_lanes(K[0]);
    t.fft_lanes[2] = wave_fft_default_lanes(K[2], h->fp32);
  }
  // FP64 beyond 128 points per axis: the Stockham passes (256^3, 98,304 atoms, each pass alone:
  // 414 against 445 us for the register line passes, whose x passes are the slower ones)
  if (!h->fp32 && std::max(K[0], std::max(K[1], K[2])) > 128) t.fft_wave = 0;
  if (t.fft_wave) {
    t.pair_cosched = 0;
    t.pair_spread_share = 0;
  }
  return t;
}

// Sets the region blocks, FFT plans, twiddle tables and plane flag of the current geometry to
// the shapes t (what a shape cannot have is corrected: lines capped by the LDS, planes only
// where a plane fits, direct write-out only on own-stream handles).  Drops captured graphs.
static int apply_tuning(ew_handle* h, const ew_tuning& t_in) {
  ew_tuning t = t_in;
  const int* K = h->prm.K;
  const size_t rs = h->fp32 ? sizeof(float) : sizeof(double);
  RegionArgs rg{};
  if (t.spread_block[0] > 0 || t.spread_block[1] > 0 || t.spread_block[2] > 0) {
    if (!region_shape(h, t.spread_block, kRegionMaxLds, &rg))
      return h->fail(EW_ERR_ARGUMENT, "spreading block does not fit the mesh or the LDS");
    // 2: owned blocks (any handle); 1: direct write-out (own-stream handles); 0: copies
    rg.direct = t.spread_direct == 2 ? 2 : (t.spread_direct && h->own_stream ? 1 : 0);
    if (rg.direct == 2 && !owned_fits(h, rg))
      return h->fail(EW_ERR_ARGUMENT, "spreading block does not fit the mesh or the LDS");
    if (rg.direct == 2) rg.items = owned_item_count(h, rg);
  } else {
    for (int a = 0; a < 3; ++a) t.spread_block[a] = 0;
  }
  t.spread_direct = rg.direct;
  t.pair_cosched = t.pair_cosched ? 1 : 0;
  t.pair_blocks = std::max(0, std::min(1 << 16, t.pair_blocks));
  t.pair_fork = t.pair_fork ? 1 : 0;
  t.pair_spread_share = std::max(0, std::min(100, t.pair_spread_share));
  for (int a = 0; a < 3; ++a)
    if (t.fft_radix[a] < 0 || t.fft_radix[a] >= kRadixVariants || t.fft_lines[a] < 1)
      return h->fail(EW_ERR_ARGUMENT, "FFT radix order must be 0..4 and lines >= 1");
  // nothing queued may still use the old shapes, copies or tables
  h->quiesce();
  h->drop_graphs();
  if (h->region.scratch) h->free_buf(h->region.scratch);
  h->region = rg;
  if (rg.region_stride > 0 && !rg.direct) {
    const size_t nblk = (size_t)rg.nb[0] * rg.nb[1] * rg.nb[2];
    unsigned char* scr;
    CHK(h->alloc(&scr, nblk * (size_t)rg.region_stride * rs));
    h->region.scratch = scr;
  }
  if (sizeof(double) * (size_t)rg.region_stride > 65536 ||
      (rg.direct == 2 && owned_smem(rg, h->fp32) > 65536))
    HIPCHK(h, prepare_region_kernels());
  h->px = make_plan(K[0], t.fft_lines[0], t.fft_radix[0]);
  h->py = make_plan(K[1], t.fft_lines[1], t.fft_radix[1]);
  h->pz = make_plan(K[2], t.fft_lines[2], t.fft_radix[2]);
  const FftPlan* pl[3] = {&h->px, &h->py, &h->pz};
  for (int a = 0; a < 3; ++a) {
    if (pl[a]->npass <= 0) return h->fail(EW_ERR_ARGUMENT, "grid not 2^a 3^b 5^c");
    t.fft_lines[a] = pl[a]->lines;
    if (h->fp32) {
      std::vector<float>& tv = h->tab_tw32[a];
      twiddles(K[a], tv, t.fft_radix[a]);
      HIPCHK(h, hipMemcpyAsync(h->tw[a], tv.data(), sizeof(float) * tv.size(),
                               hipMemcpyHostToDevice, h->stream));
    } else {
      std::vector<double>& tv = h->tab_tw64[a];
      twiddles(K[a], tv, t.fft_radix[a]);
      HIPCHK(h, hipMemcpyAsync(h->tw[a], tv.data(), sizeof(double) * tv.size(),
                               hipMemcpyHostToDevice, h->stream));
    }
  }
  h->plane = t.fft_planes && h->opt.fft_mode != 1 && plane_fits(h->px, h->py, 2 * rs);
  if (h->plane) HIPCHK(h, prepare_plane_kernels());
  t.fft_planes = h->plane ? 1 : 0;
  // register transforms where every pass has a compiled shape (the co-scheduled pair sum runs
  // only inside the Stockham launches): 2 = x-y plane kernels (square planes) + z lines, 1 =
  // x, y and z line passes; otherwise the Stockham passes
  // the handle's fft_mode holds for these too: 1 = separate passes (register lines), 2 = fused
  // planes wherever a compiled plane kernel fits
  const bool plane_ok = K[0] == K[1] && wave_plane_default_lanes(K[0]) > 0 &&
                        wave_fft_default_lanes(K[2], h->fp32) > 0;
  if (t.fft_wave == 2 && h->opt.fft_mode == 1) {
    t.fft_wave = 1;
    for (int a = 0; a < 3; ++a) t.fft_lanes[a] = wave_fft_default_lanes(K[a], h->fp32);
  } else if (t.fft_wave == 1 && h->opt.fft_mode == 2 && plane_ok) {
    t.fft_wave = 2;
    t.fft_lanes[0] = t.fft_lanes[1] = wave_plane_default_lanes(K[0]);
    t.fft_lanes[2] = wave_fft_default_lanes(K[2], h->fp32);
  }
  int wave = 0;
  if (t.fft_wave == 2 && K[0] == K[1] && wave_plane_shape(K[0], t.fft_lanes[0]) &&
      wave_fft_shape(K[2], t.fft_lanes[2])) {
    wave = 2;
    t.fft_lanes[1] = t.fft_lanes[0];
  } else if (t.fft_wave != 0) {
    wave = 1;
    for (int a = 0; a < 3; ++a)
      if (!wave_fft_shape(K[a], t.fft_lanes[a])) wave = 0;
  }
  h->wave = wave;
  t.fft_wave = wave;
  for (int a = 0; a < 3; ++a) {
    if (!wave) t.fft_lanes[a] = 0;
    h->lanes[a] = t.fft_lanes[a];
  }
  if (wave) {
    t.pair_cosched = 0;
    t.pair_spread_share = 0;
  }
  h->tuning = t;
  return EW_OK;
}

// Key of the tuning cache: device, mesh, order, precision, size class (log2 of the atoms),
// pair storage and the handle's FFT mode option.
static std::string tune_key(const ew_handle* h) {
  hipDeviceProp_t prop;
  std::string arch = "gpu";
  if (hipGetDeviceProperties(&prop, h->opt.device) == hipSuccess) arch = prop.gcnArchName;
  for (char& ch : arch)
    if (ch == ' ' || ch == ':') ch = '_';
  int nb = 0;
  while ((1 << (nb + 1)) <= h->n) ++nb;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "v5:%s:%dx%dx%d:p%d:%s:n%d:ps%d:fm%d:own%d", arch.c_str(),
                h->prm.K[0], h->prm.K[1], h->prm.K[2], h->prm.order, h->fp32 ? "mixed" : "fp64",
                nb, h->opt.pair_storage, h->opt.fft_mode, h->own_stream ? 1 : 0);
  return buf;
}

static std::mutex g_tune_mutex;
static bool g_tune_loaded = false;
static std::vector<std::pair<std::string, ew_tuning>> g_tune_entries;

// $EW_TUNE_CACHE, else ewald_tune.cache in the directory that holds this library.
static std::string tune_cache_path() {
  const char* env = std::getenv("EW_TUNE_CACHE");
  if (env && *env) return env;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&tune_cache_path), &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const size_t slash = p.rfind('/');
    return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) +
           "/ewald_tune.cache";
  }
  return "ewald_tune.cache";
}

// One line per key: key direct f0 f1 f2 planes l0 l1 l2 r0 r1 r2 cosched share eval_us wave
// lanes0 lanes1 lanes2 pair_blocks pair_fork.
static bool parse_tuning(const char* line, std::string* key, ew_tuning* t) {
  char k[200];
  ew_tuning u{};
  const int got = std::sscanf(line,
                              "%199s %d %d %d %d %d %d %d %d %d %d %d %d %d %f %d %d %d %d %d %d",
                              k, &u.spread_direct, &u.spread_block[0], &u.spread_block[1],
                              &u.spread_block[2], &u.fft_planes, &u.fft_lines[0], &u.fft_lines[1],
                              &u.fft_lines[2], &u.fft_radix[0], &u.fft_radix[1], &u.fft_radix[2],
                              &u.pair_cosched, &u.pair_spread_share, &u.eval_us, &u.fft_wave,
                              &u.fft_lanes[0], &u.fft_lanes[1], &u.fft_lanes[2], &u.pair_blocks,
                              &u.pair_fork);
  if (got != 21) return false;
  *key = k;
  *t = u;
  return true;
}

static void tune_cache_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  FILE* f = std::fopen(tune_cache_path().c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    std::string k;
    ew_tuning t;
    if (parse_tuning(line, &k, &t)) g_tune_entries.emplace_back(k, t);  // later lines win
  }
  std::fclose(f);
}

static bool tune_cache_get(const std::string& key, ew_tuning* out) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  tune_cache_load_locked();
  for (auto it = g_tune_entries.rbegin(); it != g_tune_entries.rend(); ++it)
    if (it->first == key) {
      *out = it->second;
      return true;
    }
  return false;
}

static void tune_cache_put(const std::string& key, const ew_tuning& t) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  tune_cache_load_locked();
  g_tune_entries.emplace_back(key, t);
  FILE* f = std::fopen(tune_cache_path().c_str(), "a");
  if (!f) return;  // a read-only location: the choice still holds for this process
  std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d %d %d %d %.3f %d %d %d %d %d %d\n",
               key.c_str(),
               t.spread_direct, t.spread_block[0], t.spread_block[1], t.spread_block[2],
               t.fft_planes, t.fft_lines[0], t.fft_lines[1], t.fft_lines[2], t.fft_radix[0],
               t.fft_radix[1], t.fft_radix[2], t.pair_cosched, t.pair_spread_share, t.eval_us,
               t.fft_wave, t.fft_lanes[0], t.fft_lanes[1], t.fft_lanes[2], t.pair_blocks,
               t.pair_fork);
  std::fclose(f);
}

// Device time of one evaluation in microseconds with the shapes t applied: launched as an
// ordinary evaluation on the handle's stream, on synthetic neutral charges; the smaller mean of
// two runs of about a millisecond each, timed by events.
static int time_tuning(ew_handle* h, const ew_tuning& t, const double* d_dq, double* d_out,
                       double* d_en, hipEvent_t e0, hipEvent_t e1, float* us) {
  CHK(apply_tuning(h, t));
  hipStream_t s = h->stream;
  for (int i = 0; i < 2; ++i) CHK(enqueue_eval(h, d_dq, d_out, d_en, s));
  HIPCHK(h, hipEventRecord(e0, s));
  CHK(enqueue_eval(h, d_dq, d_out, d_en, s));
  HIPCHK(h, hipEventRecord(e1, s));
  HIPCHK(h, hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
  const int reps = std::max(3, std::min(64, int(1.0f / std::max(ms, 1e-3f))));
  float best = 1e30f;
  for (int run = 0; run < 2; ++run) {
    HIPCHK(h, hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) CHK(enqueue_eval(h, d_dq, d_out, d_en, s));
    HIPCHK(h, hipEventRecord(e1, s));
    HIPCHK(h, hipEventSynchronize(e1));
    HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, 1e3f * ms / reps);
  }
  *us = best;
  return EW_OK;
}

// The autotuner: coordinate search over the spreading (write-out and block shape), the FFT
// pass structure (fused planes, lines per workgroup) and the radix order of each axis, each
// candidate timed on the device with this structure's data; the fastest shapes are applied and
// stored in the tuning cache.
static int autotune(ew_handle* h) {
  const int n = h->n;
  std::vector<double> q(n);
  double mean = 0.0;
  for (int i = 0; i < n; ++i) mean += (q[i] = std::sin(0.7 * i + 0.3));
  mean /= n;
  for (double& x : q) x -= mean;
  double *d_dq = nullptr, *d_out = nullptr, *d_en = nullptr;
  CHK(h->alloc(&d_dq, n));
  CHK(h->alloc(&d_out, n));
  CHK(h->alloc(&d_en, 1));
  HIPCHK(h, hipMemcpyAsync(d_dq, q.data(), sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev_guard{e0, e1};
  ew_tuning best = h->tuning;
  float best_us = 0.0f;
  CHK(time_tuning(h, best, d_dq, d_out, d_en, e0, e1, &best_us));
  auto consider = [&](const ew_tuning& c) -> int {
    float us = 0.0f;
    const int st = time_tuning(h, c, d_dq, d_out, d_en, e0, e1, &us);
    if (st == EW_ERR_ARGUMENT) return EW_OK;  // a shape this mesh cannot take
    if (st) return st;
    if (us < best_us) {
      best_us = us;
      best = h->tuning;
    }
    return EW_OK;
  };
  // the pair sum beside the chain (auxiliary stream) or co-scheduled in its launches, and the
  // share of its rows in the spreading launch: tried first and again after the shapes
  auto tune_pairs = [&]() -> int {
    {
      const ew_tuning base = best;
      for (int co = 0; co <= 1; ++co)
        for (int share : {0, 20, 35, 50}) {
          if (!co && share) continue;
          if (co == base.pair_cosched && share == base.pair_spread_share) continue;
          ew_tuning c = base;
          c.pair_cosched = co;
          c.pair_spread_share = share;
          if (co) c.fft_wave = 0;  // co-scheduling runs inside the Stockham launches
          CHK(consider(c));
        }
    }
    if (!best.pair_cosched) {
      // where the pair sum forks: beside the spreading or after it
      {
        ew_tuning c = best;
        c.pair_fork = best.pair_fork ? 0 : 1;
        CHK(consider(c));
      }
      // the pair sum's bounded grid beside the chain
      const ew_tuning base = best;
      for (int nb : {256, 384, 512, 768, 1024}) {
        if (nb == (base.pair_blocks > 0 ? base.pair_blocks : 512)) continue;
        ew_tuning c = base;
        c.pair_blocks = nb;
        CHK(consider(c));
      }
    }
    return EW_OK;
  };
  CHK(tune_pairs());
  // 1. spreading: write-out and block shape
  {
    const ew_tuning base = best;
    for (int direct = 2; direct >= 0; --direct)
      for (int f3 = 1; f3 <= 4; ++f3)
        for (int f2 = 1; f2 <= 4; ++f2)
          for (int f1 = 1; f1 <= 2; ++f1) {
            const int f[3] = {f1, f2, f3};
            if (!region_shape(h, f, kRegionMaxLds, nullptr)) continue;
            if (direct == 1 && !h->own_stream) continue;
            ew_tuning c = base;
            c.spread_direct = direct;
            for (int a = 0; a < 3; ++a) c.spread_block[a] = f[a];
            CHK(consider(c));
          }
  }
  // 2. register transforms: x-y planes + z lines, x, y and z line passes (lanes per line of each
  // axis), or the Stockham passes
  {
    const int* K = h->prm.K;
    bool lines = true;
    for (int a = 0; a < 3; ++a) lines = lines && wave_fft_default_lanes(K[a], h->fp32) > 0;
    const bool planes = K[0] == K[1] && wave_plane_default_lanes(K[0]) > 0 &&
                        wave_fft_default_lanes(K[2], h->fp32) > 0;
    const ew_tuning base = best;
    if (planes && base.fft_wave != 2) {
      ew_tuning c = base;
      c.fft_wave = 2;
      c.fft_lanes[0] = c.fft_lanes[1] = wave_plane_default_lanes(K[0]);
      c.fft_lanes[2] = wave_fft_default_lanes(K[2], h->fp32);
      CHK(consider(c));
    }
    if (lines && base.fft_wave != 1) {
      ew_tuning c = base;
      c.fft_wave = 1;
      for (int a = 0; a < 3; ++a) c.fft_lanes[a] = wave_fft_default_lanes(K[a], h->fp32);
      CHK(consider(c));
    }
    if (base.fft_wave != 0 && (planes || lines)) {
      ew_tuning c = base;
      c.fft_wave = 0;
      CHK(consider(c));
    }
    // lanes per line of each pass of the winner
    for (int a = 0; a < 3 && best.fft_wave; ++a) {
      if (best.fft_wave == 2 && a == 1) continue;  // the plane kernels take the x lanes
      const ew_tuning cur = best;
      for (int l : {4, 8, 16}) {
        if (l == cur.fft_lanes[a]) continue;
        const bool ok = (cur.fft_wave == 2 && a == 0) ? wave_plane_shape(K[0], l)
                                                       : wave_fft_shape(K[a], l);
        if (!ok) continue;
        ew_tuning c = cur;
        c.fft_lanes[a] = l;
        if (cur.fft_wave == 2 && a == 0) c.fft_lanes[1] = l;
        CHK(consider(c));
      }
    }
  }
  // 3. Stockham passes (when they won): fused planes or separate passes, lines per workgroup of
  // y and z
  if (!best.fft_wave) {
    const ew_tuning base = best;
    const int ly_max = std::max(1, kFftMaxElems / h->prm.K[1]);
    const int lz_max = std::max(1, kFftMaxElems / h->prm.K[2]);
    for (int planes = 0; planes <= 1; ++planes) {
      if (h->opt.fft_mode == 1 && planes) continue;
      if (h->opt.fft_mode == 2 && !planes) continue;
      for (int ly = 2; ly <= ly_max; ly *= 2) {
        if (planes && ly > 2) break;  // the y lines are unused with fused planes
        for (int lz = 2; lz <= lz_max; lz *= 2) {
          ew_tuning c = base;
          c.fft_planes = planes;
          c.fft_lines[1] = ly;
          c.fft_lines[2] = lz;
          CHK(consider(c));
        }
      }
    }
  }
  // 4. radix order per axis of the Stockham passes (orders that give the same passes are timed
  // once)
  for (int a = 0; a < 3 && !best.fft_wave; ++a) {
    const ew_tuning base = best;
    std::vector<std::vector<int>> seen;
    for (int v = 0; v < kRadixVariants; ++v) {
      int rad[kMaxRadixPasses];
      const int np = fft_plan(h->prm.K[a], rad, v);
      std::vector<int> key(rad, rad + std::max(np, 0));
      bool dup = false;
      for (auto& k : seen) dup = dup || k == key;
      seen.push_back(key);
      if (dup || v == base.fft_radix[a]) continue;
      ew_tuning c = base;
      c.fft_radix[a] = v;
      CHK(consider(c));
    }
  }
  CHK(tune_pairs());
  best.eval_us = best_us;
  CHK(apply_tuning(h, best));
  h->tuning.eval_us = best_us;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->free_buf(d_dq);
  h->free_buf(d_out);
  h->free_buf(d_en);
  tune_cache_put(tune_key(h), h->tuning);
  return EW_OK;
}

// The influence function's parameters for the current geometry (mesh, dual basis, volume, alpha,
// B-spline moduli).
static InfluenceArgs influence_args(const ew_handle* h) {
  const int* K = h->prm.K;
  InfluenceArgs inf{};
  for (int a = 0; a < 3; ++a) inf.K[a] = K[a];
  inf.Kh = K[0] / 2 + 1;
  inf.khd = make_fastdiv(inf.Kh);
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) inf.rec[3 * a + d] = h->cell.rec[a][d];
  inf.pref = 1.0 / (kPi * h->cell.volume);
  inf.a2 = kPi * kPi / (h->prm.alpha * h->prm.alpha);
  inf.bsq1 = h->bsq[0];
  inf.bsq2 = h->bsq[1];
  inf.bsq3 = h->bsq[2];
  return inf;
}

extern "C" int ew_get_tuning(const ew_handle* h, ew_tuning* out) {
  if (!h || !out) return set_global_error(EW_ERR_ARGUMENT, "ew_get_tuning: NULL argument");
  if (!h->have_geom) return set_global_error(EW_ERR_NO_GEOMETRY, "no geometry set");
  *out = h->tuning;
  return EW_OK;
}


extern "C" int ew_set_geometry(ew_handle* h, int n, const double* coords, const double* latvecs,
                               const int* species, const double* hubbard_u, int n_species,
                               const double* charges) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_set_geometry: handle is NULL");
  CHK(validate_inputs(h, n, coords, latvecs));
  HIPCHK(h, hipSetDevice(h->opt.device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->release();
  // whatever of the previous geometry's buffers the new one does not reuse is freed on return
  struct TrimOnReturn {
    ew_handle* h;
    ~TrimOnReturn() { h->trim(); }
  } trim_on_return{h};

  Cell c;
  int st = cell_metrics(latvecs, &c);
  if (st == EW_ERR_CELL) return h->fail(st, "cell volume too small");
  if (st) return h->fail(st, "non-finite lattice vector");
  double q2 = 0.0;
  if (charges)
    for (int i = 0; i < n; ++i) q2 += charges[i] * charges[i];
  Params prm;
  std::string perr;
  st = choose_parameters(c, n, q2, &h->opt.tol, h->opt.alpha, h->opt.r_cut, h->opt.k_cut,
                         h->opt.grid, h->opt.order, h->opt.alpha_rule, &prm, &perr);
  if (st) return h->fail(st, perr);
  prm.precision = h->opt.precision;

  // short-range gamma (src/dftbp/dftb/shortgamma.F90:753-775: cutoffs per species pair)
  const bool gamma = species && hubbard_u && n_species > 0;
  std::vector<double> gcut2;
  if (gamma) {
    for (int i = 0; i < n; ++i)
      if (species[i] < 0 || species[i] >= n_species)
        return h->fail(EW_ERR_ARGUMENT, "species index out of range");
    for (int s = 0; s < n_species; ++s)
      if (!(hubbard_u[s] >= 1.0e-6))  // MinHubTol, src/dftbp/common/accuracy.F90:106
        return h->fail(EW_ERR_ARGUMENT, "Hubbard U below 1e-6");
    gcut2.assign((size_t)n_species * n_species, 0.0);
    for (int a = 0; a < n_species; ++a)
      for (int b = a; b < n_species; ++b) {
        const double cut = exp_gamma_cutoff(hubbard_u[b], hubbard_u[a]);
        const double c2 = cut * cut + 1.0e-10;  // + tolSameDist^2 of the neighbour bisection
        gcut2[a * n_species + b] = gcut2[b * n_species + a] = c2;
        prm.r_list = std::max(prm.r_list, cut);
      }
  }
  h->n = n;
  h->cell = c;
  h->prm = prm;
  h->ns = gamma ? n_species : 0;
  h->self_coef = -2.0 * prm.alpha / kSqrtPi;
  // neutralising background of a charged cell (src/dftbp/dftb/coulomb.F90:1292-1296)
  h->background = -kPi / (c.volume * prm.alpha * prm.alpha);
  const int p = prm.order;
  const int* K = prm.K;

  // tiles of the SPME mesh; they also serve as the cells of the pair search
  const int Tdef[3] = {16, 8, 8};
  for (int a = 0; a < 3; ++a) {
    h->T[a] = std::min(K[a], Tdef[a]);
    h->nt[a] = (K[a] + h->T[a] - 1) / h->T[a];
  }
  h->ntmax = std::max(h->nt[0], std::max(h->nt[1], h->nt[2]));
  {
    // the interval spread needs one contiguous stencil/tile overlap per axis: K >= T + p - 1
    bool ok = true;
    for (int a = 0; a < 3; ++a)
      if (K[a] < h->T[a] + p - 1) ok = false;
    const size_t rsz = h->fp32 ? sizeof(float) : sizeof(double);
    const size_t per = 3 * (size_t)p * rsz + sizeof(int4);
    h->spread_slots = ok ? (int)std::min<size_t>(256, kSpreadSlotBytes / per) : 0;
  }
  std::vector<int> srctab((size_t)3 * h->ntmax * (kMaxSrcTiles + 1), 0);
  for (int a = 0; a < 3; ++a)
    if (!build_source_table(K[a], h->T[a], h->nt[a], p, h->ntmax, a, srctab))
      return h->fail(EW_ERR_ARGUMENT, "spreading tile table overflow (order too high)");

  const int ntiles = h->nt[0] * h->nt[1] * h->nt[2];

  // device buffers
  const size_t rs = h->fp32 ? sizeof(float) : sizeof(double);
  // privatised spreading blocks and FFT plans are set by apply_tuning below
  h->region = RegionArgs{};
  const int Kh = K[0] / 2 + 1;
  const size_t ngrid = (size_t)K[0] * K[1] * K[2], ncgrid = (size_t)Kh * K[1] * K[2];
  CHK(h->alloc(&h->perm, n));
  CHK(h->alloc(&h->k0, n));
  CHK(h->alloc(&h->frac, 3 * (size_t)n));
  {
    unsigned char* th;
    CHK(h->alloc(&th, 3 * (size_t)n * p * rs));
    h->theta = th;
  }
  CHK(h->alloc(&h->tile_of, n));
  CHK(h->alloc(&h->tile_start, ntiles + 1));
  CHK(h->alloc(&h->src_tab, srctab.size()));
  CHK(h->alloc(&h->row_ptr, (size_t)n + 1));
  CHK(h->alloc(&h->qs, n));
  CHK(h->alloc(&h->phi_real, n));
  CHK(h->alloc(&h->gmesh, n));
  CHK(h->alloc(&h->epart, std::max<size_t>((size_t)(n + 255) / 256,
                                            (size_t)h->nt[0] * h->nt[1] * h->nt[2])));
  CHK(h->alloc(&h->energy, 1));
  h->qsum = nullptr;
  if (h->opt.charged) CHK(h->alloc(&h->qsum, 1));
  CHK(h->alloc(&h->dq_io, n));
  CHK(h->alloc(&h->shift_io, n));
  CHK(h->alloc(&h->err_flag, 2));  // [0] coincident charges, [1] a partner offset beyond 16 bits
  {
    unsigned char *g, *cgp;
    CHK(h->alloc(&g, ngrid * rs));
    CHK(h->alloc(&cgp, ncgrid * 2 * rs));
    h->grid = g;
    h->cgrid = cgp;
  }
  if (gamma) {
    CHK(h->alloc(&h->species, n));
    CHK(h->alloc(&h->hub, n_species));
    CHK(h->alloc(&h->gcut2, (size_t)n_species * n_species));
    CHK(h->alloc(&h->onsite, n));
  }
  hipStream_t s = h->stream;
  HIPCHK(h, hipMemcpyAsync(h->src_tab, srctab.data(), sizeof(int) * srctab.size(),
                           hipMemcpyHostToDevice, s));
  if (gamma) {
    HIPCHK(h, hipMemcpyAsync(h->hub, hubbard_u, sizeof(double) * n_species,
                             hipMemcpyHostToDevice, s));
    HIPCHK(h, hipMemcpyAsync(h->gcut2, gcut2.data(), sizeof(double) * gcut2.size(),
                             hipMemcpyHostToDevice, s));
  }
  // fractional coordinates, stencil origins and sort tiles of the atoms, their stable sort by
  // tile and the per-atom arrays in that order, on the device (the coordinates go up once)
  int max_tile_atoms = 0;
  {
    OrderArgs og{};
    for (int a = 0; a < 3; ++a) {
      for (int d = 0; d < 3; ++d) og.rec[3 * a + d] = c.rec[a][d];
      og.K[a] = K[a];
      og.T[a] = h->T[a];
      og.nt[a] = h->nt[a];
    }
    og.order = p;
    og.n = n;
    og.ntiles = ntiles;
    OrderBufs ob{};
    double* d_coords = nullptr;
    int *key_in = nullptr, *idx_in = nullptr, *counts = nullptr, *d_max = nullptr,
        *sp_raw = nullptr;
    int4* k0_raw = nullptr;
    double* frac_raw = nullptr;
    unsigned char* temp = nullptr;
    const size_t temp_bytes = order_temp_bytes(n, ntiles);
    CHK(h->alloc(&d_coords, 3 * (size_t)n));
    CHK(h->alloc(&frac_raw, 3 * (size_t)n));
    CHK(h->alloc(&k0_raw, n));
    CHK(h->alloc(&key_in, n));
    CHK(h->alloc(&idx_in, n));
    CHK(h->alloc(&counts, ntiles + 1));
    CHK(h->alloc(&d_max, 1));
    CHK(h->alloc(&temp, temp_bytes));
    if (gamma) CHK(h->alloc(&sp_raw, n));
    HIPCHK(h, hipMemcpyAsync(d_coords, coords, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    if (gamma)
      HIPCHK(h, hipMemcpyAsync(sp_raw, species, sizeof(int) * n, hipMemcpyHostToDevice, s));
    ob.coords = d_coords;
    ob.sp_raw = sp_raw;
    ob.hub = h->hub;
    ob.frac_raw = frac_raw;
    ob.k0_raw = k0_raw;
    ob.key_in = key_in;
    ob.idx_in = idx_in;
    ob.counts = counts;
    ob.temp = temp;
    ob.temp_bytes = temp_bytes;
    ob.perm = h->perm;
    ob.tile_of = h->tile_of;
    ob.tile_start = h->tile_start;
    ob.max_tile = d_max;
    ob.frac = h->frac;
    ob.k0 = h->k0;
    ob.species = h->species;
    ob.onsite = h->onsite;
    HIPCHK(h, launch_order(og, ob, s));
    HIPCHK(h, hipMemcpyAsync(&max_tile_atoms, d_max, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    h->max_tile_atoms = max_tile_atoms;
  }
  // FFT tables, from host copies that outlive their asynchronous uploads
  for (int a = 0; a < 3; ++a) {
    std::vector<double>& bq = h->tab_bsq[a];
    bq.assign(K[a], 0.0);
    bspline_moduli(K[a], p, bq.data());
    CHK(h->alloc(&h->bsq[a], K[a]));
    HIPCHK(h, hipMemcpyAsync(h->bsq[a], bq.data(), sizeof(double) * K[a], hipMemcpyHostToDevice, s));
    unsigned char* t;
    CHK(h->alloc(&t, 2 * (size_t)K[a] * rs));
    h->tw[a] = t;  // filled by apply_tuning (the table follows the radix order)
    // W_K^m = exp(-2 pi i m / K), m < K, for the register line transforms
    CHK(h->alloc(&t, 2 * (size_t)K[a] * rs));
    h->twp[a] = t;
    if (h->fp32) {
      std::vector<float>& tv = h->tab_twp32[a];
      plain_roots(K[a], tv);
      HIPCHK(h, hipMemcpyAsync(t, tv.data(), sizeof(float) * tv.size(), hipMemcpyHostToDevice, s));
    } else {
      std::vector<double>& tv = h->tab_twp64[a];
      plain_roots(K[a], tv);
      HIPCHK(h, hipMemcpyAsync(t, tv.data(), sizeof(double) * tv.size(), hipMemcpyHostToDevice, s));
    }
  }
  {
    // the influence function on the half spectrum (read by the register z transform)
    const size_t gh = (size_t)(K[0] / 2 + 1) * K[1] * K[2];
    unsigned char* t;
    CHK(h->alloc(&t, gh * rs));
    h->inf_tab = t;
    const InfluenceArgs inf = influence_args(h);
    HIPCHK(h, h->fp32 ? launch_influence_table<float>(inf, t, s)
                      : launch_influence_table<double>(inf, t, s));
  }
  for (int a = 0; a < 3; ++a)
    if (fft_plan(K[a], h->px.radix) < 0) return h->fail(EW_ERR_ARGUMENT, "grid not 2^a 3^b 5^c");
  for (int a = 0; a < 3; ++a) {
    // a workgroup transforms whole lines in LDS: kFftMaxElems elements, 64 KiB
    const size_t lds = (h->fp32 ? 8 : 16) * ((size_t)K[a] + K[a] / 8 + 2 + K[a]);
    if (K[a] > kFftMaxElems || lds > 65536)
      return h->fail(EW_ERR_ARGUMENT, "mesh dimension " + std::to_string(K[a]) +
                                          " exceeds the single-workgroup FFT line limit");
  }
  // launch shapes: fixed rules, the tuning cache, or the caller's (the autotuner itself runs
  // once the geometry is complete, below)
  bool tune_now = false;
  {
    ew_tuning t = rule_tuning(h);
    if (h->opt.tune == 3) {
      t = h->opt.tuning;
    } else if (h->opt.tune == 1 || h->opt.tune == 2) {
      ew_tuning c;
      if (tune_cache_get(tune_key(h), &c))
        t = c;
      else
        tune_now = h->opt.tune == 1;
    }
    CHK(apply_tuning(h, t));
  }

  // B-spline weights
  BsplineArgs ba{{K[0], K[1], K[2]}, p};
  HIPCHK(h, launch_bspline(n, h->frac, ba, h->theta, h->fp32, s));

  // pair list
  PairArgs pa{};
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) pa.lat[3 * a + d] = c.lat[a][d];
  pa.alpha = prm.alpha;
  pa.rc2 = prm.r_cut * prm.r_cut;
  pa.rl2 = prm.r_list * prm.r_list + (gamma ? 1.0e-10 : 0.0);
  pa.ns = h->ns;
  pa.hub = h->hub;
  pa.gcut2 = h->gcut2;
  // Linked tiles need a unique minimum image within r_list, i.e. r_list |b_a| < 1/2 on every
  // axis (the fractional rounding of fe/fft/ewald_fft_3d_real.f90:150-157 then finds it); an
  // axis whose partner range would wrap onto itself is swept once, tile by tile.
  bool linked = true;
  for (int a = 0; a < 3; ++a) {
    pa.nt[a] = h->nt[a];
    const double wmin = double(std::min(h->T[a], K[a] - (h->nt[a] - 1) * h->T[a])) / K[a];
    pa.D[a] = int(std::ceil(prm.r_list * c.reclen[a] / wmin));
    pa.full[a] = 2 * pa.D[a] + 1 > h->nt[a] ? 1 : 0;
    if (!(prm.r_list * c.reclen[a] < 0.5)) linked = false;
  }
  h->pa = pa;
  HIPCHK(h, hipMemsetAsync(h->err_flag, 0, 2 * sizeof(int), s));
  const bool cf = h->fp32;
  const size_t cs = cf ? sizeof(float) : sizeof(double);
  h->col16 = nullptr;
  h->col = nullptr;
  if (linked) {
    h->pair_mode = 0;
    // row lengths, then their exclusive sum on the device; only the total comes back
    long long* counts = nullptr;
    unsigned char* scan_tmp = nullptr;
    const size_t scan_bytes = row_scan_temp_bytes(n);
    CHK(h->alloc(&counts, (size_t)n + 1));
    CHK(h->alloc(&scan_tmp, scan_bytes));
    HIPCHK(h, hipMemsetAsync(counts + n, 0, sizeof(long long), s));
    HIPCHK(h, launch_pair_count(n, h->frac, h->species, h->tile_of, h->tile_start, pa,
                                max_tile_atoms, counts, h->err_flag, s));
    HIPCHK(h, launch_row_scan(counts, h->row_ptr, n, scan_tmp, scan_bytes, s));
    long long run = 0;
    HIPCHK(h, hipMemcpyAsync(&run, h->row_ptr + n, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    h->nnz = run;
    // 8 spare entries: the pair sum reads aligned groups of eight.  Rows are in tile order, so
    // partners sit close to the row index: the fill writes 16-bit offsets directly (int32
    // partners when one does not fit, and for the symmetric storage, which is built from them).
    unsigned char* cp;
    CHK(h->alloc(&cp, ((size_t)run + 8) * cs));
    h->coef = cp;
    bool need32 = h->opt.pair_storage == 1;
    if (!need32) {
      short* off = nullptr;
      CHK(h->alloc(&off, (size_t)run + 8));
      HIPCHK(h, launch_pair_fill(n, h->frac, h->species, h->tile_of, h->tile_start, pa,
                                 max_tile_atoms, h->row_ptr, nullptr, off, h->err_flag + 1,
                                 h->coef, cf, h->err_flag, s));
      int too_far = 0;
      HIPCHK(h, hipMemcpyAsync(&too_far, h->err_flag + 1, sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHK(h, hipStreamSynchronize(s));
      if (too_far) {
        h->free_buf(off);
        need32 = true;
      } else {
        h->col16 = off;
      }
    }
    if (need32) {
      CHK(h->alloc(&h->col, (size_t)run + 8));
      HIPCHK(h, launch_pair_fill(n, h->frac, h->species, h->tile_of, h->tile_start, pa,
                                 max_tile_atoms, h->row_ptr, h->col, nullptr, nullptr, h->coef,
                                 cf, h->err_flag, s));
      HIPCHK(h, hipStreamSynchronize(s));
    }
  } else {
    h->pair_mode = 1;
    for (int a = 0; a < 3; ++a) pa.D[a] = int(std::ceil(prm.r_list * c.reclen[a])) + 1;
    h->pa = pa;
    const long long nn = (long long)n * n;
    if (nn > (1LL << 31)) return h->fail(EW_ERR_MEMORY, "cell too small for this many atoms");
    h->nnz = nn;
    std::vector<long long> rp((size_t)n + 1);
    for (int i = 0; i <= n; ++i) rp[i] = (long long)i * n;
    CHK(h->alloc(&h->col, (size_t)nn + 8));
    unsigned char* cp;
    CHK(h->alloc(&cp, ((size_t)nn + 8) * cs));
    h->coef = cp;
    HIPCHK(h, hipMemcpyAsync(h->row_ptr, rp.data(), sizeof(long long) * (n + 1),
                             hipMemcpyHostToDevice, s));
    HIPCHK(h, launch_pair_dense(n, h->frac, h->species, pa, h->col, h->coef, cf, h->err_flag, s));
    HIPCHK(h, hipStreamSynchronize(s));
  }
  int errflag = 0;
  HIPCHK(h, hipMemcpy(&errflag, h->err_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (errflag) {
    h->release();
    return h->fail(EW_ERR_COINCIDENT, "coincident charges");
  }
  // symmetric storage of the linked list (half the bytes per evaluation), else 16-bit partner
  // offsets (already written by the fill unless it needed int32 partners).
  h->half = false;
  h->hp = HalfPairs{};
  h->half_entries = 0;
  if (h->pair_mode == 0 && h->nnz > 0 && h->opt.pair_storage == 1) {
    std::vector<int> count(ntiles + 1);
    HIPCHK(h, hipMemcpy(count.data(), h->tile_start, sizeof(int) * (ntiles + 1),
                        hipMemcpyDeviceToHost));
    CHK(build_half_pairs(h, count, s));
  }
  if (h->pair_mode == 0 && h->nnz > 0 && !h->half && h->opt.pair_storage == 1) {
    short* off = nullptr;
    CHK(h->alloc(&off, (size_t)h->nnz + 8));
    HIPCHK(h, hipMemsetAsync(h->err_flag, 0, sizeof(int), s));
    HIPCHK(h, launch_compress_cols(n, h->row_ptr, h->col, off, h->err_flag, s));
    int too_far = 0;
    HIPCHK(h, hipMemcpyAsync(&too_far, h->err_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    if (too_far) {
      h->free_buf(off);
    } else {
      h->col16 = off;
      h->free_buf(h->col);
      h->col = nullptr;
    }
  }
  h->have_geom = true;
  if (tune_now) CHK(autotune(h));
  return EW_OK;
}

// The split gather (mesh potentials before the join, the rest after it): evaluations whose pair
// sum runs beside the chain on the auxiliary stream, with the tile gather.
static bool split_gather(const ew_handle* h) {
  return h->aux && !h->tuning.pair_cosched && h->opt.profile != 1 &&
         gather_tile_smem(h->T, h->prm.order, h->fp32 ? sizeof(float) : sizeof(double)) <=
             kGatherRegionMaxBytes;
}

extern "C" int ew_get_parameters(const ew_handle* h, ew_parameters* out) {
  if (!h || !out) return set_global_error(EW_ERR_ARGUMENT, "ew_get_parameters: NULL argument");
  if (!h->have_geom) return set_global_error(EW_ERR_NO_GEOMETRY, "no geometry set");
  std::memset(out, 0, sizeof(*out));
  out->alpha = h->prm.alpha;
  out->r_cut = h->prm.r_cut;
  out->r_list = h->prm.r_list;
  double kmin = 1e308;
  for (int a = 0; a < 3; ++a) {
    out->grid[a] = h->prm.K[a];
    kmin = std::min(kmin, kPi * h->prm.K[a] / h->cell.latlen[a]);
  }
  out->k_cut = kmin;
  out->order = h->prm.order;
  out->precision = h->prm.precision;
  out->pair_mode = h->half ? 2 : h->pair_mode;
  out->index_bytes = (h->col16 || h->half) ? 2 : 4;
  // the register transforms run in single-structure evaluations; a pool's grouped launches
  // use the Stockham passes
  out->fft_planes = h->own_stream ? ((h->wave == 2 || (h->wave == 0 && h->plane)) ? 1 : 0)
                                  : (h->plane ? 1 : 0);
  out->gather_split = split_gather(h) ? 1 : 0;
  out->n_pairs = h->half ? h->half_entries : h->nnz;
  out->device_bytes = (long long)h->device_bytes;
  return EW_OK;
}

// Launch description of the reciprocal chain of a handle with a geometry.
static SpmeLaunch make_spme_launch(const ew_handle* h) {
  const int* K = h->prm.K;
  SpmeLaunch L{};
  for (int a = 0; a < 3; ++a) L.K[a] = K[a];
  L.plane = h->plane;
  L.k0 = h->k0;
  L.theta = h->theta;
  L.q = h->qs;
  L.tile_start = h->tile_start;
  for (int a = 0; a < 3; ++a) {
    L.spread.K[a] = K[a];
    L.spread.T[a] = h->T[a];
    L.spread.nt[a] = h->nt[a];
    L.tw[a] = h->tw[a];
  }
  L.spread.ntmax = h->ntmax;
  L.spread.order = h->prm.order;
  L.spread.slots = h->spread_slots;
  L.region = h->region;
  L.spread.src = h->src_tab;
  L.px = h->px;
  L.py = h->py;
  L.pz = h->pz;
  L.inf = influence_args(h);
  L.wave = h->wave;
  for (int a = 0; a < 3; ++a) {
    L.lanes[a] = h->lanes[a];
    L.twp[a] = h->twp[a];
  }
  L.inf_tab = h->inf_tab;
  L.grid = h->grid;
  L.cgrid = h->cgrid;
  return L;
}

// The mesh part at the resident atoms plus phi_real, on-site and self terms (and the background
// of a charged cell), written in caller order; energy = 1/2 sum shift q.
static GatherArgs gather_args(const ew_handle* h) {
  GatherArgs ga{};
  for (int a = 0; a < 3; ++a) {
    ga.K[a] = h->prm.K[a];
    ga.T[a] = h->T[a];
    ga.nt[a] = h->nt[a];
  }
  ga.order = h->prm.order;
  ga.self_coef = h->self_coef;
  ga.bg = h->background;
  ga.qsum = h->qsum;
  return ga;
}

static hipError_t launch_gather(const ew_handle* h, const GatherArgs& ga, const double* q,
                                const double* phi_real, const double* onsite, double* d_out,
                                double* d_energy, hipStream_t s) {
  const size_t region = gather_tile_smem(h->T, h->prm.order, h->fp32 ? sizeof(float) : sizeof(double));
  if (region <= kGatherRegionMaxBytes)
    return h->fp32 ? launch_gather_tile<float>(h->k0, h->theta, h->grid, q, phi_real, onsite,
                                               h->perm, h->tile_start, ga, d_out, h->epart,
                                               d_energy, s)
                   : launch_gather_tile<double>(h->k0, h->theta, h->grid, q, phi_real, onsite,
                                                h->perm, h->tile_start, ga, d_out, h->epart,
                                                d_energy, s);
  return h->fp32 ? launch_gather_combine<float>(h->n, h->k0, h->theta, h->grid, q, phi_real,
                                                onsite, h->perm, ga, d_out, h->epart, d_energy, s)
                 : launch_gather_combine<double>(h->n, h->k0, h->theta, h->grid, q, phi_real,
                                                 onsite, h->perm, ga, d_out, h->epart, d_energy,
                                                 s);
}

// The root fork (see enqueue_eval): privatised spreading on an own-stream handle.  Phase
// profiling (opt.profile = 1) runs every phase alone; tracing (2) keeps the launch shape.
static bool root_fork(const ew_handle* h) {
  return h->opt.profile != 1 && h->aux && h->region.region_stride > 0;
}

static int enqueue_eval(ew_handle* h, const double* d_dq, double* d_shift, double* d_energy,
                        hipStream_t s) {
  const int n = h->n;
  const bool prof = h->opt.profile == 1;
  const bool trace = h->opt.profile == 2;
  // The real-space pair sum runs on the auxiliary stream beside the spreading and FFTs (which
  // leave CUs idle) and joins before the gather.  With privatised spreading the two branches
  // start together: the charge gather, charge sum and pair sum run on the auxiliary stream,
  // and the spreading reads the caller's charges through the sort permutation, so the chain
  // on stream s waits for nothing until the gather.  Otherwise the pair sum forks after the
  // charge gather.
  // co-scheduled pair sum: on extra workgroups of the chain's launches, one stream
  const bool co = h->tuning.pair_cosched && !h->half && !prof;
  const bool fork = !prof && h->aux && !co;
  const bool root = root_fork(h) && !co;
  const void* col = h->col16 ? static_cast<const void*>(h->col16) : h->col;
  const int ntiles = h->nt[0] * h->nt[1] * h->nt[2];
  auto pairs = [&](hipStream_t st) {
    h->record_trace(ew_handle::kTrPair0, st);
    const hipError_t e =
        h->half ? launch_pair_half(ntiles, h->tile_start, h->hp, h->fp32, h->qs, h->phi_real, st)
                : launch_pair_spmv(n, h->row_ptr, col, h->col16 != nullptr, h->coef, h->fp32,
                                   h->qs, h->phi_real, fork, st, h->tuning.pair_blocks);
    h->record_trace(ew_handle::kTrPair1, st);
    return e;
  };
  auto charges = [&](hipStream_t st) {
    h->record_trace(ew_handle::kTrChg0, st);
    HIPCHK(h, launch_gather_charges(n, h->perm, d_dq, h->qs, h->half ? h->phi_real : nullptr, st));
    if (h->qsum) HIPCHK(h, launch_charge_sum(n, h->qs, h->qsum, st));
    h->record_trace(ew_handle::kTrChg1, st);
    return EW_OK;
  };
  if (prof) HIPCHK(h, hipEventRecord(h->ev[0], s));
  h->record_trace(ew_handle::kTrStart, s);
  // pair_fork = 1: the pair sum waits for the spreading (it then runs beside the FFTs and the
  // mesh gather, which leave most CUs idle, instead of competing with the spreading for them)
  const bool late = root && h->tuning.pair_fork == 1;
  if (root) {
    HIPCHK(h, hipEventRecord(h->fork_ev, s));
    HIPCHK(h, hipStreamWaitEvent(h->aux, h->fork_ev, 0));
    CHK(charges(h->aux));
    if (!late) {
      HIPCHK(h, pairs(h->aux));
      HIPCHK(h, hipEventRecord(h->join_ev, h->aux));
    }
  } else {
    CHK(charges(s));
    if (prof) HIPCHK(h, hipEventRecord(h->ev[1], s));
    if (fork) {
      HIPCHK(h, hipEventRecord(h->fork_ev, s));
      HIPCHK(h, hipStreamWaitEvent(h->aux, h->fork_ev, 0));
      HIPCHK(h, pairs(h->aux));
      HIPCHK(h, hipEventRecord(h->join_ev, h->aux));
    }
  }
  SpmeLaunch L = make_spme_launch(h);
  if (root) {
    L.q = d_dq;
    L.qperm = h->perm;
  }
  if (co) {
    // the pair rows run inside the chain launches (their span is empty in the trace)
    h->record_trace(ew_handle::kTrPair0, s);
    h->record_trace(ew_handle::kTrPair1, s);
    L.pair.n = n;
    L.pair.r0 = 0;
    L.pair.r1 = n;
    L.pair.col16 = h->col16 != nullptr;
    L.pair.row_ptr = h->row_ptr;
    L.pair.col = col;
    L.pair.coef = h->coef;
    L.pair.q = h->qs;
    L.pair.out = h->phi_real;
    L.pair_spread_share = h->tuning.pair_spread_share;
  }
  h->record_trace(ew_handle::kTrChain0, s);
  hipEvent_t* marks = prof ? &h->ev[2] : (trace ? &h->tev[ew_handle::kTrSpread] : nullptr);
  hipEvent_t region_mark = trace ? h->tev[ew_handle::kTrRegion] : nullptr;
  if (late) L.spread_done = h->spread_ev;
  if (h->fp32)
    HIPCHK(h, launch_spme<float>(L, s, marks, region_mark));
  else
    HIPCHK(h, launch_spme<double>(L, s, marks, region_mark));
  if (late) {
    HIPCHK(h, hipStreamWaitEvent(h->aux, h->spread_ev, 0));
    HIPCHK(h, pairs(h->aux));
    HIPCHK(h, hipEventRecord(h->join_ev, h->aux));
  }
  // With the pair sum beside the chain the gather splits: the mesh potential of every atom
  // before the join (it needs nothing from the pair branch), the shift, energy and the
  // charge-dependent terms after it, so the join is followed by one short kernel instead of the
  // whole gather.
  const GatherArgs ga = gather_args(h);
  const bool split = fork && split_gather(h);
  if (split)
    HIPCHK(h, h->fp32 ? launch_gather_mesh<float>(h->k0, h->theta, h->grid, h->tile_start, ga,
                                                  h->gmesh, s)
                      : launch_gather_mesh<double>(h->k0, h->theta, h->grid, h->tile_start, ga,
                                                   h->gmesh, s));
  h->record_trace(ew_handle::kTrGathered, s);
  if (fork)
    HIPCHK(h, hipStreamWaitEvent(s, h->join_ev, 0));
  else if (!co)
    HIPCHK(h, pairs(s));
  if (prof) HIPCHK(h, hipEventRecord(h->ev[8], s));
  h->record_trace(ew_handle::kTrJoined, s);
  if (split)
    HIPCHK(h, launch_shift_combine(n, h->gmesh, h->phi_real, h->qs, h->onsite, h->perm, ga,
                                   d_shift, h->epart, d_energy, s));
  else
    HIPCHK(h, launch_gather(h, ga, h->qs, h->phi_real, h->onsite, d_shift, d_energy, s));
  if (trace) {
    h->record_trace(ew_handle::kTrEnd, s);
    HIPCHK(h, hipEventSynchronize(h->tev[ew_handle::kTrEnd]));
    using T = ew_handle;
    const int span[EW_NUM_TRACE][2] = {
        {T::kTrChg0, T::kTrChg1},    {T::kTrPair0, T::kTrPair1},  {T::kTrChain0, T::kTrRegion},
        {T::kTrRegion, T::kTrSpread}, {T::kTrSpread, T::kTrX},    {T::kTrX, T::kTrY},
        {T::kTrY, T::kTrZ},          {T::kTrZ, T::kTrYi},         {T::kTrYi, T::kTrXi},
        {T::kTrXi, T::kTrGathered},  {T::kTrGathered, T::kTrJoined},
        {T::kTrJoined, T::kTrEnd},   {T::kTrStart, T::kTrEnd}};
    for (int i = 0; i < EW_NUM_TRACE; ++i) {
      float ms = 0.0f;
      HIPCHK(h, hipEventElapsedTime(&ms, h->tev[span[i][0]], h->tev[span[i][1]]));
      h->trace_ms[i] += ms;
    }
    h->trace_calls++;
  }
  if (prof) {
    HIPCHK(h, hipEventRecord(h->ev[9], s));
    HIPCHK(h, hipEventSynchronize(h->ev[9]));
    for (int i = 0; i < EW_NUM_PHASES; ++i) {
      float ms = 0.0f;
      HIPCHK(h, hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]));
      h->phase_ms[i] += ms;
    }
    h->phase_calls++;
  }
  return EW_OK;
}

// One evaluation on stream s.  Large evaluations that fork at their start (privatised
// spreading) are launched directly: their two branches and the join replay slower from a
// captured graph than from direct launches on this runtime (98,304 atoms: 0.151 against
// 0.141 ms mixed, 0.209 against 0.200 ms FP64), and their kernels are long enough to hide the
// launch cost.  Other evaluations are
// captured once per buffer triple on the handle's own stream and replayed as a graph (the
// sequence has no memset or host work, so every replay is the same work).
constexpr int kDirectLaunchAtoms = 32768;

static int launch_eval(ew_handle* h, const double* d_dq, double* d_shift, double* d_energy,
                       hipStream_t s);

static int run_eval(ew_handle* h, const double* d_dq, double* d_shift, double* d_energy,
                    hipStream_t s) {
  HIPCHK(h, h->after_last(s));
  CHK(launch_eval(h, d_dq, d_shift, d_energy, s));
  HIPCHK(h, h->mark_last(s));
  return EW_OK;
}

static int launch_eval(ew_handle* h, const double* d_dq, double* d_shift, double* d_energy,
                       hipStream_t s) {
  if (h->opt.eager || h->opt.profile || (root_fork(h) && h->n >= kDirectLaunchAtoms))
    return enqueue_eval(h, d_dq, d_shift, d_energy, s);
  hipGraphExec_t exec = nullptr;
  for (auto& g : h->graphs)
    if (g.dq == d_dq && g.shift == d_shift && g.energy == d_energy) exec = g.exec;
  if (!exec) {
    if (h->graphs.size() >= 8) {
      // the oldest capture goes once its launches are done
      h->quiesce();
      (void)hipGraphExecDestroy(h->graphs.front().exec);
      h->graphs.erase(h->graphs.begin());
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int st = enqueue_eval(h, d_dq, d_shift, d_energy, h->stream);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->stream, &graph);
    if (st) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (ec != hipSuccess)
      return h->fail(EW_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess)
      return h->fail(EW_ERR_DEVICE, std::string("graph instantiate failed: ") + hipGetErrorString(ei));
    h->graphs.push_back({d_dq, d_shift, d_energy, exec});
  }
  HIPCHK(h, hipGraphLaunch(exec, s));
  return EW_OK;
}

extern "C" int ew_phase_times(ew_handle* h, double* ms, long long* calls, int reset) {
  if (!h || !ms) return set_global_error(EW_ERR_ARGUMENT, "ew_phase_times: NULL argument");
  if (h->opt.profile != 1)
    return h->fail(EW_ERR_ARGUMENT, "ew_phase_times needs opts.profile = 1");
  for (int i = 0; i < EW_NUM_PHASES; ++i) ms[i] = h->phase_ms[i];
  if (calls) *calls = h->phase_calls;
  if (reset) {
    for (int i = 0; i < EW_NUM_PHASES; ++i) h->phase_ms[i] = 0.0;
    h->phase_calls = 0;
  }
  return EW_OK;
}

extern "C" int ew_trace_times(ew_handle* h, double* ms, long long* calls, int reset) {
  if (!h || !ms) return set_global_error(EW_ERR_ARGUMENT, "ew_trace_times: NULL argument");
  if (h->opt.profile != 2)
    return h->fail(EW_ERR_ARGUMENT, "ew_trace_times needs opts.profile = 2");
  for (int i = 0; i < EW_NUM_TRACE; ++i) ms[i] = h->trace_ms[i];
  if (calls) *calls = h->trace_calls;
  if (reset) {
    for (int i = 0; i < EW_NUM_TRACE; ++i) h->trace_ms[i] = 0.0;
    h->trace_calls = 0;
  }
  return EW_OK;
}

extern "C" int ew_shift_device(ew_handle* h, const double* d_dq, double* d_shift,
                               double* d_energy, void* stream) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_shift_device: handle is NULL");
  if (!h->have_geom) return h->fail(EW_ERR_NO_GEOMETRY, "ew_shift_device before ew_set_geometry");
  if (!d_dq || !d_shift) return h->fail(EW_ERR_ARGUMENT, "ew_shift_device: NULL array");
  HIPCHK(h, hipSetDevice(h->opt.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
  return run_eval(h, d_dq, d_shift, d_energy ? d_energy : h->energy, s);
}

// Bytes of static scratch one evaluation of h needs: sorted charges, real-space potentials,
// energy partials and total, charge sum, mesh, spectrum and spreading copies, each on a
// 256-byte boundary.  With `base` set, points the handle's per-evaluation buffers there.
struct NoallocLayout {
  size_t qs, phi, gmesh, epart, energy, qsum, grid, cgrid, region, total;
};
static NoallocLayout noalloc_layout(const ew_handle* h) {
  const size_t rs = h->fp32 ? sizeof(float) : sizeof(double);
  const size_t n = (size_t)h->n;
  const size_t nparts = std::max<size_t>((n + 255) / 256, (size_t)h->nt[0] * h->nt[1] * h->nt[2]);
  const size_t K0 = h->prm.K[0], K1 = h->prm.K[1], K2 = h->prm.K[2];
  const size_t nblk = (size_t)h->region.nb[0] * h->region.nb[1] * h->region.nb[2];
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  NoallocLayout L{};
  size_t off = 0;
  L.qs = off;     off += up(8 * n);
  L.phi = off;    off += up(8 * n);
  L.gmesh = off;  off += up(8 * n);
  L.epart = off;  off += up(8 * nparts);
  L.energy = off; off += up(8);
  L.qsum = off;   off += up(8);
  L.grid = off;   off += up(K0 * K1 * K2 * rs);
  L.cgrid = off;  off += up((K0 / 2 + 1) * K1 * K2 * 2 * rs);
  L.region = off;
  if (h->region.region_stride > 0 && !h->region.direct)
    off += up(nblk * (size_t)h->region.region_stride * rs);
  L.total = off;
  return L;
}

extern "C" int ew_shift_device_noalloc(ew_handle* h, const double* d_dq, double* d_shift,
                                       double* d_energy, void* stream) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_shift_device_noalloc: handle is NULL");
  if (!h->have_geom)
    return h->fail(EW_ERR_NO_GEOMETRY, "ew_shift_device_noalloc before ew_set_geometry");
  if (!d_dq || !d_shift) return h->fail(EW_ERR_ARGUMENT, "ew_shift_device_noalloc: NULL array");
  if (h->opt.profile)
    return h->fail(EW_ERR_ARGUMENT, "ew_shift_device_noalloc does not run with opts.profile");
  const NoallocLayout L = noalloc_layout(h);
  if (L.total > kStaticScratchBytes)
    return h->fail(EW_ERR_MEMORY, "ew_shift_device_noalloc: this geometry needs " +
                                      std::to_string(L.total) + " bytes of scratch, above the " +
                                      std::to_string(kStaticScratchBytes) + "-byte static limit");
  void* base = g_scratch_dev[h->opt.device];
  if (!base) return h->fail(EW_ERR_DEVICE, "static scratch not resolved on this device");
  HIPCHK(h, hipSetDevice(h->opt.device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
  unsigned char* b = static_cast<unsigned char*>(base);
  // the evaluation's buffers point into the scratch for this call only (eager launches: a
  // captured graph would keep the handle's own buffers)
  struct Swap {
    ew_handle* h;
    double *qs, *phi, *gmesh, *epart, *qsum;
    void *grid, *cgrid, *region;
    ~Swap() {
      h->qs = qs;
      h->phi_real = phi;
      h->gmesh = gmesh;
      h->epart = epart;
      h->qsum = qsum;
      h->grid = grid;
      h->cgrid = cgrid;
      h->region.scratch = region;
    }
  } keep{h,       h->qs,   h->phi_real, h->gmesh,           h->epart,
         h->qsum, h->grid, h->cgrid,    h->region.scratch};
  h->qs = reinterpret_cast<double*>(b + L.qs);
  h->phi_real = reinterpret_cast<double*>(b + L.phi);
  h->gmesh = reinterpret_cast<double*>(b + L.gmesh);
  h->epart = reinterpret_cast<double*>(b + L.epart);
  if (h->qsum) h->qsum = reinterpret_cast<double*>(b + L.qsum);
  h->grid = b + L.grid;
  h->cgrid = b + L.cgrid;
  if (h->region.region_stride > 0 && !h->region.direct) h->region.scratch = b + L.region;
  double* energy = d_energy ? d_energy : reinterpret_cast<double*>(b + L.energy);
  HIPCHK(h, h->after_last(s));
  CHK(enqueue_eval(h, d_dq, d_shift, energy, s));
  HIPCHK(h, h->mark_last(s));
  return EW_OK;
}

extern "C" int ew_shift(ew_handle* h, const double* dq, double* shift, double* energy) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_shift: handle is NULL");
  if (!h->have_geom) return h->fail(EW_ERR_NO_GEOMETRY, "ew_shift before ew_set_geometry");
  if (!dq || !shift) return h->fail(EW_ERR_ARGUMENT, "ew_shift: NULL array");
  const int n = h->n;
  double net = 0.0, absq = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!std::isfinite(dq[i])) return h->fail(EW_ERR_NOT_FINITE, "non-finite charge");
    if (std::fabs(dq[i]) > kMaxCharge) return h->fail(EW_ERR_ARGUMENT, "|charge| above 1e4");
    net += dq[i];
    absq += std::fabs(dq[i]);
  }
  if (!h->opt.charged && std::fabs(net) > kNeutralityTol * absq)
    return h->fail(EW_ERR_NOT_NEUTRAL, "the cell must be charge neutral (or set opts.charged)");
  HIPCHK(h, hipSetDevice(h->opt.device));
  hipStream_t s = h->stream;
  HIPCHK(h, hipMemcpyAsync(h->dq_io, dq, sizeof(double) * n, hipMemcpyHostToDevice, s));
  CHK(run_eval(h, h->dq_io, h->shift_io, h->energy, s));
  HIPCHK(h, hipMemcpyAsync(shift, h->shift_io, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  double e = 0.0;
  HIPCHK(h, hipMemcpyAsync(&e, h->energy, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  if (energy) *energy = e;
  return EW_OK;
}

// ---- multi-vector evaluation: several charge vectors through one geometry ----

// Per-vector buffers and descriptors for at least nvec vectors; they stay until the geometry
// changes.  Every vector has its own charges, potentials, meshes and partials and shares the
// geometry's sort, weights, pair list and plans.
static int ensure_views(ew_handle* h, int nvec) {
  auto& V = h->views;
  if (V.cap >= nvec) return EW_OK;
  // the smaller buffers go: nothing queued may still read them
  h->quiesce();
  for (auto& g : h->mgraphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipFree(g.staging);
  }
  h->mgraphs.clear();
  for (void* p : {(void*)V.qs, (void*)V.phi, (void*)V.qil, (void*)V.epart, (void*)V.qsum, V.grid, V.cgrid,
                  V.region, (void*)V.dq, (void*)V.shift, (void*)V.energy, (void*)V.desc,
                  (void*)V.wio})
    if (p) h->free_buf(p);
  const void* staging = V.staging;
  V = ew_handle::Views{};
  V.staging = const_cast<void*>(staging);
  h->views_host.clear();
  const int n = h->n;
  const int* K = h->prm.K;
  const size_t rs = h->fp32 ? sizeof(float) : sizeof(double);
  const size_t ngrid = (size_t)K[0] * K[1] * K[2];
  const size_t ncgrid = (size_t)(K[0] / 2 + 1) * K[1] * K[2];
  const size_t nparts = std::max<size_t>((size_t)(n + 255) / 256,
                                         (size_t)h->nt[0] * h->nt[1] * h->nt[2]);
  const size_t nblk = (size_t)h->region.nb[0] * h->region.nb[1] * h->region.nb[2];
  const size_t rstride = h->region.region_stride > 0 && h->region.direct != 2
                             ? nblk * (size_t)h->region.region_stride * rs
                             : 0;
  const size_t cap = (size_t)nvec;
  unsigned char *g, *cg, *rg = nullptr;
  CHK(h->alloc(&V.qs, cap * n));
  CHK(h->alloc(&V.phi, cap * n));
  CHK(h->alloc(&V.qil, (size_t)spmm_stride(nvec) * n));
  CHK(h->alloc(&V.epart, cap * nparts));
  if (h->qsum) CHK(h->alloc(&V.qsum, cap));
  CHK(h->alloc(&g, cap * ngrid * rs));
  CHK(h->alloc(&cg, cap * ncgrid * 2 * rs));
  if (rstride) CHK(h->alloc(&rg, cap * rstride));
  V.grid = g;
  V.cgrid = cg;
  V.region = rg;
  CHK(h->alloc(&V.dq, cap * n));
  CHK(h->alloc(&V.shift, cap * n));
  CHK(h->alloc(&V.energy, cap));
  CHK(h->alloc(&V.desc, cap));
  std::vector<StructDesc> d(cap);
  for (size_t v = 0; v < cap; ++v) {
    fill_desc(h, &d[v]);
    d[v].qs = V.qs + v * n;
    d[v].phi_real = V.phi + v * n;
    d[v].epart = V.epart + v * nparts;
    d[v].ga.qsum = V.qsum ? V.qsum + v : nullptr;
    d[v].grid = g + v * ngrid * rs;
    d[v].cgrid = cg + v * ncgrid * 2 * rs;
    if (rg) d[v].region.scratch = rg + v * rstride;
  }
  HIPCHK(h, hipMemcpyAsync(V.desc, d.data(), sizeof(StructDesc) * cap, hipMemcpyHostToDevice,
                           h->stream));
  // the vectors' meshes, spectra and the shared influence table for the batched register FFTs
  std::vector<WaveIo> wio(cap);
  for (size_t v = 0; v < cap; ++v) wio[v] = WaveIo{d[v].grid, d[v].cgrid, h->inf_tab};
  CHK(h->alloc(&V.wio, cap));
  HIPCHK(h, hipMemcpyAsync(V.wio, wio.data(), sizeof(WaveIo) * cap, hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->views_host = std::move(d);
  V.cap = nvec;
  return EW_OK;
}

// The grouped launch of nvec vectors over a device copy (at dev) of S's words.  With the row
// list the pair sums of all vectors are one multi-vector sum; the symmetric storage keeps one
// pair-sum block set per vector.
static GroupLaunch multi_launch(const ew_handle* h, int nvec, const Staged& S, const void* dev) {
  GroupLaunch G = group_launch_of(S, h->views.desc, dev, h->prm.order, h->opt.group_grid == 1);
  if (!h->half) {
    G.has_pm = 1;
    G.pm = PairMulti{h->n,
                     nvec,
                     h->row_ptr,
                     h->col16 ? static_cast<const void*>(h->col16) : h->col,
                     h->col16 != nullptr,
                     h->coef,
                     h->views.qs,
                     h->views.phi,
                     h->views.qil,
                     0};
  }
  // profiling runs every kernel on the evaluation stream, one after the other
  G.aux = h->opt.profile == 1 ? nullptr : h->aux;
  G.pm.beside = G.aux != nullptr;
  G.fork_ev = h->fork_ev;
  G.join_ev = h->join_ev;
  if (h->wave) {
    // every vector has this geometry's mesh: one batched register-FFT sequence
    G.wave.assign(1, WaveGroup{nvec, h->views.wio, make_spme_launch(h)});
  }
  return G;
}

static int stage_multi(ew_handle* h, int nvec, const double* d_dq, double* d_shift,
                       double* d_energy, Staged* S) {
  const size_t n = (size_t)h->n;
  std::vector<IoSlot> io((size_t)nvec);
  for (int v = 0; v < nvec; ++v)
    io[v] = IoSlot{d_dq + v * n, d_shift + v * n, d_energy + v, v, 0};
  // with register transforms the views' FFTs are one batched sequence (multi_launch)
  const std::vector<unsigned char> by_wave((size_t)nvec, h->wave ? 1 : 0);
  if (!stage_group(h->views_host.data(), io.data(), nvec, h->fp32, !h->half, S, by_wave.data()))
    return h->fail(EW_ERR_ARGUMENT, "structure exceeds a workgroup's LDS");
  return EW_OK;
}

static int launch_multi(ew_handle* h, const GroupLaunch& G, hipStream_t s) {
  HIPCHK(h, h->fp32 ? launch_group<float>(G, s) : launch_group<double>(G, s));
  return EW_OK;
}

// nvec evaluations on stream s: captured once per (nvec, buffers) and replayed as a graph
// unless eager or profiling.
static int launch_multi_eval(ew_handle* h, int nvec, const double* d_dq, double* d_shift,
                             double* d_energy, hipStream_t s);

static int run_multi(ew_handle* h, int nvec, const double* d_dq, double* d_shift,
                     double* d_energy, hipStream_t s) {
  CHK(ensure_views(h, nvec));
  HIPCHK(h, h->after_last(s));
  CHK(launch_multi_eval(h, nvec, d_dq, d_shift, d_energy, s));
  HIPCHK(h, h->mark_last(s));
  return EW_OK;
}

static int launch_multi_eval(ew_handle* h, int nvec, const double* d_dq, double* d_shift,
                             double* d_energy, hipStream_t s) {
  const bool eager = h->opt.eager || h->opt.profile;
  if (!eager)
    for (auto& g : h->mgraphs)
      if (g.nvec == nvec && g.dq == d_dq && g.shift == d_shift && g.energy == d_energy) {
        HIPCHK(h, hipGraphLaunch(g.exec, s));
        return EW_OK;
      }
  Staged S;
  CHK(stage_multi(h, nvec, d_dq, d_shift, d_energy, &S));
  const size_t bytes = S.words.size() * sizeof(int);
  if (eager) {
    if (h->views_staging_cap < bytes) {
      if (h->views.staging) {
        h->quiesce();  // an earlier launch may still read it
        h->free_buf(h->views.staging);
      }
      h->views.staging = nullptr;
      h->views_staging_cap = 0;
      unsigned char* p;
      CHK(h->alloc(&p, bytes));
      h->views.staging = p;
      h->views_staging_cap = bytes;
    }
    // a pageable source: the copy has read the host image when the call returns
    HIPCHK(h, hipMemcpyAsync(h->views.staging, S.words.data(), bytes, hipMemcpyHostToDevice, s));
    return launch_multi(h, multi_launch(h, nvec, S, h->views.staging), s);
  }
  if (h->mgraphs.size() >= 8) {
    // the oldest capture goes once its launches are done
    h->quiesce();
    (void)hipGraphExecDestroy(h->mgraphs.front().exec);
    (void)hipFree(h->mgraphs.front().staging);
    h->mgraphs.erase(h->mgraphs.begin());
  }
  ew_handle::CapturedMulti c{nvec, d_dq, d_shift, d_energy, nullptr, nullptr};
  if (hipMalloc(&c.staging, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return h->fail(EW_ERR_MEMORY, "multi-vector staging allocation failed");
  }
  if (hipMemcpy(c.staging, S.words.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(c.staging);
    return h->fail(EW_ERR_DEVICE, "multi-vector staging copy failed");
  }
  const GroupLaunch G = multi_launch(h, nvec, S, c.staging);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int st = launch_multi(h, G, h->stream);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(h->stream, &graph);
  if (st != EW_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipFree(c.staging);
    if (st != EW_OK) return st;
    return h->fail(EW_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(ec));
  }
  const hipError_t ei = hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    (void)hipFree(c.staging);
    return h->fail(EW_ERR_DEVICE, std::string("graph instantiate failed: ") + hipGetErrorString(ei));
  }
  h->mgraphs.push_back(c);
  HIPCHK(h, hipGraphLaunch(c.exec, s));
  return EW_OK;
}

extern "C" int ew_shift_multi_device(ew_handle* h, int nvec, const double* d_dq, double* d_shift,
                                     double* d_energy, void* stream) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_shift_multi_device: handle is NULL");
  if (!h->have_geom)
    return h->fail(EW_ERR_NO_GEOMETRY, "ew_shift_multi_device before ew_set_geometry");
  if (!d_dq || !d_shift) return h->fail(EW_ERR_ARGUMENT, "ew_shift_multi_device: NULL array");
  if (nvec < 1 || nvec > EW_MAX_VECTORS)
    return h->fail(EW_ERR_ARGUMENT, "ew_shift_multi_device: nvec outside 1.." +
                                        std::to_string(EW_MAX_VECTORS));
  HIPCHK(h, hipSetDevice(h->opt.device));
  CHK(ensure_views(h, nvec));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
  return run_multi(h, nvec, d_dq, d_shift, d_energy ? d_energy : h->views.energy, s);
}

extern "C" int ew_shift_multi(ew_handle* h, int nvec, const double* dq, double* shift,
                              double* energy) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_shift_multi: handle is NULL");
  if (!h->have_geom) return h->fail(EW_ERR_NO_GEOMETRY, "ew_shift_multi before ew_set_geometry");
  if (!dq || !shift) return h->fail(EW_ERR_ARGUMENT, "ew_shift_multi: NULL array");
  if (nvec < 1 || nvec > EW_MAX_VECTORS)
    return h->fail(EW_ERR_ARGUMENT,
                   "ew_shift_multi: nvec outside 1.." + std::to_string(EW_MAX_VECTORS));
  const size_t n = (size_t)h->n;
  for (int v = 0; v < nvec; ++v) {
    double net = 0.0, absq = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const double q = dq[v * n + i];
      if (!std::isfinite(q)) return h->fail(EW_ERR_NOT_FINITE, "non-finite charge");
      if (std::fabs(q) > kMaxCharge) return h->fail(EW_ERR_ARGUMENT, "|charge| above 1e4");
      net += q;
      absq += std::fabs(q);
    }
    if (!h->opt.charged && std::fabs(net) > kNeutralityTol * absq)
      return h->fail(EW_ERR_NOT_NEUTRAL, "the cell must be charge neutral (or set opts.charged)");
  }
  HIPCHK(h, hipSetDevice(h->opt.device));
  CHK(ensure_views(h, nvec));
  hipStream_t s = h->stream;
  auto& V = h->views;
  HIPCHK(h, hipMemcpyAsync(V.dq, dq, sizeof(double) * n * nvec, hipMemcpyHostToDevice, s));
  CHK(run_multi(h, nvec, V.dq, V.shift, V.energy, s));
  HIPCHK(h, hipMemcpyAsync(shift, V.shift, sizeof(double) * n * nvec, hipMemcpyDeviceToHost, s));
  std::vector<double> e((size_t)nvec);
  HIPCHK(h, hipMemcpyAsync(e.data(), V.energy, sizeof(double) * nvec, hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  if (energy)
    for (int v = 0; v < nvec; ++v) energy[v] = e[v];
  return EW_OK;
}

// The resident row list when the real-space force and virial can walk it (linked mode, rows
// stored in full); otherwise they walk the tiles.
static PairRows pair_rows(const ew_handle* h) {
  PairRows r;
  if (h->pair_mode == 0 && !h->half && (h->col16 || h->col)) {
    r.row_ptr = h->row_ptr;
    r.col = h->col16 ? static_cast<const void*>(h->col16) : h->col;
    r.col16 = h->col16 != nullptr;
  }
  return r;
}

// Physical Ewald force -dU/dx (caller order, host array) at the charges of the last evaluation,
// which must have run on this handle's stream: the mesh still holds the convolved potential.
static int forces_after_eval(ew_handle* h, double* force_host) {
  const int n = h->n;
  if (!h->force) CHK(h->alloc(&h->force, 3 * (size_t)n));
  ForceArgs fa{};
  for (int a = 0; a < 3; ++a) {
    fa.K[a] = h->prm.K[a];
    for (int d = 0; d < 3; ++d) fa.rec[3 * a + d] = h->cell.rec[a][d];
  }
  fa.order = h->prm.order;
  hipStream_t s = h->stream;
  const PairRows rows = pair_rows(h);
  if (h->fp32)
    HIPCHK(h, launch_forces<float>(n, h->frac, h->k0, h->grid, h->qs, h->perm, h->tile_of,
                                   h->tile_start, fa, h->pa, h->pair_mode == 1, h->force, s,
                                   rows));
  else
    HIPCHK(h, launch_forces<double>(n, h->frac, h->k0, h->grid, h->qs, h->perm, h->tile_of,
                                    h->tile_start, fa, h->pa, h->pair_mode == 1, h->force, s,
                                    rows));
  HIPCHK(h, hipMemcpyAsync(force_host, h->force, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  return EW_OK;
}

extern "C" int ew_gradients(ew_handle* h, const double* dq, double* grad) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_gradients: handle is NULL");
  if (!grad) return h->fail(EW_ERR_ARGUMENT, "ew_gradients: NULL array");
  if (!h->have_geom) return h->fail(EW_ERR_NO_GEOMETRY, "ew_gradients before ew_set_geometry");
  std::vector<double> shift(h->n), f(3 * (size_t)h->n);
  CHK(ew_shift(h, dq, shift.data(), nullptr));
  CHK(forces_after_eval(h, f.data()));
  for (size_t k = 0; k < f.size(); ++k) grad[k] -= f[k];
  return EW_OK;
}

// Stress of the Ewald energy at charges dq, added into stress[3][3] (row-major, symmetric,
// Hartree / Bohr^3): sigma = W / V with the virial W = -dE/d(strain), the convention of
// TCoulomb%addStress (src/dftbp/dftb/coulomb.F90:721-757 with invRStress, 2777-2863):
// reciprocal part from the SPME spectrum, real-space part over the pairs within r_cut, and for
// charged cells the neutralising background (E_bg delta_ab).  Short gamma is not included
// (TShortGamma has its own stress).
extern "C" int ew_stress(ew_handle* h, const double* dq, double* stress) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_stress: handle is NULL");
  if (!dq || !stress) return h->fail(EW_ERR_ARGUMENT, "ew_stress: NULL array");
  if (!h->have_geom) return h->fail(EW_ERR_NO_GEOMETRY, "ew_stress before ew_set_geometry");
  const int n = h->n;
  double net = 0.0, absq = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!std::isfinite(dq[i])) return h->fail(EW_ERR_NOT_FINITE, "non-finite charge");
    if (std::fabs(dq[i]) > kMaxCharge) return h->fail(EW_ERR_ARGUMENT, "|charge| above 1e4");
    net += dq[i];
    absq += std::fabs(dq[i]);
  }
  if (!h->opt.charged && std::fabs(net) > kNeutralityTol * absq)
    return h->fail(EW_ERR_NOT_NEUTRAL, "the cell must be charge neutral (or set opts.charged)");
  HIPCHK(h, hipSetDevice(h->opt.device));
  hipStream_t s = h->stream;
  // an evaluation on a caller's stream may still use the charge and mesh buffers
  HIPCHK(h, h->after_last(s));
  HIPCHK(h, hipMemcpyAsync(h->dq_io, dq, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HIPCHK(h, launch_gather_charges(n, h->perm, h->dq_io, h->qs, nullptr, s));
  SpmeLaunch L = make_spme_launch(h);
  L.forward_only = 1;
  const long long gh = (long long)(h->prm.K[0] / 2 + 1) * h->prm.K[1] * h->prm.K[2];
  const PairRows rows = pair_rows(h);
  const size_t nrec = (size_t)((gh + 255) / 256),
               nreal = (size_t)real_virial_parts(n, rows.row_ptr && h->pair_mode != 1);
  double* part = nullptr;
  if (hipMalloc(&part, sizeof(double) * 6 * (nrec + nreal)) != hipSuccess) {
    (void)hipGetLastError();
    return h->fail(EW_ERR_MEMORY, "ew_stress: allocation failed");
  }
  int nb1 = 0, nb2 = 0;
  hipError_t e = h->fp32 ? launch_spme<float>(L, s, nullptr) : launch_spme<double>(L, s, nullptr);
  if (e == hipSuccess)
    e = h->fp32 ? launch_recip_virial<float>(L, part, &nb1, s)
                : launch_recip_virial<double>(L, part, &nb1, s);
  if (e == hipSuccess)
    e = launch_real_virial(n, h->frac, h->tile_of, h->tile_start, h->qs, h->pa, h->pair_mode == 1,
                           part + 6 * nrec, &nb2, s, rows);
  std::vector<double> hp(6 * (nrec + nreal), 0.0);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(part);
  if (e != hipSuccess)
    return h->fail(EW_ERR_DEVICE, std::string("ew_stress: ") + hipGetErrorString(e));
  double w[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nb1; ++b)
    for (int c = 0; c < 6; ++c) w[c] += hp[6 * (size_t)b + c];
  for (int b = 0; b < nb2; ++b)
    for (int c = 0; c < 6; ++c) w[c] += hp[6 * (nrec + (size_t)b) + c];
  if (h->opt.charged) {
    const double ebg = 0.5 * h->background * net * net;
    w[0] += ebg;
    w[1] += ebg;
    w[2] += ebg;
  }
  const double iv = 1.0 / h->cell.volume;
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  for (int c = 0; c < 6; ++c) {
    stress[3 * ia[c] + ib[c]] += w[c] * iv;
    if (ia[c] != ib[c]) stress[3 * ib[c] + ia[c]] += w[c] * iv;
  }
  return EW_OK;
}

// External point charges staged on the device for one call: sorted by the engine's sort tiles
// (coordinates, charges, widths, stencil origins, tile starts, B-spline weights), with the search
// shape of their real-space sums.  Device memory is one allocation freed with the object.
struct ExtStage {
  int m = 0;
  bool linked = true;
  ExtArgs ea{};
  std::vector<int> perm;             // sorted -> caller index
  double qsum = 0.0;
  unsigned char* mem = nullptr;
  double *frac = nullptr, *q = nullptr, *w = nullptr;
  int4* k0 = nullptr;
  int *start = nullptr, *tile = nullptr, *perm_d = nullptr;
  void* theta = nullptr;
  unsigned char* cur = nullptr;
  ~ExtStage() {
    if (mem) (void)hipFree(mem);
  }
  template <typename T>
  T* take(size_t count) {
    T* r = reinterpret_cast<T*>(cur);
    cur += (sizeof(T) * count + 255) & ~size_t(255);
    return r;
  }
};

// Validates and stages the sources; `extra` bytes of further scratch follow the staged arrays
// (taken by the caller with st.take).  The cutoff of the real-space sums is
// max(r_cut, w_max alpha r_cut): a blurred source's erfc(d / w) has decayed as far there as
// erfc(alpha d) at r_cut.
static int stage_sources(ew_handle* h, int m, const double* coords, const double* charges,
                         const double* widths, size_t extra, ExtStage& st) {
  double wmax = 0.0;
  for (int s = 0; s < m; ++s) {
    if (!std::isfinite(charges[s]) || !std::isfinite(coords[3 * s]) ||
        !std::isfinite(coords[3 * s + 1]) || !std::isfinite(coords[3 * s + 2]))
      return h->fail(EW_ERR_NOT_FINITE, "non-finite external charge or coordinate");
    if (std::fabs(charges[s]) > kMaxCharge)
      return h->fail(EW_ERR_ARGUMENT, "|external charge| above 1e4");
    if (widths) {
      if (!(widths[s] > 0.0) || !std::isfinite(widths[s]))
        return h->fail(EW_ERR_ARGUMENT, "blur widths must be finite and > 0");
      wmax = std::max(wmax, widths[s]);
    }
  }
  HIPCHK(h, hipSetDevice(h->opt.device));
  hipStream_t s = h->stream;
  // an evaluation on a caller's stream may still use the charge and mesh buffers
  h->quiesce();
  st.m = m;
  TileOrder o;
  order_by_tile(h, m, coords, o);
  std::vector<double> qs(m), ws(widths ? m : 0);
  for (int k = 0; k < m; ++k) {
    qs[k] = charges[o.perm[k]];
    if (widths) ws[k] = widths[o.perm[k]];
  }
  for (int k = 0; k < m; ++k) st.qsum += qs[k];
  st.perm = o.perm;

  const Cell& c = h->cell;
  const Params& prm = h->prm;
  const double r_ext = std::max(prm.r_cut, wmax * prm.alpha * prm.r_cut);
  ExtArgs& ea = st.ea;
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) ea.lat[3 * a + d] = c.lat[a][d];
  ea.alpha = prm.alpha;
  ea.rc2 = r_ext * r_ext;
  ea.self_coef = h->self_coef;
  ea.m = m;
  for (int a = 0; a < 3; ++a) {
    ea.nt[a] = h->nt[a];
    const int K = prm.K[a];
    const double wmin = double(std::min(h->T[a], K - (h->nt[a] - 1) * h->T[a])) / K;
    ea.D[a] = int(std::ceil(r_ext * c.reclen[a] / wmin));
    ea.full[a] = 2 * ea.D[a] + 1 > h->nt[a] ? 1 : 0;
    if (!(r_ext * c.reclen[a] < 0.5)) st.linked = false;
  }
  if (!st.linked)
    for (int a = 0; a < 3; ++a) ea.D[a] = int(std::ceil(r_ext * c.reclen[a])) + 1;

  const int ntiles = h->nt[0] * h->nt[1] * h->nt[2];
  const int p = prm.order;
  const size_t rs = h->fp32 ? sizeof(float) : sizeof(double);
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t total = up(sizeof(double) * 3 * (size_t)m) + 2 * up(sizeof(double) * m) +
                       up(sizeof(int4) * m) + up(sizeof(int) * (ntiles + 1)) +
                       2 * up(sizeof(int) * m) + up(3 * (size_t)m * p * rs) + extra;
  if (hipMalloc(&st.mem, total) != hipSuccess) {
    (void)hipGetLastError();
    return h->fail(EW_ERR_MEMORY, "external charges: device allocation failed");
  }
  st.cur = st.mem;
  st.frac = st.take<double>(3 * (size_t)m);
  st.q = st.take<double>(m);
  st.w = st.take<double>(m);
  st.k0 = st.take<int4>(m);
  st.start = st.take<int>(ntiles + 1);
  st.tile = st.take<int>(m);
  st.perm_d = st.take<int>(m);
  st.theta = st.take<unsigned char>(3 * (size_t)m * p * rs);
  HIPCHK(h, hipMemcpyAsync(st.frac, o.frac.data(), sizeof(double) * 3 * m, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemcpyAsync(st.q, qs.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
  if (widths) HIPCHK(h, hipMemcpyAsync(st.w, ws.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemcpyAsync(st.k0, o.k0.data(), sizeof(int4) * m, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemcpyAsync(st.start, o.start.data(), sizeof(int) * (ntiles + 1),
                           hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemcpyAsync(st.tile, o.tile.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemcpyAsync(st.perm_d, o.perm.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
  BsplineArgs ba{{prm.K[0], prm.K[1], prm.K[2]}, p};
  HIPCHK(h, launch_bspline(m, st.frac, ba, st.theta, h->fp32, s));
  ea.frac = st.frac;
  ea.q = st.q;
  ea.blur = widths ? st.w : nullptr;
  ea.start = st.start;
  return EW_OK;
}

// The mesh of the staged sources, convolved with the influence function (left in h->grid).
static int spme_of_sources(ew_handle* h, const ExtStage& st) {
  SpmeLaunch L = make_spme_launch(h);
  L.k0 = st.k0;
  L.theta = st.theta;
  L.q = st.q;
  L.qperm = nullptr;
  L.tile_start = st.start;
  HIPCHK(h, h->fp32 ? launch_spme<float>(L, h->stream, nullptr)
                    : launch_spme<double>(L, h->stream, nullptr));
  return EW_OK;
}

static ForceArgs force_args(const ew_handle* h) {
  ForceArgs fa{};
  for (int a = 0; a < 3; ++a) {
    fa.K[a] = h->prm.K[a];
    for (int d = 0; d < 3; ++d) fa.rec[3 * a + d] = h->cell.rec[a][d];
  }
  fa.order = h->prm.order;
  return fa;
}

// Potential of external point charges at the resident atoms (caller order, overwritten), the
// periodic form of TCoulomb%getPotential (src/dftbp/dftb/coulomb.F90:3121-3160 ->
// sumInvRPeriodicAsymm 1340-1418): potential_i = sum_s q_s ewald(r_i - r_s) with ewald of
// coulomb.F90:2311-2345 -- reciprocal part, real part (optionally Gaussian-blurred sources and a
// softened denominator), the background -pi / (V alpha^2) per source, and the self term
// -2 alpha / sqrt(pi) for a source on an atom.  The reciprocal part is the SPME convolution of
// the sources on the engine's mesh.
extern "C" int ew_external_potential(ew_handle* h, int m, const double* ext_coords,
                                     const double* ext_charges, const double* blur_widths,
                                     double eps_soften, double* potential) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_external_potential: handle is NULL");
  if (!h->have_geom)
    return h->fail(EW_ERR_NO_GEOMETRY, "ew_external_potential before ew_set_geometry");
  if (!potential || m < 0 || (m > 0 && (!ext_coords || !ext_charges)))
    return h->fail(EW_ERR_ARGUMENT, "ew_external_potential: bad arguments");
  if (!std::isfinite(eps_soften) || eps_soften < 0.0)
    return h->fail(EW_ERR_ARGUMENT, "ew_external_potential: eps_soften must be finite and >= 0");
  const int n = h->n;
  if (m == 0) {
    std::fill(potential, potential + n, 0.0);
    return EW_OK;
  }
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  ExtStage st;
  CHK(stage_sources(h, m, ext_coords, ext_charges, blur_widths,
                    3 * up(sizeof(double) * n) + 2 * up(sizeof(double)), st));
  double* d_zero = st.take<double>(n);
  double* d_phi = st.take<double>(n);
  double* d_out = st.take<double>(n);
  double* d_qsum = st.take<double>(1);
  double* d_energy = st.take<double>(1);
  hipStream_t s = h->stream;
  HIPCHK(h, hipMemcpyAsync(d_qsum, &st.qsum, sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemsetAsync(d_zero, 0, sizeof(double) * n, s));
  CHK(spme_of_sources(h, st));
  // real part at the atoms, then mesh part + real part + background per source
  st.ea.eps2 = eps_soften * eps_soften;
  HIPCHK(h, launch_ext_real(n, h->frac, h->tile_of, st.ea, !st.linked, d_phi, s));
  GatherArgs ga = gather_args(h);
  ga.self_coef = 0.0;
  ga.qsum = d_qsum;
  HIPCHK(h, launch_gather(h, ga, d_zero, d_phi, nullptr, d_out, d_energy, s));
  HIPCHK(h, hipMemcpyAsync(potential, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  return EW_OK;
}

// Derivatives of the interaction of the resident atoms with external point charges, the periodic
// forms of TCoulomb%getPotentialGradient and %addExternalPotGrad (src/dftbp/dftb/coulomb.F90:
// 3162-3235 -> addInvRPrimePeriodicAsymm 1934-2092, derivEwaldReciprocal 2388-2425,
// derivEwaldReal 2577-2620).  With atom charges dq: atom_grad[i] += dq_i sum_s q_s
// grad ewald(x_i - y_s) and, if ext_grad is given, ext_grad[s] -= the same summed over the
// atoms.  Without dq (the Hamiltonian derivative): atom_grad[i] += sum_s q_s grad ewald(x_i - y_s).
// Reciprocal parts from the B-spline derivatives of the sources' mesh at the atoms and of the
// atoms' mesh at the sources; real parts as in ew_external_potential (no softening, as in the
// reference).
extern "C" int ew_external_gradients(ew_handle* h, int m, const double* ext_coords,
                                     const double* ext_charges, const double* blur_widths,
                                     const double* dq, double* atom_grad, double* ext_grad) {
  if (!h) return set_global_error(EW_ERR_ARGUMENT, "ew_external_gradients: handle is NULL");
  if (!h->have_geom)
    return h->fail(EW_ERR_NO_GEOMETRY, "ew_external_gradients before ew_set_geometry");
  if (!atom_grad || m < 0 || (m > 0 && (!ext_coords || !ext_charges)))
    return h->fail(EW_ERR_ARGUMENT, "ew_external_gradients: bad arguments");
  const int n = h->n;
  if (dq)
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(dq[i])) return h->fail(EW_ERR_NOT_FINITE, "non-finite charge");
      else if (std::fabs(dq[i]) > kMaxCharge)
        return h->fail(EW_ERR_ARGUMENT, "|charge| above 1e4");
  if (m == 0) return EW_OK;
  const bool want_ext = dq && ext_grad;
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  ExtStage st;
  CHK(stage_sources(h, m, ext_coords, ext_charges, blur_widths,
                    2 * up(sizeof(double) * 3 * (size_t)n) + up(sizeof(double) * n) +
                        (want_ext ? 2 * up(sizeof(double) * 3 * (size_t)m) : 0),
                    st));
  double* d_frec = st.take<double>(3 * (size_t)n);
  double* d_greal = st.take<double>(3 * (size_t)n);
  double* d_scale = st.take<double>(n);
  double* d_srec = want_ext ? st.take<double>(3 * (size_t)m) : nullptr;
  double* d_sreal = want_ext ? st.take<double>(3 * (size_t)m) : nullptr;
  hipStream_t s = h->stream;
  // atom charges in sorted order (ones for the Hamiltonian derivative)
  if (dq) {
    HIPCHK(h, hipMemcpyAsync(h->dq_io, dq, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIPCHK(h, launch_gather_charges(n, h->perm, h->dq_io, h->qs, nullptr, s));
    HIPCHK(h, hipMemcpyAsync(d_scale, h->qs, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  } else {
    const std::vector<double> ones(n, 1.0);
    HIPCHK(h, hipMemcpyAsync(d_scale, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIPCHK(h, hipStreamSynchronize(s));  // ones is a host temporary
  }
  // atoms: -dq grad(sources' mesh potential), and dq times the real-space gradient
  CHK(spme_of_sources(h, st));
  const ForceArgs fa = force_args(h);
  HIPCHK(h, h->fp32 ? launch_force_recip<float>(n, h->frac, h->k0, h->grid, d_scale, h->perm, fa,
                                                 d_frec, s)
                    : launch_force_recip<double>(n, h->frac, h->k0, h->grid, d_scale, h->perm,
                                                 fa, d_frec, s));
  HIPCHK(h, launch_ext_real_grad(n, h->frac, h->tile_of, h->perm, d_scale, st.ea, !st.linked,
                                 d_greal, s));
  if (want_ext) {
    // sources: -q_s grad(atoms' mesh potential), and q_s times the real-space gradient of the
    // atoms' potential (the Gaussian width is the source's own)
    SpmeLaunch L = make_spme_launch(h);
    HIPCHK(h, h->fp32 ? launch_spme<float>(L, s, nullptr) : launch_spme<double>(L, s, nullptr));
    HIPCHK(h, h->fp32 ? launch_force_recip<float>(m, st.frac, st.k0, h->grid, st.q, st.perm_d,
                                                   fa, d_srec, s)
                      : launch_force_recip<double>(m, st.frac, st.k0, h->grid, st.q, st.perm_d,
                                                   fa, d_srec, s));
    ExtArgs ea = st.ea;
    ea.m = n;
    ea.frac = h->frac;
    ea.q = h->qs;
    ea.blur = nullptr;
    ea.tblur = blur_widths ? st.w : nullptr;
    ea.start = h->tile_start;
    HIPCHK(h, launch_ext_real_grad(m, st.frac, st.tile, st.perm_d, st.q, ea, !st.linked, d_sreal,
                                   s));
  }
  std::vector<double> frec(3 * (size_t)n), greal(3 * (size_t)n);
  std::vector<double> srec(want_ext ? 3 * (size_t)m : 0), sreal(want_ext ? 3 * (size_t)m : 0);
  HIPCHK(h, hipMemcpyAsync(frec.data(), d_frec, sizeof(double) * frec.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipMemcpyAsync(greal.data(), d_greal, sizeof(double) * greal.size(),
                           hipMemcpyDeviceToHost, s));
  if (want_ext) {
    HIPCHK(h, hipMemcpyAsync(srec.data(), d_srec, sizeof(double) * srec.size(),
                             hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipMemcpyAsync(sreal.data(), d_sreal, sizeof(double) * sreal.size(),
                             hipMemcpyDeviceToHost, s));
  }
  HIPCHK(h, hipStreamSynchronize(s));
  for (size_t k = 0; k < frec.size(); ++k) atom_grad[k] += greal[k] - frec[k];
  // d/dy_s ewald(x - y_s) = -grad ewald(x - y_s): ext_grad -= dq q_s grad ewald = q_s grad phi_atoms
  for (size_t k = 0; k < srec.size(); ++k) ext_grad[k] += sreal[k] - srec[k];
  return EW_OK;
}

extern "C" int ew_potential_force(int n, const double* coords, const double* charges,
                                  const double* latvecs, const ew_options* opts, double* pot,
                                  double* force) {
  if (!charges || !pot) return set_global_error(EW_ERR_ARGUMENT, "ew_potential_force: NULL array");
  ew_handle* h = nullptr;
  int st = ew_create(opts, &h);
  if (st) return st;
  st = ew_set_geometry(h, n, coords, latvecs, nullptr, nullptr, 0, charges);
  if (!st) st = ew_shift(h, charges, pot, nullptr);
  if (!st && force) st = forces_after_eval(h, force);
  if (st) global_error() = h->err;
  ew_destroy(h);
  return st;
}

extern "C" int ew_parameters_for(int n, const double* charges, const double* latvecs,
                                 const ew_options* opts, ew_parameters* out) {
  if (!latvecs || !out || n <= 0)
    return set_global_error(EW_ERR_ARGUMENT, "ew_parameters_for: bad argument");
  ew_options o{};
  if (opts) o = *opts;
  Cell c;
  int st = cell_metrics(latvecs, &c);
  if (st) return set_global_error(st, "invalid cell");
  double q2 = 0.0;
  if (charges)
    for (int i = 0; i < n; ++i) q2 += charges[i] * charges[i];
  Params p;
  std::string err;
  st = choose_parameters(c, n, q2, &o.tol, o.alpha, o.r_cut, o.k_cut, o.grid, o.order,
                         o.alpha_rule, &p, &err);
  if (st) return set_global_error(st, err);
  std::memset(out, 0, sizeof(*out));
  out->alpha = p.alpha;
  out->r_cut = p.r_cut;
  out->r_list = p.r_list;
  double kmin = 1e308;
  for (int a = 0; a < 3; ++a) {
    out->grid[a] = p.K[a];
    kmin = std::min(kmin, kPi * p.K[a] / c.latlen[a]);
  }
  out->k_cut = kmin;
  out->order = p.order;
  out->precision = o.precision;
  out->n_pairs = estimate_pairs(n, c, p.r_list);
  return EW_OK;
}

extern "C" int ew_grid_memory(int n, const double* charges, const double* latvecs,
                              const ew_options* opts, double* gigabytes) {
  if (!gigabytes) return set_global_error(EW_ERR_ARGUMENT, "ew_grid_memory: NULL output");
  ew_parameters p;
  int st = ew_parameters_for(n, charges, latvecs, opts, &p);
  if (st) return st;
  const double rs = p.precision == EW_PRECISION_MIXED ? 4.0 : 8.0;
  const double K3 = double(p.grid[0]) * p.grid[1] * p.grid[2];
  const double Kh = double(p.grid[0] / 2 + 1) * p.grid[1] * p.grid[2];
  const double grids = K3 * rs + Kh * 2.0 * rs;
  const double atoms = double(n) * (4 + 16 + 24 + 3.0 * p.order * rs + 4 + 8 * 6 + 8);
  const double pairs = double(p.n_pairs) * (4.0 + rs);
  *gigabytes = (grids + atoms + pairs) / (1024.0 * 1024.0 * 1024.0);
  return EW_OK;
}

// ------------------------------------------------------------------------------------------
// grouped-kernel view of a handle
// ------------------------------------------------------------------------------------------

SpmeLaunch ewamd::spme_launch_of(const ew_handle* h) { return make_spme_launch(h); }

void ewamd::fill_desc(const ew_handle* h, StructDesc* d) {
  *d = StructDesc{};
  const int* K = h->prm.K;
  const int p = h->prm.order;
  d->n = h->n;
  for (int a = 0; a < 3; ++a) d->K[a] = K[a];
  d->ntiles = h->nt[0] * h->nt[1] * h->nt[2];
  d->spread_mode = h->region.region_stride > 0 ? (h->region.direct == 2 ? 3 : 0)
                                                : (h->spread_slots > 0 ? 1 : 2);
  d->gather_tile =
      gather_tile_smem(h->T, p, h->fp32 ? sizeof(float) : sizeof(double)) <= kGatherRegionMaxBytes;
  d->nparts = d->gather_tile ? d->ntiles : (h->n + 255) / 256;
  d->col16 = h->col16 != nullptr;
  d->perm = h->perm;
  d->k0 = h->k0;
  d->theta = h->theta;
  d->tile_start = h->tile_start;
  d->qs = h->qs;
  d->phi_real = h->phi_real;
  d->onsite = h->onsite;
  d->epart = h->epart;
  d->row_ptr = h->row_ptr;
  d->col = h->col16 ? static_cast<const void*>(h->col16) : h->col;
  d->half = h->half;
  d->hp = h->hp;
  d->coef = h->coef;
  d->grid = h->grid;
  d->cgrid = h->cgrid;
  for (int a = 0; a < 3; ++a) {
    d->tw[a] = h->tw[a];
    d->spread.K[a] = K[a];
    d->spread.T[a] = h->T[a];
    d->spread.nt[a] = h->nt[a];
    d->ga.K[a] = K[a];
    d->ga.T[a] = h->T[a];
    d->ga.nt[a] = h->nt[a];
    d->inf.K[a] = K[a];
  }
  d->spread.ntmax = h->ntmax;
  d->spread.order = p;
  d->spread.slots = h->spread_slots;
  d->spread.src = h->src_tab;
  d->region = h->region;
  // the grouped kernels take owned blocks as they are; any other blocks assemble copies
  d->region.direct = h->region.direct == 2 ? 2 : 0;
  d->ga.order = p;
  d->ga.self_coef = h->self_coef;
  d->ga.bg = h->background;
  d->ga.qsum = h->qsum;
  const int Kh = K[0] / 2 + 1;
  d->inf.Kh = Kh;
  d->inf.khd = make_fastdiv(Kh);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) d->inf.rec[3 * a + b] = h->cell.rec[a][b];
  d->inf.pref = 1.0 / (kPi * h->cell.volume);
  d->inf.a2 = kPi * kPi / (h->prm.alpha * h->prm.alpha);
  d->inf.bsq1 = h->bsq[0];
  d->inf.bsq2 = h->bsq[1];
  d->inf.bsq3 = h->bsq[2];
  d->px = h->px;
  d->py = h->py;
  d->pz = h->pz;
  d->plane = h->plane;
  d->yd = make_fastdiv(Kh);
  d->zd = make_fastdiv(Kh * K[1]);
}
