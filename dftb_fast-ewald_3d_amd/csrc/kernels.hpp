// kernels.hpp -- kernel argument blocks and launchers (host <-> device contract).
#pragma once

#include <vector>
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace ewamd {

constexpr int kSpreadTileMax = 2048;  // mesh points per spreading tile (LDS: 16 KiB in FP64)

struct SpreadArgs {
  int K[3];
  int T[3];        // tile extents
  int nt[3];       // tiles per axis
  int ntmax;       // max(nt) -- row length of the source-tile table
  int order;
  int slots;       // atoms staged per chunk by k_spread (0: use k_spread_checked)
  const int* src;  // [3][ntmax][1 + kMaxSrcTiles]: count, then source tiles per output tile
};

// LDS budget of k_spread's staged atoms; rows (z, y) per tile for its row bitmasks
constexpr size_t kSpreadSlotBytes = 40 * 1024;
constexpr int kSpreadRows = 64;
constexpr int kSpreadThreads = 1024;  // one thread per point of a 16x8x8 tile

// Symmetric ("half") pair storage of the linked-tile list: every unordered pair once, grouped
// by the tile A of its lower atom and by chunks of <= 256 atoms of A's partner tiles B >= A.
// An entry is (i in A) << 8 | (j in chunk), with the coefficient beside it.
constexpr int kHalfMaxChunks = 64;  // chunks per tile
struct HalfPairs {
  const int* tile_chunk;   // [ntiles + 1]: chunks of tile A
  const int* chunk_tile;   // [nchunks + 1]: partner tiles of a chunk (index into ptile)
  const int* chunk_entry;  // [nchunks + 1]: entries of a chunk
  const int* chunk_atoms;  // [nchunks]: atoms staged by a chunk (<= 256)
  const int* ptile;        // partner tile ids, ascending within a tile's list
  const int* ptile_off;    // atom offset of each partner tile inside its chunk
  const int* ptile_chunk;  // chunk of each partner tile, relative to its tile's first chunk
  const int* rc_base;      // [ntiles + 1]: per-(chunk, row) table base of a tile
  const unsigned short* entry;
  const void* coef;
};

struct InfluenceArgs {
  int K[3];
  int Kh;
  FastDiv khd;     // / Kh
  double rec[9];   // dual basis rows b_i (no 2 pi)
  double pref;     // 1 / (pi V)
  double a2;       // pi^2 / alpha^2
  const double* bsq1;
  const double* bsq2;
  const double* bsq3;
};

struct GatherArgs {
  int K[3];
  int order;
  double self_coef;  // -2 alpha / sqrt(pi)
  double bg;            // neutralising background -pi / (V alpha^2) (charged cells)
  const double* qsum;   // total charge (charged cells), null: neutral, no background term
  int T[3];          // tile extents (tile gather)
  int nt[3];         // tiles per axis (tile gather)
};

// LDS budget of the tile gather: staged region (tile + (p-1) halo per axis) and the weights of
// 64 atoms
constexpr size_t kGatherRegionMaxBytes = 60 * 1024;
inline size_t gather_tile_smem(const int* T, int p, size_t rsz) {
  return rsz * ((size_t)(T[0] + p - 1) * (T[1] + p - 1) * (T[2] + p - 1) + 64 * 3 * (size_t)p);
}

struct BsplineArgs {
  int K[3];
  int order;
};

struct PairArgs {
  double lat[9];   // rows a_i
  double alpha;
  double rc2;      // r_cut^2
  double rl2;      // r_list^2
  int D[3];        // partner tiles (linked mode) or image shells (dense mode) per side
  int full[3];     // linked mode: 1 = sweep every tile of this axis once (2D+1 > nt)
  int nt[3];
  int ns;          // species count (0: no short gamma)
  const double* hub;    // [ns]
  const double* gcut2;  // [ns*ns] squared short-gamma cutoffs
};

// The resident pair list of the linked mode (rows in sort order, each partner once, minimum
// image within r_list): the real-space force and virial walk it instead of the tiles.
struct PairRows {
  const long long* row_ptr = nullptr;
  const void* col = nullptr;  // int16 offsets (col16) or int32 partners
  bool col16 = false;
};

// Per-geometry ordering of the atoms on the device (order.hip).
struct OrderArgs {
  double rec[9];  // dual basis rows b_i (b_i . a_j = delta_ij)
  int K[3], T[3], nt[3];
  int order;
  int n, ntiles;
};
struct OrderBufs {
  const double* coords;   // [n][3] caller order
  const int* sp_raw;      // species in caller order, or null (no short gamma)
  const double* hub;      // Hubbard U per species (with sp_raw)
  double* frac_raw;       // [n][3] scratch
  int4* k0_raw;           // [n] scratch
  int *key_in, *idx_in;   // [n] scratch
  int* counts;            // [ntiles + 1] scratch
  void* temp;             // hipCUB scratch of order_temp_bytes(n, ntiles)
  size_t temp_bytes;
  // results: sorted position -> caller index, tile of each sorted atom, tile starts, most atoms
  // in one tile, and the per-atom arrays in sorted order
  int *perm, *tile_of, *tile_start, *max_tile;
  double* frac;
  int4* k0;
  int* species;
  double* onsite;
};
size_t order_temp_bytes(int n, int ntiles);
hipError_t launch_order(const OrderArgs& g, const OrderBufs& o, hipStream_t st);
// row_ptr[0..n] = exclusive sum of counts[0..n] (counts[n] = 0): the rows of the pair list.
size_t row_scan_temp_bytes(int n);
hipError_t launch_row_scan(const long long* counts, long long* row_ptr, int n, void* temp,
                           size_t temp_bytes, hipStream_t st);

// Where the pair-list fill writes the partners of a row (see store_partner).
struct PartnerOut {
  int* col = nullptr;      // partner indices, or null for 16-bit offsets
  short* off = nullptr;
  int* too_far = nullptr;  // set when an offset does not fit 16 bits
  int n = 0;
};

struct ForceArgs {
  int K[3];
  int order;
  double rec[9];  // dual basis rows b_i
};

// Privatised spreading: blocks of f[0] x f[1] x f[2] sort tiles, S = f * T mesh points each.
struct RegionArgs {
  int K[3];
  int S[3];        // block extents (mesh points)
  int nb[3];       // blocks per axis
  int f[3];        // sort tiles per block per axis
  int nt[3];       // sort tiles per axis
  int order;
  long long region_stride;  // elements per block copy: (S0+p-1)(S1+p-1)(S2+p-1)
  void* scratch;   // nb0*nb1*nb2 copies; null: no privatised spreading
  int direct;      // 1: each block adds its LDS copy into the (cleared) mesh with device-scope
                   // atomics instead of writing it to scratch for the assembly (single-structure
                   // evaluations; the grouped kernels keep the copies)
  int items;       // owned blocks (direct 2): entries of the (atom, plane) item list, 0 for
                   // kOwnedItems
};

// LDS budget of a block copy: the fixed rule keeps copies within 60 KiB (several workgroups per
// CU); tuned shapes may use up to the gfx950 workgroup limit
constexpr size_t kRegionMaxBytes = 60 * 1024;
constexpr int kSpreadBatch = 3;  // (atom, z plane) items whose loads a thread issues together
constexpr size_t kRegionMaxLds = 160 * 1024;
// Owned blocks: the (atom, z plane) items of one pass live in an LDS list after the block copy
// (4 bytes an entry; at most kOwnedItems, sized by the host to the atoms a block can reach so
// that more workgroups share a CU; a block with more atoms takes several passes); the tile runs
// (at most kOwnedMaxRuns) in LDS tables.  A shape whose runs exceed the table is refused by the
// host.
constexpr int kOwnedItems = 6144;
constexpr int kOwnedMaxRuns = 128;
constexpr size_t kOwnedStatic = 4096;  // LDS kept for the kernel's static tables
// The block copy has odd row and plane pitches (in FP64 elements), so that the planes of one
// atom, which neighbouring lanes add at once, fall in different LDS banks.
inline __host__ __device__ int owned_row_pitch(const int* S) { return S[0] | 1; }
inline __host__ __device__ int owned_plane_pitch(const int* S) {
  return (owned_row_pitch(S) * S[1]) | 1;
}
inline __host__ __device__ int owned_items(const RegionArgs& g) {
  return g.items > 0 ? g.items : kOwnedItems;
}
// Bytes of an owned block's fixed-point accumulator per mesh point: an int64 (FP32 mesh), or
// an int64 remainder and an int32 leading part (FP64 mesh).
constexpr size_t owned_point_bytes(bool fp32) { return fp32 ? 8 : 12; }
inline size_t owned_smem(const RegionArgs& g, bool fp32) {
  return owned_point_bytes(fp32) * (size_t)owned_plane_pitch(g.S) * g.S[2] +
         sizeof(int) * (size_t)owned_items(g);
}
// Raises the dynamic LDS limit of the region spreading kernels to kRegionMaxLds (once).
hipError_t prepare_region_kernels();
// Assembly of the block copies: workgroups per mesh block, mesh points in flight per thread
#ifndef EW_ASSEMBLE_SPLIT
#define EW_ASSEMBLE_SPLIT 1
#endif
constexpr int kAssembleSplit = EW_ASSEMBLE_SPLIT;
constexpr int kAssembleUnroll = 4;

// Fused plane transforms: one workgroup of kPlaneBlock threads holds a whole x-y plane of the
// half spectrum in LDS (up to kPlaneMaxLds, the gfx950 workgroup limit) and runs the x R2C and
// y passes of that plane (and the inverse y and x C2R passes) in one launch.
constexpr int kPlaneBlock = 1024;
constexpr int kPlaneMinPoints = 4096;  // half-spectrum points of a plane worth a fused workgroup
constexpr int kPlaneEpt = 10;  // complex elements per thread: planes up to 10240 elements
constexpr size_t kPlaneMaxLds = 160 * 1024;
inline size_t plane_smem(const FftPlan& px, const FftPlan& py, size_t cs) {
  const size_t nl = (size_t)(py.n + 1) / 2, Kh = (size_t)px.n / 2 + 1;
  const size_t lines = nl * px.pitch > Kh * py.pitch ? nl * px.pitch : Kh * py.pitch;
  return cs * (lines + px.n + py.n);
}
inline bool plane_fits(const FftPlan& px, const FftPlan& py, size_t cs) {
  const long long nl = (py.n + 1) / 2, Kh = px.n / 2 + 1;
  const long long cap = (long long)kPlaneBlock * kPlaneEpt;
  return nl * px.n <= cap && Kh * py.n <= cap && plane_smem(px, py, cs) <= kPlaneMaxLds;
}
// Raises the dynamic LDS limit of the plane kernels to kPlaneMaxLds (once per process).
hipError_t prepare_plane_kernels();

// Rows [r0, r1) of the real-space pair sum run by the extra blocks of one chain kernel launch
// (blocks >= first), so that the pair sum fills the CUs the latency-bound chain kernel leaves
// idle, on the same stream (co-scheduling; no fork and join across streams).
constexpr int kCoWaves = 2048;  // waves of pair blocks appended to one chain kernel launch
struct PairCo {
  int first = 0x7fffffff;  // chain blocks of the launch; the blocks after them run pair rows
  int r0 = 0, r1 = 0;      // rows of this launch
  int nblk = 0;            // pair blocks of this launch
  int n = 0;
  int col16 = 0;
  const long long* row_ptr = nullptr;
  const void* col = nullptr;
  const void* coef = nullptr;
  const double* q = nullptr;
  double* out = nullptr;
};

// One structure of a batched register-transform launch (launch_wave_fft_batch).
struct WaveIo {
  void* grid;
  void* cgrid;
  const void* inf;
};

struct SpmeLaunch {
  int K[3];
  int plane;  // 1: fused x-y plane transforms (plane_fits), 0: separate x and y passes
  int forward_only;  // 1: stop after the forward transform (cgrid = spectrum of the charges)
  const int4* k0;
  const void* theta;
  const double* q;
  const int* qperm;  // privatised spreading only: q is in caller order, read through qperm
  const int* tile_start;
  SpreadArgs spread;
  RegionArgs region;
  FftPlan px, py, pz;
  const void* tw[3];
  InfluenceArgs inf;
  void* grid;
  void* cgrid;
  // co-scheduled pair sum (pair.r1 > 0): its rows are split over the spreading launch
  // (pair_spread_share percent) and the FFT launches (the rest, in equal parts)
  PairCo pair;
  int pair_spread_share;
  // register transforms (wfft.hip) in place of the Stockham passes (1: x, y, z line passes; 2:
  // x-y plane kernels and the z line pass): lanes per line of each axis (planes: lanes[0]), the
  // tables W_K^m of each axis and the influence function on the half spectrum
  int wave;
  int lanes[3];
  const void* twp[3];
  const void* inf_tab;
  hipEvent_t spread_done;  // optional: recorded after the spreading, before the FFTs
};

// Register line transforms (wfft.hip): whether an axis of length n has a compiled transform
// with `lanes` lanes per line; the default lanes for a length (0: none).
bool wave_fft_shape(int n, int lanes);
int wave_fft_default_lanes(int n, bool fp32);
// Register x-y plane transforms (square planes): whether a plane side has a compiled kernel with
// `lanes` lanes per line; the default lanes for a side (0: none).
bool wave_plane_shape(int n, int lanes);
int wave_plane_default_lanes(int n);
// x, y, z (with the influence table), inverse y and inverse x passes of a full evaluation
// (L.wave, every pass a compiled shape); ev (optional): events 1..5 after each pass.
template <typename R>
hipError_t launch_wave_fft(const SpmeLaunch& L, hipStream_t st, hipEvent_t* ev);
// The same passes over `count` structures of one mesh shape and lanes at once (blockIdx.y picks
// the structure): bt[count] in device memory holds their meshes, spectra and influence tables;
// L gives the mesh, mode, lanes and twiddle tables.
template <typename R>
hipError_t launch_wave_fft_batch(const SpmeLaunch& L, const WaveIo* bt, int count,
                                 hipStream_t st);
// The influence function on the half spectrum ((kz K2 + ky) Kh + kx), as R, once per geometry.
template <typename R>
hipError_t launch_influence_table(const InfluenceArgs& inf, void* tab, hipStream_t st);

// ---------------------------------------------------------------------------------------
// Grouped evaluation: many resident structures in one launch per kernel.
// ---------------------------------------------------------------------------------------

// One resident structure as the grouped kernels see it (device copy, built per geometry).
struct StructDesc {
  int n;
  int K[3];
  int ntiles;
  int nparts;       // energy partials written by the gather
  int spread_mode;  // 0: privatised blocks + assembly, 1: staged (slots), 2: checked,
                    // 3: owned blocks
  int gather_tile;  // 1: tile gather, 0: per-atom gather
  int col16;        // 1: 16-bit partner offsets
  const int* perm;
  const int4* k0;
  const void* theta;
  const int* tile_start;
  double* qs;
  double* phi_real;
  const double* onsite;
  double* epart;
  const long long* row_ptr;
  const void* col;
  const void* coef;
  int half;        // 1: symmetric pair storage (hp), one pair-sum block per tile
  HalfPairs hp;
  void* grid;
  void* cgrid;
  const void* tw[3];
  SpreadArgs spread;
  RegionArgs region;
  GatherArgs ga;
  InfluenceArgs inf;
  FftPlan px, py, pz;
  int plane;   // 1: fused x-y plane transforms
  FastDiv yd;  // / Kh            (y lines: L = kz * Kh + kx)
  FastDiv zd;  // / (Kh * K2)     (z lines: L = ky * Kh + kx)
};

// One evaluation of a launch: structure sid, charges in, shift (caller order) and energy out.
struct IoSlot {
  const double* dq;
  double* shift;
  double* energy;
  int sid;
  int pad;
};

// Blocks of one grouped kernel: entry i covers blocks [prefix[i], prefix[i+1]) for the
// evaluation slot[i] (an index into the IoSlot array).
struct GroupMap {
  const int* prefix;
  const int* slot;
  int count;
  int total;  // work items (blocks) of the launch: prefix[count]
};

// Persistent grouped launches (GroupLaunch::persistent): a kind with more work items than
// this many workgroups per CU (256 CUs) runs on a bounded grid whose workgroups walk the items.
constexpr int kGroupWgPerCu = 4;
constexpr int kNumCus = 256;

enum GroupKind {
  kGkCharges = 0,
  kGkRegion,
  kGkAssemble,
  kGkSlotted,
  kGkChecked,
  kGkFftX,
  kGkFftY,
  kGkFftZ,
  kGkPairs,
  kGkGatherTile,
  kGkGatherAtom,
  kGkEnergy,
  kGkChargeSum,
  kGkPlane,
  kGkOwned,
  kGroupKinds
};

// Pair sums of nvec charge vectors through one row list (the coefficients are read once):
// out[v * n + i] = sum_j coef_ij q[v * n + j].
struct PairMulti {
  int n;
  int nvec;  // 1..8
  const long long* row_ptr;
  const void* col;
  int col16;
  const void* coef;
  const double* q;
  double* out;
  double* qil;  // [n][spmm_stride(nvec)] scratch: the charges interleaved per atom
  int beside;   // 1: runs next to other kernels (bounded grid)
};

// Row length of the interleaved charges: nvec rounded up to an even count (16-byte rows).
constexpr int spmm_stride(int nvec) { return nvec == 1 ? 1 : (nvec + 1) & ~1; }

// One batched register-transform FFT sequence of a grouped launch: n evaluations of one mesh
// shape, mode and lanes, their meshes, spectra and influence tables at io (device); l: the mesh,
// mode, lanes and twiddle tables.
struct WaveGroup {
  int n;
  const WaveIo* io;
  SpmeLaunch l;
};

struct GroupLaunch {
  const StructDesc* desc;
  const IoSlot* io;
  int order;
  // optional: the pair sums of all evaluations as one multi-vector sum (all evaluations are
  // views of one structure); the per-structure pair kind then has no blocks
  int has_pm;
  PairMulti pm;
  // optional: the pair sums run on `aux` beside the SPME chain (fork after the charges are
  // gathered, join before the gather)
  hipStream_t aux;
  hipEvent_t fork_ev, join_ev;
  GroupMap map[kGroupKinds];
  int blocks[kGroupKinds];
  size_t smem[kGroupKinds];  // dynamic LDS: the largest of the structures in the kind
  int persistent;            // 1: bounded grids walking the work items (kGroupWgPerCu per CU)
  // optional: the FFTs of the evaluations whose structures take register transforms, as one
  // batched sequence per (mesh, mode, lanes); the FFT kinds cover the other evaluations, so an
  // evaluation's transforms do not depend on what else is in the launch
  std::vector<WaveGroup> wave;
};

template <typename R>
hipError_t launch_group(const GroupLaunch& G, hipStream_t st);
hipError_t launch_pair_spmm(const PairMulti& pm, bool fp32, hipStream_t st);

// ev (optional, 6 events): recorded after spread, x, y, z, inverse y, inverse x; ev_region
// (optional): after the privatised spreading's region kernel, before its assembly (at the start
// of the spreading when it has no region kernel).
template <typename R>
hipError_t launch_spme(const SpmeLaunch& L, hipStream_t st, hipEvent_t* ev,
                       hipEvent_t ev_region = nullptr);

// qs[k] = dq[perm[k]]; zero (when not null) is cleared for the symmetric pair sum.
hipError_t launch_gather_charges(int n, const int* perm, const double* dq, double* qs,
                                 double* zero, hipStream_t st);
// Total charge of qs[0..n) in a fixed order (one workgroup) into *qsum.
hipError_t launch_charge_sum(int n, const double* qs, double* qsum, hipStream_t st);
// beside: the sum runs next to other kernels on a bounded grid of beside_blocks workgroups
// (0: 512), which leaves them CU room and bandwidth; otherwise it has the device to itself and
// takes a workgroup per four rows.
hipError_t launch_pair_spmv(int n, const long long* row_ptr, const void* col, bool col16,
                            const void* coef, bool fp32, const double* q, double* out,
                            bool beside, hipStream_t st, int beside_blocks = 0);
hipError_t launch_compress_cols(int n, const long long* row_ptr, const int* col, short* off,
                                int* too_far, hipStream_t st);

// Geometry time: count the entries per (chunk, row) of the full list (row_ptr, col) into rc,
// then write them (rc holding the starting positions) with their coefficients.
hipError_t launch_half_count(int n, const long long* row_ptr, const int* col, const int* tile_of,
                             const int* tile_start, const HalfPairs& hp, int* rc,
                             hipStream_t st);
hipError_t launch_half_fill(int n, const long long* row_ptr, const int* col, const void* coef,
                            bool fp32, const int* tile_of, const int* tile_start,
                            const HalfPairs& hp, const int* rc, unsigned short* entry,
                            void* coef_out, hipStream_t st);
// Per evaluation: out (zeroed beforehand) += the pair sum, one workgroup per tile.
hipError_t launch_pair_half(int ntiles, const int* tile_start, const HalfPairs& hp, bool fp32,
                            const double* q, double* out, hipStream_t st);
template <typename R>
hipError_t launch_gather_combine(int n, const int4* k0, const void* theta, const void* grid,
                                 const double* q, const double* phi_real, const double* onsite,
                                 const int* perm, const GatherArgs& g, double* shift,
                                 double* epart, double* energy, hipStream_t st);
// Per-tile gather with the mesh region staged in LDS; needs
// (T1+p-1)(T2+p-1)(T3+p-1) * sizeof(R) <= kGatherRegionMaxBytes.  epart: one per tile.
// The split gather of an evaluation whose pair sum runs beside the chain: the mesh potential of
// every atom (sorted order) before the join, then the shift in caller order with the pair sum,
// on-site, self and background terms and the energy after it.
template <typename R>
hipError_t launch_gather_mesh(const int4* k0, const void* theta, const void* grid,
                              const int* tile_start, const GatherArgs& g, double* mesh,
                              hipStream_t st);
hipError_t launch_shift_combine(int n, const double* mesh, const double* phi_real,
                                const double* q, const double* onsite, const int* perm,
                                const GatherArgs& g, double* shift, double* epart,
                                double* energy, hipStream_t st);
template <typename R>
hipError_t launch_gather_tile(const int4* k0, const void* theta, const void* grid,
                              const double* q, const double* phi_real, const double* onsite,
                              const int* perm, const int* tile_start, const GatherArgs& g,
                              double* shift, double* epart, double* energy, hipStream_t st);
hipError_t launch_bspline(int n, const double* frac, const BsplineArgs& g, void* theta, bool fp32,
                          hipStream_t st);
// Linked-tile pair list: row lengths, then the rows (max_tile_atoms: most atoms in a sort tile).
hipError_t launch_pair_count(int n, const double* frac, const int* sp, const int* tile_of,
                             const int* tile_start, const PairArgs& g, int max_tile_atoms,
                             long long* counts, int* err, hipStream_t st);
// Partners go to col (int32) or, when col is null, as 16-bit offsets to off, with *too_far set
// when one does not fit.
hipError_t launch_pair_fill(int n, const double* frac, const int* sp, const int* tile_of,
                            const int* tile_start, const PairArgs& g, int max_tile_atoms,
                            long long* row_ptr, int* col, short* off, int* too_far, void* coef,
                            bool fp32, int* err, hipStream_t st);
// Virial W_ab = -dE/d(strain_ab) of the Ewald energy, as per-block partials of 6 components
// (xx, yy, zz, xy, xz, yz) that the caller sums: the reciprocal part from the forward spectrum
// left in L.cgrid by a forward_only launch_spme (nparts = number of blocks written), the
// real-space part over the pairs within r_cut.
template <typename R>
hipError_t launch_recip_virial(const SpmeLaunch& L, double* partials, int* nparts,
                               hipStream_t st);
hipError_t launch_real_virial(int n, const double* frac, const int* tile_of,
                              const int* tile_start, const double* q, const PairArgs& pa,
                              bool dense, double* partials, int* nparts, hipStream_t st,
                              const PairRows& rows = {});

// Reciprocal force (overwrites force, caller order) then the real-space Ewald force (adds);
// with rows.row_ptr set (linked mode) the real-space part walks the pair list.
template <typename R>
hipError_t launch_forces(int n, const double* frac, const int4* k0, const void* grid,
                         const double* q, const int* perm, const int* tile_of,
                         const int* tile_start, const ForceArgs& fa, const PairArgs& pa,
                         bool dense, double* force, hipStream_t st, const PairRows& rows = {});
// Workgroups (partials) the real-space virial writes.
inline int real_virial_parts(int n, bool pair_rows) { return pair_rows ? (n + 3) / 4 : (n + 255) / 256; }
hipError_t launch_pair_dense(int n, const double* frac, const int* sp, const PairArgs& g,
                             int* col, void* coef, bool fp32, int* err, hipStream_t st);

// External point charges (sources) acting on the resident atoms (targets): the real-space part
// of the periodic Ewald sum of TCoulomb%getPotential (src/dftbp/dftb/coulomb.F90:3121-3160 ->
// sumInvRPeriodicAsymm 1340-1418 -> ewald 2311-2345, ewaldReal 2523-2575).
struct ExtArgs {
  double lat[9];        // rows a_i
  double alpha;
  double rc2;           // squared cutoff of the source sum
  double eps2;          // squared softening length (0: none)
  double self_coef;     // -2 alpha / sqrt(pi): a source on a target (or on one of its images)
  int D[3];             // source tiles (linked mode) or image shells (dense mode) per side
  int full[3];          // linked mode: 1 = sweep every tile of this axis once
  int nt[3];
  int m;                // sources
  const double* frac;   // sources in tile order, fractional [m][3]
  const double* q;      // their charges
  const double* blur;   // their Gaussian widths, or null (point charges)
  const double* tblur;  // gradient only: the targets' Gaussian widths (used in place of blur)
  const int* start;     // [ntiles + 1] first source of each tile
};
// phi[k] = real-space potential of the sources at sorted target k (overwritten).
hipError_t launch_ext_real(int n, const double* frac, const int* tile_of, const ExtArgs& g,
                           bool dense, double* phi, hipStream_t st);
// out[perm[k]][3] = scale[k] * (gradient of that real-space sum at sorted target k), scale null: 1.
hipError_t launch_ext_real_grad(int n, const double* frac, const int* tile_of, const int* perm,
                                const double* scale, const ExtArgs& g, bool dense, double* out,
                                hipStream_t st);
// Reciprocal force only: force[perm[k]][3] = -q_k grad(mesh potential) at sorted point k.
template <typename R>
hipError_t launch_force_recip(int n, const double* frac, const int4* k0, const void* grid,
                              const double* q, const int* perm, const ForceArgs& fa,
                              double* force, hipStream_t st);

}  // namespace ewamd
