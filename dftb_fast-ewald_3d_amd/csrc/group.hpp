// group.hpp -- block counts and LDS sizes of the grouped kernels for one structure (shared by
// the structure pool and the multi-vector evaluation of one structure).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>

#include "kernels.hpp"

namespace ewamd {

inline size_t smem_of(const StructDesc& d, int kind, bool fp32) {
  const size_t rs = fp32 ? sizeof(float) : sizeof(double);
  const size_t cs = 2 * rs;
  const int p = d.spread.order;
  switch (kind) {
    case kGkRegion:
      return sizeof(double) * (size_t)(d.region.S[0] + p - 1) * (d.region.S[1] + p - 1) *
             (d.region.S[2] + p - 1);
    case kGkSlotted:
      return sizeof(unsigned long long) * kSpreadRows * 4 + sizeof(int4) * (size_t)d.spread.slots +
             rs * 3 * (size_t)d.spread.slots * p;
    case kGkFftX:
      return cs * ((size_t)d.px.lines * d.px.pitch + d.px.n);
    case kGkFftY:
      return cs * ((size_t)d.py.lines * d.py.pitch + d.py.n);
    case kGkFftZ:
      return cs * ((size_t)d.pz.lines * d.pz.pitch + d.pz.n);
    case kGkGatherTile:
      return gather_tile_smem(d.spread.T, p, rs);
    case kGkPlane:
      return d.plane ? plane_smem(d.px, d.py, cs) : 0;
    case kGkOwned:
      return d.spread_mode == 3 ? owned_smem(d.region, fp32) : 0;
    default:
      return 0;
  }
}

inline int blocks_of(const StructDesc& d, int kind) {
  const int Kh = d.K[0] / 2 + 1;
  switch (kind) {
    case kGkCharges: return (d.n + 255) / 256;
    case kGkRegion: return d.spread_mode == 0 ? d.region.nb[0] * d.region.nb[1] * d.region.nb[2] : 0;
    case kGkAssemble:
      return d.spread_mode == 0
                 ? d.region.nb[0] * d.region.nb[1] * d.region.nb[2] * kAssembleSplit
                 : 0;
    case kGkOwned:
      return d.spread_mode == 3 ? d.region.nb[0] * d.region.nb[1] * d.region.nb[2] : 0;
    case kGkSlotted: return d.spread_mode == 1 ? d.ntiles : 0;
    case kGkChecked: return d.spread_mode == 2 ? d.ntiles : 0;
    case kGkFftX: return d.plane ? 0 : (d.K[1] * d.K[2] + 2 * d.px.lines - 1) / (2 * d.px.lines);
    case kGkFftY: return d.plane ? 0 : (Kh * d.K[2] + d.py.lines - 1) / d.py.lines;
    case kGkPlane: return d.plane ? d.K[2] : 0;
    case kGkFftZ: return (Kh * d.K[1] + d.pz.lines - 1) / d.pz.lines;
    case kGkPairs: return d.half ? d.ntiles : (d.n + 3) / 4;
    case kGkGatherTile: return d.gather_tile ? d.ntiles : 0;
    case kGkGatherAtom: return d.gather_tile ? 0 : (d.n + 255) / 256;
    case kGkEnergy: return 1;
    case kGkChargeSum: return d.ga.qsum ? 1 : 0;
    default: return 0;
  }
}

// Host image of the map and io arrays of one grouped launch (per kind: prefix[count+1] and
// slot[count], then the IoSlots).
struct Staged {
  std::vector<int> words;
  size_t io_offset = 0;  // in bytes
  // batched register-transform FFT groups: each group's table (in bytes), its evaluations and a
  // structure of the group (its mesh, lanes and twiddle tables)
  std::vector<size_t> wave_offset;
  std::vector<int> wave_n, wave_sid;
  int offs[kGroupKinds] = {}, counts[kGroupKinds] = {}, blocks[kGroupKinds] = {};
  size_t smem[kGroupKinds] = {};
};

inline bool fft_kind(int k) {
  return k == kGkFftX || k == kGkFftY || k == kGkFftZ || k == kGkPlane;
}

// Stages `count` evaluations; descs is indexed by IoSlot::sid.  With skip_pairs the pair kind
// gets no blocks (the launch's multi-vector pair sum covers it); evaluations with wave_fft[i] set
// get no blocks in the FFT kinds (a batched register-transform group covers them).  False when a
// kind needs more LDS than a workgroup has.
inline bool stage_group(const StructDesc* descs, const IoSlot* io, int count, bool fp32,
                        bool skip_pairs, Staged* S, const unsigned char* wave_fft = nullptr) {
  S->words.clear();
  for (int k = 0; k < kGroupKinds; ++k) {
    std::vector<int> prefix, slot;
    int run = 0;
    size_t sm = 0;
    for (int i = 0; i < count; ++i) {
      const StructDesc& d = descs[io[i].sid];
      const bool skip = (skip_pairs && k == kGkPairs) || (wave_fft && wave_fft[i] && fft_kind(k));
      const int b = skip ? 0 : blocks_of(d, k);
      if (b <= 0) continue;
      prefix.push_back(run);
      slot.push_back(i);
      run += b;
      sm = std::max(sm, smem_of(d, k, fp32));
    }
    prefix.push_back(run);
    S->offs[k] = (int)S->words.size();
    S->counts[k] = (int)slot.size();
    S->blocks[k] = run;
    S->smem[k] = sm;
    S->words.insert(S->words.end(), prefix.begin(), prefix.end());
    S->words.insert(S->words.end(), slot.begin(), slot.end());
  }
  while ((S->words.size() * sizeof(int)) % alignof(IoSlot)) S->words.push_back(0);
  S->io_offset = S->words.size() * sizeof(int);
  const size_t nw = (S->io_offset + sizeof(IoSlot) * (size_t)count + sizeof(int) - 1) / sizeof(int);
  S->words.resize(nw, 0);
  std::memcpy(reinterpret_cast<char*>(S->words.data()) + S->io_offset, io,
              sizeof(IoSlot) * (size_t)count);
  for (int k = 0; k < kGroupKinds; ++k)
    if (S->smem[k] > (k == kGkPlane    ? kPlaneMaxLds
                      : k == kGkRegion ? kRegionMaxLds
                      : k == kGkOwned  ? kRegionMaxLds - kOwnedStatic
                                       : 65536))
      return false;
  return true;
}

// Appends one batched register-FFT group: the meshes, spectra and influence tables of its
// evaluations; sid: a structure of the group.
inline void stage_wave_io(const std::vector<WaveIo>& w, int sid, Staged* S) {
  while ((S->words.size() * sizeof(int)) % alignof(WaveIo)) S->words.push_back(0);
  const size_t off = S->words.size() * sizeof(int);
  S->wave_offset.push_back(off);
  S->wave_n.push_back((int)w.size());
  S->wave_sid.push_back(sid);
  const size_t nw = (off + sizeof(WaveIo) * w.size() + sizeof(int) - 1) / sizeof(int);
  S->words.resize(nw, 0);
  std::memcpy(reinterpret_cast<char*>(S->words.data()) + off, w.data(), sizeof(WaveIo) * w.size());
}

// The launch over a device copy (at dev) of S's words, with the descriptor table at desc_dev.
inline GroupLaunch group_launch_of(const Staged& S, const StructDesc* desc_dev, const void* dev,
                                   int order, bool persistent) {
  GroupLaunch G{};
  G.persistent = persistent ? 1 : 0;
  G.desc = desc_dev;
  const int* w = static_cast<const int*>(dev);
  G.io = reinterpret_cast<const IoSlot*>(static_cast<const char*>(dev) + S.io_offset);
  G.order = order;
  for (int k = 0; k < kGroupKinds; ++k) {
    G.map[k] = GroupMap{w + S.offs[k], w + S.offs[k] + S.counts[k] + 1, S.counts[k],
                        S.blocks[k]};
    G.blocks[k] = S.blocks[k];
    G.smem[k] = S.smem[k];
  }
  // (the caller fills each group's launch parameters, wave[g].l)
  G.wave.resize(S.wave_n.size());
  for (size_t g = 0; g < S.wave_n.size(); ++g) {
    G.wave[g].n = S.wave_n[g];
    G.wave[g].io =
        reinterpret_cast<const WaveIo*>(static_cast<const char*>(dev) + S.wave_offset[g]);
  }
  return G;
}

}  // namespace ewamd
