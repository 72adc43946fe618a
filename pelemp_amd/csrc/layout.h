// layout.h -- droplet storage layout and cell ordering shared by all kernels.
//
// Droplets are stored field-major (structure of arrays): field f of the
// droplet at storage position i is base[f*cap + i].  Cells are grouped in
// 4x4x4 tiles and ordered by the tile-major key
//   key = tile_linear * 64 + (k%4 * 4 + j%4) * 4 + i%4,
// tiles in x-fastest order, so the droplets of one tile are contiguous once
// sorted by key.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pelemp_spray.h"
#include "pm_math.h"

namespace pelemp {

constexpr uint32_t kDeadKey = 0xffffffffu;
constexpr int kTile = 4;
constexpr int kTileCells = kTile * kTile * kTile;

// SoA field indices.  For the FP64 form pos is the absolute position; for
// the FP32 form it is the position inside the droplet's cell in cell units,
// in [0,1).
template <int NF>
struct Fields {
  static constexpr int POS = 0;
  static constexpr int VEL = 3;
  static constexpr int T = 6;
  static constexpr int DIA = 7;
  static constexpr int Y = 8;
  static constexpr int ND = 8 + NF;
  static constexpr int FILM = 9 + NF;
  static constexpr int BM1 = 10 + NF;
  static constexpr int BM2 = 11 + NF;
  static constexpr int N0 = 12 + NF;
  static constexpr int NHOT = 10 + NF;      // without breakup state
  static constexpr int NHOT_BRK = 13 + NF;  // with breakup state
};

// FP32 form: single-precision fields plus nf double-precision liquid species
// masses per droplet in a separate array (mass bookkeeping stays FP64; the
// diameter is written for output only and recomputed from the mass).  The
// FP32 form has no walls and so no wall films: there is no film-height
// field (droplets with a film height are refused on input).
struct Fields32 {
  static constexpr int POS = 0;  // position inside the cell, [0,1)
  static constexpr int VEL = 3;
  static constexpr int T = 6;
  static constexpr int DIA = 7;
  static constexpr int ND = 8;
  static constexpr int NHOT = 9;
};

// Heavy tiles.  A tile with more than kSegDroplets droplets is updated by
// several workgroups, one per segment of kSegDroplets cell-sorted droplets
// (the first by the tile's own workgroup, the others listed as work items).
// A stayer's rank is then local to its segment: arrival = (segment <<
// kSegShift) | rank, and seg_stay[(slot + segment) * 64 + cell] holds the
// stayers of the cell in the earlier segments (the split_combine kernel
// turns the per-segment counts into these offsets).
constexpr int32_t kSegDroplets = PELEMP_SEGMENT_DROPLETS;
constexpr int kSegShift = 13;
static_assert(kSegDroplets <= (1 << kSegShift), "segment rank field too narrow");
static_assert(kSegDroplets % 256 == 0, "segments hold whole chunks of the update workgroup");
constexpr int32_t kSegRankMask = (1 << kSegShift) - 1;

// Split state of the tiles for the next update pass.  counters: [items,
// split tiles, segment slots].
struct SplitDev {
  int32_t* tile_slot;  // per tile: first segment slot, or -1 (one segment)
  int2* items;         // (tile, segment >= 1) work items
  int4* list;          // (tile, first slot, segments, -) of every split tile
  int32_t* counters;
  double* part;        // per slot: 64 cells x NSRC partial source sums
  int32_t* seg_stay;   // per slot: 64 cells: stayers, then their offsets
};

PM_HD uint32_t tile_key(int i, int j, int k, const int* ntile) {
  const uint32_t tile = (static_cast<uint32_t>(k / kTile) * ntile[1] + (j / kTile)) * ntile[0] + (i / kTile);
  return tile * kTileCells + ((k % kTile) * kTile + (j % kTile)) * kTile + (i % kTile);
}

}  // namespace pelemp
