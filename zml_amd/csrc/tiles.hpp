// Workgroup -> tile maps shared by the GEMM kernels.
#pragma once

#include <cstdint>

#include "zmlhip_internal.hpp"

namespace zmlhip {

// Tiles of a triangular update only (tri lower with ntm >= ntn, or upper with
// ntn >= ntm): the active tiles are enumerated compactly, in groups of 8 tile
// columns (lower; 8 tile rows for upper) so the group's B (A) panels stay in L2,
// and split into equal contiguous ranges per XCD (workgroup ids are dealt
// round-robin over the 8 XCDs). In group h
// (width g), the first g rows (columns) form a triangle of 1, 2, .., g tiles,
// the rest are full rows (columns) of g tiles.
__host__ __device__ inline int64_t tri_group_tiles(int64_t h, int64_t nlong, int64_t nshort) {
  const int64_t g = nshort - 8 * h < 8 ? nshort - 8 * h : 8;
  return g * (g + 1) / 2 + (nlong - 8 * h - g) * g;
}

inline __device__ void tri_tile_coords(int64_t bid, int64_t nt, int64_t ntm, int64_t ntn, int tri, int64_t& tm,
                                       int64_t& tn) {
  const int64_t x = bid & 7, q = nt >> 3, r8 = nt & 7;
  int64_t pid = x * q + (x < r8 ? x : r8) + (bid >> 3);
  const int64_t nlong = tri == kLower ? ntm : ntn, nshort = tri == kLower ? ntn : ntm;
  int64_t h = 0;
  for (;; ++h) {
    const int64_t c = tri_group_tiles(h, nlong, nshort);
    if (pid < c) break;
    pid -= c;
  }
  const int64_t g = nshort - 8 * h < 8 ? nshort - 8 * h : 8;
  int64_t lo, sh;   // position along the long side, along the short side
  if (pid < g * (g + 1) / 2) {
    int64_t k = 0;
    while ((k + 1) * (k + 2) / 2 <= pid) ++k;
    lo = 8 * h + k;
    sh = 8 * h + (pid - k * (k + 1) / 2);
  } else {
    const int64_t r = pid - g * (g + 1) / 2;
    lo = 8 * h + g + r / g;
    sh = 8 * h + r % g;
  }
  if (tri == kLower) { tm = lo; tn = sh; }
  else { tm = sh; tn = lo; }
}

// Number of tiles tri_tile_coords enumerates, or 0 when the shape is not one
// it covers (then every tile is launched and the kernel skips the inactive ones).
inline int64_t tri_tile_count(int tri, int64_t ntm, int64_t ntn) {
  if (!((tri == kLower && ntm >= ntn) || (tri == kUpper && ntn >= ntm))) return 0;
  const int64_t nlong = tri == kLower ? ntm : ntn, nshort = tri == kLower ? ntn : ntm;
  int64_t t = 0;
  for (int64_t h = 0; 8 * h < nshort; ++h) t += tri_group_tiles(h, nlong, nshort);
  return t;
}

}  // namespace zmlhip
