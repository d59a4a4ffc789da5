// Host-side construction of the series -> chunk -> tile decomposition used by
// every chunked kernel (see bn_lanes.h).  Pure C++, shared by the HIP library
// and its host build.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <string>
#include <vector>

namespace bn {

static constexpr int kTile = 128;  // chunks per level-1 tile, and tiles per series at level 2

struct PlanHost {
  int32_t n_series = 0, n_chunks = 0, n_chunks_pad = 0, n_tiles = 0, chunk_len = 0, gshift = 6;
  int64_t n_slots = 0, n_steps = 0;
  bool needs_scan = false;  // some series spans more than one chunk
  std::vector<int32_t> chunk_series, chunk_len_, chunk_tile, tile_first, tile_count, tile_series;
  std::vector<int32_t> series_first_tile, series_ntiles, series_first_chunk, series_nchunks;
  std::vector<uint8_t> chunk_flags, tile_flags;
  std::vector<int64_t> chunk_start;
  // Scan groups of the tile scans: workgroup g scans the up to kGroupMax tiles
  // group_tiles[g * kGroupMax + h] (-1 for an empty place) side by side, each on
  // 1 / group_w[g] of the lanes, so a tile of a group of width w has at most kTile / w
  // chunks.  Tiles of a one-chunk series need no scan and are in no group.
  std::vector<int32_t> group_tiles;
  std::vector<uint8_t> group_w;
  // Longest chunk among each 64 consecutive chunks (the lanes of one wavefront of the
  // lane-per-chunk kernels): the backward walks run their steps in step with it.
  std::vector<int32_t> wave_maxlen;
};

constexpr int kGroupMax = 8;

// Packs the tiles into scan groups by size: width 8 for tiles of at most kTile / 8
// chunks, then 4 and 2, width 1 for the rest.  The width-1 groups (the deepest scans)
// come first so they start in the first round of workgroups.
inline void build_scan_groups(PlanHost& p) {
  p.group_tiles.clear();
  p.group_w.clear();
  std::vector<int32_t> bucket[4];  // widths 1, 2, 4, 8
  for (int32_t t = 0; t < (int32_t)p.tile_count.size(); ++t) {
    const int32_t c = p.tile_count[t];
    if (p.tile_flags[t] == 3 && c == 1) continue;  // first and last tile of a one-chunk series
    int b = 0;
    while (b < 3 && c <= kTile / (2 << b)) ++b;
    bucket[b].push_back(t);
  }
  for (int b = 0; b < 4; ++b) {
    const int w = 1 << b;
    for (size_t i = 0; i < bucket[b].size(); i += w) {
      for (int h = 0; h < kGroupMax; ++h)
        p.group_tiles.push_back(h < w && i + h < bucket[b].size() ? bucket[b][i + h] : -1);
      p.group_w.push_back((uint8_t)w);
    }
  }
}
inline bool build_plan(const int64_t* offsets, int32_t n_series, int32_t chunk_len, PlanHost& p, std::string& err) {
  if (n_series < 1) { err = "n_series must be >= 1"; return false; }
  if (chunk_len < 1 || chunk_len > (1 << 24)) { err = "chunk_len must be in [1, 2^24]"; return false; }
  if (offsets[0] != 0) { err = "offsets[0] must be 0"; return false; }
  p = PlanHost();
  p.n_series = n_series;
  p.chunk_len = chunk_len;
  p.series_first_tile.assign(n_series, -1);
  p.series_ntiles.assign(n_series, 0);
  p.series_first_chunk.assign(n_series, 0);
  p.series_nchunks.assign(n_series, 0);
  for (int32_t b = 0; b < n_series; ++b) {
    const int64_t n = offsets[b + 1] - offsets[b];
    if (n < 0) { err = "offsets must be non-decreasing"; return false; }
    const int64_t nc = (n + chunk_len - 1) / chunk_len;
    if (nc > (int64_t)kTile * kTile) {
      err = "series " + std::to_string(b) + " needs " + std::to_string(nc) +
            " chunks (> kTile^2): raise chunk_len";
      return false;
    }
    p.series_first_chunk[b] = (int32_t)p.chunk_series.size();
    p.series_nchunks[b] = (int32_t)nc;
    if (nc > 1) p.needs_scan = true;
    if (nc == 0) continue;
    const int32_t nt = (int32_t)((nc + kTile - 1) / kTile);
    p.series_first_tile[b] = (int32_t)p.tile_first.size();
    p.series_ntiles[b] = nt;
    for (int32_t t = 0; t < nt; ++t) {
      const int32_t cfirst = (int32_t)p.chunk_series.size() + t * kTile;
      const int32_t cnt = (int32_t)std::min<int64_t>(kTile, nc - (int64_t)t * kTile);
      p.tile_first.push_back(cfirst);
      p.tile_count.push_back(cnt);
      p.tile_series.push_back(b);
      p.tile_flags.push_back((uint8_t)((t == 0 ? 1 : 0) | (t == nt - 1 ? 2 : 0)));
    }
    for (int64_t c = 0; c < nc; ++c) {
      const int64_t start = offsets[b] + c * chunk_len;
      const int64_t len = std::min<int64_t>(chunk_len, offsets[b + 1] - start);
      p.chunk_series.push_back(b);
      p.chunk_len_.push_back((int32_t)len);
      p.chunk_start.push_back(start);
      p.chunk_flags.push_back((uint8_t)((c == 0 ? 1 : 0) | (c == nc - 1 ? 2 : 0)));
      p.chunk_tile.push_back(p.series_first_tile[b] + (int32_t)(c / kTile));
    }
  }
  p.n_chunks = (int32_t)p.chunk_series.size();
  p.n_tiles = (int32_t)p.tile_first.size();
  build_scan_groups(p);
  p.wave_maxlen.assign((size_t)(p.n_chunks + 63) / 64, 0);
  for (int32_t j = 0; j < p.n_chunks; ++j)
    p.wave_maxlen[(size_t)(j >> 6)] = std::max(p.wave_maxlen[(size_t)(j >> 6)], p.chunk_len_[(size_t)j]);
  // Empty series inside a batch are fine (no chunk, log-likelihood 0); a batch
  // with no step at all would launch empty grids, so it is refused here.
  if (p.n_chunks == 0) { err = "offsets describe no steps (every series is empty)"; return false; }
  // Lanes are grouped 64 to a wavefront; with fewer than 64 chunks the group
  // shrinks to the next power of two so the tiled layout does not pad 64x.
  p.gshift = 6;
  while (p.gshift > 0 && (1 << (p.gshift - 1)) >= p.n_chunks) --p.gshift;
  const int32_t G = 1 << p.gshift;
  p.n_chunks_pad = ((p.n_chunks + G - 1) / G) * G;
  p.n_slots = (int64_t)(p.n_chunks_pad / G) * chunk_len * G;
  p.n_steps = offsets[n_series];
  return true;
}

}  // namespace bn
