// Internal declarations of the batched persistent renderer.
#pragma once

#include "hyd_internal.hpp"

namespace hyd {

constexpr int kXcds = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin over them

struct BatchViewDev {
  float trans[16];
  float ctg;
  float ray_o[3];
  int32_t W, H;
  int32_t volume;
  int32_t pad_;
  int64_t out_offset;
};

struct BatchVolDev {
  int32_t grid[3];
  int32_t cell_off;
  int32_t bs, apron;
  float bmax[3];
  float step;
  float value_scale;
};

struct BatchArgs {
  const uint32_t *tiles;       // view << 16 | ty << 8 | tx
  int32_t chunk_start[kXcds + 1];
  unsigned *counters;          // kXcds work counters, zeroed by reset_counters
  const BatchViewDev *views;
  const BatchVolDev *vols;
  const int32_t *cells;        // concatenated cell codes of every volume
  const void *pools[3];         // the pool in its three layouts (x, y, z fastest)
  int32_t oriented;             // 0: only the x-fastest layout is held (every tile reads layout 0)
  int32_t pairs;                // 1: pools[] are the row-pair layouts (2 voxels per element); 2: the quad layouts (4)
  int32_t pad;
  uint32_t page_elems;         // pad^3
  const float *tf;
  int32_t tf_n;
  float density;
  float ert;
  int32_t max_steps;
  float4 *out;
  unsigned long long *samples;
};

hipError_t launch_batch(const BatchArgs &a, int fmt, int tile_w, int unroll, int group, int grid, hipStream_t s);

// Writes pages [first_page, first_page + n_pages) of the three row-pair
// layouts of an x-fastest FP16 or u8 pool of total_pages pages (layout p at
// dst + p * 2 e pad^3 total_pages bytes, e the voxel's size).
hipError_t launch_orient_pairs(int fmt, const void *src, void *dst, int pad, int total_pages, int first_page,
                               int n_pages, hipStream_t s);

// Writes pages [first_page, first_page + n_pages) of the three quad layouts
// of an x-fastest FP16 or u8 pool of total_pages pages (layout p at dst + p *
// 4 e pad^3 total_pages bytes, e the voxel's size).
hipError_t launch_orient_quads(int fmt, const void *src, void *dst, int pad, int total_pages, int first_page,
                               int n_pages, hipStream_t s);

// Writes pages [first_page, first_page + n_pages) of the y-fastest (at dst) and
// z-fastest (at dst + pad^3 * total_pages elements) layouts of an x-fastest
// pool of total_pages pages.
hipError_t launch_orient_pool(int fmt, const void *src, void *dst, int pad, int total_pages, int first_page,
                              int n_pages, hipStream_t s);

}  // namespace hyd
