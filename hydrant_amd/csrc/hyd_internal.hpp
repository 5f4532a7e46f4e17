// Internal declarations shared by the C ABI implementation and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/hydrant_hip.h"
#include "hyd_math.hpp"

namespace hyd {

void set_error(const char *fmt, ...);

#define HYD_CHECK_HIP(expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      ::hyd::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                       __LINE__);                                                    \
      return HYD_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

#define HYD_REQUIRE(cond, code, ...)   \
  do {                                 \
    if (!(cond)) {                     \
      ::hyd::set_error(__VA_ARGS__);   \
      return (code);                   \
    }                                  \
  } while (0)

// Device-resident paging state.  page[i] = {k, b.x, b.y, b.z}.
struct Paging {
  const int32_t *vaddr = nullptr;
  int32_t vdim[3] = {0, 0, 0};
  const float4 *page = nullptr;
  int32_t n_pages = 0;
  int32_t lowest_blkcnt = 0;
};

struct Pool {
  const void *data = nullptr;
  int32_t fmt = -1;
  int32_t pad = 0;       // padded brick edge
  int32_t n_pages = 0;
  bool owned = false;
};

struct Grid {
  const int32_t *data = nullptr;
  int32_t dim[3] = {0, 0, 0};
};

struct GridF {
  const float *data = nullptr;
  int32_t dim[3] = {0, 0, 0};
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  Pool pool;
  Paging paging;
  Grid cheb;
  GridF mean;
  const float *tf = nullptr;
  int32_t tf_n = 0;
  uint64_t pool_gen = 0;          // bumped by every pool upload / bind
  // bumped whenever device memory a captured batch graph reads may move
  // (pool upload / bind, transfer function, volume set)
  uint64_t batch_state_gen = 0;
  // owned device allocations (freed at destroy)
  void *own_pool = nullptr;
  void *own_oriented = nullptr;   // the pool's oriented layouts (batched renderer)
  size_t oriented_cap = 0;
  int oriented_pairs = 0;          // own_oriented holds the y- and z-fastest single-voxel layouts (0),
                                   // the three row-pair layouts (1) or the three quad layouts (2)
  uint64_t oriented_gen = ~0ull;  // pool_gen the oriented layouts were made from
  int32_t *own_vaddr = nullptr;
  float4 *own_page = nullptr;
  size_t vaddr_cap = 0, page_cap = 0;   // bytes of own_vaddr / own_page
  // host copies of the page table and chebyshev grid, combined into one code per
  // brick (see VolumeArgs::codes) while their dims agree
  std::vector<int32_t> host_vaddr, host_cheb;
  int32_t *own_codes = nullptr;
  size_t codes_cap = 0;
  bool codes_valid = false;
  int32_t *own_cheb = nullptr;
  float *own_mean = nullptr;
  float *own_tf = nullptr;
  const float *u8_lut = nullptr;  // 256-entry float(v)/255 table
  const float *srgb8_t = nullptr; // 256 sRGB byte thresholds (srgb8_thresholds)
  float *own_lut = nullptr;       // both tables, 512 floats
  // multi-volume set of the batched renderer
  void *own_vols = nullptr;       // BatchVolDev[n_vols]
  int32_t *own_cells = nullptr;   // concatenated cell codes
  int32_t n_vols = 0;
  int32_t vol_pad = 0;            // padded brick edge shared by every volume
  int32_t vol_max_page = -1;   // largest page a cell of the volume set names (-1: none)
  int32_t n_cus = 0;
  // differentiable renderer: device copy of the views
  void *own_diff_views = nullptr;
  int32_t diff_view_cap = 0;
  void *own_diff_scratch = nullptr;   // adjoint scratch (per-sample dL/ds), grown on demand
  size_t diff_scratch_cap = 0;
};

// Flattened kernel arguments for the volume march (passed by value).
// Tile order of a frame launch (one 8x8 tile per workgroup, grid x = tile
// column): row r of the tile grid is rotated by HYD_TILE_ROTATE * r tiles, a
// bijection on each row.  Workgroups are dealt round-robin over the 8 XCDs and
// then over each XCD's 32 CUs, so one CU receives workgroups 256 apart; in a
// 64-tile-wide grid (512 pixels) those are the tiles of one image column,
// every fourth tile row, and the CUs that hold the centre columns, whose rays
// are the longest, set the frame.  The rotation spreads a CU's tiles over 16
// columns.  512^2 frames: plain FP32 0.551 -> 0.528 ms, paged u8 0.351 ->
// 0.317 ms (rotations 3, 5, 9, 11, 23: 0.531-0.557 and 0.316-0.319 ms).
#ifndef HYD_TILE_ROTATE
#define HYD_TILE_ROTATE 7
#endif
__device__ __forceinline__ int rotated_tile_x() {
  return (int)((blockIdx.x + (unsigned)HYD_TILE_ROTATE * blockIdx.y) % gridDim.x);
}

struct VolumeArgs {
  // shader
  f3 bmin, bmax;
  float step;
  int32_t max_steps;
  int32_t mode;
  float rank;
  float density;
  int32_t schedule;
  float ert;
  // view
  float trans[16];
  float ctg;
  f3 ray_o;
  int32_t W, H;
  // tables
  const float *tf;
  int32_t tf_n;
  const int32_t *cheb;
  int32_t cdim[3];
  const int32_t *vaddr;
  int32_t vdim[3];
  // one code per brick when the chebyshev and page-table grids have equal dims
  // (else null): page id >= 0, -1 not resident, <= -2 empty at distance -1 - code
  const int32_t *codes;
  int32_t lowest_blkcnt;
  const float4 *page;
  const void *pool;
  uint64_t pool_elems;    // voxels in the pool (n_pages * pad^3)
  int32_t pad;
  const float *u8_lut;
  // outputs
  hyd_volume_pixel *film;
  float4 *rgba;
  unsigned long long *samples;
  // sort-last segments: slots[ord * W*H + ray], nslots[ray]
  float4 *slots;
  int32_t *nslots;
  int32_t max_slots;
};

hipError_t launch_volume(const VolumeArgs &a, int fmt, bool resume, hipStream_t s);
hipError_t launch_segments_clear(float4 *slots, size_t n, hipStream_t s);
hipError_t launch_segments_pass(const VolumeArgs &a, int fmt, hipStream_t s);
// Residency batches rendered by one segment launch (grid z = batch): each
// batch's pool, per-brick codes and page descriptors replace the arguments'.
constexpr int kMaxSegBatches = 16;
struct SegBatchTable {
  const void *pool[kMaxSegBatches];
  const int32_t *codes[kMaxSegBatches];
  const float4 *page[kMaxSegBatches];
  int32_t n;
};
hipError_t launch_segments_batches(const VolumeArgs &a, const SegBatchTable &tb, int fmt, hipStream_t s);
hipError_t launch_segments_composite(const float4 *slots, const int32_t *nslots, int max_slots, int P, float4 *rgba,
                                     hipStream_t s);
hipError_t launch_volume_trace(const VolumeArgs &a, int fmt, int x, int y, hyd_trace_rec *out_dev, int max_recs,
                               int *n_dev, hipStream_t s);
hipError_t launch_pixel_pass(const float4 *rgba, int W, int H, uchar3 clear, uint8_t *rgb, hipStream_t s);
hipError_t launch_fetch_pass(const hyd_volume_pixel *film, int W, int H, hyd_fetch_pixel *out, hipStream_t s);
hipError_t launch_composite(int n, const hyd_fetch_pixel *partials, const int32_t *z_order_dev, int W, int H,
                            float4 *rgba, uint8_t *rgb, const float *srgb_t, hipStream_t s);
hipError_t launch_iso_fetch(const hyd_iso_pixel *px, int W, int H, hyd_iso_fetch_pixel *out, hipStream_t s);
hipError_t launch_iso_composite(int n, const hyd_iso_fetch_pixel *partials, const int32_t *z_order_dev, int W, int H,
                                uint8_t *rgb, float *depth, hipStream_t s);
// t[0] = -inf, t[k] (k = 1..255) = least float whose saturated sRGB byte is >= k
// (volume.cc:175-181,284-288 evaluated on the host).
void srgb8_thresholds(float t[256]);

struct BlocksArgs {
  f3 bmin, bmax;
  float step;
  int32_t max_steps;
  int32_t mode;
  float density;
  float trans[16];
  float ctg;
  f3 ray_o;
  int32_t W, H;
  const int32_t *cheb;
  const float *mean;
  int32_t dim[3];
  const int32_t *vaddr;   // page table (HYD_BLOCKS_PAGING)
  int32_t vdim[3];
  int32_t lowest;         // pages at or above this id are level-0 pages
  float4 *rgba;
};
hipError_t launch_blocks(const BlocksArgs &a, hipStream_t s);

struct IsoArgs {
  f3 bmin, bmax;
  float step, du;
  int32_t max_steps;
  int32_t mode;
  float to_world[16];
  f3 eye_pos, light_pos, surface_color;
  float isovalue;
  int32_t corrected, bisect_steps;
  float trans[16];
  float ctg;
  f3 ray_o;
  int32_t W, H;
  const int32_t *cheb;
  int32_t cdim[3];
  const int32_t *vaddr;
  int32_t vdim[3];
  const int32_t *codes;   // as VolumeArgs::codes
  const float4 *page;
  const void *pool;
  uint64_t pool_elems;    // voxels in the pool (n_pages * pad^3)
  int32_t pad;
  const float *u8_lut;
  hyd_iso_pixel *out;
  unsigned long long *samples;
};
hipError_t launch_iso(const IsoArgs &a, int fmt, hipStream_t s);

int32_t activate(Ctx *c);
// After pool pages [first, first + n) were overwritten in place: updates the
// batched renderer's oriented copies of those pages if they are current.
int32_t refresh_oriented_pages(Ctx *c, int32_t first, int32_t n);

}  // namespace hyd

// The opaque handle of the C ABI.
struct hyd_ctx : hyd::Ctx {};
