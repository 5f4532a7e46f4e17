// Internal declarations of the differentiable renderer.
#pragma once

#include "hyd_internal.hpp"

namespace hyd {

struct DiffView {
  float trans[16];
  float ctg;
  float ray_o[3];
  float inv[12];               // rows 0-2 of trans^-1 (world -> camera), row-major
};

struct DiffArgs {
  const float *vol;
  int32_t nx, ny, nz;
  const float *tf;
  int32_t tf_n;
  float sigma;
  int32_t K;
  int32_t ckpt;
  int32_t nck;                 // checkpoints per ray: ceil(K / ckpt) + 1
  const DiffView *views;
  int32_t n_views;
  int32_t W, H;
  float4 *rgba;                // [view][y][x]
  float *ckpts;                // [view][y][x][nck]
  const float4 *grad_rgba;     // adjoint input
  float *grad_vol;
  float *grad_tf;
  unsigned long long *samples;
  // adjoint scratch
  float *ds;                   // [view][y][x][K] dL/ds per sample
  float *dsmax;                // [view][y][x] max |dL/ds| of the ray (0: nothing to scatter)
  float4 *rays;                // [view][y][x][2] setup of active rays: (dx, dy, dz, tnear), (dt, 0, 0, 0)
  float *blocks;               // [view group][brick][10^3] voxel-gradient blocks (brick + one-voxel halo)
  int32_t view_groups;         // scatter workgroups per brick (views split into contiguous groups)
  float *bound;                // [kDiffBoundParts] partial bounds of one sample's TF-gradient term
  unsigned long long *gtf_acc; // [4][256] fixed-point TF-gradient sums of the launch
  int32_t *tile_counter;       // [1] work queue of the ray-order pass
  int32_t headroom;            // extra right shift of the TF fixed-point scale for very large launches
  int32_t n_workers;           // workgroups of the ray-order pass
};

constexpr int kDiffBoundParts = 256;   // workgroups (partial maxima) of the adjoint's bound prepass
constexpr int kDiffBlock = 1000;       // floats per voxel-gradient block of an 8^3 brick (10^3 corner slots)

inline size_t diff_bricks(int32_t nx, int32_t ny, int32_t nz) {
  return (size_t)((nx + 7) / 8) * (size_t)((ny + 7) / 8) * (size_t)((nz + 7) / 8);
}

// view groups of the scatter: about kDiffScatterWgs workgroups in all (128 per CU
// of MI355X, so that the slow bricks' workgroups are spread over a long launch),
// at most one group per view
constexpr size_t kDiffScatterWgs = 32768;

inline int32_t diff_view_groups(int32_t nx, int32_t ny, int32_t nz, int32_t n_views) {
  const size_t nb = diff_bricks(nx, ny, nz);
  const size_t g = (kDiffScatterWgs + nb - 1) / nb;
  return (int32_t)(g < 1 ? 1 : (g > (size_t)n_views ? (size_t)n_views : g));
}

// floats of the dL/ds buffer ([view][y][x][k])
inline size_t diff_ds_floats(int32_t n_views, int32_t W, int32_t H, int32_t K) {
  return (size_t)n_views * (size_t)W * (size_t)H * (size_t)K;
}
hipError_t launch_diff_forward(const DiffArgs &a, hipStream_t s);
hipError_t launch_diff_backward(const DiffArgs &a, hipStream_t s);

}  // namespace hyd
