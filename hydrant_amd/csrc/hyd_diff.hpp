// Internal declarations of the differentiable renderer.
#pragma once

#include "hyd_internal.hpp"

namespace hyd {

struct DiffView {
  float trans[16];
  float ctg;
  float ray_o[3];
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
};

hipError_t launch_diff_forward(const DiffArgs &a, hipStream_t s);
hipError_t launch_diff_backward(const DiffArgs &a, hipStream_t s);

}  // namespace hyd
