// C ABI of the differentiable renderer (forward with checkpoints + adjoint).
#include <cstring>
#include <vector>

#include "hyd_diff.hpp"

using namespace hyd;

namespace {

int32_t nck_of(const hyd_diff_params *p) { return (p->n_steps + p->ckpt_every - 1) / p->ckpt_every + 1; }

int32_t check_params(const hyd_diff_params *p) {
  HYD_REQUIRE(p, HYD_ERR_ARG, "null diff params");
  HYD_REQUIRE(p->dims[0] > 0 && p->dims[1] > 0 && p->dims[2] > 0, HYD_ERR_ARG, "volume dims must be > 0");
  HYD_REQUIRE(p->n_steps > 0 && p->ckpt_every > 0, HYD_ERR_ARG, "n_steps and ckpt_every must be > 0");
  HYD_REQUIRE(p->sigma >= 0.f, HYD_ERR_ARG, "sigma must be >= 0");
  return HYD_OK;
}

int32_t fill(Ctx *c, const hyd_diff_params *p, const hyd_view *views, int32_t n_views, int32_t width,
             int32_t height, const float *vol, const float *tf, int32_t tf_n, DiffArgs &a) {
  int32_t st = check_params(p);
  if (st) return st;
  HYD_REQUIRE(views && n_views >= 1 && n_views <= 65535 && width > 0 && height > 0, HYD_ERR_ARG,
              "invalid views or image size");
  HYD_REQUIRE(vol && tf && tf_n >= 1, HYD_ERR_ARG, "null volume or transfer function");
  HYD_REQUIRE(tf_n <= 256, HYD_ERR_RANGE, "transfer function has %d entries; the limit is 256", tf_n);
  // the views go to a context-owned device buffer, grown on demand
  if (c->diff_view_cap < n_views) {
    if (c->own_diff_views) (void)hipFree(c->own_diff_views);
    c->own_diff_views = nullptr;
    c->diff_view_cap = 0;
    HYD_CHECK_HIP(hipMalloc(&c->own_diff_views, (size_t)n_views * sizeof(DiffView)));
    c->diff_view_cap = n_views;
  }
  DiffView *dviews = static_cast<DiffView *>(c->own_diff_views);
  std::vector<DiffView> hv(n_views);
  for (int i = 0; i < n_views; ++i) {
    std::memcpy(hv[i].trans, views[i].trans, sizeof(hv[i].trans));
    hv[i].ctg = views[i].ctg_fovy_2;
    std::memcpy(hv[i].ray_o, views[i].ray_o, sizeof(hv[i].ray_o));
  }
  HYD_CHECK_HIP(hipMemcpyAsync(dviews, hv.data(), hv.size() * sizeof(DiffView), hipMemcpyHostToDevice, c->stream));
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  std::memset(&a, 0, sizeof(a));
  a.vol = vol;
  a.nx = p->dims[0]; a.ny = p->dims[1]; a.nz = p->dims[2];
  a.tf = tf;
  a.tf_n = tf_n;
  a.sigma = p->sigma;
  a.K = p->n_steps;
  a.ckpt = p->ckpt_every;
  a.nck = nck_of(p);
  a.views = dviews;
  a.n_views = n_views;
  a.W = width;
  a.H = height;
  return HYD_OK;
}

}  // namespace

extern "C" {

int32_t hyd_diff_ckpt_floats(const hyd_diff_params *p, int32_t *per_ray) {
  int32_t st = check_params(p);
  if (st) return st;
  HYD_REQUIRE(per_ray, HYD_ERR_ARG, "null output");
  *per_ray = nck_of(p);
  return HYD_OK;
}

int32_t hyd_diff_forward(hyd_ctx *c, const hyd_diff_params *p, const hyd_view *views, int32_t n_views,
                         int32_t width, int32_t height, const float *volume_dev, const float *tf_dev, int32_t tf_n,
                         float *rgba_dev, float *ckpt_dev, unsigned long long *samples_dev) {
  int32_t st = activate(c);
  if (st) return st;
  DiffArgs a;
  if ((st = fill(c, p, views, n_views, width, height, volume_dev, tf_dev, tf_n, a))) return st;
  HYD_REQUIRE(rgba_dev && ckpt_dev, HYD_ERR_ARG, "null rgba or checkpoint buffer");
  a.rgba = reinterpret_cast<float4 *>(rgba_dev);
  a.ckpts = ckpt_dev;
  a.samples = samples_dev;
  HYD_CHECK_HIP(launch_diff_forward(a, c->stream));
  return HYD_OK;
}

int32_t hyd_diff_backward(hyd_ctx *c, const hyd_diff_params *p, const hyd_view *views, int32_t n_views,
                          int32_t width, int32_t height, const float *volume_dev, const float *tf_dev, int32_t tf_n,
                          const float *ckpt_dev, const float *grad_rgba_dev, float *grad_volume_dev,
                          float *grad_tf_dev) {
  int32_t st = activate(c);
  if (st) return st;
  DiffArgs a;
  if ((st = fill(c, p, views, n_views, width, height, volume_dev, tf_dev, tf_n, a))) return st;
  HYD_REQUIRE(ckpt_dev && grad_rgba_dev && grad_volume_dev && grad_tf_dev, HYD_ERR_ARG, "null adjoint buffer");
  a.ckpts = const_cast<float *>(ckpt_dev);
  a.grad_rgba = reinterpret_cast<const float4 *>(grad_rgba_dev);
  a.grad_vol = grad_volume_dev;
  a.grad_tf = grad_tf_dev;
  HYD_CHECK_HIP(launch_diff_backward(a, c->stream));
  return HYD_OK;
}

}  // extern "C"
