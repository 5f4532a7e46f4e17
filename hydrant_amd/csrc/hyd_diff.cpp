// C ABI of the differentiable renderer (forward with checkpoints + adjoint).
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "hyd_diff.hpp"

using namespace hyd;

namespace {

// Inverse of a column-major 4x4 (m[4*col + row]) in double; rows 0-2 of the
// result, row-major, into out[12].  Returns false for a singular matrix.
bool inverse_rows(const float *m, float *out) {
  double a[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[4 * c + r];
      a[r][4 + c] = (r == c) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    if (std::fabs(a[piv][c]) < 1e-300) return false;
    if (piv != c)
      for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[piv][k]);
    const double inv = 1.0 / a[c][c];
    for (int k = 0; k < 8; ++k) a[c][k] *= inv;
    for (int r = 0; r < 4; ++r)
      if (r != c) {
        const double f = a[r][c];
        for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
      }
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) out[4 * r + c] = (float)a[r][4 + c];
  return true;
}

int32_t nck_of(const hyd_diff_params *p) { return (p->n_steps + p->ckpt_every - 1) / p->ckpt_every + 1; }

int32_t check_params(const hyd_diff_params *p) {
  HYD_REQUIRE(p, HYD_ERR_ARG, "null diff params");
  HYD_REQUIRE(p->dims[0] >= 2 && p->dims[1] > 0 && p->dims[2] > 0, HYD_ERR_ARG,
              "volume dims must be > 0, and >= 2 along x (taps load x-adjacent pairs)");
  HYD_REQUIRE(p->dims[0] < (1 << 22), HYD_ERR_RANGE, "volume rows of %d voxels exceed 2^22", p->dims[0]);
  HYD_REQUIRE(p->n_steps > 0 && p->ckpt_every > 0, HYD_ERR_ARG, "n_steps and ckpt_every must be > 0");
  HYD_REQUIRE(p->sigma >= 0.f, HYD_ERR_ARG, "sigma must be >= 0");
  HYD_REQUIRE((int64_t)p->dims[0] * p->dims[1] * p->dims[2] * 4 <= 2147483647ll, HYD_ERR_RANGE,
              "volume of %d x %d x %d voxels exceeds 2^31 - 1 bytes", p->dims[0], p->dims[1], p->dims[2]);
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
    HYD_REQUIRE(inverse_rows(views[i].trans, hv[i].inv), HYD_ERR_ARG, "view %d has a singular camera matrix", i);
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
  // adjoint scratch: per-sample dL/ds (padded 8x8 tiles), per-ray maxima, bound partials, counter
  const int64_t n_rays = (int64_t)n_views * width * height;
  HYD_REQUIRE(n_rays < (1ll << 31), HYD_ERR_RANGE, "adjoint launch of %lld rays; the limit is 2^31", (long long)n_rays);
  const size_t ds_floats = diff_ds_floats(n_views, width, height, a.K);
  const size_t rays_off = (ds_floats * sizeof(float) + 255) & ~(size_t)255;
  a.view_groups = diff_view_groups(a.nx, a.ny, a.nz, n_views);
  const size_t blocks_floats = (size_t)a.view_groups * diff_bricks(a.nx, a.ny, a.nz) * kDiffBlock;
  const size_t gtf_off = (rays_off + sizeof(float) * (8 * (size_t)n_rays + blocks_floats + (size_t)n_rays +
                                                     kDiffBoundParts) + 255) & ~(size_t)255;
  const size_t need = gtf_off + sizeof(unsigned long long) * 4 * 256 + 64;
  if (c->diff_scratch_cap < need) {
    if (c->own_diff_scratch) (void)hipFree(c->own_diff_scratch);
    c->own_diff_scratch = nullptr;
    c->diff_scratch_cap = 0;
    HYD_CHECK_HIP(hipMalloc(&c->own_diff_scratch, need));
    c->diff_scratch_cap = need;
  }
  char *base = static_cast<char *>(c->own_diff_scratch);
  a.ds = reinterpret_cast<float *>(base);
  a.rays = reinterpret_cast<float4 *>(base + rays_off);
  a.blocks = reinterpret_cast<float *>(a.rays + 2 * n_rays);
  a.dsmax = a.blocks + blocks_floats;
  a.bound = a.dsmax + n_rays;
  a.gtf_acc = reinterpret_cast<unsigned long long *>(base + gtf_off);
  a.tile_counter = reinterpret_cast<int32_t *>(a.gtf_acc + 4 * 256);
  // fixed-point headroom: the int64 accumulators hold 2^33 terms of up to 2^30 each
  const double terms = 2.0 * (double)n_rays * (double)a.K;
  a.headroom = terms > 8589934592.0 ? (int32_t)std::ceil(std::log2(terms / 8589934592.0)) : 0;
  // 4-wave workgroups, one per SIMD's wave slot: the ray-order pass holds 124
  // registers, so 4 waves per SIMD are resident (a fifth, at 96 registers with
  // spills, measured 2% faster and is not used)
  a.n_workers = c->n_cus > 0 ? c->n_cus * 4 : 1024;
  HYD_CHECK_HIP(launch_diff_backward(a, c->stream));
  return HYD_OK;
}

}  // extern "C"
