// JSON-configured renderers: Volume, Blocks, Isosurface (the reference's
// VolumeRenderer src/slave/renderers/volume/volume.cc, BlocksRenderer
// blocks/blocks.cc, IsosurfaceRenderer isosurface/isosurface.cc), each a
// thin host driver over the C ABI of this library.  Offline rendering follows
// the reference's lossless path: the bricks a view needs are made resident in
// batches (nearest first, at most mem_limit_mb of pool and 4096 pages per
// batch, lossless_block_paging.cc:116-174); the first batch runs the emit pass,
// later batches resume the film (volume.cc:95-134); then the pixel pass writes
// RGB8 (pixel_template.hpp:11-20) and the image is copied to the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>

#include "../../include/hydrant_dataset.h"
#include "hyd_internal.hpp"
#include "renderer.hpp"

namespace hyd {

hipError_t launch_pixel_pass_strided(const float *v, int stride_floats, int W, int H, uchar3 clear, uint8_t *rgb,
                                     hipStream_t s);

namespace front {
namespace {

void check(int32_t st) {
  if (st != HYD_OK) throw Error(st, hyd_last_error());
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(HYD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

std::array<float, 3> vec3(const json::Value &v, const char *what) {
  const auto &it = v.items(what);
  if (it.size() != 3) throw Error(HYD_ERR_ARG, std::string(what) + ": expected 3 numbers");
  return {(float)it[0].as_number(what), (float)it[1].as_number(what), (float)it[2].as_number(what)};
}

float num_or(const json::Value &obj, const char *key, float dflt) {
  const json::Value *v = obj.find(key);
  return (v && !v->is_null()) ? (float)v->as_number(key) : dflt;
}

std::string str_or(const json::Value &obj, const char *key, const std::string &dflt) {
  const json::Value *v = obj.find(key);
  return (v && !v->is_null()) ? v->as_string(key) : dflt;
}

int enum_of(const json::Value &obj, const char *key, std::initializer_list<std::pair<const char *, int>> names,
            int dflt) {
  const json::Value *v = obj.find(key);
  if (!v || v->is_null()) return dflt;
  if (v->is_number()) return (int)v->as_number(key);
  const std::string &s = v->as_string(key);
  for (const auto &n : names)
    if (s == n.first) return n.second;
  std::string all;
  for (const auto &n : names) all += std::string(all.empty() ? "" : ", ") + n.first;
  throw Error(HYD_ERR_ARG, std::string(key) + ": unknown value '" + s + "' (expected one of " + all + ")");
}

// TransferFnConfig.values (transfer_fn.schema.hpp:10-14): a flat list of RGBA
// quadruples, or (as the README writes it) a list of [r, g, b, a] arrays.
std::vector<float> tf_values(const json::Value &tf) {
  const json::Value *vals = tf.find("values");
  if (!vals) return {};
  std::vector<float> out;
  for (const auto &e : vals->items("transfer_fn.values")) {
    if (e.is_array()) {
      const auto &q = e.items("transfer_fn.values");
      if (q.size() != 4) throw Error(HYD_ERR_ARG, "transfer_fn.values: nested entries must be [r, g, b, a]");
      for (const auto &c : q) out.push_back((float)c.as_number("transfer_fn.values"));
    } else {
      out.push_back((float)e.as_number("transfer_fn.values"));
    }
  }
  if (out.empty() || out.size() % 4) throw Error(HYD_ERR_ARG, "transfer_fn.values: need a multiple of 4 numbers");
  if (out.size() / 4 > 256) throw Error(HYD_ERR_RANGE, "transfer_fn.values: at most 256 entries");
  return out;
}

template <class T> void dev_grow(T **p, size_t &cap, size_t bytes) {
  if (cap >= bytes) return;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  cap = 0;
  hip_check(hipMalloc(reinterpret_cast<void **>(p), bytes), "hipMalloc");
  cap = bytes;
}

}  // namespace

// ------------------------------------------------------------ config ----
hyd_camera parse_camera(const json::Value &in) {
  const json::Value *c = in.find("camera");
  const json::Value &cfg = c ? *c : in;
  if (!cfg.is_object()) throw Error(HYD_ERR_ARG, "camera: expected an object");
  float fovy = 60.f;
  if (const json::Value *p = cfg.find("perspective")) fovy = num_or(*p, "fovy", 60.f);
  hyd_camera cam{};
  if (const json::Value *o = cfg.find("orbit")) {
    const json::Value *arm = o->find("arm");
    if (!arm) throw Error(HYD_ERR_ARG, "camera.orbit: missing arm");
    const auto a = vec3(*arm, "camera.orbit.arm");
    std::array<float, 3> ctr{0.f, 0.f, 0.f};
    if (const json::Value *cc = o->find("center")) ctr = vec3(*cc, "camera.orbit.center");
    check(hyd_orbit_camera(a.data(), ctr.data(), fovy, &cam));
    return cam;
  }
  if (const json::Value *p = cfg.find("ptu")) {
    const json::Value *pos = p->find("position");
    if (!pos) throw Error(HYD_ERR_ARG, "camera.ptu: missing position");
    const auto ps = vec3(*pos, "camera.ptu.position");
    std::array<float, 3> tg{0.f, 0.f, 0.f}, up{0.f, 1.f, 0.f};
    if (const json::Value *t = p->find("target")) tg = vec3(*t, "camera.ptu.target");
    if (const json::Value *u = p->find("up")) up = vec3(*u, "camera.ptu.up");
    std::memcpy(cam.position, ps.data(), sizeof(cam.position));
    std::memcpy(cam.target, tg.data(), sizeof(cam.target));
    std::memcpy(cam.up, up.data(), sizeof(cam.up));
    cam.fovy = fovy;
    return cam;
  }
  throw Error(HYD_ERR_ARG, "camera: needs \"orbit\" or \"ptu\"");
}

RendererConfig parse_renderer_config(const json::Value &in) {
  const json::Value *r = in.find("render");
  const json::Value &cfg = r ? *r : in;
  if (!cfg.is_object()) throw Error(HYD_ERR_ARG, "renderer config: expected an object");
  RendererConfig out;
  if (const json::Value *res = cfg.find("resolution")) {
    const auto &it = res->items("resolution");
    if (it.size() != 2) throw Error(HYD_ERR_ARG, "resolution: expected [width, height]");
    out.width = (int32_t)it[0].as_number("resolution");
    out.height = (int32_t)it[1].as_number("resolution");
  }
  if (out.width <= 0 || out.height <= 0 || (int64_t)out.width * out.height > (1ll << 28))
    throw Error(HYD_ERR_RANGE, "resolution out of range");
  const json::Value *name = cfg.find("renderer");
  if (!name) throw Error(HYD_ERR_ARG, "renderer config: missing \"renderer\"");
  out.renderer = name->as_string("renderer");
  if (const json::Value *p = cfg.find("params")) out.params = *p;
  return out;
}

// ---------------------------------------------------------- IRenderer ----
void IRenderer::init(std::shared_ptr<Dataset> dataset, const RendererConfig &cfg) {
  dataset_ = std::move(dataset);
  width_ = cfg.width;
  height_ = cfg.height;
}

// BasicRenderer::update (basic_renderer.hpp:74-84, basic_renderer.schema.hpp:10-16).
void IRenderer::update(const json::Value &params) {
  if (!params.is_object() && !params.is_null()) throw Error(HYD_ERR_ARG, "params: expected an object");
  if (params.is_null()) return;
  const std::string device = str_or(params, "device", "Cuda");
  if (device == "Cpu")
    throw Error(HYD_ERR_ARG, "device \"Cpu\": this library renders on the GPU only (\"Cuda\" or \"Hip\")");
  if (device != "Cuda" && device != "Hip") throw Error(HYD_ERR_ARG, "device: unknown value '" + device + "'");
  sample_rate_ = num_or(params, "sample_rate", sample_rate_);
  if (!(sample_rate_ > 0.f)) throw Error(HYD_ERR_ARG, "sample_rate must be > 0");
  max_steps_ = (int32_t)num_or(params, "max_steps", (float)max_steps_);
  if (max_steps_ <= 0) throw Error(HYD_ERR_ARG, "max_steps must be > 0");
  if (const json::Value *cc = params.find("clear_color")) {
    const auto c = vec3(*cc, "clear_color");
    // vec3 clear colour in [0, 1] (basic_renderer.schema.hpp:16), saturated to bytes
    for (int i = 0; i < 3; ++i) clear_[i] = (uint8_t)(int)std::min(255.f, std::max(0.f, c[i] * 255.f));
  }
  compute_geometry();
}

// BasicRenderer::init/update (basic_renderer.hpp:57-84): bbox [0, raw/bs],
// exhibit centre f_dim/2, du = 1/bs, step = du/4/sample_rate, binary32.
void IRenderer::compute_geometry() {
  const float blk = (float)dataset_->block_size;
  for (int i = 0; i < 3; ++i) {
    geom_.f_dim[i] = (float)dataset_->dims[i] / blk;
    geom_.center[i] = geom_.f_dim[i] / 2.f;
  }
  geom_.du = 1.f / blk;
  geom_.step = (geom_.du / 4.f) / sample_rate_;
}

hyd_view IRenderer::view_of(const hyd_camera &camera) const {
  hyd_view v{};
  check(hyd_make_view(&camera, geom_.center, geom_.f_dim, &v));
  return v;
}

float *IRenderer::rgba_buffer(size_t pixels) {
  dev_grow(&rgba_dev_, rgba_cap_, pixels * 16);
  return static_cast<float *>(rgba_dev_);
}

uint8_t *IRenderer::rgb_buffer(size_t pixels) {
  dev_grow(&rgb_dev_, rgb_cap_, pixels * 3);
  return static_cast<uint8_t *>(rgb_dev_);
}

void IRenderer::download_rgb(uint8_t *rgb_host) {
  check(hyd_synchronize(ctx_));
  hip_check(hipMemcpy(rgb_host, rgb_dev_, (size_t)width_ * height_ * 3, hipMemcpyDeviceToHost), "hipMemcpy");
}

void IRenderer::release() {
  if (rgba_dev_) (void)hipFree(rgba_dev_);
  if (rgb_dev_) (void)hipFree(rgb_dev_);
  rgba_dev_ = rgb_dev_ = nullptr;
  rgba_cap_ = rgb_cap_ = 0;
}

std::map<std::string, RendererRegistry::Make> &RendererRegistry::types() {
  static std::map<std::string, Make> t;
  return t;
}

std::unique_ptr<IRenderer> RendererFactory::create(hyd_ctx *ctx, const RendererConfig &cfg) {
  auto it = RendererRegistry::types().find(cfg.renderer);
  if (it == RendererRegistry::types().end()) {
    std::string all;
    for (const auto &n : list_candidates()) all += (all.empty() ? "" : ", ") + n;
    throw Error(HYD_ERR_ARG, "unknown renderer '" + cfg.renderer + "' (registered: " + all + ")");
  }
  std::unique_ptr<IRenderer> r(it->second());
  r->attach(ctx);
  r->init(dataset_, cfg);
  return r;
}

std::vector<std::string> RendererFactory::list_candidates() {
  std::vector<std::string> out;
  for (const auto &e : RendererRegistry::types()) out.push_back(e.first);
  return out;
}

namespace {

// ---------------------------------------------------- brick residency ----
// Resident-set helper shared by the Volume and Isosurface renderers: packs the
// listed bricks of the dataset into a pool, uploads it with a page table that
// maps exactly those bricks (every other brick -1), BlockSamplerMapping
// k = bs/pad, b = (apron/bs)*k (lossless_block_paging.cc:21-22).
void make_resident(hyd_ctx *ctx, const Dataset &ds, int32_t pool_fmt, int32_t normalize,
                   const std::vector<int32_t> &bricks) {
  const int32_t pad = ds.block_size + 2 * ds.apron;
  const size_t vox = (size_t)pad * pad * pad;
  const size_t esz = pool_fmt == HYD_FMT_U8 ? 1 : (pool_fmt == HYD_FMT_F16 ? 2 : 4);
  const int32_t n = (int32_t)bricks.size();
  std::vector<uint8_t> pool(std::max<size_t>(1, (size_t)n) * vox * esz, 0);
  if (n > 0)
    check(hyd_pack_bricks(ds.voxels.data(), ds.fmt, ds.dims.data(), ds.block_size, ds.apron, pool_fmt, normalize,
                          bricks.data(), n, pool.data()));
  check(hyd_upload_pool(ctx, pool_fmt, pad, std::max(1, n), pool.data()));
  std::vector<int32_t> vaddr(ds.n_bricks(), -1);
  for (int32_t i = 0; i < n; ++i) vaddr[bricks[i]] = i;
  const float k = (float)ds.block_size / (float)pad;
  const float b = ((float)ds.apron / (float)ds.block_size) * k;
  std::vector<float> pk(std::max(1, n), k), pb(3 * std::max(1, n), b);
  check(hyd_set_page_table(ctx, ds.grid.data(), vaddr.data(), std::max(1, n), pk.data(), pb.data(), 0));
}

// Non-empty bricks ordered by the distance of their centre to the camera
// (LosslessBlockPagingServer sorts the culled set nearest first).
std::vector<int32_t> nearest_first(const Dataset &ds, const std::vector<uint8_t> &empty, const hyd_view &view) {
  std::vector<int32_t> ids;
  std::vector<float> d2;
  for (size_t i = 0; i < ds.n_bricks(); ++i) {
    if (empty[i]) continue;
    const int x = (int)(i % ds.grid[0]), y = (int)((i / ds.grid[0]) % ds.grid[1]),
              z = (int)(i / ((size_t)ds.grid[0] * ds.grid[1]));
    const float cx = x + 0.5f - view.ray_o[0], cy = y + 0.5f - view.ray_o[1], cz = z + 0.5f - view.ray_o[2];
    ids.push_back((int32_t)i);
    d2.push_back(cx * cx + cy * cy + cz * cz);
  }
  std::vector<size_t> ord(ids.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return d2[a] < d2[b]; });
  std::vector<int32_t> out(ids.size());
  for (size_t i = 0; i < ord.size(); ++i) out[i] = ids[ord[i]];
  return out;
}

int32_t pool_format_of(const Dataset &ds) {
  // u8 data stays u8 (the sampler normalises v/255 exactly as the reference's
  // NormalizedFloat read); f16 / f32 data keep their format
  return ds.fmt;
}

size_t pages_per_batch(const Dataset &ds, int32_t pool_fmt, double mem_limit_mb) {
  const size_t pad = (size_t)(ds.block_size + 2 * ds.apron);
  const size_t esz = pool_fmt == HYD_FMT_U8 ? 1 : (pool_fmt == HYD_FMT_F16 ? 2 : 4);
  const double bytes = mem_limit_mb * 1048576.0;
  const size_t by_mem = (size_t)std::max(1.0, std::floor(bytes / (double)(pad * pad * pad * esz)));
  return std::min<size_t>(by_mem, 4096);
}

// ------------------------------------------------------ Volume renderer ----
// VolumeRenderer (volume.cc:12-134; VolumeRendererParams ui/volume.schema.hpp:16-21).
class VolumeRenderer : public IRenderer {
 public:
  void init(std::shared_ptr<Dataset> dataset, const RendererConfig &cfg) override {
    IRenderer::init(std::move(dataset), cfg);
    compute_geometry();
    tf_ = {0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 1.f, 1.f};   // TransferFnConfig default ramp
    update(cfg.params);
    if (!tf_set_) apply_tf();
  }

  void update(const json::Value &params) override {
    IRenderer::update(params);
    if (params.is_null()) return;
    mode_ = enum_of(params, "mode", {{"Default", HYD_MODE_DEFAULT}, {"Volume", HYD_MODE_DEFAULT},
                                     {"Partition", HYD_MODE_PARTITION}, {"Paging", HYD_MODE_PAGING}}, mode_);
    density_ = num_or(params, "density", density_);
    mem_limit_mb_ = num_or(params, "mem_limit_mb", mem_limit_mb_);
    if (!(mem_limit_mb_ > 0.f)) throw Error(HYD_ERR_ARG, "mem_limit_mb must be > 0");
    if (const json::Value *tf = params.find("transfer_fn")) {
      std::vector<float> v = tf_values(*tf);
      if (!v.empty()) {
        tf_ = std::move(v);
        apply_tf();
      }
    }
  }

  void offline_render(const hyd_camera &camera, uint8_t *rgb_host) override {
    const hyd_view view = view_of(camera);
    const Dataset &ds = *dataset_;
    const int32_t fmt = pool_format_of(ds);
    const std::vector<int32_t> order = nearest_first(ds, empty_, view);
    const size_t per = pages_per_batch(ds, fmt, mem_limit_mb_);
    const size_t n_batches = std::max<size_t>(1, (order.size() + per - 1) / per);
    const size_t P = (size_t)width_ * height_;
    float *rgba = rgba_buffer(P);
    hyd_volume_pixel *film = nullptr;
    if (n_batches > 1) {
      dev_grow(&film_dev_, film_cap_, P * sizeof(hyd_volume_pixel));
      film = static_cast<hyd_volume_pixel *>(film_dev_);
    }
    hyd_volume_params p{};
    for (int i = 0; i < 3; ++i) {
      p.bbox_min[i] = 0.f;
      p.bbox_max[i] = geom_.f_dim[i];
    }
    p.step = geom_.step;
    p.du = geom_.du;
    p.max_steps = max_steps_;
    p.mode = mode_;
    p.rank = 0.f;
    p.density = density_;
    p.schedule = HYD_SCHED_REFERENCE;
    p.ert = 0.f;
    for (size_t b = 0; b < n_batches; ++b) {
      const size_t lo = b * per, hi = std::min(order.size(), lo + per);
      std::vector<int32_t> batch(order.begin() + (std::ptrdiff_t)lo, order.begin() + (std::ptrdiff_t)hi);
      make_resident(ctx_, ds, fmt, 1, batch);
      if (b == 0) check(hyd_render_volume(ctx_, &p, &view, width_, height_, film, rgba, nullptr));
      else check(hyd_resume_volume(ctx_, &p, width_, height_, film, rgba, nullptr));
    }
    check(hyd_pixel_pass(ctx_, rgba, width_, height_, clear_, rgb_buffer(P)));
    download_rgb(rgb_host);
  }

  void release() override {
    if (film_dev_) (void)hipFree(film_dev_);
    film_dev_ = nullptr;
    film_cap_ = 0;
    IRenderer::release();
  }

 private:
  // TF table and the brick emptiness it implies (TF alpha 0 over the padded
  // brick's [min, max]), and the chebyshev grid from it.
  void apply_tf() {
    const int32_t n = (int32_t)(tf_.size() / 4);
    check(hyd_set_tf(ctx_, n, tf_.data()));
    const Dataset &ds = *dataset_;
    empty_.assign(ds.n_bricks(), 0);
    if (!ds.cheb_thumb.empty()) {
      // a package's chebyshev thumbnail, as VolumeRenderer::init reads it (volume.cc:47-51)
      for (size_t i = 0; i < ds.n_bricks(); ++i) empty_[i] = ds.cheb_thumb[i] != 0;
      check(hyd_set_empty_grid(ctx_, ds.grid.data(), ds.cheb_thumb.data()));
      tf_set_ = true;
      return;
    }
    for (size_t i = 0; i < ds.n_bricks(); ++i) empty_[i] = (uint8_t)hyd_tf_alpha_zero_on(tf_.data(), n, ds.bmin[i], ds.bmax[i]);
    std::vector<int32_t> cheb(ds.n_bricks());
    check(hyd_chebyshev(ds.grid.data(), empty_.data(), cheb.data()));
    check(hyd_set_empty_grid(ctx_, ds.grid.data(), cheb.data()));
    tf_set_ = true;
  }

  std::vector<float> tf_;
  std::vector<uint8_t> empty_;
  bool tf_set_ = false;
  int32_t mode_ = HYD_MODE_DEFAULT;
  float density_ = 1.f;
  float mem_limit_mb_ = 2048.f;
  void *film_dev_ = nullptr;
  size_t film_cap_ = 0;
};

// ------------------------------------------------------ Blocks renderer ----
// BlocksRenderer (blocks.cc:30-72; BlocksRendererParams ui/blocks.schema.hpp:14-18):
// a grid-rate march over the mean and chebyshev thumbnails.  A brick is empty
// when its maximum is 0 (no data).
class BlocksRenderer : public IRenderer {
 public:
  void init(std::shared_ptr<Dataset> dataset, const RendererConfig &cfg) override {
    IRenderer::init(std::move(dataset), cfg);
    compute_geometry();
    const Dataset &ds = *dataset_;
    if (!ds.cheb_thumb.empty()) {
      // a package's chebyshev and mean thumbnails (blocks.cc:35-48)
      if (ds.mean_thumb.empty()) throw Error(HYD_ERR_STATE, "Blocks: the package has no \"mean\" thumbnail");
      check(hyd_set_empty_grid(ctx_, ds.grid.data(), ds.cheb_thumb.data()));
      check(hyd_set_mean_grid(ctx_, ds.grid.data(), ds.mean_thumb.data()));
    } else {
      std::vector<uint8_t> empty(ds.n_bricks());
      for (size_t i = 0; i < ds.n_bricks(); ++i) empty[i] = ds.bmax[i] <= 0.f;
      std::vector<int32_t> cheb(ds.n_bricks());
      check(hyd_chebyshev(ds.grid.data(), empty.data(), cheb.data()));
      check(hyd_set_empty_grid(ctx_, ds.grid.data(), cheb.data()));
      check(hyd_set_mean_grid(ctx_, ds.grid.data(), ds.bmean.data()));
    }
    update(cfg.params);
  }

  void update(const json::Value &params) override {
    IRenderer::update(params);
    if (params.is_null()) return;
    mode_ = enum_of(params, "mode", {{"Volume", HYD_BLOCKS_VOLUME}, {"Solid", HYD_BLOCKS_SOLID}}, mode_);
    density_ = num_or(params, "density", density_);
  }

  void offline_render(const hyd_camera &camera, uint8_t *rgb_host) override {
    const hyd_view view = view_of(camera);
    const size_t P = (size_t)width_ * height_;
    hyd_blocks_params p{};
    for (int i = 0; i < 3; ++i) {
      p.bbox_min[i] = 0.f;
      p.bbox_max[i] = geom_.f_dim[i];
    }
    p.step = geom_.step;
    p.max_steps = max_steps_;
    p.mode = mode_;
    p.density = density_;
    float *rgba = rgba_buffer(P);
    check(hyd_render_blocks(ctx_, &p, &view, width_, height_, rgba));
    check(hyd_pixel_pass(ctx_, rgba, width_, height_, clear_, rgb_buffer(P)));
    download_rgb(rgb_host);
  }

 private:
  int32_t mode_ = HYD_BLOCKS_VOLUME;
  float density_ = 1e-2f;
};

// -------------------------------------------------- Isosurface renderer ----
// IsosurfaceRenderer (isosurface.cc:95-140; IsosurfaceRendererParams
// ui/isosurface.schema.hpp:16-22).  A brick is empty when the iso value lies
// outside its [min, max]; the non-empty bricks must fit mem_limit_mb.
class IsosurfaceRenderer : public IRenderer {
 public:
  void init(std::shared_ptr<Dataset> dataset, const RendererConfig &cfg) override {
    IRenderer::init(std::move(dataset), cfg);
    compute_geometry();
    update(cfg.params);
    if (!resident_) refresh();
  }

  void update(const json::Value &params) override {
    IRenderer::update(params);
    if (params.is_null()) return;
    mode_ = enum_of(params, "mode", {{"Color", HYD_ISO_COLOR}, {"Normal", HYD_ISO_NORMAL},
                                     {"Position", HYD_ISO_POSITION}}, mode_);
    if (const json::Value *c = params.find("surface_color")) color_ = vec3(*c, "surface_color");
    const float iso = num_or(params, "isovalue", isovalue_);
    mem_limit_mb_ = num_or(params, "mem_limit_mb", mem_limit_mb_);
    corrected_ = (int32_t)num_or(params, "corrected", (float)corrected_);
    bisect_ = (int32_t)num_or(params, "bisect_steps", (float)bisect_);
    if (iso != isovalue_ || !resident_) {
      isovalue_ = iso;
      if (ctx_ && dataset_) refresh();
    }
  }

  void offline_render(const hyd_camera &camera, uint8_t *rgb_host) override {
    const hyd_view view = view_of(camera);
    const size_t P = (size_t)width_ * height_;
    hyd_iso_params p{};
    for (int i = 0; i < 3; ++i) {
      p.bbox_min[i] = 0.f;
      p.bbox_max[i] = geom_.f_dim[i];
    }
    p.step = geom_.step;
    p.du = geom_.du;
    p.max_steps = max_steps_;
    p.mode = mode_;
    float iet[16];
    check(hyd_exhibit_iet(geom_.center, geom_.f_dim, iet));
    check(hyd_mat4_inverse(iet, p.to_world));
    // light and eye (isosurface.cc:108-113): light = pos + target + up + cross(target, up)
    const float *pos = camera.position, *tg = camera.target, *up = camera.up;
    const float cr[3] = {tg[1] * up[2] - up[1] * tg[2], tg[2] * up[0] - up[2] * tg[0],
                         tg[0] * up[1] - up[0] * tg[1]};
    for (int i = 0; i < 3; ++i) {
      p.eye_pos[i] = pos[i];
      p.light_pos[i] = ((pos[i] + tg[i]) + up[i]) + cr[i];
      p.surface_color[i] = color_[i];
    }
    p.isovalue = isovalue_;
    p.corrected = corrected_;
    p.bisect_steps = bisect_;
    dev_grow(&iso_dev_, iso_cap_, P * sizeof(hyd_iso_pixel));
    hyd_iso_pixel *px = static_cast<hyd_iso_pixel *>(iso_dev_);
    check(hyd_render_iso(ctx_, &p, &view, width_, height_, px, nullptr));
    hip_check(launch_pixel_pass_strided(reinterpret_cast<const float *>(px), (int)(sizeof(hyd_iso_pixel) / 4),
                                        width_, height_, make_uchar3(clear_[0], clear_[1], clear_[2]),
                                        rgb_buffer(P), ctx_->stream),
              "isosurface pixel pass");
    download_rgb(rgb_host);
  }

  void release() override {
    if (iso_dev_) (void)hipFree(iso_dev_);
    iso_dev_ = nullptr;
    iso_cap_ = 0;
    IRenderer::release();
  }

 private:
  void refresh() {
    const Dataset &ds = *dataset_;
    std::vector<uint8_t> empty(ds.n_bricks());
    std::vector<int32_t> bricks;
    // a package's chebyshev thumbnail (isosurface.cc:50-53), else the iso value's
    // brick range
    for (size_t i = 0; i < ds.n_bricks(); ++i) {
      empty[i] = ds.cheb_thumb.empty() ? (ds.bmin[i] > isovalue_) || (ds.bmax[i] < isovalue_)
                                       : ds.cheb_thumb[i] != 0;
      if (!empty[i]) bricks.push_back((int32_t)i);
    }
    const int32_t fmt = pool_format_of(ds);
    if (bricks.size() > pages_per_batch(ds, fmt, mem_limit_mb_))
      throw Error(HYD_ERR_RANGE, "isosurface: " + std::to_string(bricks.size()) +
                                     " non-empty bricks exceed mem_limit_mb / 4096 pages");
    if (!ds.cheb_thumb.empty()) {
      check(hyd_set_empty_grid(ctx_, ds.grid.data(), ds.cheb_thumb.data()));
    } else {
      std::vector<int32_t> cheb(ds.n_bricks());
      check(hyd_chebyshev(ds.grid.data(), empty.data(), cheb.data()));
      check(hyd_set_empty_grid(ctx_, ds.grid.data(), cheb.data()));
    }
    make_resident(ctx_, ds, fmt, 1, bricks);
    resident_ = true;
  }

  int32_t mode_ = HYD_ISO_COLOR;
  std::array<float, 3> color_{1.f, 1.f, 1.f};
  float isovalue_ = 0.5f;
  float mem_limit_mb_ = 2048.f;
  int32_t corrected_ = 0;
  int32_t bisect_ = 8;
  bool resident_ = false;
  void *iso_dev_ = nullptr;
  size_t iso_cap_ = 0;
};

// ------------------------------------------------------ Paging renderer ----
// PagingRenderer (paging.cc:10-56; PagingRendererConfig paging.schema.hpp:9-13).
// Its offline path renders nothing: offline_render_ctxed returns the film as
// created (paging.cc:51-56), untouched pixels, which the pixel rule writes as
// the clear colour.  The residency view its realtime path draws is
// hyd_render_blocks with HYD_BLOCKS_PAGING over the caller's page table.
class PagingRenderer : public IRenderer {
 public:
  void init(std::shared_ptr<Dataset> dataset, const RendererConfig &cfg) override {
    IRenderer::init(std::move(dataset), cfg);
    compute_geometry();
    update(cfg.params);
  }

  void update(const json::Value &params) override {
    IRenderer::update(params);
    if (params.is_null()) return;
    mem_limit_mb_ = num_or(params, "mem_limit_mb", mem_limit_mb_);
    if (!(mem_limit_mb_ > 0.f)) throw Error(HYD_ERR_ARG, "mem_limit_mb must be > 0");
  }

  void offline_render(const hyd_camera &camera, uint8_t *rgb_host) override {
    (void)view_of(camera);   // the camera is checked as by the other renderers
    const size_t P = (size_t)width_ * height_;
    for (size_t i = 0; i < P; ++i) std::memcpy(rgb_host + 3 * i, clear_, 3);
  }

 private:
  float mem_limit_mb_ = 2048.f;   // used by the reference's realtime path only
};

HYD_REGISTER_RENDERER(VolumeRenderer, "Volume");
HYD_REGISTER_RENDERER(BlocksRenderer, "Blocks");
HYD_REGISTER_RENDERER(IsosurfaceRenderer, "Isosurface");
HYD_REGISTER_RENDERER(PagingRenderer, "Paging");

}  // namespace

// ------------------------------------------------------ dataset package ----
// The package this library reads (varch's own layout is not available here,
// so its field names are kept and its binary files are plain arrays):
//   package_meta.json  {"block_size": bs, "padding": p,
//                       "sample_levels": [{"raw": {"x","y","z"}, "dim": {"x","y","z"},
//                         "thumbnails": {"chebyshev": file, "mean": file},
//                         "voxels": {"path": file, "format": "u8"|"f16"|"f32"}}, ...]}
//   level-0 voxels     decoded voxels, x fastest (the reference decodes H.264
//                      brick archives instead, which this library leaves out)
//   thumbnails         dim.x * dim.y * dim.z little-endian int32 (chebyshev: 0 for a
//                      non-empty brick, else the L-inf brick distance to the nearest
//                      non-empty one) and float32 (mean), x fastest
// Relative file names resolve against the package directory.
namespace {
std::string resolve(const std::string &root, const std::string &name) {
  if (!name.empty() && name[0] == '/') return name;
  return root.empty() ? name : root + "/" + name;
}

// The file's bytes, after checking that it holds exactly `expect` of them; with
// load false only the check is made.
std::vector<uint8_t> read_file(const std::string &path, size_t expect, const char *what, bool load = true) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(HYD_ERR_ARG, std::string(what) + ": cannot open " + path);
  f.seekg(0, std::ios::end);
  const std::streamoff n = f.tellg();
  if (n < 0 || (size_t)n != expect)
    throw Error(HYD_ERR_ARG, std::string(what) + ": " + path + " holds " + std::to_string((long long)n) +
                                 " bytes, expected " + std::to_string(expect));
  if (!load) return {};
  std::vector<uint8_t> out(expect);
  f.seekg(0);
  if (expect && !f.read(reinterpret_cast<char *>(out.data()), (std::streamsize)expect))
    throw Error(HYD_ERR_ARG, std::string(what) + ": read failed for " + path);
  return out;
}

int32_t int_field(const json::Value &v, const char *key, const char *what) {
  const json::Value *f = v.find(key);
  if (!f) throw Error(HYD_ERR_ARG, std::string(what) + ": missing \"" + key + "\"");
  const double d = f->as_number(what);
  if (d != std::floor(d) || d < -2147483648.0 || d > 2147483647.0)
    throw Error(HYD_ERR_ARG, std::string(what) + "." + key + ": expected an integer");
  return (int32_t)d;
}

std::array<int32_t, 3> xyz_field(const json::Value &lvl, const char *key) {
  const json::Value *v = lvl.find(key);
  if (!v || !v->is_object()) throw Error(HYD_ERR_ARG, std::string("sample_levels[0].") + key + ": expected {x, y, z}");
  return {int_field(*v, "x", key), int_field(*v, "y", key), int_field(*v, "z", key)};
}
}  // namespace

std::shared_ptr<Dataset> read_package(const std::string &root, bool load_voxels) {
  const std::string meta_path = resolve(root, "package_meta.json");
  std::ifstream mf(meta_path);
  if (!mf) throw Error(HYD_ERR_ARG, "dataset package: cannot open " + meta_path);
  std::stringstream ss;
  ss << mf.rdbuf();
  const json::Value meta = json::parse(ss.str());
  auto d = std::make_shared<Dataset>();
  d->block_size = int_field(meta, "block_size", "package_meta");
  d->apron = int_field(meta, "padding", "package_meta");
  if (d->block_size < 1 || d->apron < 0 || d->apron > d->block_size || d->block_size + 2 * d->apron > 1024)
    throw Error(HYD_ERR_ARG, "package_meta: invalid block_size / padding");
  const json::Value *levels = meta.find("sample_levels");
  if (!levels || !levels->is_array() || levels->items("sample_levels").empty())
    throw Error(HYD_ERR_ARG, "package_meta: sample_levels must be a non-empty array");
  const json::Value &lvl = levels->items("sample_levels")[0];
  const std::array<int32_t, 3> raw = xyz_field(lvl, "raw"), dim = xyz_field(lvl, "dim");
  for (int i = 0; i < 3; ++i) {
    if (raw[i] < 1) throw Error(HYD_ERR_ARG, "sample_levels[0].raw must be positive");
    if (dim[i] != (raw[i] + d->block_size - 1) / d->block_size)
      throw Error(HYD_ERR_ARG, "sample_levels[0].dim does not match raw / block_size");
  }
  d->dims = raw;
  d->grid = dim;
  const json::Value *vox = lvl.find("voxels");
  if (!vox || !vox->is_object()) throw Error(HYD_ERR_ARG, "sample_levels[0].voxels: expected {path, format}");
  const json::Value *fmt = vox->find("format");
  const std::string fs = fmt ? fmt->as_string("voxels.format") : "";
  if (fs == "u8") d->fmt = HYD_FMT_U8;
  else if (fs == "f16") d->fmt = HYD_FMT_F16;
  else if (fs == "f32") d->fmt = HYD_FMT_F32;
  else throw Error(HYD_ERR_ARG, "sample_levels[0].voxels.format must be u8, f16 or f32");
  const json::Value *vpath = vox->find("path");
  if (!vpath) throw Error(HYD_ERR_ARG, "sample_levels[0].voxels: missing \"path\"");
  const size_t nvox = (size_t)raw[0] * raw[1] * raw[2];
  const size_t esz = d->fmt == HYD_FMT_U8 ? 1 : (d->fmt == HYD_FMT_F16 ? 2 : 4);
  d->voxels = read_file(resolve(root, vpath->as_string("voxels.path")), nvox * esz, "voxels", load_voxels);
  const size_t nb = d->n_bricks();
  const json::Value *th = lvl.find("thumbnails");
  if (!th || !th->is_object()) throw Error(HYD_ERR_ARG, "sample_levels[0].thumbnails: expected an object");
  const json::Value *cheb = th->find("chebyshev");
  if (!cheb) throw Error(HYD_ERR_ARG, "sample_levels[0].thumbnails: missing \"chebyshev\"");
  const std::vector<uint8_t> cb = read_file(resolve(root, cheb->as_string("thumbnails.chebyshev")), nb * 4,
                                            "chebyshev thumbnail");
  d->cheb_thumb.resize(nb);
  std::memcpy(d->cheb_thumb.data(), cb.data(), nb * 4);
  for (int32_t c : d->cheb_thumb)
    if (c < 0) throw Error(HYD_ERR_ARG, "chebyshev thumbnail holds a negative distance");
  if (const json::Value *mean = th->find("mean")) {
    const std::vector<uint8_t> mb = read_file(resolve(root, mean->as_string("thumbnails.mean")), nb * 4,
                                              "mean thumbnail");
    d->mean_thumb.resize(nb);
    std::memcpy(d->mean_thumb.data(), mb.data(), nb * 4);
  }
  return d;
}
}  // namespace front
}  // namespace hyd

// ------------------------------------------------------------- C ABI ----
struct hyd_renderer {
  hyd_ctx *ctx = nullptr;
  std::unique_ptr<hyd::front::IRenderer> impl;
};

namespace {

template <class F> int32_t guarded(F &&f) {
  try {
    f();
    return HYD_OK;
  } catch (const hyd::front::Error &e) {
    hyd::set_error("%s", e.what());
    return e.code;
  } catch (const hyd::json::Error &e) {
    hyd::set_error("%s", e.what());
    return HYD_ERR_ARG;
  } catch (const std::bad_alloc &) {
    hyd::set_error("host allocation failed");
    return HYD_ERR_RANGE;
  } catch (const std::exception &e) {
    hyd::set_error("%s", e.what());
    return HYD_ERR_ARG;
  }
}

}  // namespace

extern "C" {

int32_t hyd_renderer_candidates(char *buf, int32_t cap, int32_t *len) {
  std::string all;
  for (const auto &n : hyd::front::RendererFactory::list_candidates()) all += (all.empty() ? "" : ",") + n;
  if (len) *len = (int32_t)all.size();
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(all.size(), (size_t)cap - 1);
    std::memcpy(buf, all.data(), n);
    buf[n] = '\0';
  }
  return HYD_OK;
}

int32_t hyd_renderer_create(hyd_ctx *ctx, const hyd_dataset *ds, const char *config_json, hyd_renderer **out) {
  HYD_REQUIRE(ctx && ds && config_json && out, HYD_ERR_ARG, "null argument");
  *out = nullptr;
  HYD_REQUIRE(ds->voxels, HYD_ERR_ARG, "dataset has no voxels");
  HYD_REQUIRE(ds->dims[0] > 0 && ds->dims[1] > 0 && ds->dims[2] > 0, HYD_ERR_ARG, "dataset dims must be > 0");
  HYD_REQUIRE(ds->fmt == HYD_FMT_U8 || ds->fmt == HYD_FMT_F16 || ds->fmt == HYD_FMT_F32, HYD_ERR_ARG,
              "unknown dataset format %d", ds->fmt);
  HYD_REQUIRE(ds->block_size >= 1 && ds->apron >= 0 && ds->apron <= ds->block_size, HYD_ERR_ARG,
              "invalid block size / apron");
  std::unique_ptr<hyd_renderer> r(new hyd_renderer);
  r->ctx = ctx;
  const int32_t st = guarded([&] {
    auto d = std::make_shared<hyd::front::Dataset>();
    d->dims = {ds->dims[0], ds->dims[1], ds->dims[2]};
    d->fmt = ds->fmt;
    d->block_size = ds->block_size;
    d->apron = ds->apron;
    const size_t esz = ds->fmt == HYD_FMT_U8 ? 1 : (ds->fmt == HYD_FMT_F16 ? 2 : 4);
    const size_t n = (size_t)ds->dims[0] * ds->dims[1] * ds->dims[2];
    d->voxels.assign(static_cast<const uint8_t *>(ds->voxels), static_cast<const uint8_t *>(ds->voxels) + n * esz);
    for (int i = 0; i < 3; ++i) d->grid[i] = (ds->dims[i] + ds->block_size - 1) / ds->block_size;
    d->bmin.resize(d->n_bricks());
    d->bmax.resize(d->n_bricks());
    d->bmean.resize(d->n_bricks());
    hyd::front::check(hyd_brick_stats(d->voxels.data(), d->fmt, d->dims.data(), d->block_size, d->apron,
                                      d->bmin.data(), d->bmax.data(), d->bmean.data()));
    const hyd::json::Value cfg = hyd::json::parse(config_json);
    r->impl = hyd::front::RendererFactory(d).create(ctx, hyd::front::parse_renderer_config(cfg));
  });
  if (st != HYD_OK) return st;
  *out = r.release();
  return HYD_OK;
}

int32_t hyd_package_read_info(const char *dataset_path, hyd_package_info *out) {
  HYD_REQUIRE(dataset_path && out, HYD_ERR_ARG, "null argument");
  return guarded([&] {
    const auto d = hyd::front::read_package(dataset_path, false);
    hyd_package_info info{};
    for (int i = 0; i < 3; ++i) {
      info.raw[i] = d->dims[i];
      info.dim[i] = d->grid[i];
    }
    info.block_size = d->block_size;
    info.padding = d->apron;
    info.fmt = d->fmt;
    info.has_mean = d->mean_thumb.empty() ? 0 : 1;
    *out = info;
  });
}

int32_t hyd_package_read_thumbnails(const char *dataset_path, int32_t *chebyshev, float *mean, int64_t n) {
  HYD_REQUIRE(dataset_path && chebyshev && n >= 0, HYD_ERR_ARG, "null argument");
  return guarded([&] {
    const auto d = hyd::front::read_package(dataset_path, false);
    if ((int64_t)d->n_bricks() != n)
      throw hyd::front::Error(HYD_ERR_ARG, "thumbnail buffers hold " + std::to_string((long long)n) +
                                               " entries, the package has " + std::to_string(d->n_bricks()));
    std::memcpy(chebyshev, d->cheb_thumb.data(), d->n_bricks() * 4);
    if (mean) {
      if (d->mean_thumb.empty()) throw hyd::front::Error(HYD_ERR_ARG, "the package has no \"mean\" thumbnail");
      std::memcpy(mean, d->mean_thumb.data(), d->n_bricks() * 4);
    }
  });
}

int32_t hyd_renderer_create_package(hyd_ctx *ctx, const char *dataset_path, const char *config_json,
                                    hyd_renderer **out) {
  HYD_REQUIRE(ctx && dataset_path && config_json && out, HYD_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<hyd_renderer> r(new hyd_renderer);
  r->ctx = ctx;
  const int32_t st = guarded([&] {
    const auto d = hyd::front::read_package(dataset_path, true);
    d->bmin.resize(d->n_bricks());
    d->bmax.resize(d->n_bricks());
    d->bmean.resize(d->n_bricks());
    hyd::front::check(hyd_brick_stats(d->voxels.data(), d->fmt, d->dims.data(), d->block_size, d->apron,
                                      d->bmin.data(), d->bmax.data(), d->bmean.data()));
    const hyd::json::Value cfg = hyd::json::parse(config_json);
    r->impl = hyd::front::RendererFactory(d).create(ctx, hyd::front::parse_renderer_config(cfg));
  });
  if (st != HYD_OK) return st;
  *out = r.release();
  return HYD_OK;
}

int32_t hyd_renderer_update(hyd_renderer *r, const char *params_json) {
  HYD_REQUIRE(r && r->impl && params_json, HYD_ERR_ARG, "null argument");
  return guarded([&] {
    const hyd::json::Value v = hyd::json::parse(params_json);
    const hyd::json::Value *p = v.find("params");
    r->impl->update(p ? *p : v);
  });
}

int32_t hyd_renderer_resolution(const hyd_renderer *r, int32_t *width, int32_t *height) {
  HYD_REQUIRE(r && r->impl && width && height, HYD_ERR_ARG, "null argument");
  *width = r->impl->width();
  *height = r->impl->height();
  return HYD_OK;
}

int32_t hyd_renderer_offline_render(hyd_renderer *r, const char *camera_json, uint8_t *rgb_host) {
  HYD_REQUIRE(r && r->impl && camera_json && rgb_host, HYD_ERR_ARG, "null argument");
  return guarded([&] {
    const hyd_camera cam = hyd::front::parse_camera(hyd::json::parse(camera_json));
    r->impl->offline_render(cam, rgb_host);
  });
}

int32_t hyd_renderer_destroy(hyd_renderer *r) {
  if (!r) return HYD_OK;
  if (r->impl) r->impl->release();
  delete r;
  return HYD_OK;
}

int32_t hyd_renderer_parse_config(const char *config_json, char *name, int32_t cap, int32_t *width,
                                  int32_t *height) {
  HYD_REQUIRE(config_json && width && height, HYD_ERR_ARG, "null argument");
  return guarded([&] {
    const hyd::front::RendererConfig c = hyd::front::parse_renderer_config(hyd::json::parse(config_json));
    if (!hyd::front::RendererRegistry::types().count(c.renderer))
      throw hyd::front::Error(HYD_ERR_ARG, "unknown renderer '" + c.renderer + "'");
    *width = c.width;
    *height = c.height;
    if (name && cap > 0) {
      const size_t n = std::min<size_t>(c.renderer.size(), (size_t)cap - 1);
      std::memcpy(name, c.renderer.data(), n);
      name[n] = '\0';
    }
  });
}

int32_t hyd_renderer_parse_camera(const char *camera_json, hyd_camera *out) {
  HYD_REQUIRE(camera_json && out, HYD_ERR_ARG, "null argument");
  return guarded([&] { *out = hyd::front::parse_camera(hyd::json::parse(camera_json)); });
}

}  // extern "C"
