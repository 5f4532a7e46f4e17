// JSON-configured renderers over the C ABI: the counterpart of the reference's
// IRenderer / RendererFactory / REGISTER_RENDERER (include/hydrant/core/renderer.hpp:30-89)
// and BasicRenderer (include/hydrant/basic_renderer.hpp:37-84).
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hydrant_hip.h"
#include "../../include/hydrant_renderer.h"
#include "json.hpp"

namespace hyd {
namespace front {

struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Raw voxels plus the brick layout (the reference's Dataset is a varch package,
// renderer.hpp:17-21).
struct Dataset {
  std::array<int32_t, 3> dims{};
  int32_t fmt = HYD_FMT_U8;
  std::vector<uint8_t> voxels;   // x fastest
  int32_t block_size = 64;
  int32_t apron = 1;
  // per-brick statistics of normalised values (hyd_brick_stats), grid order
  std::array<int32_t, 3> grid{};
  std::vector<float> bmin, bmax, bmean;
  // thumbnails of a dataset package (sample_levels[0].thumbnails, grid order);
  // empty for a dataset given as raw voxels, whose grids are derived instead
  std::vector<int32_t> cheb_thumb;
  std::vector<float> mean_thumb;
  size_t n_bricks() const { return (size_t)grid[0] * grid[1] * grid[2]; }
};

// A dataset package (RendererFactory(dataset_path), renderer.hpp:53-59): reads
// package_meta.json under `root` and the files it names.  Throws Error.
std::shared_ptr<Dataset> read_package(const std::string &root, bool load_voxels);

// RendererConfig (core/renderer.schema.hpp:8-13).
struct RendererConfig {
  int32_t width = 512, height = 512;
  std::string renderer;
  json::Value params = json::Value::object();
};

// CameraConfig (core/scene.schema.hpp:10-35) resolved to a hyd_camera.
hyd_camera parse_camera(const json::Value &cfg);
RendererConfig parse_renderer_config(const json::Value &cfg);

struct RGBImage {
  int32_t width = 0, height = 0;
  std::vector<uint8_t> rgb;   // width * height * 3
};

class IRenderer {
 public:
  virtual ~IRenderer() = default;
  virtual void init(std::shared_ptr<Dataset> dataset, const RendererConfig &cfg);
  virtual void update(const json::Value &params);
  virtual void offline_render(const hyd_camera &camera, uint8_t *rgb_host) = 0;
  int32_t width() const { return width_; }
  int32_t height() const { return height_; }
  void attach(hyd_ctx *ctx) { ctx_ = ctx; }

 protected:
  // BasicRenderer geometry (basic_renderer.hpp:57-84), binary32.
  struct Geometry {
    float f_dim[3];    // raw / block size: the bbox [0, f_dim]
    float center[3];   // exhibit centre
    float du;          // 1 / block size
    float step;        // du / 4 / sample_rate
  };
  void compute_geometry();
  hyd_view view_of(const hyd_camera &camera) const;
  // device buffers grown on demand
  float *rgba_buffer(size_t pixels);
  uint8_t *rgb_buffer(size_t pixels);
  void download_rgb(uint8_t *rgb_host);

  hyd_ctx *ctx_ = nullptr;
  std::shared_ptr<Dataset> dataset_;
  int32_t width_ = 512, height_ = 512;
  float sample_rate_ = 1.f;
  int32_t max_steps_ = 4000000;
  uint8_t clear_[3] = {0, 0, 0};
  Geometry geom_{};
  void *rgba_dev_ = nullptr;
  size_t rgba_cap_ = 0;
  void *rgb_dev_ = nullptr;
  size_t rgb_cap_ = 0;

 public:
  virtual void release();
};

class RendererRegistry {
 public:
  using Make = std::function<IRenderer *()>;
  static std::map<std::string, Make> &types();
  static int add(const char *name, Make make) {
    types()[name] = std::move(make);
    return 0;
  }
};

// REGISTER_RENDERER (renderer.hpp:70-89): a static registration by name.
#define HYD_REGISTER_RENDERER_CAT2(a, b) a##b
#define HYD_REGISTER_RENDERER_CAT(a, b) HYD_REGISTER_RENDERER_CAT2(a, b)
#define HYD_REGISTER_RENDERER(T, name)                                                       \
  static const int HYD_REGISTER_RENDERER_CAT(hyd_renderer_registrar_, __COUNTER__) =         \
      ::hyd::front::RendererRegistry::add(name, []() -> ::hyd::front::IRenderer * { return new T; })

class RendererFactory {
 public:
  explicit RendererFactory(std::shared_ptr<Dataset> dataset) : dataset_(std::move(dataset)) {}
  std::unique_ptr<IRenderer> create(hyd_ctx *ctx, const RendererConfig &cfg);
  static std::vector<std::string> list_candidates();

 private:
  std::shared_ptr<Dataset> dataset_;
};

}  // namespace front
}  // namespace hyd
