// C ABI of the batched renderer: multi-volume set, batch plans, HIP-graph
// capture/replay and the launch-shape autotuner.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "hyd_batch.hpp"

namespace hyd {
namespace {

int32_t dev_upload(void **dst, const void *src, size_t bytes, hipStream_t s) {
  if (*dst) (void)hipFree(*dst);
  *dst = nullptr;
  HYD_CHECK_HIP(hipMalloc(dst, bytes));
  HYD_CHECK_HIP(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s));
  HYD_CHECK_HIP(hipStreamSynchronize(s));
  return HYD_OK;
}

int group_of(const hyd_batch_config *c) { return c->group == 4 ? 4 : 1; }

bool valid_config(const hyd_batch_config *c) {
  return c && (c->tile_w == 8 || c->tile_w == 16 || c->tile_w == 32) &&
         (c->unroll == 1 || c->unroll == 2 || c->unroll == 4) && c->waves_per_cu >= 1 && c->waves_per_cu <= 32 &&
         (c->group == 0 || c->group == 1 || (c->group == 4 && c->tile_w == 8));
}

// tile plans: tile_w 8 / 16 / 32 (TW x 8 tiles) and 16 x 16 blocks of four 8x8 tiles
constexpr int kPlans = 4;
int plan_slot(int tw, int group) { return group == 4 ? 3 : (tw == 8 ? 0 : tw == 16 ? 1 : 2); }

}  // namespace
}  // namespace hyd

using namespace hyd;

struct hyd_batch {
  std::vector<hyd_batch_view> views;
  BatchViewDev *d_views = nullptr;
  uint32_t *d_tiles[kPlans] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<int32_t> chunks[kPlans];
  unsigned *d_counters = nullptr;
  int64_t total_px = 0;
  int32_t max_volume = 0;   // largest volume index the views name
  // captured graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t captured_gen = 0;   // the context's batch_state_gen the graph's pointers belong to
  hipStream_t capture_stream = nullptr;
};

namespace {

// Interleaves the bits of x and y (< 256 each): x in the even bits.
uint32_t morton2(uint32_t x, uint32_t y) {
  uint32_t k = 0;
  for (int i = 0; i < 8; ++i) k |= ((x >> i) & 1u) << (2 * i) | ((y >> i) & 1u) << (2 * i + 1);
  return k;
}

int32_t ensure_tiles(hyd_batch *b, int tw, int group, hipStream_t s) {
  const int slot = plan_slot(tw, group);
  if (b->d_tiles[slot]) return HYD_OK;
  const int bw = group == 4 ? 16 : tw, bh = group == 4 ? 16 : 8;
  std::vector<uint32_t> tiles;
  for (size_t v = 0; v < b->views.size(); ++v) {
    const int nx = (b->views[v].width + bw - 1) / bw, ny = (b->views[v].height + bh - 1) / bh;
    HYD_REQUIRE(nx <= 256 && ny <= 256, HYD_ERR_RANGE, "view %zu: %d x %d tiles exceeds 256 per axis", v, nx, ny);
    // Z order (bits of tx and ty interleaved): the tiles that run together on
    // one XCD form a compact block of the view, whose rays cross fewer bricks
    // than a strip of whole rows does, so more of their taps hit the L2
    std::vector<std::pair<uint32_t, uint32_t>> keyed;
    keyed.reserve((size_t)nx * ny);
    for (int ty = 0; ty < ny; ++ty)
      for (int tx = 0; tx < nx; ++tx) keyed.emplace_back(morton2(tx, ty), (uint32_t)ty << 8 | (uint32_t)tx);
    std::sort(keyed.begin(), keyed.end());
    for (const auto &k : keyed) tiles.push_back((uint32_t)v << 16 | k.second);
  }
  std::vector<int32_t> &ch = b->chunks[slot];
  ch.resize(kXcds + 1);
  for (int i = 0; i <= kXcds; ++i) ch[i] = (int32_t)((int64_t)tiles.size() * i / kXcds);
  void *p = nullptr;
  int32_t st = dev_upload(&p, tiles.data(), tiles.size() * sizeof(uint32_t), s);
  if (st) return st;
  b->d_tiles[slot] = static_cast<uint32_t *>(p);
  return HYD_OK;
}

size_t layout_bytes(const Ctx *c) {
  const size_t elem = c->pool.fmt == HYD_FMT_U8 ? 1 : c->pool.fmt == HYD_FMT_F16 ? 2 : 4;
  return (size_t)c->pool.pad * c->pool.pad * c->pool.pad * (size_t)c->pool.n_pages * elem;
}

// Holds at least `bytes` of oriented-layout memory; false (and nothing held)
// when the device cannot give it.
bool reserve_oriented(Ctx *c, size_t bytes) {
  if (c->oriented_cap >= bytes) return true;
  if (c->own_oriented) (void)hipFree(c->own_oriented);
  c->own_oriented = nullptr;
  c->oriented_cap = 0;
  if (hipMalloc(&c->own_oriented, bytes) != hipSuccess) {
    (void)hipGetLastError();   // clear the allocation error: the fallbacks need less or no memory
    c->own_oriented = nullptr;
    return false;
  }
  c->oriented_cap = bytes;
  return true;
}

bool env_is_1(const char *name) {
  const char *v = std::getenv(name);
  return v && std::strcmp(v, "1") == 0;
}

// The pool's oriented layouts, made once per pool upload / bind (the stream is
// synchronised so that graphs captured on another stream see them).  FP16
// and u8 pools get the three quad layouts (12x the pool's bytes), or, when
// the device cannot hold them, the three row-pair layouts (6x, while their
// byte offsets fit 32 bits); otherwise, or when the device cannot hold those,
// the y- and z-fastest layouts (2x).  When the device cannot hold those
// either, or when the environment sets HYD_BATCH_SINGLE_LAYOUT=1, the
// renderer runs on the x-fastest layout alone.  HYD_BATCH_NO_QUADS=1 skips the
// quad layouts, HYD_BATCH_NO_PAIRS=1 the quads and the row pairs.  Every
// choice renders the same images.
int32_t ensure_oriented(Ctx *c) {
  if (c->oriented_gen == c->pool_gen) return HYD_OK;
  const size_t bytes = layout_bytes(c);
  // until the copies below are complete the context holds none (a failed
  // launch leaves the generation stale, so the next launch tries again)
  c->oriented_pairs = 0;
  int pairs = 0;
  if (env_is_1("HYD_BATCH_SINGLE_LAYOUT")) {
    if (c->own_oriented) (void)hipFree(c->own_oriented);
    c->own_oriented = nullptr;
    c->oriented_cap = 0;
    c->oriented_gen = c->pool_gen;
    return HYD_OK;
  }
  const bool pairs_fit = (c->pool.fmt == HYD_FMT_F16 || c->pool.fmt == HYD_FMT_U8) && 2 * bytes <= (size_t)UINT32_MAX;
  // the quads' offsets are 64-bit in the kernel
  const bool quads = (c->pool.fmt == HYD_FMT_F16 || c->pool.fmt == HYD_FMT_U8) && !env_is_1("HYD_BATCH_NO_QUADS");
  if (quads && !env_is_1("HYD_BATCH_NO_PAIRS") && reserve_oriented(c, 12 * bytes)) {
    HYD_CHECK_HIP(launch_orient_quads(c->pool.fmt, c->pool.data, c->own_oriented, c->pool.pad, c->pool.n_pages, 0,
                                      c->pool.n_pages, c->stream));
    pairs = 2;
  } else if (pairs_fit && !env_is_1("HYD_BATCH_NO_PAIRS") && reserve_oriented(c, 6 * bytes)) {
    HYD_CHECK_HIP(launch_orient_pairs(c->pool.fmt, c->pool.data, c->own_oriented, c->pool.pad, c->pool.n_pages, 0,
                                      c->pool.n_pages, c->stream));
    pairs = 1;
  } else if (reserve_oriented(c, 2 * bytes)) {
    HYD_CHECK_HIP(launch_orient_pool(c->pool.fmt, c->pool.data, c->own_oriented, c->pool.pad, c->pool.n_pages, 0,
                                     c->pool.n_pages, c->stream));
  } else {
    c->oriented_gen = c->pool_gen;   // no memory for copies: the pool alone
    return HYD_OK;
  }
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  c->oriented_pairs = pairs;
  c->oriented_gen = c->pool_gen;
  return HYD_OK;
}

}  // namespace

// Pages [first, first + n) of the pool were overwritten in place: bring the
// current oriented copies up to date for those pages (nothing to do when the
// copies are stale anyway, or when the renderer reads the pool alone).
int32_t hyd::refresh_oriented_pages(Ctx *c, int32_t first, int32_t n) {
  if (c->oriented_gen != c->pool_gen || !c->own_oriented || n <= 0) return HYD_OK;
  if (c->oriented_pairs == 2) {
    HYD_CHECK_HIP(launch_orient_quads(c->pool.fmt, c->pool.data, c->own_oriented, c->pool.pad, c->pool.n_pages,
                                      first, n, c->stream));
  } else if (c->oriented_pairs) {
    HYD_CHECK_HIP(launch_orient_pairs(c->pool.fmt, c->pool.data, c->own_oriented, c->pool.pad, c->pool.n_pages,
                                      first, n, c->stream));
  } else {
    HYD_CHECK_HIP(launch_orient_pool(c->pool.fmt, c->pool.data, c->own_oriented, c->pool.pad, c->pool.n_pages,
                                     first, n, c->stream));
  }
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  return HYD_OK;
}

namespace {

int32_t fill_args(Ctx *c, hyd_batch *b, const hyd_batch_config *cfg, float density, float ert, float *rgba_dev,
                  unsigned long long *samples_dev, BatchArgs &a, int &grid) {
  HYD_REQUIRE(valid_config(cfg), HYD_ERR_ARG,
              "invalid batch config (tile_w 8/16/32, unroll 1/2/4, waves_per_cu 1..32, group 1 or 4 with tile_w 8)");
  HYD_REQUIRE(rgba_dev, HYD_ERR_ARG, "null output");
  HYD_REQUIRE(c->own_cells && c->n_vols > 0, HYD_ERR_STATE, "no volume set: call hyd_set_volumes first");
  HYD_REQUIRE(c->vol_max_page < c->pool.n_pages, HYD_ERR_STATE,
              "the volume set names page %d but the pool holds %d pages", c->vol_max_page, c->pool.n_pages);
  HYD_REQUIRE(b->max_volume < c->n_vols, HYD_ERR_STATE,
              "the batch's views name volume %d but the volume set holds %d volumes", b->max_volume, c->n_vols);
  HYD_REQUIRE(c->pool.data && c->tf, HYD_ERR_STATE, "batched render needs a pool and a transfer function");
  HYD_REQUIRE(c->pool.pad == c->vol_pad, HYD_ERR_STATE, "pool pad edge %d differs from the volume set's %d",
              c->pool.pad, c->vol_pad);
  // the kernel addresses voxels by 32-bit byte offsets from the pool base and
  // forms a tap's index in a page in float (exact up to 2^24 = 256^3)
  HYD_REQUIRE(layout_bytes(c) <= (size_t)UINT32_MAX && c->pool.pad <= 256, HYD_ERR_RANGE,
              "batched render needs a pool of at most 4 GiB with a padded brick edge of at most 256 "
              "(%zu bytes, edge %d)", layout_bytes(c), c->pool.pad);
  const int group = group_of(cfg);
  int32_t st = ensure_tiles(b, cfg->tile_w, group, c->stream);
  if (st) return st;
  if ((st = ensure_oriented(c))) return st;
  const int slot = plan_slot(cfg->tile_w, group);
  std::memset(&a, 0, sizeof(a));
  a.tiles = b->d_tiles[slot];
  for (int i = 0; i <= kXcds; ++i) a.chunk_start[i] = b->chunks[slot][i];
  a.counters = b->d_counters;
  a.views = b->d_views;
  a.vols = static_cast<const BatchVolDev *>(c->own_vols);
  a.cells = c->own_cells;
  a.oriented = c->own_oriented != nullptr;
  a.pairs = c->oriented_pairs;
  if (a.pairs) {
    const char *o = static_cast<const char *>(c->own_oriented);
    for (int p = 0; p < 3; ++p) a.pools[p] = o + p * 2 * a.pairs * layout_bytes(c);
  } else {
    a.pools[0] = c->pool.data;
    a.pools[1] = a.oriented ? c->own_oriented : c->pool.data;
    a.pools[2] = a.oriented ? static_cast<const char *>(c->own_oriented) + layout_bytes(c) : c->pool.data;
  }
  a.pad = c->pool.pad;
  a.page_elems = (uint32_t)c->pool.pad * c->pool.pad * c->pool.pad;
  a.tf = c->tf;
  a.tf_n = c->tf_n;
  a.density = density;
  a.ert = ert;
  a.max_steps = 4000000;
  a.out = reinterpret_cast<float4 *>(rgba_dev);
  a.samples = samples_dev;
  grid = (c->n_cus * cfg->waves_per_cu + group - 1) / group;
  return HYD_OK;
}

}  // namespace

extern "C" {

int32_t hyd_set_volumes(hyd_ctx *c, int32_t n, const hyd_volume_desc *vols) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(n >= 1 && vols, HYD_ERR_ARG, "empty volume set");
  std::vector<BatchVolDev> dv(n);
  std::vector<int32_t> cells;
  int pad = -1, max_page = -1;
  for (int v = 0; v < n; ++v) {
    const hyd_volume_desc &d = vols[v];
    HYD_REQUIRE(d.grid[0] > 0 && d.grid[1] > 0 && d.grid[2] > 0 && d.vaddr && d.chebyshev, HYD_ERR_ARG,
                "volume %d: invalid grid", v);
    HYD_REQUIRE(d.bs > 0 && d.apron >= 1, HYD_ERR_ARG, "volume %d: block size > 0 and apron >= 1 required", v);
    HYD_REQUIRE(d.step > 0.f, HYD_ERR_ARG, "volume %d: step must be > 0", v);
    const int p = d.bs + 2 * d.apron;
    HYD_REQUIRE(pad < 0 || pad == p, HYD_ERR_ARG, "volume %d: every volume must share one padded brick edge", v);
    pad = p;
    BatchVolDev &o = dv[v];
    std::memcpy(o.grid, d.grid, sizeof(o.grid));
    o.cell_off = (int32_t)cells.size();
    o.bs = d.bs;
    o.apron = d.apron;
    std::memcpy(o.bmax, d.bbox_max, sizeof(o.bmax));
    o.step = d.step;
    HYD_REQUIRE(d.value_scale > 0.f, HYD_ERR_ARG, "volume %d: value_scale must be > 0", v);
    o.value_scale = d.value_scale;
    const size_t nc = (size_t)d.grid[0] * d.grid[1] * d.grid[2];
    for (size_t i = 0; i < nc; ++i) {
      const int32_t cd = d.chebyshev[i], pg = d.vaddr[i];
      HYD_REQUIRE(cd >= 0 && pg >= -1, HYD_ERR_RANGE, "volume %d cell %zu: invalid chebyshev/vaddr", v, i);
      if (cd == 0 && c->pool.data)
        HYD_REQUIRE(pg < c->pool.n_pages, HYD_ERR_RANGE, "volume %d cell %zu: page %d beyond the pool", v, i, pg);
      if (cd == 0) max_page = std::max(max_page, pg);
      cells.push_back(cd > 0 ? -1 - cd : pg);
    }
  }
  if (c->own_vols) (void)hipFree(c->own_vols);
  if (c->own_cells) (void)hipFree(c->own_cells);
  c->own_vols = nullptr;
  c->own_cells = nullptr;
  c->n_vols = 0;
  ++c->batch_state_gen;
  void *pv = nullptr;
  if ((st = dev_upload(&pv, dv.data(), dv.size() * sizeof(BatchVolDev), c->stream))) return st;
  c->own_vols = pv;
  void *pcells = nullptr;
  if ((st = dev_upload(&pcells, cells.data(), cells.size() * sizeof(int32_t), c->stream))) return st;
  c->own_cells = static_cast<int32_t *>(pcells);
  c->n_vols = n;
  c->vol_pad = pad;
  c->vol_max_page = max_page;
  return HYD_OK;
}

int32_t hyd_batch_create(hyd_ctx *c, const hyd_batch_view *views, int32_t n_views, hyd_batch **out) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(out && views && n_views >= 1 && n_views <= 65535, HYD_ERR_ARG, "1..65535 views required");
  HYD_REQUIRE(c->n_vols > 0, HYD_ERR_STATE, "no volume set: call hyd_set_volumes first");
  hyd_batch *b = new hyd_batch();
  b->views.assign(views, views + n_views);
  std::vector<BatchViewDev> dv(n_views);
  for (int i = 0; i < n_views; ++i) {
    const hyd_batch_view &v = views[i];
    if (!(v.width > 0 && v.height > 0 && v.volume >= 0 && v.volume < c->n_vols && v.out_offset >= 0)) {
      delete b;
      set_error("view %d: invalid size, volume or output offset", i);
      return HYD_ERR_ARG;
    }
    std::memcpy(dv[i].trans, v.view.trans, sizeof(dv[i].trans));
    dv[i].ctg = v.view.ctg_fovy_2;
    std::memcpy(dv[i].ray_o, v.view.ray_o, sizeof(dv[i].ray_o));
    dv[i].W = v.width;
    dv[i].H = v.height;
    dv[i].volume = v.volume;
    dv[i].out_offset = v.out_offset;
    b->total_px = std::max<int64_t>(b->total_px, v.out_offset + (int64_t)v.width * v.height);
    b->max_volume = std::max(b->max_volume, v.volume);
  }
  void *p = nullptr;
  if ((st = dev_upload(&p, dv.data(), dv.size() * sizeof(BatchViewDev), c->stream))) { delete b; return st; }
  b->d_views = static_cast<BatchViewDev *>(p);
  if (hipMalloc((void **)&b->d_counters, kXcds * sizeof(unsigned)) != hipSuccess) {
    (void)hipFree(b->d_views);
    delete b;
    set_error("hipMalloc of the work counters failed");
    return HYD_ERR_HIP;
  }
  *out = b;
  return HYD_OK;
}

int32_t hyd_batch_destroy(hyd_batch *b) {
  if (!b) return HYD_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  (void)hipFree(b->d_views);
  for (auto *t : b->d_tiles) if (t) (void)hipFree(t);
  (void)hipFree(b->d_counters);
  delete b;
  return HYD_OK;
}

int32_t hyd_batch_output_pixels(const hyd_batch *b, int64_t *n_out) {
  HYD_REQUIRE(b && n_out, HYD_ERR_ARG, "null argument");
  *n_out = b->total_px;
  return HYD_OK;
}

int32_t hyd_batch_render(hyd_ctx *c, hyd_batch *b, const hyd_batch_config *cfg, float density, float ert,
                         float *rgba_dev, unsigned long long *samples_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(b, HYD_ERR_ARG, "null batch");
  BatchArgs a;
  int grid = 0;
  if ((st = fill_args(c, b, cfg, density, ert, rgba_dev, samples_dev, a, grid))) return st;
  HYD_CHECK_HIP(launch_batch(a, c->pool.fmt, cfg->tile_w, cfg->unroll, group_of(cfg), grid, c->stream));
  return HYD_OK;
}

int32_t hyd_batch_capture(hyd_ctx *c, hyd_batch *b, const hyd_batch_config *cfg, float density, float ert,
                          float *rgba_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(b, HYD_ERR_ARG, "null batch");
  BatchArgs a;
  int grid = 0;
  if ((st = fill_args(c, b, cfg, density, ert, rgba_dev, nullptr, a, grid))) return st;
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  if (!b->capture_stream) HYD_CHECK_HIP(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  HYD_CHECK_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_batch(a, c->pool.fmt, cfg->tile_w, cfg->unroll, group_of(cfg), grid, b->capture_stream);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->capture_stream, &g);
  HYD_CHECK_HIP(e);
  HYD_CHECK_HIP(e2);
  b->graph = g;
  HYD_CHECK_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->captured_gen = c->batch_state_gen;
  return HYD_OK;
}

int32_t hyd_batch_replay(hyd_ctx *c, hyd_batch *b) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(b && b->exec, HYD_ERR_STATE, "no captured graph: call hyd_batch_capture first");
  // the graph holds device pointers of the state it was captured with
  HYD_REQUIRE(b->captured_gen == c->batch_state_gen, HYD_ERR_STATE,
              "pool, transfer function or volume set changed since the capture: call hyd_batch_capture again");
  HYD_CHECK_HIP(hipGraphLaunch(b->exec, c->stream));
  return HYD_OK;
}

int32_t hyd_batch_autotune_schedule(hyd_ctx *c, hyd_batch *b, float density, float max_error, float *rgba_dev,
                                    int32_t reps, hyd_batch_config *best, float *best_ert, float *best_ms) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(b && best && best_ert && reps >= 1, HYD_ERR_ARG, "invalid autotune arguments");
  HYD_REQUIRE(max_error > 0.f, HYD_ERR_ARG, "max_error must be > 0");
  // termination thresholds whose truncation (at most 1 - ert of any channel)
  // takes at most half the error budget; 0 (no termination) is always admissible
  std::vector<float> erts;
  for (float ert : {0.99f, 0.995f, 0.999f, 0.9995f, 0.9999f})
    if (1.f - ert <= 0.5f * max_error) erts.push_back(ert);
  erts.push_back(0.f);
  // coordinate search: launch shapes at the earliest admissible termination,
  // then every admissible threshold (that one again too) at that shape, timed
  // back to back so that the thresholds are compared under the same conditions
  hyd_batch_config cfg;
  float t_shape = 0.f;
  if ((st = hyd_batch_autotune(c, b, density, erts[0], rgba_dev, reps, &cfg, &t_shape))) return st;
  float t_best = 1e30f, e_best = erts[0];
  hipEvent_t e0, e1;
  HYD_CHECK_HIP(hipEventCreate(&e0));
  HYD_CHECK_HIP(hipEventCreate(&e1));
  for (size_t i = 0; i < erts.size() && st == HYD_OK; ++i) {
    BatchArgs a;
    int grid = 0;
    if ((st = fill_args(c, b, &cfg, density, erts[i], rgba_dev, nullptr, a, grid))) break;
    const int g = group_of(&cfg);
    if (launch_batch(a, c->pool.fmt, cfg.tile_w, cfg.unroll, g, grid, c->stream) != hipSuccess) {
      st = HYD_ERR_HIP;
      set_error("autotune launch failed");
      break;
    }
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) (void)launch_batch(a, c->pool.fmt, cfg.tile_w, cfg.unroll, g, grid, c->stream);
    (void)hipEventRecord(e1, c->stream);
    if (hipEventSynchronize(e1) != hipSuccess) { st = HYD_ERR_HIP; set_error("autotune launch failed"); break; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)reps;
    if (ms < t_best) { t_best = ms; e_best = erts[i]; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  *best = cfg;
  *best_ert = e_best;
  if (best_ms) *best_ms = t_best;
  return HYD_OK;
}

int32_t hyd_batch_autotune(hyd_ctx *c, hyd_batch *b, float density, float ert, float *rgba_dev, int32_t reps,
                           hyd_batch_config *best, float *best_ms) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(b && best && reps >= 1, HYD_ERR_ARG, "invalid autotune arguments");
  hipEvent_t e0, e1;
  HYD_CHECK_HIP(hipEventCreate(&e0));
  HYD_CHECK_HIP(hipEventCreate(&e1));
  float best_t = 1e30f;
  hyd_batch_config best_c{8, 4, 8, 1};
  // the search space: measured optima sit at 8x8 tiles, unroll 4 and 2-4 waves
  // per SIMD; neighbouring shapes and the 4-wave shared-tile form are timed too
  std::vector<hyd_batch_config> space;
  for (int tw : {8, 16, 32})
    for (int un : {2, 4})
      for (int wpc : {8, 12, 16}) space.push_back(hyd_batch_config{tw, un, wpc, 1});
  for (int wpc : {6, 10, 14}) space.push_back(hyd_batch_config{8, 4, wpc, 1});
  for (int un : {2, 4})
    for (int wpc : {8, 12, 16}) space.push_back(hyd_batch_config{8, un, wpc, 4});
  for (const hyd_batch_config &cfg : space) {
    BatchArgs a;
    int grid = 0;
    if ((st = fill_args(c, b, &cfg, density, ert, rgba_dev, nullptr, a, grid))) goto done;
    const int g = group_of(&cfg);
    // one untimed launch (first-use costs), then reps timed launches
    if (launch_batch(a, c->pool.fmt, cfg.tile_w, cfg.unroll, g, grid, c->stream) != hipSuccess) {
      st = HYD_ERR_HIP;
      goto done;
    }
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) (void)launch_batch(a, c->pool.fmt, cfg.tile_w, cfg.unroll, g, grid, c->stream);
    (void)hipEventRecord(e1, c->stream);
    if (hipEventSynchronize(e1) != hipSuccess) { st = HYD_ERR_HIP; set_error("autotune launch failed"); goto done; }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)reps;
    if (ms < best_t) { best_t = ms; best_c = cfg; }
  }
done:
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  *best = best_c;
  if (best_ms) *best_ms = best_t;
  return HYD_OK;
}

}  // extern "C"
