// C ABI implementation: context, resident state and pass dispatch.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "hyd_internal.hpp"

namespace hyd {

namespace {
thread_local char g_err[1024] = "";
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

namespace {

constexpr int kMaxTf = 256;

int fmt_bytes(int fmt) { return fmt == HYD_FMT_U8 ? 1 : fmt == HYD_FMT_F16 ? 2 : fmt == HYD_FMT_F32 ? 4 : 0; }

template <class T> void free_dev(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

template <class T> int32_t upload(T *&dst, const void *src, size_t bytes, hipStream_t s) {
  free_dev(dst);
  HYD_CHECK_HIP(hipMalloc((void **)&dst, bytes));
  HYD_CHECK_HIP(hipMemcpyAsync((void *)dst, src, bytes, hipMemcpyHostToDevice, s));
  HYD_CHECK_HIP(hipStreamSynchronize(s));
  return HYD_OK;
}

// upload() into the existing allocation when it holds `bytes` (the copy is
// ordered after the kernels already queued on the stream that read it).
template <class T> int32_t upload_reuse(T *&dst, size_t &cap, const void *src, size_t bytes, hipStream_t s) {
  if (!dst || cap < bytes) {
    cap = 0;
    int32_t st = upload(dst, src, bytes, s);
    if (st == HYD_OK) cap = bytes;
    return st;
  }
  HYD_CHECK_HIP(hipMemcpyAsync((void *)dst, src, bytes, hipMemcpyHostToDevice, s));
  HYD_CHECK_HIP(hipStreamSynchronize(s));
  return HYD_OK;
}

// One code per brick from the host copies of the page table and chebyshev grid
// (when both are set with equal dims): the marches then read one entry per
// brick instead of two dependent ones.
int32_t rebuild_codes(Ctx *c) {
  c->codes_valid = false;
  if (c->host_vaddr.empty() || c->host_cheb.empty() ||
      std::memcmp(c->paging.vdim, c->cheb.dim, sizeof(c->cheb.dim)) != 0)
    return HYD_OK;
  const size_t n = c->host_vaddr.size();
  std::vector<int32_t> codes(n);
  for (size_t i = 0; i < n; ++i) {
    const int32_t cd = c->host_cheb[i];
    codes[i] = cd > 0 ? -1 - cd : c->host_vaddr[i];
  }
  int32_t st = upload_reuse(c->own_codes, c->codes_cap, codes.data(), n * sizeof(int32_t), c->stream);
  if (st) return st;
  c->codes_valid = true;
  return HYD_OK;
}

int32_t check_volume_state(const Ctx *c) {
  HYD_REQUIRE(c->pool.data, HYD_ERR_STATE, "no brick pool: call hyd_upload_pool first");
  HYD_REQUIRE(c->paging.vaddr, HYD_ERR_STATE, "no page table: call hyd_set_page_table first");
  HYD_REQUIRE(c->cheb.data, HYD_ERR_STATE, "no empty-space grid: call hyd_set_empty_grid first");
  HYD_REQUIRE(c->tf, HYD_ERR_STATE, "no transfer function: call hyd_set_tf first");
  HYD_REQUIRE(c->paging.n_pages <= c->pool.n_pages, HYD_ERR_STATE,
              "page table names %d pages but the pool holds %d", c->paging.n_pages, c->pool.n_pages);
  return HYD_OK;
}

int32_t fill_volume_args(const Ctx *c, const hyd_volume_params *p, VolumeArgs &a) {
  HYD_REQUIRE(p, HYD_ERR_ARG, "null params");
  HYD_REQUIRE(p->step > 0.f, HYD_ERR_ARG, "step must be > 0");
  HYD_REQUIRE(p->schedule == HYD_SCHED_REFERENCE || p->schedule == HYD_SCHED_LATTICE, HYD_ERR_ARG,
              "unknown schedule %d", p->schedule);
  HYD_REQUIRE(p->mode >= HYD_MODE_DEFAULT && p->mode <= HYD_MODE_PAGING, HYD_ERR_ARG, "unknown mode %d", p->mode);
  std::memset(&a, 0, sizeof(a));
  a.bmin = f3{p->bbox_min[0], p->bbox_min[1], p->bbox_min[2]};
  a.bmax = f3{p->bbox_max[0], p->bbox_max[1], p->bbox_max[2]};
  a.step = p->step;
  a.max_steps = p->max_steps;
  a.mode = p->mode;
  a.rank = p->rank;
  a.density = p->density;
  a.schedule = p->schedule;
  a.ert = p->ert;
  a.tf = c->tf;
  a.tf_n = c->tf_n;
  a.cheb = c->cheb.data;
  std::memcpy(a.cdim, c->cheb.dim, sizeof(a.cdim));
  a.vaddr = c->paging.vaddr;
  std::memcpy(a.vdim, c->paging.vdim, sizeof(a.vdim));
  a.codes = c->codes_valid ? c->own_codes : nullptr;
  a.lowest_blkcnt = c->paging.lowest_blkcnt;
  a.page = c->paging.page;
  a.pool = c->pool.data;
  a.pool_elems = (uint64_t)c->pool.n_pages * c->pool.pad * c->pool.pad * c->pool.pad;
  a.pad = c->pool.pad;
  a.u8_lut = c->u8_lut;
  return HYD_OK;
}

void fill_view(const hyd_view *v, float trans[16], float &ctg, f3 &ray_o) {
  std::memcpy(trans, v->trans, 16 * sizeof(float));
  ctg = v->ctg_fovy_2;
  ray_o = f3{v->ray_o[0], v->ray_o[1], v->ray_o[2]};
}

}  // namespace

namespace {
// Device staging of the z order for hyd_composite_noalloc: a file-scope device
// array (each device holds its own copy for the life of the process) whose
// address is resolved once per device at context creation.  In host code the
// array's name is not a device address; only the resolved pointer is used.
constexpr int32_t kNoAllocMaxPartials = 4096;
static __device__ __align__(256) int32_t scratch[kNoAllocMaxPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[kMaxDevices] = {};

// Resolves this device's scratch address (the device must be current).
int32_t resolve_scratch(int device) {
  HYD_REQUIRE(device >= 0 && device < kMaxDevices, HYD_ERR_RANGE, "device %d beyond the %d the library tracks",
              device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return HYD_OK;
  void *ptr = nullptr;
  HYD_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  HYD_REQUIRE(ptr, HYD_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
  g_scratch_ptr[device] = static_cast<int32_t *>(ptr);
  return HYD_OK;
}

int32_t *scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// linear_to_srgb (volume.cc:175-181) then saturate (glm_math.hpp:75-81), binary32
// on the host as the reference evaluates it.
uint8_t srgb8_host(float x) {
  const float s = x <= 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.f / 2.4f) - 0.055f;
  return (uint8_t)(int)gclamp(s * 255.f, 0.f, 255.f);
}
}  // namespace

void srgb8_thresholds(float t[256]) {
  // the byte is non-decreasing in x: bisect the bit patterns of [0, +inf] for
  // the first float of each byte value (x = 1 gives 254: 1.055f - 0.055f rounds
  // below 1)
  t[0] = -INFINITY;
  for (int k = 1; k < 256; ++k) {
    uint32_t lo = 0, hi = 0x7f800000u;   // byte(lo) < k <= byte(hi)
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      float x;
      std::memcpy(&x, &mid, 4);
      if (srgb8_host(x) >= k) hi = mid; else lo = mid;
    }
    std::memcpy(&t[k], &hi, 4);
  }
}

int32_t activate(Ctx *c) {
  HYD_REQUIRE(c, HYD_ERR_ARG, "null context");
  HYD_CHECK_HIP(hipSetDevice(c->device));
  return HYD_OK;
}

}  // namespace hyd

using namespace hyd;

extern "C" {

const char *hyd_last_error(void) { return g_err; }

int32_t hyd_version(void) { return 1; }

int32_t hyd_create(int32_t device, hyd_ctx **out) {
  HYD_REQUIRE(out, HYD_ERR_ARG, "hyd_create: null out");
  *out = nullptr;
  int n = 0;
  HYD_CHECK_HIP(hipGetDeviceCount(&n));
  HYD_REQUIRE(device >= 0 && device < n, HYD_ERR_ARG, "device %d out of range (%d devices)", device, n);
  HYD_CHECK_HIP(hipSetDevice(device));
  hyd_ctx *c = new hyd_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cus = prop.multiProcessorCount;
  if (c->n_cus <= 0) c->n_cus = 256;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    set_error("hipStreamCreateWithFlags failed");
    return HYD_ERR_HIP;
  }
  c->stream = c->own_stream;
  // float(v) / 255.f for every u8 value: the normalisation Texture3D<u8>
  // applies on upload (texture_3d.hpp:78-86, glm_math.hpp:83-89).
  // Then the sRGB byte thresholds of the composite.
  float lut[512];
  for (int i = 0; i < 256; ++i) lut[i] = float(i) / 255.f;
  srgb8_thresholds(lut + 256);
  int32_t st = upload(c->own_lut, lut, sizeof(lut), c->stream);
  if (st != HYD_OK) { (void)hipStreamDestroy(c->own_stream); delete c; return st; }
  c->u8_lut = c->own_lut;
  c->srgb8_t = c->own_lut + 256;
  if ((st = resolve_scratch(device)) != HYD_OK) {
    (void)hipFree(c->own_lut);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return st;
  }
  *out = c;
  return HYD_OK;
}

int32_t hyd_destroy(hyd_ctx *c) {
  if (!c) return HYD_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_dev(c->own_pool);
  free_dev(c->own_oriented);
  free_dev(c->own_vaddr);
  free_dev(c->own_page);
  free_dev(c->own_codes);
  free_dev(c->own_cheb);
  free_dev(c->own_mean);
  free_dev(c->own_tf);
  free_dev(c->own_lut);
  free_dev(c->own_vols);
  free_dev(c->own_cells);
  free_dev(c->own_diff_views);
  free_dev(c->own_diff_scratch);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return HYD_OK;
}

int32_t hyd_set_stream(hyd_ctx *c, void *stream) {
  HYD_REQUIRE(c, HYD_ERR_ARG, "null context");
  c->stream = static_cast<hipStream_t>(stream);  // NULL = the device's null stream
  return HYD_OK;
}

int32_t hyd_synchronize(hyd_ctx *c) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  return HYD_OK;
}

int32_t hyd_upload_pool(hyd_ctx *c, int32_t fmt, int32_t pad_edge, int32_t n_pages, const void *host) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(fmt_bytes(fmt) > 0, HYD_ERR_ARG, "unknown brick format %d", fmt);
  HYD_REQUIRE(pad_edge >= 2 && pad_edge <= 1024, HYD_ERR_ARG, "pad_edge %d out of range [2, 1024]", pad_edge);
  HYD_REQUIRE(n_pages > 0 && host, HYD_ERR_ARG, "empty pool");
  const size_t bytes = (size_t)n_pages * pad_edge * pad_edge * pad_edge * fmt_bytes(fmt);
  st = upload(c->own_pool, host, bytes, c->stream);
  if (st) return st;
  c->pool = Pool{c->own_pool, fmt, pad_edge, n_pages, true};
  ++c->pool_gen;
  ++c->batch_state_gen;
  return HYD_OK;
}

int32_t hyd_bind_pool_device(hyd_ctx *c, int32_t fmt, int32_t pad_edge, int32_t n_pages, const void *pool_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(fmt_bytes(fmt) > 0, HYD_ERR_ARG, "unknown brick format %d", fmt);
  HYD_REQUIRE(pad_edge >= 2 && pad_edge <= 1024 && n_pages > 0 && pool_dev, HYD_ERR_ARG, "invalid pool");
  free_dev(c->own_pool);
  c->pool = Pool{pool_dev, fmt, pad_edge, n_pages, false};
  ++c->pool_gen;
  ++c->batch_state_gen;
  return HYD_OK;
}

int32_t hyd_write_pool_pages(hyd_ctx *c, int32_t first_page, int32_t n_pages, const void *host) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(c->pool.data, HYD_ERR_STATE, "no brick pool: call hyd_upload_pool or hyd_bind_pool_device first");
  HYD_REQUIRE(host && n_pages >= 0 && first_page >= 0 && (int64_t)first_page + n_pages <= c->pool.n_pages,
              HYD_ERR_RANGE, "pages [%d, %d) outside the pool of %d pages", first_page, first_page + n_pages,
              c->pool.n_pages);
  const size_t page = (size_t)c->pool.pad * c->pool.pad * c->pool.pad * fmt_bytes(c->pool.fmt);
  char *dst = static_cast<char *>(const_cast<void *>(c->pool.data)) + (size_t)first_page * page;
  HYD_CHECK_HIP(hipMemcpyAsync(dst, host, (size_t)n_pages * page, hipMemcpyHostToDevice, c->stream));
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  // the batched renderer's oriented copies of these pages follow (in place, so
  // captured graphs stay valid)
  return refresh_oriented_pages(c, first_page, n_pages);
}

int32_t hyd_set_page_table(hyd_ctx *c, const int32_t dims[3], const int32_t *vaddr, int32_t n_pages,
                           const float *page_k, const float *page_b, int32_t lowest_blkcnt) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(dims && vaddr && page_k && page_b && n_pages > 0, HYD_ERR_ARG, "invalid page table");
  HYD_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, HYD_ERR_ARG, "page table dims must be > 0");
  const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
  for (size_t i = 0; i < cells; ++i)
    HYD_REQUIRE(vaddr[i] >= -1 && vaddr[i] < n_pages, HYD_ERR_RANGE, "vaddr[%zu] = %d outside [-1, %d)", i,
                vaddr[i], n_pages);
  std::vector<float4> pages(n_pages);
  for (int i = 0; i < n_pages; ++i)
    pages[i] = make_float4(page_k[i], page_b[3 * i], page_b[3 * i + 1], page_b[3 * i + 2]);
  st = upload_reuse(c->own_vaddr, c->vaddr_cap, vaddr, cells * sizeof(int32_t), c->stream);
  if (st) return st;
  st = upload_reuse(c->own_page, c->page_cap, pages.data(), pages.size() * sizeof(float4), c->stream);
  if (st) return st;
  c->paging.vaddr = c->own_vaddr;
  std::memcpy(c->paging.vdim, dims, 3 * sizeof(int32_t));
  c->paging.page = c->own_page;
  c->paging.n_pages = n_pages;
  c->paging.lowest_blkcnt = lowest_blkcnt;
  c->host_vaddr.assign(vaddr, vaddr + cells);
  return rebuild_codes(c);
}

int32_t hyd_set_empty_grid(hyd_ctx *c, const int32_t dims[3], const int32_t *chebyshev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(dims && chebyshev && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, HYD_ERR_ARG, "invalid empty grid");
  const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
  for (size_t i = 0; i < cells; ++i)
    HYD_REQUIRE(chebyshev[i] >= 0, HYD_ERR_RANGE, "chebyshev[%zu] = %d is negative", i, chebyshev[i]);
  st = upload(c->own_cheb, chebyshev, cells * sizeof(int32_t), c->stream);
  if (st) return st;
  c->cheb.data = c->own_cheb;
  std::memcpy(c->cheb.dim, dims, 3 * sizeof(int32_t));
  c->host_cheb.assign(chebyshev, chebyshev + cells);
  return rebuild_codes(c);
}

int32_t hyd_set_mean_grid(hyd_ctx *c, const int32_t dims[3], const float *mean) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(dims && mean && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, HYD_ERR_ARG, "invalid mean grid");
  const size_t cells = (size_t)dims[0] * dims[1] * dims[2];
  st = upload(c->own_mean, mean, cells * sizeof(float), c->stream);
  if (st) return st;
  c->mean.data = c->own_mean;
  std::memcpy(c->mean.dim, dims, 3 * sizeof(int32_t));
  return HYD_OK;
}

int32_t hyd_set_tf(hyd_ctx *c, int32_t n, const float *rgba) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(rgba && n >= 1, HYD_ERR_ARG, "invalid transfer function");
  HYD_REQUIRE(n <= kMaxTf, HYD_ERR_RANGE, "transfer function has %d entries; the limit is %d", n, kMaxTf);
  st = upload(c->own_tf, rgba, (size_t)n * 4 * sizeof(float), c->stream);
  if (st) return st;
  c->tf = c->own_tf;
  c->tf_n = n;
  ++c->batch_state_gen;
  return HYD_OK;
}

int32_t hyd_render_volume(hyd_ctx *c, const hyd_volume_params *p, const hyd_view *view, int32_t width,
                          int32_t height, hyd_volume_pixel *film_dev, float *rgba_dev,
                          unsigned long long *samples_dev) {
  int32_t st = activate(c);
  if (st) return st;
  if ((st = check_volume_state(c))) return st;
  HYD_REQUIRE(view && width > 0 && height > 0, HYD_ERR_ARG, "invalid view or image size");
  VolumeArgs a;
  if ((st = fill_volume_args(c, p, a))) return st;
  fill_view(view, a.trans, a.ctg, a.ray_o);
  a.W = width;
  a.H = height;
  a.film = film_dev;
  a.rgba = reinterpret_cast<float4 *>(rgba_dev);
  a.samples = samples_dev;
  HYD_CHECK_HIP(launch_volume(a, c->pool.fmt, false, c->stream));
  return HYD_OK;
}

int32_t hyd_resume_volume(hyd_ctx *c, const hyd_volume_params *p, int32_t width, int32_t height,
                          hyd_volume_pixel *film_dev, float *rgba_dev, unsigned long long *samples_dev) {
  int32_t st = activate(c);
  if (st) return st;
  if ((st = check_volume_state(c))) return st;
  HYD_REQUIRE(film_dev && width > 0 && height > 0, HYD_ERR_ARG, "resume needs a film");
  VolumeArgs a;
  if ((st = fill_volume_args(c, p, a))) return st;
  a.W = width;
  a.H = height;
  a.film = film_dev;
  a.rgba = reinterpret_cast<float4 *>(rgba_dev);
  a.samples = samples_dev;
  HYD_CHECK_HIP(launch_volume(a, c->pool.fmt, true, c->stream));
  return HYD_OK;
}

int32_t hyd_pixel_pass(hyd_ctx *c, const float *rgba_dev, int32_t width, int32_t height,
                       const uint8_t clear_color[3], uint8_t *rgb_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(rgba_dev && rgb_dev && width > 0 && height > 0, HYD_ERR_ARG, "invalid pixel pass arguments");
  uchar3 clear = make_uchar3(0, 0, 0);
  if (clear_color) clear = make_uchar3(clear_color[0], clear_color[1], clear_color[2]);
  HYD_CHECK_HIP(launch_pixel_pass(reinterpret_cast<const float4 *>(rgba_dev), width, height, clear, rgb_dev,
                                  c->stream));
  return HYD_OK;
}

int32_t hyd_fetch_pass(hyd_ctx *c, const hyd_volume_pixel *film_dev, int32_t width, int32_t height,
                       hyd_fetch_pixel *out_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(film_dev && out_dev && width > 0 && height > 0, HYD_ERR_ARG, "invalid fetch pass arguments");
  HYD_CHECK_HIP(launch_fetch_pass(film_dev, width, height, out_dev, c->stream));
  return HYD_OK;
}

int32_t hyd_composite(hyd_ctx *c, int32_t n, const hyd_fetch_pixel *partials_dev, const int32_t *z_order,
                      int32_t width, int32_t height, float *rgba_dev, uint8_t *rgb_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(n >= 1 && partials_dev && z_order && width > 0 && height > 0, HYD_ERR_ARG,
              "invalid composite arguments");
  std::vector<char> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    HYD_REQUIRE(z_order[i] >= 0 && z_order[i] < n && !seen[z_order[i]], HYD_ERR_ARG,
                "z_order must be a permutation of 0..%d", n - 1);
    seen[z_order[i]] = 1;
  }
  int32_t *zdev = nullptr;
  HYD_CHECK_HIP(hipMallocAsync((void **)&zdev, n * sizeof(int32_t), c->stream));
  HYD_CHECK_HIP(hipMemcpyAsync(zdev, z_order, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  hipError_t e = launch_composite(n, partials_dev, zdev, width, height, reinterpret_cast<float4 *>(rgba_dev),
                                  rgb_dev, c->srgb8_t, c->stream);
  (void)hipFreeAsync(zdev, c->stream);
  HYD_CHECK_HIP(e);
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  return HYD_OK;
}

int32_t hyd_composite_noalloc(hyd_ctx *c, int32_t n, const hyd_fetch_pixel *partials_dev, const int32_t *z_order,
                              int32_t width, int32_t height, float *rgba_dev, uint8_t *rgb_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(n >= 1 && partials_dev && z_order && width > 0 && height > 0, HYD_ERR_ARG,
              "invalid composite arguments");
  HYD_REQUIRE(n <= kNoAllocMaxPartials, HYD_ERR_RANGE,
              "hyd_composite_noalloc composites at most %d partial images (got %d)", kNoAllocMaxPartials, n);
  // a permutation check on the host, in a fixed bitset (no allocation)
  uint64_t seen[kNoAllocMaxPartials / 64] = {};
  for (int i = 0; i < n; ++i) {
    const int32_t z = z_order[i];
    HYD_REQUIRE(z >= 0 && z < n && !(seen[z >> 6] >> (z & 63) & 1), HYD_ERR_ARG,
                "z_order must be a permutation of 0..%d", n - 1);
    seen[z >> 6] |= 1ull << (z & 63);
  }
  int32_t *zdev = scratch_ptr(c->device);
  HYD_REQUIRE(zdev, HYD_ERR_STATE, "no scratch address resolved for device %d", c->device);
  HYD_CHECK_HIP(hipMemcpyAsync(zdev, z_order, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HYD_CHECK_HIP(launch_composite(n, partials_dev, zdev, width, height, reinterpret_cast<float4 *>(rgba_dev), rgb_dev,
                                 c->srgb8_t, c->stream));
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  return HYD_OK;
}

int32_t hyd_srgb8_thresholds(float out[256]) {
  HYD_REQUIRE(out, HYD_ERR_ARG, "null output");
  srgb8_thresholds(out);
  return HYD_OK;
}

int32_t hyd_iso_fetch_pass(hyd_ctx *c, const hyd_iso_pixel *px_dev, int32_t width, int32_t height,
                           hyd_iso_fetch_pixel *out_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(px_dev && out_dev && width > 0 && height > 0, HYD_ERR_ARG, "invalid isosurface fetch arguments");
  HYD_CHECK_HIP(launch_iso_fetch(px_dev, width, height, out_dev, c->stream));
  return HYD_OK;
}

int32_t hyd_composite_iso(hyd_ctx *c, int32_t n, const hyd_iso_fetch_pixel *partials_dev, const int32_t *z_order,
                          int32_t width, int32_t height, uint8_t *rgb_dev, float *depth_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(n >= 1 && partials_dev && z_order && width > 0 && height > 0 && (rgb_dev || depth_dev), HYD_ERR_ARG,
              "invalid isosurface composite arguments");
  std::vector<char> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    HYD_REQUIRE(z_order[i] >= 0 && z_order[i] < n && !seen[z_order[i]], HYD_ERR_ARG,
                "z_order must be a permutation of 0..%d", n - 1);
    seen[z_order[i]] = 1;
  }
  int32_t *zdev = nullptr;
  HYD_CHECK_HIP(hipMallocAsync((void **)&zdev, n * sizeof(int32_t), c->stream));
  HYD_CHECK_HIP(hipMemcpyAsync(zdev, z_order, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  hipError_t e = launch_iso_composite(n, partials_dev, zdev, width, height, rgb_dev, depth_dev, c->stream);
  (void)hipFreeAsync(zdev, c->stream);
  HYD_CHECK_HIP(e);
  HYD_CHECK_HIP(hipStreamSynchronize(c->stream));
  return HYD_OK;
}

int32_t hyd_segments_begin(hyd_ctx *c, int32_t width, int32_t height, int32_t max_slots, float *slots_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(slots_dev && width > 0 && height > 0 && max_slots > 0, HYD_ERR_ARG, "invalid segment buffers");
  HYD_CHECK_HIP(launch_segments_clear(reinterpret_cast<float4 *>(slots_dev), (size_t)width * height * max_slots,
                                      c->stream));
  return HYD_OK;
}

int32_t hyd_segments_pass(hyd_ctx *c, const hyd_volume_params *p, const hyd_view *view, int32_t width,
                          int32_t height, int32_t max_slots, float *slots_dev, int32_t *nslots_dev,
                          unsigned long long *samples_dev) {
  int32_t st = activate(c);
  if (st) return st;
  if ((st = check_volume_state(c))) return st;
  HYD_REQUIRE(view && slots_dev && nslots_dev && width > 0 && height > 0, HYD_ERR_ARG, "invalid segment pass");
  const int32_t need = c->paging.vdim[0] + c->paging.vdim[1] + c->paging.vdim[2];
  HYD_REQUIRE(max_slots >= need, HYD_ERR_RANGE, "max_slots %d < %d (sum of the brick grid dims)", max_slots, need);
  VolumeArgs a;
  if ((st = fill_volume_args(c, p, a))) return st;
  HYD_REQUIRE(a.schedule == HYD_SCHED_LATTICE, HYD_ERR_ARG, "segments need the lattice schedule");
  fill_view(view, a.trans, a.ctg, a.ray_o);
  a.W = width;
  a.H = height;
  a.slots = reinterpret_cast<float4 *>(slots_dev);
  a.nslots = nslots_dev;
  a.max_slots = max_slots;
  a.samples = samples_dev;
  HYD_CHECK_HIP(launch_segments_pass(a, c->pool.fmt, c->stream));
  return HYD_OK;
}

}  // extern "C"

struct hyd_segment_set {
  int device = 0;
  int32_t n = 0, fmt = 0, pad = 0, max_pages = 0;
  int32_t vdim[3] = {0, 0, 0};
  uint64_t max_pool_elems = 0;
  const void *pool[hyd::kMaxSegBatches] = {};
  int32_t *codes = nullptr;   // [n][cells]
  float4 *pages = nullptr;    // [n][max_pages]
};

extern "C" {

int32_t hyd_segment_set_destroy(hyd_segment_set *set) {
  if (!set) return HYD_OK;
  if (hipSetDevice(set->device) == hipSuccess) {
    if (set->codes) (void)hipFree(set->codes);
    if (set->pages) (void)hipFree(set->pages);
  }
  delete set;
  return HYD_OK;
}

int32_t hyd_segment_set_create(hyd_ctx *c, int32_t n_batches, const hyd_segment_batch *b, hyd_segment_set **out) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(b && out, HYD_ERR_ARG, "null argument");
  *out = nullptr;
  HYD_REQUIRE(n_batches >= 1 && n_batches <= kMaxSegBatches, HYD_ERR_RANGE, "n_batches %d outside [1, %d]",
              n_batches, kMaxSegBatches);
  HYD_REQUIRE(!c->host_cheb.empty(), HYD_ERR_STATE, "no empty-space grid: call hyd_set_empty_grid first");
  const size_t cells = c->host_cheb.size();
  std::unique_ptr<hyd_segment_set> set(new hyd_segment_set);
  set->device = c->device;
  set->n = n_batches;
  set->fmt = b[0].fmt;
  set->pad = b[0].pad_edge;
  std::memcpy(set->vdim, c->cheb.dim, sizeof(set->vdim));
  for (int i = 0; i < n_batches; ++i) {
    HYD_REQUIRE(fmt_bytes(b[i].fmt) > 0 && b[i].fmt == set->fmt, HYD_ERR_ARG,
                "batch %d: format %d (every batch needs the first batch's known format)", i, b[i].fmt);
    HYD_REQUIRE(b[i].pad_edge == set->pad && set->pad >= 2 && set->pad <= 1024, HYD_ERR_ARG,
                "batch %d: page edge %d (every batch needs the first batch's, in [2, 1024])", i, b[i].pad_edge);
    HYD_REQUIRE(b[i].n_pages > 0 && b[i].pool_dev && b[i].vaddr && b[i].page_k && b[i].page_b, HYD_ERR_ARG,
                "batch %d: invalid pool or page table", i);
    set->max_pages = std::max(set->max_pages, b[i].n_pages);
    set->max_pool_elems =
        std::max(set->max_pool_elems, (uint64_t)b[i].n_pages * set->pad * set->pad * set->pad);
    set->pool[i] = b[i].pool_dev;
  }
  std::vector<int32_t> codes((size_t)n_batches * cells);
  std::vector<float4> pages((size_t)n_batches * set->max_pages, make_float4(0.f, 0.f, 0.f, 0.f));
  for (int i = 0; i < n_batches; ++i) {
    for (size_t j = 0; j < cells; ++j) {
      const int32_t v = b[i].vaddr[j];
      HYD_REQUIRE(v >= -1 && v < b[i].n_pages, HYD_ERR_RANGE, "batch %d: vaddr[%zu] = %d outside [-1, %d)", i, j,
                  v, b[i].n_pages);
      const int32_t cd = c->host_cheb[j];
      codes[(size_t)i * cells + j] = cd > 0 ? -1 - cd : v;
    }
    for (int k = 0; k < b[i].n_pages; ++k)
      pages[(size_t)i * set->max_pages + k] =
          make_float4(b[i].page_k[k], b[i].page_b[3 * k], b[i].page_b[3 * k + 1], b[i].page_b[3 * k + 2]);
  }
  HYD_CHECK_HIP(hipMalloc((void **)&set->codes, codes.size() * sizeof(int32_t)));
  HYD_CHECK_HIP(hipMalloc((void **)&set->pages, pages.size() * sizeof(float4)));
  HYD_CHECK_HIP(hipMemcpy(set->codes, codes.data(), codes.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HYD_CHECK_HIP(hipMemcpy(set->pages, pages.data(), pages.size() * sizeof(float4), hipMemcpyHostToDevice));
  *out = set.release();
  return HYD_OK;
}

int32_t hyd_segments_pass_set(hyd_ctx *c, const hyd_segment_set *set, const hyd_volume_params *p,
                              const hyd_view *view, int32_t width, int32_t height, int32_t max_slots,
                              float *slots_dev, int32_t *nslots_dev, unsigned long long *samples_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(set && view && slots_dev && nslots_dev && width > 0 && height > 0, HYD_ERR_ARG,
              "invalid segment pass");
  HYD_REQUIRE(set->device == c->device, HYD_ERR_ARG, "the set was made on device %d, the context uses %d",
              set->device, c->device);
  HYD_REQUIRE(c->tf, HYD_ERR_STATE, "no transfer function: call hyd_set_tf first");
  HYD_REQUIRE(c->cheb.data, HYD_ERR_STATE, "no empty-space grid: call hyd_set_empty_grid first");
  const int32_t need = set->vdim[0] + set->vdim[1] + set->vdim[2];
  HYD_REQUIRE(max_slots >= need, HYD_ERR_RANGE, "max_slots %d < %d (sum of the brick grid dims)", max_slots, need);
  VolumeArgs a;
  if ((st = fill_volume_args(c, p, a))) return st;
  HYD_REQUIRE(a.schedule == HYD_SCHED_LATTICE, HYD_ERR_ARG, "segments need the lattice schedule");
  std::memcpy(a.vdim, set->vdim, sizeof(a.vdim));
  a.pool = set->pool[0];
  a.codes = set->codes;
  a.page = set->pages;
  a.pool_elems = set->max_pool_elems;
  a.pad = set->pad;
  fill_view(view, a.trans, a.ctg, a.ray_o);
  a.W = width;
  a.H = height;
  a.slots = reinterpret_cast<float4 *>(slots_dev);
  a.nslots = nslots_dev;
  a.max_slots = max_slots;
  a.samples = samples_dev;
  SegBatchTable tb{};
  tb.n = set->n;
  const size_t cells = (size_t)set->vdim[0] * set->vdim[1] * set->vdim[2];
  for (int i = 0; i < set->n; ++i) {
    tb.pool[i] = set->pool[i];
    tb.codes[i] = set->codes + (size_t)i * cells;
    tb.page[i] = set->pages + (size_t)i * set->max_pages;
  }
  HYD_CHECK_HIP(launch_segments_batches(a, tb, set->fmt, c->stream));
  return HYD_OK;
}

int32_t hyd_segments_composite(hyd_ctx *c, int32_t width, int32_t height, int32_t max_slots,
                               const float *slots_dev, const int32_t *nslots_dev, float *rgba_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(slots_dev && nslots_dev && rgba_dev && width > 0 && height > 0 && max_slots > 0, HYD_ERR_ARG,
              "invalid segment composite");
  HYD_CHECK_HIP(launch_segments_composite(reinterpret_cast<const float4 *>(slots_dev), nslots_dev, max_slots,
                                          width * height, reinterpret_cast<float4 *>(rgba_dev), c->stream));
  return HYD_OK;
}

int32_t hyd_trace_volume(hyd_ctx *c, const hyd_volume_params *p, const hyd_view *view, int32_t width,
                         int32_t height, int32_t x, int32_t y, hyd_trace_rec *out, int32_t max_recs,
                         int32_t *n_out) {
  int32_t st = activate(c);
  if (st) return st;
  if ((st = check_volume_state(c))) return st;
  HYD_REQUIRE(view && out && n_out && max_recs > 0, HYD_ERR_ARG, "invalid trace arguments");
  HYD_REQUIRE(x >= 0 && x < width && y >= 0 && y < height, HYD_ERR_ARG, "pixel outside the image");
  VolumeArgs a;
  if ((st = fill_volume_args(c, p, a))) return st;
  HYD_REQUIRE(a.schedule == HYD_SCHED_REFERENCE, HYD_ERR_ARG, "trace supports the reference schedule only");
  fill_view(view, a.trans, a.ctg, a.ray_o);
  a.W = width;
  a.H = height;
  hyd_trace_rec *dout = nullptr;
  int *dn = nullptr;
  HYD_CHECK_HIP(hipMalloc((void **)&dout, (size_t)max_recs * sizeof(hyd_trace_rec)));
  if (hipMalloc((void **)&dn, sizeof(int)) != hipSuccess) {
    (void)hipFree(dout);
    set_error("hipMalloc failed");
    return HYD_ERR_HIP;
  }
  hipError_t e = launch_volume_trace(a, c->pool.fmt, x, y, dout, max_recs, dn, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  int n = 0;
  if (e == hipSuccess) e = hipMemcpy(&n, dn, sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && n > 0) e = hipMemcpy(out, dout, (size_t)n * sizeof(hyd_trace_rec), hipMemcpyDeviceToHost);
  (void)hipFree(dout);
  (void)hipFree(dn);
  HYD_CHECK_HIP(e);
  *n_out = n;
  return HYD_OK;
}

int32_t hyd_render_blocks(hyd_ctx *c, const hyd_blocks_params *p, const hyd_view *view, int32_t width,
                          int32_t height, float *rgba_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(p && view && rgba_dev && width > 0 && height > 0, HYD_ERR_ARG, "invalid blocks arguments");
  HYD_REQUIRE(p->step > 0.f, HYD_ERR_ARG, "step must be > 0");
  HYD_REQUIRE(p->mode == HYD_BLOCKS_VOLUME || p->mode == HYD_BLOCKS_SOLID || p->mode == HYD_BLOCKS_PAGING,
              HYD_ERR_ARG, "unknown blocks mode %d", p->mode);
  HYD_REQUIRE(c->cheb.data, HYD_ERR_STATE, "no empty-space grid: call hyd_set_empty_grid first");
  HYD_REQUIRE(p->mode != HYD_BLOCKS_VOLUME || c->mean.data, HYD_ERR_STATE,
              "Blocks.Volume needs a mean grid: call hyd_set_mean_grid first");
  HYD_REQUIRE(p->mode != HYD_BLOCKS_VOLUME || std::memcmp(c->mean.dim, c->cheb.dim, sizeof(c->cheb.dim)) == 0,
              HYD_ERR_STATE, "mean grid and empty-space grid dims differ");
  HYD_REQUIRE(p->mode != HYD_BLOCKS_PAGING || c->paging.vaddr, HYD_ERR_STATE,
              "the paging view needs a page table: call hyd_set_page_table first");
  BlocksArgs a;
  std::memset(&a, 0, sizeof(a));
  a.bmin = f3{p->bbox_min[0], p->bbox_min[1], p->bbox_min[2]};
  a.bmax = f3{p->bbox_max[0], p->bbox_max[1], p->bbox_max[2]};
  a.step = p->step;
  a.max_steps = p->max_steps;
  a.mode = p->mode;
  a.density = p->density;
  fill_view(view, a.trans, a.ctg, a.ray_o);
  a.W = width;
  a.H = height;
  a.cheb = c->cheb.data;
  a.mean = c->mean.data;
  std::memcpy(a.dim, c->cheb.dim, sizeof(a.dim));
  if (p->mode == HYD_BLOCKS_PAGING) {
    a.vaddr = c->paging.vaddr;
    std::memcpy(a.vdim, c->paging.vdim, sizeof(a.vdim));
    a.lowest = c->paging.lowest_blkcnt;
  }
  a.rgba = reinterpret_cast<float4 *>(rgba_dev);
  HYD_CHECK_HIP(launch_blocks(a, c->stream));
  return HYD_OK;
}

int32_t hyd_render_iso(hyd_ctx *c, const hyd_iso_params *p, const hyd_view *view, int32_t width,
                       int32_t height, hyd_iso_pixel *out_dev, unsigned long long *samples_dev) {
  int32_t st = activate(c);
  if (st) return st;
  HYD_REQUIRE(c->pool.data && c->paging.vaddr && c->cheb.data, HYD_ERR_STATE,
              "isosurface needs pool, page table and empty-space grid");
  HYD_REQUIRE(c->paging.n_pages <= c->pool.n_pages, HYD_ERR_STATE,
              "page table names %d pages but the pool holds %d", c->paging.n_pages, c->pool.n_pages);
  HYD_REQUIRE(p && view && out_dev && width > 0 && height > 0, HYD_ERR_ARG, "invalid isosurface arguments");
  HYD_REQUIRE(p->step > 0.f, HYD_ERR_ARG, "step must be > 0");
  HYD_REQUIRE(p->bisect_steps >= 0 && p->bisect_steps <= 64, HYD_ERR_ARG, "bisect_steps out of range");
  IsoArgs a;
  std::memset(&a, 0, sizeof(a));
  a.bmin = f3{p->bbox_min[0], p->bbox_min[1], p->bbox_min[2]};
  a.bmax = f3{p->bbox_max[0], p->bbox_max[1], p->bbox_max[2]};
  a.step = p->step;
  a.du = p->du;
  a.max_steps = p->max_steps;
  a.mode = p->mode;
  std::memcpy(a.to_world, p->to_world, sizeof(a.to_world));
  a.eye_pos = f3{p->eye_pos[0], p->eye_pos[1], p->eye_pos[2]};
  a.light_pos = f3{p->light_pos[0], p->light_pos[1], p->light_pos[2]};
  a.surface_color = f3{p->surface_color[0], p->surface_color[1], p->surface_color[2]};
  a.isovalue = p->isovalue;
  a.corrected = p->corrected;
  a.bisect_steps = p->bisect_steps;
  fill_view(view, a.trans, a.ctg, a.ray_o);
  a.W = width;
  a.H = height;
  a.cheb = c->cheb.data;
  std::memcpy(a.cdim, c->cheb.dim, sizeof(a.cdim));
  a.vaddr = c->paging.vaddr;
  std::memcpy(a.vdim, c->paging.vdim, sizeof(a.vdim));
  a.codes = c->codes_valid ? c->own_codes : nullptr;
  a.page = c->paging.page;
  a.pool = c->pool.data;
  a.pool_elems = (uint64_t)c->pool.n_pages * c->pool.pad * c->pool.pad * c->pool.pad;
  a.pad = c->pool.pad;
  a.u8_lut = c->u8_lut;
  a.out = out_dev;
  a.samples = samples_dev;
  HYD_CHECK_HIP(launch_iso(a, c->pool.fmt, c->stream));
  return HYD_OK;
}

}  // extern "C"
