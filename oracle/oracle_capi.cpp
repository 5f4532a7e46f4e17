// C entry points of the oracle (TEST INFRASTRUCTURE ONLY; see oracle.hpp).
// Every output is written as double: FP32 results convert exactly, so the
// caller recovers the binary32 values with a float32 cast.
#include "oracle.hpp"

#include <cstdlib>

using namespace orc;

extern "C" {

typedef struct {
  double trans[16];  // column-major
  double ctg_fovy_2;
  double ray_o[3];
} orc_view;

typedef struct {
  int W, H;
  double bbox_min[3], bbox_max[3];
  double step, du;
  int max_steps, mode;
  double rank, density;
  int schedule;
  double ert;
  int tf_n;
  const float *tf;
  int cheb_dim[3];
  const int *cheb;
  int vaddr_dim[3];
  const int *vaddr;
  int lowest_blkcnt;
  const float *page_k;
  const float *page_b;
  const float *pool;
  int pad;
  int map_bs;   // > 0: lattice schedule uses the exact mapping for this block size
} orc_volume_args;

typedef struct {
  int W, H;
  double bbox_min[3], bbox_max[3];
  double step;
  int max_steps, mode;
  double density;
  int dim[3];
  const int *cheb;
  const float *mean;
  const int *vaddr;   // mode 2
  int vdim[3];
  int lowest;
} orc_blocks_args;

typedef struct {
  int W, H;
  double bbox_min[3], bbox_max[3];
  double step, du;
  int max_steps, mode;
  double to_world[16];
  double eye_pos[3], light_pos[3], surface_color[3];
  double isovalue;
  int cheb_dim[3];
  const int *cheb;
  int vaddr_dim[3];
  const int *vaddr;
  const float *page_k;
  const float *page_b;
  const float *pool;
  int pad;
  int corrected;
  int bisect_steps;
} orc_iso_args;

typedef struct { int ix, iy, iz, pgid, tix, tiy, tiz, tjx, tjy, tjz; } orc_trace_rec;
}

namespace {

int resolve_threads(int n) {
  if (n > 0) return n;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

template <class T> View<T> to_view(const orc_view *v) {
  View<T> r;
  for (int c = 0; c < 4; ++c)
    r.trans.c[c] = V4<T>{T(v->trans[4 * c]), T(v->trans[4 * c + 1]), T(v->trans[4 * c + 2]), T(v->trans[4 * c + 3])};
  r.ctg_fovy_2 = T(v->ctg_fovy_2);
  r.ray_o = V3<T>{T(v->ray_o[0]), T(v->ray_o[1]), T(v->ray_o[2])};
  return r;
}

template <class T> void from_view(const View<T> &r, orc_view *v) {
  for (int c = 0; c < 4; ++c) {
    v->trans[4 * c] = r.trans.c[c].x; v->trans[4 * c + 1] = r.trans.c[c].y;
    v->trans[4 * c + 2] = r.trans.c[c].z; v->trans[4 * c + 3] = r.trans.c[c].w;
  }
  v->ctg_fovy_2 = r.ctg_fovy_2;
  v->ray_o[0] = r.ray_o.x; v->ray_o[1] = r.ray_o.y; v->ray_o[2] = r.ray_o.z;
}

template <class T> V3<T> v3(const double *p) { return V3<T>{T(p[0]), T(p[1]), T(p[2])}; }

template <class T> VolumeShader<T> to_shader(const orc_volume_args *a) {
  VolumeShader<T> s{};
  s.bbox = Box<T>{v3<T>(a->bbox_min), v3<T>(a->bbox_max)};
  s.step = T(a->step); s.du = T(a->du);
  s.max_steps = a->max_steps; s.mode = a->mode; s.rank = T(a->rank); s.density = T(a->density);
  s.tf = a->tf; s.tf_n = a->tf_n;
  s.cheb = a->cheb; s.cheb_dim = I3{a->cheb_dim[0], a->cheb_dim[1], a->cheb_dim[2]};
  s.paging = Paging{I3{a->vaddr_dim[0], a->vaddr_dim[1], a->vaddr_dim[2]}, a->vaddr, a->lowest_blkcnt,
                    a->page_k, a->page_b, a->pool, a->pad, a->map_bs};
  s.schedule = a->schedule; s.ert = T(a->ert);
  return s;
}

// film layout per pixel (16 doubles): o[3] d[3] nsteps step_size v[4] theta[3] phi
// (lattice schedule: f[6] holds the ray extent lf instead of nsteps)
template <class T> void store_pixel(const VolumePixel<T> &p, double *f, bool lattice) {
  f[0] = p.o.x; f[1] = p.o.y; f[2] = p.o.z; f[3] = p.d.x; f[4] = p.d.y; f[5] = p.d.z;
  f[6] = lattice ? double(p.lf) : double(p.nsteps); f[7] = p.step_size;
  f[8] = p.v.x; f[9] = p.v.y; f[10] = p.v.z; f[11] = p.v.w;
  f[12] = p.theta.x; f[13] = p.theta.y; f[14] = p.theta.z; f[15] = p.phi;
}
template <class T> VolumePixel<T> load_pixel(const double *f, bool lattice) {
  VolumePixel<T> p;
  p.o = V3<T>{T(f[0]), T(f[1]), T(f[2])}; p.d = V3<T>{T(f[3]), T(f[4]), T(f[5])};
  p.nsteps = lattice ? 0 : int(f[6]); p.lf = lattice ? T(f[6]) : T(0); p.step_size = int(f[7]);
  p.v = V4<T>{T(f[8]), T(f[9]), T(f[10]), T(f[11])};
  p.theta = V3<T>{T(f[12]), T(f[13]), T(f[14])}; p.phi = T(f[15]);
  return p;
}

template <class T>
int volume_render(int nthreads, const orc_view *view, const orc_volume_args *a, double *film, double *rgba,
                  long long *counts, long long *total) {
  View<T> v = to_view<T>(view);
  VolumeShader<T> sh = to_shader<T>(a);
  std::atomic<long long> tot{0};
  row_interleave(a->H, resolve_threads(nthreads), [&](int y) {
    long long row = 0;
    for (int x = 0; x < a->W; ++x) {
      Ray<T> r = gen_ray(v, x, y, a->W, a->H);
      VolumePixel<T> px;
      long long ns = 0;
      volume_emit(sh, r, px, &ns, nullptr);
      size_t i = (size_t)y * a->W + x;
      if (film) store_pixel(px, film + 16 * i, sh.schedule == SCHED_LATTICE);
      if (rgba) { rgba[4 * i] = px.v.x; rgba[4 * i + 1] = px.v.y; rgba[4 * i + 2] = px.v.z; rgba[4 * i + 3] = px.v.w; }
      if (counts) counts[i] = ns;
      row += ns;
    }
    tot += row;
  });
  if (total) *total = tot.load();
  return 0;
}

template <class T>
int volume_resume(int nthreads, const orc_volume_args *a, double *film, long long *total) {
  VolumeShader<T> sh = to_shader<T>(a);
  std::atomic<long long> tot{0};
  row_interleave(a->H, resolve_threads(nthreads), [&](int y) {
    long long row = 0;
    for (int x = 0; x < a->W; ++x) {
      size_t i = (size_t)y * a->W + x;
      const bool lat = sh.schedule == SCHED_LATTICE;
      VolumePixel<T> px = load_pixel<T>(film + 16 * i, lat);
      long long ns = 0;
      // ray_march_shader_impl, shader.hpp:146-157
      if (lat) volume_main_lattice(sh, px, &ns);
      else volume_main(sh, px, &ns, nullptr);
      store_pixel(px, film + 16 * i, lat);
      row += ns;
    }
    tot += row;
  });
  if (total) *total = tot.load();
  return 0;
}

template <class T> BlocksShader<T> to_blocks(const orc_blocks_args *a) {
  BlocksShader<T> s{};
  s.bbox = Box<T>{v3<T>(a->bbox_min), v3<T>(a->bbox_max)};
  s.step = T(a->step); s.max_steps = a->max_steps; s.mode = a->mode; s.density = T(a->density);
  s.cheb = a->cheb; s.dim = I3{a->dim[0], a->dim[1], a->dim[2]}; s.mean = a->mean;
  s.vaddr = a->vaddr; s.vdim = I3{a->vdim[0], a->vdim[1], a->vdim[2]}; s.lowest = a->lowest;
  return s;
}

template <class T> IsoShader<T> to_iso(const orc_iso_args *a) {
  IsoShader<T> s{};
  s.bbox = Box<T>{v3<T>(a->bbox_min), v3<T>(a->bbox_max)};
  s.step = T(a->step); s.du = T(a->du); s.max_steps = a->max_steps; s.mode = a->mode;
  for (int c = 0; c < 4; ++c)
    s.to_world.c[c] = V4<T>{T(a->to_world[4 * c]), T(a->to_world[4 * c + 1]), T(a->to_world[4 * c + 2]),
                            T(a->to_world[4 * c + 3])};
  s.eye_pos = v3<T>(a->eye_pos); s.light_pos = v3<T>(a->light_pos); s.surface_color = v3<T>(a->surface_color);
  s.isovalue = T(a->isovalue);
  s.cheb = a->cheb; s.cheb_dim = I3{a->cheb_dim[0], a->cheb_dim[1], a->cheb_dim[2]};
  s.paging = Paging{I3{a->vaddr_dim[0], a->vaddr_dim[1], a->vaddr_dim[2]}, a->vaddr, 0, a->page_k, a->page_b,
                    a->pool, a->pad};
  s.corrected = a->corrected; s.bisect_steps = a->bisect_steps;
  return s;
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

// Camera orbit position (Camera::update_params(CameraOrbit), scene.hpp:68-81).
int orc_orbit_position(int prec, const double arm[3], double out[3]) {
  if (prec == 32) {
    V3<float> p = orbit_position(V3<float>{float(arm[0]), float(arm[1]), float(arm[2])});
    out[0] = p.x; out[1] = p.y; out[2] = p.z;
  } else {
    V3<double> p = orbit_position(V3<double>{arm[0], arm[1], arm[2]});
    out[0] = p.x; out[1] = p.y; out[2] = p.z;
  }
  return 0;
}

int orc_make_view(int prec, const double pos[3], const double target[3], const double up[3], double fovy,
                  const double ex_center[3], const double ex_size[3], orc_view *out) {
  if (prec == 32) {
    from_view(make_view(v3<float>(pos), v3<float>(target), v3<float>(up), float(fovy), v3<float>(ex_center),
                        v3<float>(ex_size)),
              out);
  } else {
    from_view(make_view(v3<double>(pos), v3<double>(target), v3<double>(up), fovy, v3<double>(ex_center),
                        v3<double>(ex_size)),
              out);
  }
  return 0;
}

// inverse(iet) for the isosurface shader's to_world (isosurface.cc:108).
int orc_inverse(int prec, const double m[16], double out[16]) {
  if (prec == 32) {
    M4<float> a;
    for (int c = 0; c < 4; ++c) a.c[c] = V4<float>{float(m[4 * c]), float(m[4 * c + 1]), float(m[4 * c + 2]), float(m[4 * c + 3])};
    M4<float> r = inverse(a);
    for (int c = 0; c < 4; ++c) { out[4 * c] = r.c[c].x; out[4 * c + 1] = r.c[c].y; out[4 * c + 2] = r.c[c].z; out[4 * c + 3] = r.c[c].w; }
  } else {
    M4<double> a;
    for (int c = 0; c < 4; ++c) a.c[c] = V4<double>{m[4 * c], m[4 * c + 1], m[4 * c + 2], m[4 * c + 3]};
    M4<double> r = inverse(a);
    for (int c = 0; c < 4; ++c) { out[4 * c] = r.c[c].x; out[4 * c + 1] = r.c[c].y; out[4 * c + 2] = r.c[c].z; out[4 * c + 3] = r.c[c].w; }
  }
  return 0;
}

int orc_exhibit_iet(int prec, const double center[3], const double size[3], double out[16]) {
  if (prec == 32) {
    M4<float> r = exhibit_iet(v3<float>(center), v3<float>(size));
    for (int c = 0; c < 4; ++c) { out[4 * c] = r.c[c].x; out[4 * c + 1] = r.c[c].y; out[4 * c + 2] = r.c[c].z; out[4 * c + 3] = r.c[c].w; }
  } else {
    M4<double> r = exhibit_iet(v3<double>(center), v3<double>(size));
    for (int c = 0; c < 4; ++c) { out[4 * c] = r.c[c].x; out[4 * c + 1] = r.c[c].y; out[4 * c + 2] = r.c[c].z; out[4 * c + 3] = r.c[c].w; }
  }
  return 0;
}

// Ray of one pixel: o[3], d[3].
int orc_gen_ray(int prec, const orc_view *view, int x, int y, int W, int H, double out[6]) {
  if (prec == 32) {
    Ray<float> r = gen_ray(to_view<float>(view), x, y, W, H);
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z; out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
  } else {
    Ray<double> r = gen_ray(to_view<double>(view), x, y, W, H);
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z; out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
  }
  return 0;
}

int orc_intersect(int prec, const double o[3], const double d[3], const double bmin[3], const double bmax[3],
                  double out[2]) {
  int hit;
  if (prec == 32) {
    Ray<float> r{v3<float>(o), v3<float>(d)};
    float tn, tf;
    hit = r.intersect(Box<float>{v3<float>(bmin), v3<float>(bmax)}, tn, tf);
    out[0] = tn; out[1] = tf;
  } else {
    Ray<double> r{v3<double>(o), v3<double>(d)};
    double tn, tf;
    hit = r.intersect(Box<double>{v3<double>(bmin), v3<double>(bmax)}, tn, tf);
    out[0] = tn; out[1] = tf;
  }
  return hit;
}

double orc_brick_sample(int prec, const float *data, int n, const double u[3], int taps_out[6]) {
  Taps t;
  double r;
  if (prec == 32) r = brick_sample(data, n, v3<float>(u), &t);
  else r = brick_sample(data, n, v3<double>(u), &t);
  if (taps_out) { taps_out[0] = t.ix; taps_out[1] = t.iy; taps_out[2] = t.iz; taps_out[3] = t.jx; taps_out[4] = t.jy; taps_out[5] = t.jz; }
  return r;
}

int orc_tf_sample(int prec, const float *rgba, int n, double s, double out[4]) {
  if (prec == 32) {
    V4<float> v = tf_sample(rgba, n, float(s));
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
  } else {
    V4<double> v = tf_sample(rgba, n, s);
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
  }
  return 0;
}

int orc_volume_render(int prec, int nthreads, const orc_view *view, const orc_volume_args *a, double *film,
                      double *rgba, long long *counts, long long *total) {
  if (prec == 32) return volume_render<float>(nthreads, view, a, film, rgba, counts, total);
  return volume_render<double>(nthreads, view, a, film, rgba, counts, total);
}

// Render only the listed pixels (xy: n pairs); rgba gets 4 doubles per listed pixel.
int orc_volume_render_pixels(int prec, int nthreads, const orc_view *view, const orc_volume_args *a, int n,
                             const int *xy, double *rgba, long long *total) {
  std::atomic<long long> tot{0};
  auto run = [&](auto tag) {
    using T = decltype(tag);
    View<T> v = to_view<T>(view);
    VolumeShader<T> sh = to_shader<T>(a);
    const int nt = resolve_threads(nthreads);
    const int rows = (n + 63) / 64;
    row_interleave(rows, nt, [&](int r) {
      long long acc = 0;
      for (int i = r * 64; i < std::min(n, r * 64 + 64); ++i) {
        VolumePixel<T> px;
        long long ns = 0;
        volume_emit(sh, gen_ray(v, xy[2 * i], xy[2 * i + 1], a->W, a->H), px, &ns, nullptr);
        rgba[4 * i] = px.v.x; rgba[4 * i + 1] = px.v.y; rgba[4 * i + 2] = px.v.z; rgba[4 * i + 3] = px.v.w;
        acc += ns;
      }
      tot += acc;
    });
  };
  if (prec == 32) run(float{}); else run(double{});
  if (total) *total = tot.load();
  return 0;
}

int orc_volume_resume(int prec, int nthreads, const orc_volume_args *a, double *film, long long *total) {
  if (prec == 32) return volume_resume<float>(nthreads, a, film, total);
  return volume_resume<double>(nthreads, a, film, total);
}

// Per-sample page/tap trace of one pixel (FP32 reference schedule).
int orc_volume_trace(const orc_view *view, const orc_volume_args *a, int x, int y, orc_trace_rec *out, int max_recs) {
  View<float> v = to_view<float>(view);
  VolumeShader<float> sh = to_shader<float>(a);
  Ray<float> r = gen_ray(v, x, y, a->W, a->H);
  std::vector<TraceRec> recs;
  TraceBuf tb{&recs, (size_t)max_recs};
  VolumePixel<float> px;
  volume_emit(sh, r, px, nullptr, &tb);
  for (size_t i = 0; i < recs.size(); ++i) {
    const TraceRec &t = recs[i];
    out[i] = orc_trace_rec{t.ix, t.iy, t.iz, t.pgid, t.t.ix, t.t.iy, t.t.iz, t.t.jx, t.t.jy, t.t.jz};
  }
  return (int)recs.size();
}

// Pixel pass (StdVec4Pixel::write_to) over an RGBA float image.
int orc_pixel_pass(int W, int H, const float *rgba, const unsigned char clear[3], unsigned char *out) {
  for (int i = 0; i < W * H; ++i)
    write_rgb8(V4<float>{rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], rgba[4 * i + 3]}, clear, out + 3 * i);
  return 0;
}

// Sort-last composite of N fetched partial images (VolumeFetchPixel: theta[3],
// phi, val[4]; 8 floats per pixel) in z_order, then sRGB + saturate
// (src/slave/renderers/volume/volume.cc:267-292).  out_rgba gets the
// composited linear RGBA, out_rgb8 the sRGB bytes.
int orc_composite(int N, int W, int H, const float *partials, const int *z_order, float *out_rgba,
                  unsigned char *out_rgb8) {
  size_t P = (size_t)W * H;
  for (size_t i = 0; i < P; ++i) {
    const float *f = partials + ((size_t)z_order[0] * P + i) * 8;
    float fv[4] = {f[4], f[5], f[6], f[7]};
    for (int k = 1; k < N; ++k) {
      const float *b = partials + ((size_t)z_order[k] * P + i) * 8;
      float vx = fv[0] + b[4] - fv[3] * b[0];
      float vy = fv[1] + b[5] - fv[3] * b[1];
      float vz = fv[2] + b[6] - fv[3] * b[2];
      float an = b[3] * fv[3] + b[7];
      fv[0] = vx; fv[1] = vy; fv[2] = vz; fv[3] = an;
    }
    if (out_rgba) { for (int c = 0; c < 4; ++c) out_rgba[4 * i + c] = fv[c]; }
    if (out_rgb8) {
      float s[3] = {srgb_f(fv[0]), srgb_f(fv[1]), srgb_f(fv[2])};
      for (int c = 0; c < 3; ++c) out_rgb8[3 * i + c] = (unsigned char)gclamp(s[c] * 255.f, 0.f, 255.f);
    }
  }
  return 0;
}

// Checks a table t[0..255] of sRGB byte thresholds against saturate(linear_to_srgb(x))
// (volume.cc:175-181,284-288): for every float x in [0, 2] and a sample of negative,
// large and special values, the byte equals the number of t[1..255] at or below x.
// Returns the number of mismatches; *first_bad gets the bits of the first one.
long long orc_srgb8_check(const float *t, int nthreads, unsigned *first_bad) {
  auto byte_ref = [](float x) { return (unsigned)(unsigned char)gclamp(srgb_f(x) * 255.f, 0.f, 255.f); };
  auto byte_tab = [&](float x) {
    unsigned k = 0;
    for (int i = 1; i < 256; ++i) k += t[i] <= x;
    return k;
  };
  const int n = resolve_threads(nthreads);
  const uint32_t end = 0x40000000u;   // 2.0f
  std::vector<long long> bad(n, 0);
  std::vector<unsigned> first(n, ~0u);
  std::vector<std::thread> th;
  for (int w = 0; w < n; ++w)
    th.emplace_back([&, w] {
      // thresholds are increasing: walk the range keeping the table byte incrementally
      const uint32_t lo = (uint32_t)((uint64_t)end * w / n), hi = (uint32_t)((uint64_t)end * (w + 1) / n);
      float x0;
      std::memcpy(&x0, &lo, 4);
      unsigned k = byte_tab(x0);
      for (uint32_t u = lo; u < hi; ++u) {
        float x;
        std::memcpy(&x, &u, 4);
        while (k < 255 && t[k + 1] <= x) ++k;
        if (byte_ref(x) != k) {
          if (!bad[w]) first[w] = u;
          ++bad[w];
        }
      }
    });
  for (auto &x : th) x.join();
  long long total = 0;
  unsigned fb = ~0u;
  for (int w = 0; w < n; ++w) {
    total += bad[w];
    if (bad[w] && fb == ~0u) fb = first[w];
  }
  const float extra[] = {-0.0f, -1e-30f, -0.5f, -1.f, -1e30f, 2.5f, 10.f, 1e30f, INFINITY, -INFINITY};
  for (float x : extra)
    if (byte_ref(x) != byte_tab(x)) {
      if (!total) std::memcpy(&fb, &x, 4);
      ++total;
    }
  if (first_bad) *first_bad = fb;
  return total;
}

// IsosurfaceShaderKernel::fetch (isosurface_shader.cu:42-49): saturate(v) bytes
// and depth of each isosurface pixel (6 doubles: rgba, depth, hit).
int orc_iso_fetch(int n, const double *px, unsigned char *val, float *depth) {
  for (int i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) val[3 * i + c] = (unsigned char)gclamp(float(px[6 * i + c]) * 255.f, 0.f, 255.f);
    depth[i] = float(px[6 * i + 4]);
  }
  return 0;
}

// Isosurface sort-last composite (isosurface.cc:272-289): the front is the
// partial of z_order[0]; a later partial in z order replaces a front pixel when
// its depth is strictly smaller.  val: N*P*3 bytes, depth: N*P floats.
int orc_composite_iso(int N, int W, int H, const unsigned char *val, const float *depth, const int *z_order,
                      unsigned char *out_rgb, float *out_depth) {
  const size_t P = (size_t)W * H;
  for (size_t i = 0; i < P; ++i) {
    size_t f = (size_t)z_order[0] * P + i;
    for (int k = 1; k < N; ++k) {
      const size_t b = (size_t)z_order[k] * P + i;
      if (depth[b] < depth[f]) f = b;
    }
    for (int c = 0; c < 3; ++c) out_rgb[3 * i + c] = val[3 * f + c];
    out_depth[i] = depth[f];
  }
  return 0;
}

int orc_blocks_render(int prec, int nthreads, const orc_view *view, const orc_blocks_args *a, double *rgba) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    View<T> v = to_view<T>(view);
    BlocksShader<T> sh = to_blocks<T>(a);
    row_interleave(a->H, resolve_threads(nthreads), [&](int y) {
      for (int x = 0; x < a->W; ++x) {
        VolumePixel<T> px;
        blocks_emit(sh, gen_ray(v, x, y, a->W, a->H), px);
        size_t i = (size_t)y * a->W + x;
        rgba[4 * i] = px.v.x; rgba[4 * i + 1] = px.v.y; rgba[4 * i + 2] = px.v.z; rgba[4 * i + 3] = px.v.w;
      }
    });
  };
  if (prec == 32) run(float{}); else run(double{});
  return 0;
}

// out per pixel: rgba[4], depth, hit  (6 doubles)
int orc_iso_render(int prec, int nthreads, const orc_view *view, const orc_iso_args *a, double *out,
                   long long *total) {
  std::atomic<long long> tot{0};
  auto run = [&](auto tag) {
    using T = decltype(tag);
    View<T> v = to_view<T>(view);
    IsoShader<T> sh = to_iso<T>(a);
    row_interleave(a->H, resolve_threads(nthreads), [&](int y) {
      long long row = 0;
      for (int x = 0; x < a->W; ++x) {
        IsoPixel<T> px;
        long long ns = 0;
        iso_emit(sh, gen_ray(v, x, y, a->W, a->H), px, &ns);
        size_t i = (size_t)y * a->W + x;
        double *o = out + 6 * i;
        o[0] = px.v.x; o[1] = px.v.y; o[2] = px.v.z; o[3] = px.v.w; o[4] = px.depth; o[5] = px.hit;
        row += ns;
      }
      tot += row;
    });
  };
  if (prec == 32) run(float{}); else run(double{});
  if (total) *total = tot.load();
  return 0;
}

}  // extern "C"
