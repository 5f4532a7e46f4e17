// Oracle for the Hydrant render hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a CPU restatement of the reference's
// multi-thread CPU render backend (xlnx/hydrant), written from its sources and
// templated on the scalar type:  T = float reproduces the reference's FP32
// arithmetic operation by operation, T = double gives the FP64 ground truth.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and only as the checker or as the timed CPU baseline; the product path
// (hydrant_amd/) never links, loads or calls it.
//
// Build with g++ -O2 -ffp-contract=off (no -ffast-math, no FMA-enabling -march):
// every float operation below is then one IEEE-754 binary32 operation, which is
// what the reference's g++/nvcc host compile performs.
//
// Parity pins (tests/test_oracle_*.py): closed-form orbit position, centre-pixel
// ray direction, FP64 brute-force slab test, affine-field trilinear exactness,
// TF knot values and Border behaviour, constant-field compositing closed form,
// pause/resume bit-exactness, sort-last associativity, isosurface analytic
// sphere.  glm's exact operation order for matrix inverse / lookAt is restated
// from upstream glm 0.9.9 (the reference's glm-cuda fork is an unpinned
// dependency); that part is "parity unpinned" beyond the closed-form checks.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- vectors --
// glm semantics (glm 0.9.9 func_common.inl): min(x,y) = (y < x) ? y : x,
// max(x,y) = (x < y) ? y : x, clamp = min(max(x, lo), hi).
template <class T> inline T gmin(T x, T y) { return (y < x) ? y : x; }
template <class T> inline T gmax(T x, T y) { return (x < y) ? y : x; }
template <class T> inline T gclamp(T x, T lo, T hi) { return gmin(gmax(x, lo), hi); }

template <class T> struct V2 { T x, y; };
template <class T> struct V3 {
  T x, y, z;
  V3 operator+(V3 b) const { return {x + b.x, y + b.y, z + b.z}; }
  V3 operator-(V3 b) const { return {x - b.x, y - b.y, z - b.z}; }
  V3 operator*(V3 b) const { return {x * b.x, y * b.y, z * b.z}; }
  V3 operator/(V3 b) const { return {x / b.x, y / b.y, z / b.z}; }
  V3 operator*(T s) const { return {x * s, y * s, z * s}; }
  V3 operator/(T s) const { return {x / s, y / s, z / s}; }
  V3 operator-(T s) const { return {x - s, y - s, z - s}; }
  V3 operator+(T s) const { return {x + s, y + s, z + s}; }
  V3 &operator+=(V3 b) { x += b.x; y += b.y; z += b.z; return *this; }
};
template <class T> inline V3<T> operator*(T s, V3<T> v) { return {s * v.x, s * v.y, s * v.z}; }
template <class T> struct V4 {
  T x, y, z, w;
  V4 operator+(V4 b) const { return {x + b.x, y + b.y, z + b.z, w + b.w}; }
  V4 operator*(V4 b) const { return {x * b.x, y * b.y, z * b.z, w * b.w}; }
  V4 operator*(T s) const { return {x * s, y * s, z * s, w * s}; }
  V4 &operator+=(V4 b) { x += b.x; y += b.y; z += b.z; w += b.w; return *this; }
  V4 &operator*=(V4 b) { x *= b.x; y *= b.y; z *= b.z; w *= b.w; return *this; }
};
struct I3 { int x, y, z; };

template <class T> inline V3<T> vfloor(V3<T> v) { return {std::floor(v.x), std::floor(v.y), std::floor(v.z)}; }
template <class T> inline V3<T> vceil(V3<T> v) { return {std::ceil(v.x), std::ceil(v.y), std::ceil(v.z)}; }
template <class T> inline V3<T> vabs(V3<T> v) { return {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)}; }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
// glm compMax / compMin: sequential fold with glm::max / glm::min.
template <class T> inline T comp_max(V3<T> v) { T r = v.x; r = gmax(r, v.y); r = gmax(r, v.z); return r; }
template <class T> inline T comp_min(V3<T> v) { T r = v.x; r = gmin(r, v.y); r = gmin(r, v.z); return r; }
// glm dot for vec3: tmp = a*b; return (tmp.x + tmp.y) + tmp.z.
template <class T> inline T dot(V3<T> a, V3<T> b) { V3<T> t = a * b; return t.x + t.y + t.z; }
template <class T> inline V3<T> cross(V3<T> x, V3<T> y) {
  return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
// glm normalize = v * inversesqrt(dot(v, v)), inversesqrt(x) = 1 / sqrt(x).
template <class T> inline V3<T> normalize(V3<T> v) { return v * (T(1) / std::sqrt(dot(v, v))); }
template <class T> inline T length(V3<T> v) { return std::sqrt(dot(v, v)); }
// glm mod(x, y) = x - y * floor(x / y)
template <class T> inline V3<T> vmod(V3<T> x, V3<T> y) { return x - y * vfloor(x / y); }

// Column-major 4x4 (glm layout: c[col] is a column).
template <class T> struct M4 {
  V4<T> c[4];
  static M4 identity() { return {{{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}}; }
};
// glm mat4 * vec4: (m0*x + m1*y) + (m2*z + m3*w).
template <class T> inline V4<T> mul(const M4<T> &m, V4<T> v) {
  V4<T> a0 = m.c[0] * v.x + m.c[1] * v.y;
  V4<T> a1 = m.c[2] * v.z + m.c[3] * v.w;
  return a0 + a1;
}
// glm mat4 * mat4: Result[i] = ((A0*B[i][0] + A1*B[i][1]) + A2*B[i][2]) + A3*B[i][3].
template <class T> inline M4<T> mul(const M4<T> &a, const M4<T> &b) {
  M4<T> r;
  for (int i = 0; i < 4; ++i) {
    const V4<T> &s = b.c[i];
    r.c[i] = a.c[0] * s.x + a.c[1] * s.y + a.c[2] * s.z + a.c[3] * s.w;
  }
  return r;
}
template <class T> inline V3<T> xyz(V4<T> v) { return {v.x, v.y, v.z}; }

// glm compute_inverse<4,4> (cofactor expansion, glm 0.9.9 func_matrix.inl).
template <class T> M4<T> inverse(const M4<T> &M) {
  auto m = [&](int c, int r) -> T { const V4<T> &v = M.c[c]; return r == 0 ? v.x : r == 1 ? v.y : r == 2 ? v.z : v.w; };
  T c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
  T c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
  T c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
  T c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
  T c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
  T c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
  T c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
  T c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
  T c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
  T c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
  T c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
  T c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
  T c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
  T c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
  T c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
  T c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
  T c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
  T c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
  V4<T> f0{c00, c00, c02, c03}, f1{c04, c04, c06, c07}, f2{c08, c08, c10, c11};
  V4<T> f3{c12, c12, c14, c15}, f4{c16, c16, c18, c19}, f5{c20, c20, c22, c23};
  V4<T> v0{m(1, 0), m(0, 0), m(0, 0), m(0, 0)};
  V4<T> v1{m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
  V4<T> v2{m(1, 2), m(0, 2), m(0, 2), m(0, 2)};
  V4<T> v3{m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
  auto sub = [](V4<T> a, V4<T> b) { return V4<T>{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; };
  V4<T> i0 = sub(v1 * f0, v2 * f1) + v3 * f2;
  V4<T> i1 = sub(v0 * f0, v2 * f3) + v3 * f4;
  V4<T> i2 = sub(v0 * f1, v1 * f3) + v3 * f5;
  V4<T> i3 = sub(v0 * f2, v1 * f4) + v2 * f5;
  V4<T> sa{+1, -1, +1, -1}, sb{-1, +1, -1, +1};
  M4<T> inv{{i0 * sa, i1 * sb, i2 * sa, i3 * sb}};
  V4<T> row0{inv.c[0].x, inv.c[1].x, inv.c[2].x, inv.c[3].x};
  V4<T> d0 = M.c[0] * row0;
  T d1 = (d0.x + d0.y) + (d0.z + d0.w);
  T one_over = T(1) / d1;
  for (auto &c : inv.c) c = c * one_over;
  return inv;
}

// glm lookAtRH.
template <class T> M4<T> look_at(V3<T> eye, V3<T> center, V3<T> up) {
  V3<T> f = normalize(center - eye);
  V3<T> s = normalize(cross(f, up));
  V3<T> u = cross(s, f);
  M4<T> r = M4<T>::identity();
  r.c[0].x = s.x; r.c[1].x = s.y; r.c[2].x = s.z;
  r.c[0].y = u.x; r.c[1].y = u.y; r.c[2].y = u.z;
  r.c[0].z = -f.x; r.c[1].z = -f.y; r.c[2].z = -f.z;
  r.c[3].x = -dot(s, eye); r.c[3].y = -dot(u, eye); r.c[3].z = dot(f, eye);
  return r;
}

template <class T> inline T radians(T deg) { return deg * static_cast<T>(0.01745329251994329576923690768489); }

// glm rotate(m, angle, axis) with m = identity (glm 0.9.9 matrix_transform.inl).
template <class T> M4<T> rotate_identity(T angle, V3<T> v) {
  T c = std::cos(angle), s = std::sin(angle);
  V3<T> axis = normalize(v);
  V3<T> temp = axis * (T(1) - c);
  T r00 = c + temp.x * axis.x, r01 = temp.x * axis.y + s * axis.z, r02 = temp.x * axis.z - s * axis.y;
  T r10 = temp.y * axis.x - s * axis.z, r11 = c + temp.y * axis.y, r12 = temp.y * axis.z + s * axis.x;
  T r20 = temp.z * axis.x + s * axis.y, r21 = temp.z * axis.y - s * axis.x, r22 = c + temp.z * axis.z;
  M4<T> m = M4<T>::identity();
  M4<T> r;
  r.c[0] = m.c[0] * r00 + m.c[1] * r01 + m.c[2] * r02;
  r.c[1] = m.c[0] * r10 + m.c[1] * r11 + m.c[2] * r12;
  r.c[2] = m.c[0] * r20 + m.c[1] * r21 + m.c[2] * r22;
  r.c[3] = m.c[3];
  return r;
}

// ------------------------------------------------------------ scene / view --
// Camera::update_params(CameraOrbit), include/hydrant/core/scene.hpp:68-81.
template <class T> V3<T> orbit_position(V3<T> arm) {
  V4<T> p = mul(rotate_identity(radians(arm.y), V3<T>{0, 0, 1}), V4<T>{arm.z, 0, 0, 1});
  V4<T> q = mul(rotate_identity(radians(arm.x), V3<T>{0, 1, 0}), V4<T>{p.x, p.y, p.z, 1});
  return xyz(q);
}

// Exhibit::get_iet, include/hydrant/core/scene.hpp:20-29.
template <class T> M4<T> exhibit_iet(V3<T> center, V3<T> size) {
  V3<T> d = vmax(vabs(center - size), vabs(center));
  T scale = comp_max(d);
  return {{{scale, 0, 0, 0}, {0, scale, 0, 0}, {0, 0, scale, 0}, {center.x, center.y, center.z, 1}}};
}

// ViewArgs (include/hydrant/core/shader.hpp:63-68) filled as
// Raycaster::fill_ray_emit_args, include/hydrant/core/raycaster.hpp:126-135.
template <class T> struct View {
  M4<T> trans;
  T ctg_fovy_2;
  V3<T> ray_o;
};

template <class T> View<T> make_view(V3<T> pos, V3<T> target, V3<T> up, T fovy_deg, V3<T> ex_center, V3<T> ex_size) {
  View<T> v;
  M4<T> iet = exhibit_iet(ex_center, ex_size);
  M4<T> ivt = inverse(look_at(pos, target, up));
  v.trans = mul(iet, ivt);
  // Camera(CameraConfig): ctg_fovy_2 = 1.f / tan(radians(fovy / 2.f)), scene.hpp:53
  v.ctg_fovy_2 = T(1) / std::tan(radians(fovy_deg / T(2)));
  v.ray_o = xyz(mul(iet, V4<T>{pos.x, pos.y, pos.z, 1}));
  return v;
}

// ------------------------------------------------------------------- rays --
template <class T> struct Box { V3<T> min, max; };
template <class T> struct Ray {
  V3<T> o, d;
  // Ray::intersect, include/hydrant/core/glm_math.hpp:58-72.
  bool intersect(const Box<T> &box, T &tnear, T &tfar) const {
    V3<T> invr = V3<T>{1, 1, 1} / d;
    V3<T> tbot = invr * (box.min - o);
    V3<T> ttop = invr * (box.max - o);
    V3<T> tmin = vmin(ttop, tbot);
    V3<T> tmax = vmax(ttop, tbot);
    tnear = comp_max(tmin);
    tfar = comp_min(tmax);
    return tfar > tnear;
  }
};

// Ray generation, src/slave/core/cpu_shading.cc:16-26 (integer pixel corners,
// uv scaled by res.y, y flipped).
template <class T> Ray<T> gen_ray(const View<T> &v, int x, int y, int W, int H) {
  T ccx = T(W) / T(2), ccy = T(H) / T(2);
  T ux = ((T(x) - ccx) * T(2)) / T(H);
  T uy = ((T(y) - ccy) * T(2)) / T(H);
  V4<T> p = mul(v.trans, V4<T>{ux, -uy, -v.ctg_fovy_2, T(1)});
  return Ray<T>{v.ray_o, normalize(xyz(p) - v.ray_o)};
}

// --------------------------------------------------------------- samplers --
// CpuSampler, include/hydrant/bridge/cpu_sampler.hpp:22-173.
enum class Addr { Clamp, Border, Wrap };
enum class Filt { None, Linear };

// Integer grids (chebyshev, vaddr): ThumbnailTexture / vaddr Texture3D<int> use
// Options::as_array() + Clamp (basic_renderer.hpp:147-159,
// lossless_block_paging.cc:46-52): unnormalised coordinates, point sampling.
template <class T> inline int grid_lookup(const int *g, I3 dim, V3<T> p) {
  V3<T> fd{T(dim.x), T(dim.y), T(dim.z)};
  V3<T> fx{gclamp(p.x, T(0), fd.x), gclamp(p.y, T(0), fd.y), gclamp(p.z, T(0), fd.z)};
  int ix = gclamp(int(std::floor(fx.x)), 0, dim.x - 1);
  int iy = gclamp(int(std::floor(fx.y)), 0, dim.y - 1);
  int iz = gclamp(int(std::floor(fx.z)), 0, dim.z - 1);
  return g[iz * dim.x * dim.y + iy * dim.x + ix];
}
template <class T> inline int grid_index(I3 dim, V3<T> p) {
  V3<T> fd{T(dim.x), T(dim.y), T(dim.z)};
  V3<T> fx{gclamp(p.x, T(0), fd.x), gclamp(p.y, T(0), fd.y), gclamp(p.z, T(0), fd.z)};
  int ix = gclamp(int(std::floor(fx.x)), 0, dim.x - 1);
  int iy = gclamp(int(std::floor(fx.y)), 0, dim.y - 1);
  int iz = gclamp(int(std::floor(fx.z)), 0, dim.z - 1);
  return iz * dim.x * dim.y + iy * dim.x + ix;
}
template <class T> inline T grid_lookup_f(const float *g, I3 dim, V3<T> p) {
  return T(g[grid_index(dim, p)]);
}

// Tap record for the per-sample trace (CpuSampler::filter_linear tap indices).
struct Taps { int ix, iy, iz, jx, jy, jz; };

// 3-D brick sampler: Wrap + Linear + normalised coordinates over a padded brick
// of edge n (the lossless paging storage options, volume.cc:86-90).
template <class T> inline T brick_sample(const float *data, int n, V3<T> u, Taps *taps = nullptr) {
  V3<T> fd{T(n), T(n), T(n)};
  V3<T> fx = fd * u;              // normalize_coords: fd * x
  fx = vmod(fx, fd);              // AddressMode::Wrap: mod(fx, fd)
  fx = fx - T(0.5f);              // filter_linear(fx - .5f)
  V3<T> flr = vfloor(fx);
  V3<T> a = fx - flr;
  I3 i{gclamp(int(flr.x), 0, n - 1), gclamp(int(flr.y), 0, n - 1), gclamp(int(flr.z), 0, n - 1)};
  V3<T> c = vceil(fx);
  I3 j{gclamp(int(c.x), 0, n - 1), gclamp(int(c.y), 0, n - 1), gclamp(int(c.z), 0, n - 1)};
  if (taps) *taps = Taps{i.x, i.y, i.z, j.x, j.y, j.z};
  auto at = [&](int x, int y, int z) -> T { return T(data[(size_t)z * n * n + (size_t)y * n + x]); };
  auto lerp = [](T x, T y, T t) { return x * (T(1) - t) + y * t; };
  T x00 = lerp(at(i.x, i.y, i.z), at(j.x, i.y, i.z), a.x);
  T x10 = lerp(at(i.x, j.y, i.z), at(j.x, j.y, i.z), a.x);
  T x01 = lerp(at(i.x, i.y, j.z), at(j.x, i.y, j.z), a.x);
  T x11 = lerp(at(i.x, j.y, j.z), at(j.x, j.y, j.z), a.x);
  T y0 = lerp(x00, x10, a.y);
  T y1 = lerp(x01, x11, a.y);
  return lerp(y0, y1, a.z);
}

// Transfer function: Texture1D<vec4>, Border + Linear + Raw + normalised
// (include/hydrant/transfer_fn.hpp:14-28; cpu_sampler.hpp:117-124,142-166).
template <class T> inline V4<T> tf_sample(const float *rgba, int n, T s) {
  T fd = T(n);
  T fx = fd * s;
  if (!(T(0) <= fx && fx <= fd)) return V4<T>{0, 0, 0, 0};
  fx = fx - T(0.5f);
  T flr = std::floor(fx);
  T a = fx - flr;
  int ix = gclamp(int(flr), 0, n - 1);
  int jx = gclamp(int(std::ceil(fx)), 0, n - 1);
  const float *p = rgba + 4 * ix, *q = rgba + 4 * jx;
  T w = T(1) - a;
  return V4<T>{T(p[0]) * w + T(q[0]) * a, T(p[1]) * w + T(q[1]) * a, T(p[2]) * w + T(q[2]) * a,
               T(p[3]) * w + T(q[3]) * a};
}

// ----------------------------------------------------------------- paging --
// BlockPaging (include/hydrant/paging/block_paging.hpp:11-44): vaddr grid ->
// page id -> BlockSampler{mapping k, b} over a padded brick in the pool.
struct Paging {
  I3 dim;                 // brick grid
  const int *vaddr;       // dim.x*dim.y*dim.z, -1 = not resident
  int lowest_blkcnt;
  const float *page_k;    // per page
  const float *page_b;    // per page, 3 floats
  const float *pool;      // n_pages * pad^3 normalised voxel values
  int pad;                // padded brick edge
  int map_bs = 0;         // > 0: lattice uses the exact mapping for block size map_bs
};

// ---------------------------------------------------------------- shaders --
enum Schedule { SCHED_REFERENCE = 0, SCHED_LATTICE = 1 };

template <class T> struct VolumeShader {
  Box<T> bbox;
  T step;
  T du;
  int max_steps;
  int mode;        // 0 Default, 1 Partition, 2 Paging (ui/volume.schema.hpp:11-14)
  T rank;
  T density;       // opacity scale on the TF alpha; 1 reproduces the reference literally
  const float *tf; int tf_n;
  const int *cheb; I3 cheb_dim;
  Paging paging;
  int schedule;    // SCHED_REFERENCE (escalation) or SCHED_LATTICE (fixed step + ERT)
  T ert;           // lattice early-termination alpha (0 disables)
};

// VolumePixel (src/slave/renderers/volume/volume_shader.hpp:11-15;
// IPixel shader.hpp:26-34; StdVec4Pixel pixel_template.hpp:9-24).
template <class T> struct VolumePixel {
  V3<T> o, d;
  int nsteps, step_size;
  V4<T> v;
  V3<T> theta;
  T phi;
  T lf;      // lattice schedule: ray extent in steps
};

struct TraceRec { int ix, iy, iz, pgid; Taps t; };
struct TraceBuf {
  std::vector<TraceRec> *out = nullptr;
  size_t limit = 0;
};

// skip_nblock_steps, src/slave/renderers/volume/volume_shader.cu:5-14.
template <class T> inline int skip_nblock_steps(Ray<T> &ray, V3<T> ip, int nblocks, T cdu, T step) {
  T tnear, tfar;
  ray.intersect(Box<T>{ip, ip + T(1)}, tnear, tfar);
  T di = std::ceil((tfar + T(nblocks - 1) * cdu) / step);
  ray.o += ray.d * di * step;
  return (int)di;
}

// Escalation thresholds: `pixel.v.w > 0.93` / `> 0.85` compare a float against
// double literals (volume_shader.cu:71-75).
template <class T> inline bool gt_093(T a) { return double(a) > 0.93; }
template <class T> inline bool gt_085(T a) { return double(a) > 0.85; }

// VolumeShaderKernel::main, src/slave/renderers/volume/volume_shader.cu:32-81,
// plus the build-defined density scale on the TF alpha.
template <class T>
void volume_main(const VolumeShader<T> &sh, VolumePixel<T> &px_io, long long *nsamples, TraceBuf *trace) {
  const T cdu = T(1) / comp_max(vabs(px_io.d));
  VolumePixel<T> px = px_io;
  Ray<T> ray{px.o, px.d};
  int &nsteps = px.nsteps;
  int &step_size = px.step_size;
  long long ns = 0;
  while (nsteps > 0) {
    V3<T> ip = vfloor(ray.o);
    int cd = grid_lookup(sh.cheb, sh.cheb_dim, ip);
    if (cd) {
      nsteps -= skip_nblock_steps(ray, ip, cd, cdu, sh.step * T(step_size));
    } else {
      int pgid = grid_lookup(sh.paging.vaddr, sh.paging.dim, ip);
      if (pgid == -1) break;
      V3<T> x = ray.o - ip;
      T k = T(sh.paging.page_k[pgid]);
      V3<T> b{T(sh.paging.page_b[3 * pgid]), T(sh.paging.page_b[3 * pgid + 1]), T(sh.paging.page_b[3 * pgid + 2])};
      V3<T> u = x * k + b;   // BlockSamplerMapping::mapped: k * x + b
      int pad = sh.paging.pad;
      Taps taps;
      T s = brick_sample(sh.paging.pool + (size_t)pgid * pad * pad * pad, pad, u, trace ? &taps : nullptr);
      if (trace && trace->out && trace->out->size() < trace->limit) {
        I3 g{gclamp(int(std::floor(gclamp(ip.x, T(0), T(sh.paging.dim.x)))), 0, sh.paging.dim.x - 1),
             gclamp(int(std::floor(gclamp(ip.y, T(0), T(sh.paging.dim.y)))), 0, sh.paging.dim.y - 1),
             gclamp(int(std::floor(gclamp(ip.z, T(0), T(sh.paging.dim.z)))), 0, sh.paging.dim.z - 1)};
        trace->out->push_back(TraceRec{g.x, g.y, g.z, pgid, taps});
      }
      V4<T> ub = tf_sample(sh.tf, sh.tf_n, s);
      ub.w = ub.w * sh.density;
      if (sh.mode == 1) {
        V3<T> lo{1, 0, 0}, hi{0, 0, 1};
        T l = length(V3<T>{ub.x, ub.y, ub.z});
        V3<T> m = (lo * (T(1) - sh.rank) + hi * sh.rank) * l;
        ub = V4<T>{m.x, m.y, m.z, ub.w};
      } else if (sh.mode == 2) {
        ub = pgid >= sh.paging.lowest_blkcnt ? V4<T>{0, 1, 0, ub.w} : V4<T>{1, 0, 0, ub.w};
      }
      ub *= V4<T>{ub.w, ub.w, ub.w, T(1)};
      px.theta += V3<T>{ub.x, ub.y, ub.z} * px.phi;
      px.phi *= T(1) - ub.w;
      px.v += ub * (T(1) - px.v.w);
      ++ns;
      if (gt_093(px.v.w)) {
        step_size = 16;
      } else if (gt_085(px.v.w)) {
        step_size = 4;
      }
    }
    ray.o += ray.d * sh.step * T(step_size);
    nsteps -= step_size;
  }
  px.o = ray.o;
  px.d = ray.d;
  px_io = px;
  if (nsamples) *nsamples += ns;
}

// Lattice schedule (build-defined; used by brick-order-independent sort-last
// compositing and by the batched form).  With o0 the clipped entry point and
// L = (tfar - tnear) / step the ray extent in steps:
//   sample k covers [k, min(k + 1, L)] steps: coverage w_k = min(1, L - k),
//   it sits at the covered segment's midpoint o0 + d * ((k + w_k / 2) * step)
//   (inside the box even for the last, partial segment), k = 0 .. ceil(L) - 1,
//   and its opacity is scaled by w_k,
//   empty bricks are skipped to the first lattice point past the empty region,
//   no step escalation, early termination once v.w > ert (ert > 0).
// Every decision is continuous in the ray geometry (a sample that appears or is
// skipped carries ~zero opacity), so FP32 and FP64 renders agree to rounding.
// Film: o = o0, lf = L, step_size = next lattice index k.
template <class T>
void volume_main_lattice(const VolumeShader<T> &sh, VolumePixel<T> &px_io, long long *nsamples) {
  VolumePixel<T> px = px_io;
  const V3<T> o0 = px.o;
  const T lf = px.lf;
  long long ns = 0;
  int k = px.step_size;
  const int kend = (int)std::ceil(lf);
  const T cdu = T(1) / comp_max(vabs(px.d));
  while (k < kend) {
    const T w = std::min(T(1), lf - T(k));         // coverage of sample k
    const T t = (T(k) + T(0.5f) * w) * sh.step;   // midpoint of the covered segment
    V3<T> o = o0 + px.d * t;
    V3<T> ip = vfloor(o);
    int cd = grid_lookup(sh.cheb, sh.cheb_dim, ip);
    if (cd) {
      Ray<T> r{o, px.d};
      T tn, tf;
      r.intersect(Box<T>{ip, ip + T(1)}, tn, tf);
      T t_exit = t + tf + T(cd - 1) * cdu;
      int kn = (int)std::ceil(t_exit / sh.step - T(0.5f));
      k = std::max(kn, k + 1);
      continue;
    }
    int pgid = grid_lookup(sh.paging.vaddr, sh.paging.dim, ip);
    if (pgid == -1) break;
    V3<T> x = o - ip;
    T kk, bx, by, bz;
    if (sh.paging.map_bs > 0) {   // exact mapping k = bs/pad, b = (apron/bs) * k in T
      const int pad = sh.paging.pad;
      kk = T(sh.paging.map_bs) / T(pad);
      bx = by = bz = (T((pad - sh.paging.map_bs) / 2) / T(sh.paging.map_bs)) * kk;
    } else {
      kk = T(sh.paging.page_k[pgid]);
      bx = T(sh.paging.page_b[3 * pgid]); by = T(sh.paging.page_b[3 * pgid + 1]); bz = T(sh.paging.page_b[3 * pgid + 2]);
    }
    V3<T> u = x * kk + V3<T>{bx, by, bz};
    int pad = sh.paging.pad;
    T s = brick_sample(sh.paging.pool + (size_t)pgid * pad * pad * pad, pad, u);
    V4<T> ub = tf_sample(sh.tf, sh.tf_n, s);
    ub.w = ub.w * sh.density * w;
    ub *= V4<T>{ub.w, ub.w, ub.w, T(1)};
    px.theta += V3<T>{ub.x, ub.y, ub.z} * px.phi;
    px.phi *= T(1) - ub.w;
    px.v += ub * (T(1) - px.v.w);
    ++ns;
    ++k;
    if (sh.ert > T(0) && px.v.w > sh.ert) { k = kend; break; }
  }
  px.step_size = k;
  px.nsteps = kend - k;
  px_io = px;
  if (nsamples) *nsamples += ns;
}

// ray_emit_shader_impl, include/hydrant/core/shader.hpp:113-137, with
// VolumeShaderKernel::init (volume_shader.cu:16-21).
template <class T>
void volume_emit(const VolumeShader<T> &sh, const Ray<T> &ray_in, VolumePixel<T> &out, long long *nsamples,
                 TraceBuf *trace) {
  VolumePixel<T> px{};
  px.theta = V3<T>{0, 0, 0};
  px.phi = T(1);
  Ray<T> ray = ray_in;
  T tnear, tfar;
  if (ray.intersect(sh.bbox, tnear, tfar)) {
    if (tnear < 0) tnear = 0;
    if (tfar < 0) tfar = 0;
    ray.o += ray.d * tnear;
    px.o = ray.o;
    px.d = ray.d;
    if (sh.schedule == SCHED_LATTICE) {
      px.lf = std::min((tfar - tnear) / sh.step, T(sh.max_steps));
      px.step_size = 0;
      volume_main_lattice(sh, px, nsamples);
    } else {
      px.nsteps = std::min(sh.max_steps, int((tfar - tnear) / sh.step));
      px.step_size = 1;
      volume_main(sh, px, nsamples, trace);
    }
  } else {
    px.o = ray.o;
    px.d = ray.d;
    px.nsteps = 0;
    px.lf = T(0);
    px.step_size = sh.schedule == SCHED_LATTICE ? 0 : 1;  // lattice: next index 0
  }
  out = px;
}

// ------------------------------------------------------- blocks (A24) ------
template <class T> struct BlocksShader {
  Box<T> bbox;
  T step;
  int max_steps;
  int mode;          // 0 Volume, 1 Solid (ui/blocks.schema.hpp:11-12), 2 the Paging renderer's view
  T density;
  const int *cheb; I3 dim;
  const float *mean;
  const int *vaddr; I3 vdim;   // mode 2: page table
  int lowest;                  // mode 2: BlockPaging::lowest_blkcnt
};

// VisBlocksShaderKernel::main, src/slave/renderers/blocks/blocks_shader.cu:16-50;
// mode 2: PagingShaderKernel::main, src/slave/renderers/paging/paging_shader.cu:24-61.
template <class T> void blocks_emit(const BlocksShader<T> &sh, const Ray<T> &ray_in, VolumePixel<T> &out) {
  VolumePixel<T> px{};
  Ray<T> ray = ray_in;
  T tnear, tfar;
  if (!ray.intersect(sh.bbox, tnear, tfar)) {
    px.o = ray.o; px.d = ray.d; px.nsteps = 0; px.step_size = 1; out = px; return;
  }
  if (tnear < 0) tnear = 0;
  if (tfar < 0) tfar = 0;
  ray.o += ray.d * tnear;
  px.nsteps = std::min(sh.max_steps, int((tfar - tnear) / sh.step));
  px.step_size = 1;
  const T cdu = T(1) / comp_max(vabs(ray.d));
  const T opacity_threshold = T(0.95f);
  int &nsteps = px.nsteps;
  while (nsteps > 0) {
    V3<T> ip = vfloor(ray.o);
    int cd = grid_lookup(sh.cheb, sh.dim, ip);
    if (cd) {
      nsteps -= skip_nblock_steps(ray, ip, cd, cdu, sh.step);
    } else if (sh.mode == 2) {
      int pgid = grid_lookup(sh.vaddr, sh.vdim, ip);
      if (pgid == -1) break;   // page fault
      V4<T> col = pgid >= sh.lowest ? V4<T>{T(0), T(1), T(0), T(1)} * T(.5f) : V4<T>{T(1), T(0), T(0), T(1)} * T(.2f);
      nsteps -= skip_nblock_steps(ray, ip, 1, cdu, sh.step);
      px.v += col * (T(1) - px.v.w);
      if (px.v.w > opacity_threshold) { nsteps = 0; break; }
    } else {
      V3<T> lo = sh.bbox.min, hi = sh.bbox.max - T(1);
      V3<T> c{gclamp(ip.x, lo.x, hi.x), gclamp(ip.y, lo.y, hi.y), gclamp(ip.z, lo.z, hi.z)};
      V3<T> rip = c / (sh.bbox.max - sh.bbox.min);
      if (sh.mode == 0) {
        T mean = grid_lookup_f(sh.mean, sh.dim, ip);
        V4<T> col = V4<T>{rip.x, rip.y, rip.z, T(1)} * sh.density * mean;
        px.v += col * (T(1) - px.v.w);
        if (px.v.w > opacity_threshold) { nsteps = 0; break; }
      } else {
        px.v = V4<T>{rip.x, rip.y, rip.z, T(1)};
        nsteps = 0;
        break;
      }
    }
    ray.o += ray.d * sh.step;
    nsteps -= 1;
  }
  px.o = ray.o; px.d = ray.d;
  out = px;
}

// ----------------------------------------------------- isosurface (A25) ----
template <class T> struct IsoShader {
  Box<T> bbox;
  T step, du;
  int max_steps;
  int mode;             // 0 Color, 1 Normal, 2 Position (ui/isosurface.schema.hpp:11-12)
  M4<T> to_world;
  V3<T> eye_pos, light_pos, surface_color;
  T isovalue;
  const int *cheb; I3 cheb_dim;
  Paging paging;
  int corrected;        // 0: literal reference; 1: bracketing + bisection
  int bisect_steps;
};

template <class T> struct IsoPixel {
  V3<T> o, d;
  int nsteps, step_size;
  V4<T> v;
  V3<T> origin;
  T depth;
  int hit;
};

template <class T> inline T gsign(T x) { return T((T(0) < x) - (x < T(0))); }

template <class T> inline T iso_sample(const IsoShader<T> &sh, int pgid, V3<T> x) {
  T k = T(sh.paging.page_k[pgid]);
  V3<T> b{T(sh.paging.page_b[3 * pgid]), T(sh.paging.page_b[3 * pgid + 1]), T(sh.paging.page_b[3 * pgid + 2])};
  int pad = sh.paging.pad;
  return brick_sample(sh.paging.pool + (size_t)pgid * pad * pad * pad, pad, x * k + b);
}

// IsosurfaceShaderKernel::linear_to_srgb (double arithmetic on float input),
// isosurface_shader.cu:25-32.
template <class T> inline T iso_srgb(T linear) {
  if (double(linear) <= 0.0031308) return T(12.92 * double(linear));
  return T((1.0 + 0.055) * std::pow(double(linear), 1.0 / 2.4) - 0.055);
}

// IsosurfaceShaderKernel::main, isosurface_shader.cu:50-125 (corrected == 0),
// or the build-defined bracketing + bisection form (corrected == 1).
template <class T> void iso_emit(const IsoShader<T> &sh, const Ray<T> &ray_in, IsoPixel<T> &out, long long *nsamples) {
  IsoPixel<T> px{};
  px.origin = ray_in.o;
  px.depth = T(INFINITY);
  Ray<T> ray = ray_in;
  T tnear, tfar;
  if (!ray.intersect(sh.bbox, tnear, tfar)) {
    px.o = ray.o; px.d = ray.d; px.nsteps = 0; px.step_size = 1; out = px; return;
  }
  if (tnear < 0) tnear = 0;
  if (tfar < 0) tfar = 0;
  ray.o += ray.d * tnear;
  px.nsteps = std::min(sh.max_steps, int((tfar - tnear) / sh.step));
  px.step_size = 1;
  const T cdu = T(1) / comp_max(vabs(ray.d));
  T prev_value = T(0);
  bool have_prev = false;
  long long ns = 0;
  int &nsteps = px.nsteps;
  while (nsteps > 0) {
    V3<T> ip = vfloor(ray.o);
    int cd = grid_lookup(sh.cheb, sh.cheb_dim, ip);
    if (cd) {
      nsteps -= skip_nblock_steps(ray, ip, cd, cdu, sh.step);
      have_prev = false;
    } else {
      int pgid = grid_lookup(sh.paging.vaddr, sh.paging.dim, ip);
      if (pgid == -1) break;
      T value = iso_sample(sh, pgid, ray.o - ip);
      ++ns;
      bool crossing;
      if (sh.corrected) {
        crossing = have_prev && gsign(value - sh.isovalue) != gsign(prev_value - sh.isovalue);
      } else {
        crossing = gsign(value - sh.isovalue) != gsign(prev_value - sh.isovalue);
      }
      if (crossing) {
        V3<T> p = ray.o;
        V3<T> d = ray.d;
        T dt = sh.step;
        px.v = V4<T>{sh.surface_color.x, sh.surface_color.y, sh.surface_color.z, T(1)};
        V3<T> inter_p;
        V3<T> nn;
        V3<T> gp_ip = ip;
        if (!sh.corrected) {
          T a = (sh.isovalue - prev_value) / (value - prev_value);
          inter_p = (T(1) - a) * (p - dt * d) + a * p;
          int pg = pgid;
          V3<T> q = inter_p - ip;
          T h = sh.du * T(2);
          auto s = [&](V3<T> y) { return iso_sample(sh, pg, y); };
          nn = V3<T>{s(q - V3<T>{h, 0, 0}) - s(q + V3<T>{h, 0, 0}), s(q - V3<T>{0, h, 0}) - s(q + V3<T>{0, h, 0}),
                     s(q - V3<T>{0, 0, h}) - s(q + V3<T>{0, 0, h})};
        } else {
          // bisection on [p - dt*d, p]; the low end has the sign of prev_value
          T lo = T(0), hi = dt;            // distance back from p: hi = prev sample
          V3<T> plo = p, phi_ = p - dt * d; // plo: current (crossed) side, phi_: previous side
          (void)lo; (void)hi;
          T sprev = gsign(prev_value - sh.isovalue);
          for (int it = 0; it < sh.bisect_steps; ++it) {
            V3<T> mid = (plo + phi_) * T(0.5);
            V3<T> mip = vfloor(mid);
            int mpg = grid_lookup(sh.paging.vaddr, sh.paging.dim, mip);
            T mv = mpg == -1 ? value : iso_sample(sh, mpg, mid - mip);
            if (gsign(mv - sh.isovalue) == sprev) phi_ = mid; else plo = mid;
          }
          inter_p = (plo + phi_) * T(0.5);
          gp_ip = vfloor(inter_p);
          int gpg = grid_lookup(sh.paging.vaddr, sh.paging.dim, gp_ip);
          if (gpg == -1) gpg = pgid, gp_ip = ip;
          V3<T> q = inter_p - gp_ip;
          T h = sh.du;
          auto s = [&](V3<T> y) { return iso_sample(sh, gpg, y); };
          // outward normal: gradient oriented toward the previous (outside) side
          V3<T> g{s(q + V3<T>{h, 0, 0}) - s(q - V3<T>{h, 0, 0}), s(q + V3<T>{0, h, 0}) - s(q - V3<T>{0, h, 0}),
                  s(q + V3<T>{0, 0, h}) - s(q - V3<T>{0, 0, h})};
          nn = g * sprev;
        }
        V3<T> world_p = xyz(mul(sh.to_world, V4<T>{inter_p.x, inter_p.y, inter_p.z, T(1)}));
        V3<T> n = normalize(xyz(mul(sh.to_world, V4<T>{nn.x, nn.y, nn.z, T(0)})));
        V3<T> light_dir = normalize(sh.light_pos - world_p);
        V3<T> h = sh.corrected ? normalize(light_dir + normalize(sh.eye_pos - world_p)) : normalize(light_dir - world_p);
        const T ambient = T(0.2f);
        T diffuse = T(.6f) * gclamp(dot(light_dir, n), T(0), T(1));
        T specular = T(.2f) * std::pow(gclamp(dot(h, n), T(0), T(1)), T(100));
        T distance = length(world_p - sh.eye_pos) / T(2);
        if (sh.mode == 0) {
          px.v = px.v * (ambient + (diffuse + specular) / distance);
          px.v = V4<T>{iso_srgb(px.v.x), iso_srgb(px.v.y), iso_srgb(px.v.z), px.v.w};
        } else if (sh.mode == 2) {
          V3<T> q = inter_p / sh.bbox.max;
          px.v = V4<T>{q.x, q.y, q.z, T(1)};
        } else {
          px.v = V4<T>{n.x, n.y, n.z, T(1)};
        }
        px.depth = sh.corrected ? length(inter_p - px.origin) : length(px.origin - ray.o);
        px.hit = 1;
        nsteps = 0;
        break;
      }
      if (sh.corrected) { prev_value = value; have_prev = true; }
    }
    ray.o += ray.d * sh.step;
    nsteps -= 1;
  }
  px.o = ray.o; px.d = ray.d;
  out = px;
  if (nsamples) *nsamples += ns;
}

// ------------------------------------------------------ pixel / fetch ------
// StdVec4Pixel::write_to, include/hydrant/pixel_template.hpp:11-20.
template <class T> inline void write_rgb8(V4<T> v, const uint8_t clear[3], uint8_t *dst) {
  if (v.w != T(0)) {
    T c[3] = {v.x, v.y, v.z};
    for (int i = 0; i < 3; ++i) dst[i] = (uint8_t)gclamp(c[i] * T(255.f), T(0), T(255));
  } else {
    dst[0] = clear[0]; dst[1] = clear[1]; dst[2] = clear[2];
  }
}

// linear_to_srgb, src/slave/renderers/volume/volume.cc:175-181 (float).
inline float srgb_f(float x) {
  if (x <= 0.0031308f) return 12.92f * x;
  return 1.055f * std::pow(x, 1.f / 2.4f) - 0.055f;
}

// --------------------------------------------------------- dispatch --------
// CPU dispatch: nthreads threads, thread i renders rows y = i (mod nthreads),
// created and joined per pass (src/slave/core/cpu_shading.cc:9-31).
template <class F> void row_interleave(int H, int nthreads, F &&fn) {
  if (nthreads <= 1) { for (int y = 0; y < H; ++y) fn(y); return; }
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i)
    th.emplace_back([&, i] { for (int y = i; y < H; y += nthreads) fn(y); });
  for (auto &t : th) t.join();
}

}  // namespace orc
