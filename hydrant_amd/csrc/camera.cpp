// Camera, exhibit and view setup on the host (binary32, glm 0.9.9 operation
// order, compiled with -ffp-contract=off).
//   Camera::update_params(CameraOrbit)   include/hydrant/core/scene.hpp:68-81
//   Camera(CameraConfig) ctg_fovy_2      scene.hpp:53
//   Camera::get_ivt = inverse(lookAt)    scene.hpp:58
//   Exhibit::get_iet                     scene.hpp:20-29
//   Raycaster::fill_ray_emit_args        include/hydrant/core/raycaster.hpp:126-135
#include <cmath>
#include <cstring>

#include "hyd_internal.hpp"

namespace hyd {
namespace {

struct M4 { float m[16]; };  // column-major

M4 identity() {
  M4 r{};
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.f;
  return r;
}

// glm mat4 * mat4: column i = ((A0*b0 + A1*b1) + A2*b2) + A3*b3.
M4 matmul(const M4 &a, const M4 &b) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int row = 0; row < 4; ++row) {
      float acc = a.m[row] * b.m[4 * i];
      acc = acc + a.m[4 + row] * b.m[4 * i + 1];
      acc = acc + a.m[8 + row] * b.m[4 * i + 2];
      acc = acc + a.m[12 + row] * b.m[4 * i + 3];
      r.m[4 * i + row] = acc;
    }
  return r;
}

float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

// glm rotate(identity, angle, axis).
M4 rotate_identity(float angle, f3 v) {
  const float c = std::cos(angle), s = std::sin(angle);
  const f3 axis = normalize3(v);
  const f3 t = muls(axis, 1.f - c);
  const float r[3][3] = {{c + t.x * axis.x, t.x * axis.y + s * axis.z, t.x * axis.z - s * axis.y},
                         {t.y * axis.x - s * axis.z, c + t.y * axis.y, t.y * axis.z + s * axis.x},
                         {t.z * axis.x + s * axis.y, t.z * axis.y - s * axis.x, c + t.z * axis.z}};
  const M4 id = identity();
  M4 out;
  for (int i = 0; i < 3; ++i)
    for (int row = 0; row < 4; ++row)
      out.m[4 * i + row] = id.m[row] * r[i][0] + id.m[4 + row] * r[i][1] + id.m[8 + row] * r[i][2];
  for (int row = 0; row < 4; ++row) out.m[12 + row] = id.m[12 + row];
  return out;
}

// glm lookAtRH.
M4 look_at(f3 eye, f3 center, f3 up) {
  const f3 f = normalize3(sub(center, eye));
  const f3 s = normalize3(cross3(f, up));
  const f3 u = cross3(s, f);
  M4 r = identity();
  r.m[0] = s.x; r.m[4] = s.y; r.m[8] = s.z;
  r.m[1] = u.x; r.m[5] = u.y; r.m[9] = u.z;
  r.m[2] = -f.x; r.m[6] = -f.y; r.m[10] = -f.z;
  r.m[12] = -dot3(s, eye);
  r.m[13] = -dot3(u, eye);
  r.m[14] = dot3(f, eye);
  return r;
}

// glm compute_inverse<4,4>: cofactors, sign pattern, 1/det from row 0.
M4 inverse(const M4 &M) {
  auto m = [&](int c, int r) { return M.m[4 * c + r]; };
  const float c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
  const float c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
  const float c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
  const float c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
  const float c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
  const float c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
  const float c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
  const float c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
  const float c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
  const float c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
  const float c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
  const float c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
  const float c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
  const float c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
  const float c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
  const float c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
  const float c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
  const float c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
  const float fac[6][4] = {{c00, c00, c02, c03}, {c04, c04, c06, c07}, {c08, c08, c10, c11},
                           {c12, c12, c14, c15}, {c16, c16, c18, c19}, {c20, c20, c22, c23}};
  const float vec[4][4] = {{m(1, 0), m(0, 0), m(0, 0), m(0, 0)},
                           {m(1, 1), m(0, 1), m(0, 1), m(0, 1)},
                           {m(1, 2), m(0, 2), m(0, 2), m(0, 2)},
                           {m(1, 3), m(0, 3), m(0, 3), m(0, 3)}};
  // Inv0 = Vec1*Fac0 - Vec2*Fac1 + Vec3*Fac2, Inv1 = Vec0*Fac0 - Vec2*Fac3 + Vec3*Fac4,
  // Inv2 = Vec0*Fac1 - Vec1*Fac3 + Vec3*Fac5, Inv3 = Vec0*Fac2 - Vec1*Fac4 + Vec2*Fac5
  const int terms[4][6] = {{1, 0, 2, 1, 3, 2}, {0, 0, 2, 3, 3, 4}, {0, 1, 1, 3, 3, 5}, {0, 2, 1, 4, 2, 5}};
  const float sign[4][4] = {{+1, -1, +1, -1}, {-1, +1, -1, +1}, {+1, -1, +1, -1}, {-1, +1, -1, +1}};
  M4 inv;
  for (int c = 0; c < 4; ++c) {
    const int *t = terms[c];
    for (int r = 0; r < 4; ++r) {
      const float v = (vec[t[0]][r] * fac[t[1]][r] - vec[t[2]][r] * fac[t[3]][r]) + vec[t[4]][r] * fac[t[5]][r];
      inv.m[4 * c + r] = v * sign[c][r];
    }
  }
  const float row0[4] = {inv.m[0], inv.m[4], inv.m[8], inv.m[12]};
  const float d0[4] = {M.m[0] * row0[0], M.m[1] * row0[1], M.m[2] * row0[2], M.m[3] * row0[3]};
  const float det = (d0[0] + d0[1]) + (d0[2] + d0[3]);
  const float one_over = 1.f / det;
  for (float &x : inv.m) x = x * one_over;
  return inv;
}

M4 exhibit_iet(f3 center, f3 size) {
  const f3 a = abs3(sub(center, size)), b = abs3(center);
  const f3 d = {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)};
  const float scale = comp_max(d);
  M4 r{};
  r.m[0] = r.m[5] = r.m[10] = scale;
  r.m[12] = center.x; r.m[13] = center.y; r.m[14] = center.z; r.m[15] = 1.f;
  return r;
}

f3 load3(const float *p) { return f3{p[0], p[1], p[2]}; }

}  // namespace
}  // namespace hyd

using namespace hyd;

extern "C" {

int32_t hyd_orbit_camera(const float arm[3], const float center[3], float fovy, hyd_camera *out) {
  HYD_REQUIRE(arm && out, HYD_ERR_ARG, "hyd_orbit_camera: null argument");
  float p4[4], q4[4];
  M4 rz = rotate_identity(radians(arm[1]), f3{0.f, 0.f, 1.f});
  mat4_vec4(rz.m, arm[2], 0.f, 0.f, 1.f, p4);
  M4 ry = rotate_identity(radians(arm[0]), f3{0.f, 1.f, 0.f});
  mat4_vec4(ry.m, p4[0], p4[1], p4[2], 1.f, q4);
  out->position[0] = q4[0]; out->position[1] = q4[1]; out->position[2] = q4[2];
  const float zero[3] = {0.f, 0.f, 0.f};
  const float *c = center ? center : zero;
  out->target[0] = c[0]; out->target[1] = c[1]; out->target[2] = c[2];
  out->up[0] = 0.f; out->up[1] = 1.f; out->up[2] = 0.f;
  out->fovy = fovy;
  return HYD_OK;
}

int32_t hyd_make_view(const hyd_camera *cam, const float ex_center[3], const float ex_size[3], hyd_view *out) {
  HYD_REQUIRE(cam && ex_center && ex_size && out, HYD_ERR_ARG, "hyd_make_view: null argument");
  HYD_REQUIRE(cam->fovy > 0.f && cam->fovy < 180.f, HYD_ERR_ARG, "fovy %g outside (0, 180) degrees", cam->fovy);
  const M4 iet = exhibit_iet(load3(ex_center), load3(ex_size));
  const M4 ivt = inverse(look_at(load3(cam->position), load3(cam->target), load3(cam->up)));
  const M4 trans = matmul(iet, ivt);
  hyd_view v{};
  std::memcpy(v.trans, trans.m, sizeof(trans.m));
  v.ctg_fovy_2 = 1.f / std::tan(radians(cam->fovy / 2.f));
  float o4[4];
  mat4_vec4(iet.m, cam->position[0], cam->position[1], cam->position[2], 1.f, o4);
  v.ray_o[0] = o4[0]; v.ray_o[1] = o4[1]; v.ray_o[2] = o4[2];
  // the reference renders NaN rays here; a degenerate camera (position on the
  // target, up along the view direction) is refused instead
  bool finite = std::isfinite(v.ctg_fovy_2);
  for (int i = 0; i < 16; ++i) finite = finite && std::isfinite(v.trans[i]);
  for (int i = 0; i < 3; ++i) finite = finite && std::isfinite(v.ray_o[i]);
  HYD_REQUIRE(finite, HYD_ERR_ARG,
              "degenerate camera: position equals target, or up is parallel to the view direction");
  *out = v;
  return HYD_OK;
}

int32_t hyd_exhibit_iet(const float ex_center[3], const float ex_size[3], float out[16]) {
  HYD_REQUIRE(ex_center && ex_size && out, HYD_ERR_ARG, "hyd_exhibit_iet: null argument");
  const M4 r = exhibit_iet(load3(ex_center), load3(ex_size));
  std::memcpy(out, r.m, sizeof(r.m));
  return HYD_OK;
}

int32_t hyd_mat4_inverse(const float m[16], float out[16]) {
  HYD_REQUIRE(m && out, HYD_ERR_ARG, "hyd_mat4_inverse: null argument");
  M4 a;
  std::memcpy(a.m, m, sizeof(a.m));
  const M4 r = inverse(a);
  std::memcpy(out, r.m, sizeof(r.m));
  return HYD_OK;
}

}  // extern "C"
