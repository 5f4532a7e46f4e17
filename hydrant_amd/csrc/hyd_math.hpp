// Host/device math for the render path.
//
// Every helper is one IEEE binary32 operation per component in the order the
// reference's glm calls evaluate them (glm 0.9.9 semantics as used by
// include/hydrant/core/glm_math.hpp and scene.hpp).  All sources are compiled
// with -ffp-contract=off, so nothing here is fused into an FMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD __host__ __device__ __forceinline__

namespace hyd {

struct f3 { float x, y, z; };
struct f4v { float x, y, z, w; };

// glm min/max/clamp: min(x,y) = (y < x) ? y : x, max(x,y) = (x < y) ? y : x.
HD float gmin(float x, float y) { return (y < x) ? y : x; }
HD float gmax(float x, float y) { return (x < y) ? y : x; }
HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
// Integer clamp for lo <= hi (one v_med3_i32 on the device).
HD int iclamp(int x, int lo, int hi) { return min(max(x, lo), hi); }

// glm mod(x, d) = x - d * floor(x / d) for d > 0 (the sampler's Wrap mode).
// For 0 <= x <= d (1 - 2^-20) the correctly rounded quotient is below 1, so
// the result is x itself and the division is skipped; otherwise the full
// expression is evaluated.  Bit-identical to the expression for every x.
HD float wrap_mod(float x, float d) {
  if (x >= 0.f && x <= d * 0.99999904632568359375f) return x;
  return x - d * floorf(x / d);
}

HD f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HD f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HD f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
HD f3 muls(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
HD f3 floor3(f3 a) { return {floorf(a.x), floorf(a.y), floorf(a.z)}; }
HD f3 abs3(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
HD float comp_max(f3 v) { float r = v.x; r = gmax(r, v.y); r = gmax(r, v.z); return r; }
HD float comp_min(f3 v) { float r = v.x; r = gmin(r, v.y); r = gmin(r, v.z); return r; }
// glm dot(vec3): (x*x' + y*y') + z*z'
HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HD f3 cross3(f3 x, f3 y) { return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y}; }
// glm normalize: v * (1 / sqrt(dot(v, v))), sqrt and divide correctly rounded.
HD f3 normalize3(f3 v) { return muls(v, 1.f / sqrtf(dot3(v, v))); }
HD float length3(f3 v) { return sqrtf(dot3(v, v)); }

// Ray/box slab test (Ray::intersect, include/hydrant/core/glm_math.hpp:58-72).
// invr = 1/d, tbot/ttop, min/max with glm NaN semantics, compMax/compMin.
HD bool intersect_box(f3 o, f3 d, f3 bmin, f3 bmax, float &tnear, float &tfar) {
  f3 invr = {1.f / d.x, 1.f / d.y, 1.f / d.z};
  f3 tbot = mul(invr, sub(bmin, o));
  f3 ttop = mul(invr, sub(bmax, o));
  f3 tmin = {gmin(ttop.x, tbot.x), gmin(ttop.y, tbot.y), gmin(ttop.z, tbot.z)};
  f3 tmax = {gmax(ttop.x, tbot.x), gmax(ttop.y, tbot.y), gmax(ttop.z, tbot.z)};
  tnear = comp_max(tmin);
  tfar = comp_min(tmax);
  return tfar > tnear;
}

// intersect_box with the ray's reciprocal direction invr = 1 / d (componentwise,
// correctly rounded) computed once per ray: the same operations on the same
// values, so the same result.
HD void intersect_box_inv(f3 o, f3 invr, f3 bmin, f3 bmax, float &tnear, float &tfar) {
  f3 tbot = mul(invr, sub(bmin, o));
  f3 ttop = mul(invr, sub(bmax, o));
  f3 tmin = {gmin(ttop.x, tbot.x), gmin(ttop.y, tbot.y), gmin(ttop.z, tbot.z)};
  f3 tmax = {gmax(ttop.x, tbot.x), gmax(ttop.y, tbot.y), gmax(ttop.z, tbot.z)};
  tnear = comp_max(tmin);
  tfar = comp_min(tmax);
}
HD f3 recip3(f3 d) { return {1.f / d.x, 1.f / d.y, 1.f / d.z}; }

// Column-major 4x4, m[4*col + row].
// glm mat4 * vec4: (m0*x + m1*y) + (m2*z + m3*w).
HD void mat4_vec4(const float *m, float x, float y, float z, float w, float out[4]) {
  for (int r = 0; r < 4; ++r) {
    float a0 = m[r] * x + m[4 + r] * y;
    float a1 = m[8 + r] * z + m[12 + r] * w;
    out[r] = a0 + a1;
  }
}

// Ray direction of pixel (x, y) (ray_emit_kernel_impl, cuda_shading.cu:17-29;
// cpu_shading.cc:16-26): integer pixel corners, uv scaled by the height.
HD f3 pixel_ray_dir(const float *trans, float ctg, f3 ray_o, int x, int y, int W, int H) {
  float ccx = float(W) / 2.f, ccy = float(H) / 2.f;
  float ux = ((float(x) - ccx) * 2.f) / float(H);
  float uy = ((float(y) - ccy) * 2.f) / float(H);
  float p[4];
  mat4_vec4(trans, ux, -uy, -ctg, 1.f, p);
  return normalize3(sub(f3{p[0], p[1], p[2]}, ray_o));
}

}  // namespace hyd
