// The reference CPU sampler (BlockSamplerMapping + CpuSampler Wrap/Linear over
// a padded brick, block_paging.hpp:11-36, cpu_sampler.hpp:88-104,142-166)
// split into the steps a march pipelines: tap offsets and weights of a
// sample, the loads of its taps, their conversion to float, and the trilinear
// interpolation.  Shared by the volume and isosurface marches.
#pragma once

#include <hip/hip_fp16.h>

#include "hyd_internal.hpp"

namespace hyd {

// Reference sampler taps of one sample: the element offsets of the four rows
// (y, z) = (i, i), (j, i), (i, j), (j, j) of the padded brick, the x taps and the
// weights (BlockSamplerMapping + CpuSampler Wrap/Linear, block_paging.hpp:11-36,
// cpu_sampler.hpp:88-104,142-166).
template <class OFF> struct RefTap {
  OFF r00, r10, r01, r11;
  int ix, jx;
  float ax, ay, az;
};

// floor(fx) and ceil(fx) are integers, so clamping them to [0, pad - 1] as floats
// (one med3) and converting gives the reference's clamped int taps.  Row
// offsets below 2^24 voxels per page use 24-bit multiplies (pad <= 1024, so
// z * pad + y < 2^20 and a row start < 2^30).
template <class OFF>
__device__ __forceinline__ RefTap<OFF> ref_tap(f3 x, float4 pg, int pad, OFF page_base) {
  const float fd = float(pad);
  const float hi = fd - 1.f;
  const float u[3] = {x.x * pg.x + pg.y, x.y * pg.x + pg.z, x.z * pg.x + pg.w};
  float a[3];
  uint32_t i[3], j[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float fx = fd * u[k];
    fx = wrap_mod(fx, fd);   // Wrap: glm mod(fx, fd)
    fx = fx - 0.5f;
    const float flr = floorf(fx);
    a[k] = fx - flr;
    i[k] = (uint32_t)__builtin_amdgcn_fmed3f(flr, 0.f, hi);
    j[k] = (uint32_t)__builtin_amdgcn_fmed3f(ceilf(fx), 0.f, hi);
  }
  RefTap<OFF> t;
  const uint32_t p = (uint32_t)pad;
  t.r00 = page_base + (OFF)__umul24(__umul24(i[2], p) + i[1], p);
  t.r10 = page_base + (OFF)__umul24(__umul24(i[2], p) + j[1], p);
  t.r01 = page_base + (OFF)__umul24(__umul24(j[2], p) + i[1], p);
  t.r11 = page_base + (OFF)__umul24(__umul24(j[2], p) + j[1], p);
  t.ix = (int)i[0];
  t.jx = (int)j[0];
  t.ax = a[0];
  t.ay = a[1];
  t.az = a[2];
  return t;
}

// Raw taps as loaded.  Every row is read as an x pair: the two voxels at
// m = min(ix, pad - 2) and m + 1 in one load (8 bytes at 4-byte alignment for
// FP32, 4 bytes at 2-byte alignment for FP16, 2 bytes at any alignment for
// u8), which the vector memory unit serves at the per-lane cost of one of the
// two single loads; the taps are then v(ix) = ix == m ? lo : hi and
// v(jx) = jx == m + 1 ? hi : lo, which also covers the clamped and integral
// cases (ix == jx).  Each load has a register of its own (pairs of rows packed
// into one register would make each load wait for the previous one).  A
// grouped u8 march issues 8 samples x 4 row loads per group; with one load per
// tap it issued 64, and the vector memory unit's per-instruction cost (one
// cache-line lookup per distinct line a wave's lanes touch) set the time of a
// group, not the memory latency.  Values are converted to float once, when a
// sample uses them.
template <int FMT> struct RawVox;
template <> struct RawVox<HYD_FMT_F32> { using T = float; static constexpr int N = 8; };
template <> struct RawVox<HYD_FMT_F16> { using T = uint16_t; static constexpr int N = 4; };
template <> struct RawVox<HYD_FMT_U8> { using T = uint8_t; static constexpr int N = 4; };

template <int FMT> struct RawTaps { uint32_t v[RawVox<FMT>::N]; };

// Byte offsets are formed in OFF, so with 32-bit offsets a load is the pool's
// scalar base plus one 32-bit register.
template <int FMT, class OFF>
__device__ __forceinline__ void ref_fetch(const void *pool, const RefTap<OFF> &t, int pad, RawTaps<FMT> &r) {
  using T = typename RawVox<FMT>::T;
  const char *b = static_cast<const char *>(pool);
  constexpr OFF sz = (OFF)sizeof(T);
  const OFF rows[4] = {t.r00 * sz, t.r10 * sz, t.r01 * sz, t.r11 * sz};
  const OFF xm = (OFF)min(t.ix, pad - 2) * sz;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if constexpr (FMT == HYD_FMT_F32) {
      __builtin_memcpy(&r.v[2 * k], b + (rows[k] + xm), 8);
    } else if constexpr (FMT == HYD_FMT_F16) {
      __builtin_memcpy(&r.v[k], b + (rows[k] + xm), 4);
    } else {
      uint16_t pair;
      __builtin_memcpy(&pair, b + (rows[k] + xm), 2);
      r.v[k] = pair;
    }
  }
}

// The eight tap values of a fresh sample, x-y-z order.
template <int FMT, class OFF>
__device__ __forceinline__ void ref_values(const RawTaps<FMT> &r, const RefTap<OFF> &t, int pad, const float *lut,
                                           float (&v)[8]) {
  const int m = min(t.ix, pad - 2);
  const bool i_lo = t.ix == m, j_hi = t.jx == m + 1;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if constexpr (FMT == HYD_FMT_U8) {
      // select the bytes first: two table reads per row, as with single loads
      const uint32_t lo = r.v[k] & 0xffu, hi = r.v[k] >> 8;
      v[2 * k] = lut[i_lo ? lo : hi];
      v[2 * k + 1] = lut[j_hi ? hi : lo];
    } else {
      float lo, hi;
      if constexpr (FMT == HYD_FMT_F32) {
        lo = __uint_as_float(r.v[2 * k]);
        hi = __uint_as_float(r.v[2 * k + 1]);
      } else {
        lo = __half2float(__ushort_as_half((unsigned short)(r.v[k] & 0xffffu)));
        hi = __half2float(__ushort_as_half((unsigned short)(r.v[k] >> 16)));
      }
      v[2 * k] = i_lo ? lo : hi;
      v[2 * k + 1] = j_hi ? hi : lo;
    }
  }
}

// x -> y -> z lerps with x * (1 - t) + y * t (cpu_sampler.hpp:126-129).
template <class OFF>
__device__ __forceinline__ float ref_interp(const float (&v)[8], const RefTap<OFF> &t) {
  const float bx = 1.f - t.ax;
  const float x00 = v[0] * bx + v[1] * t.ax;
  const float x10 = v[2] * bx + v[3] * t.ax;
  const float x01 = v[4] * bx + v[5] * t.ax;
  const float x11 = v[6] * bx + v[7] * t.ax;
  const float by = 1.f - t.ay;
  const float y0 = x00 * by + x10 * t.ay;
  const float y1 = x01 * by + x11 * t.ay;
  const float bz = 1.f - t.az;
  return y0 * bz + y1 * t.az;
}


}  // namespace hyd
