// Volume ray-march kernels for gfx950.
//
// One lane per pixel; a 64-lane workgroup covers an 8x8 pixel tile so a
// wavefront marches a coherent ray packet.  Ray emit, box clip and the march
// loop are fused (the reference splits them over ray_emit_kernel_impl and a
// device-function-pointer call to VolumeShaderKernel::main,
// src/slave/core/cuda_shading.cu:14-37 and volume_shader.cu:32-81).
//
// Sampling is done in software with the reference CPU backend's exact
// arithmetic (CpuSampler, include/hydrant/bridge/cpu_sampler.hpp:88-166):
// hardware texture filtering would not reproduce it.  The transfer function
// and the u8 normalisation table are staged in LDS per workgroup.
#include <hip/hip_fp16.h>

#include <cstdlib>
#include <type_traits>
#include <utility>

#include "hyd_internal.hpp"
#include "hyd_sampler.hpp"

namespace hyd {
namespace {

constexpr int kTile = 8;           // 8x8 pixel packet per wavefront
constexpr int kMaxTf = 256;        // LDS capacity of the TF table (entries)

template <int FMT> struct Voxel;
template <> struct Voxel<HYD_FMT_F32> {
  static __device__ __forceinline__ float at(const void *base, size_t i, const float *) {
    return static_cast<const float *>(base)[i];
  }
};
template <> struct Voxel<HYD_FMT_F16> {
  static __device__ __forceinline__ float at(const void *base, size_t i, const float *) {
    return __half2float(static_cast<const __half *>(base)[i]);
  }
};
template <> struct Voxel<HYD_FMT_U8> {
  // float(v) / 255.f, correctly rounded, from the LDS table
  static __device__ __forceinline__ float at(const void *base, size_t i, const float *lut) {
    return lut[static_cast<const uint8_t *>(base)[i]];
  }
};

// Point lookup in an int grid with Clamp addressing (Options::as_array,
// basic_renderer.hpp:147-159; CpuSampler filter_none).
__device__ __forceinline__ int grid_index(f3 p, const int32_t dim[3]) {
  float fx = gclamp(p.x, 0.f, float(dim[0]));
  float fy = gclamp(p.y, 0.f, float(dim[1]));
  float fz = gclamp(p.z, 0.f, float(dim[2]));
  int ix = iclamp((int)floorf(fx), 0, dim[0] - 1);
  int iy = iclamp((int)floorf(fy), 0, dim[1] - 1);
  int iz = iclamp((int)floorf(fz), 0, dim[2] - 1);
  return iz * dim[0] * dim[1] + iy * dim[0] + ix;
}

struct TapIdx { int ix, iy, iz, jx, jy, jz; };

// The last march sample's taps: first and last tap index (which fix all eight)
// and their values.
struct VolTaps {
  size_t id0 = ~(size_t)0, id7 = ~(size_t)0;
  float v[8];
};

// Brick sample: BlockSamplerMapping (k*x + b) then CpuSampler Wrap + Linear +
// normalised coordinates over a pad^3 brick (block_paging.hpp:11-36,
// cpu_sampler.hpp:88-104,142-166), for consecutive march samples of one ray:
// the eight taps are loaded only when they change (at half-voxel steps about
// half the samples share their predecessor's cell); the value does not depend
// on whether they were reloaded.
template <int FMT>
__device__ __forceinline__ float brick_sample_reuse(const void *pool, int pad, int pgid, float4 pg, f3 x,
                                                    const float *lut, TapIdx *taps, VolTaps &c) {
  const float fd = float(pad);
  float u[3] = {x.x * pg.x + pg.y, x.y * pg.x + pg.z, x.z * pg.x + pg.w};
  float a[3];
  int i[3], j[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float fx = fd * u[k];
    fx = wrap_mod(fx, fd);   // Wrap: glm mod(fx, fd)
    fx = fx - 0.5f;
    float flr = floorf(fx);
    a[k] = fx - flr;
    i[k] = iclamp((int)flr, 0, pad - 1);
    j[k] = iclamp((int)ceilf(fx), 0, pad - 1);
  }
  if (taps) *taps = TapIdx{i[0], i[1], i[2], j[0], j[1], j[2]};
  const size_t base = (size_t)pgid * pad * pad * pad;
  const size_t row_ii = base + ((size_t)i[2] * pad + i[1]) * pad;
  const size_t row_jj = base + ((size_t)j[2] * pad + j[1]) * pad;
  if (row_ii + i[0] != c.id0 || row_jj + j[0] != c.id7) {
    const size_t row_ji = base + ((size_t)i[2] * pad + j[1]) * pad;
    const size_t row_ij = base + ((size_t)j[2] * pad + i[1]) * pad;
    c.v[0] = Voxel<FMT>::at(pool, row_ii + i[0], lut); c.v[1] = Voxel<FMT>::at(pool, row_ii + j[0], lut);
    c.v[2] = Voxel<FMT>::at(pool, row_ji + i[0], lut); c.v[3] = Voxel<FMT>::at(pool, row_ji + j[0], lut);
    c.v[4] = Voxel<FMT>::at(pool, row_ij + i[0], lut); c.v[5] = Voxel<FMT>::at(pool, row_ij + j[0], lut);
    c.v[6] = Voxel<FMT>::at(pool, row_jj + i[0], lut); c.v[7] = Voxel<FMT>::at(pool, row_jj + j[0], lut);
    c.id0 = row_ii + i[0];
    c.id7 = row_jj + j[0];
  }
  const float ax = a[0], bx = 1.f - ax;
  float x00 = c.v[0] * bx + c.v[1] * ax;
  float x10 = c.v[2] * bx + c.v[3] * ax;
  float x01 = c.v[4] * bx + c.v[5] * ax;
  float x11 = c.v[6] * bx + c.v[7] * ax;
  const float ay = a[1], by = 1.f - ay;
  float y0 = x00 * by + x10 * ay;
  float y1 = x01 * by + x11 * ay;
  const float az = a[2], bz = 1.f - az;
  return y0 * bz + y1 * az;
}

// Transfer function: Border + Linear + normalised 1-D lookup
// (transfer_fn.hpp:14-28, cpu_sampler.hpp:117-124,142-166).
__device__ __forceinline__ float4 tf_sample(const float *tf, int n, float s) {
  const float fd = float(n);
  float fx = fd * s;
  if (!(0.f <= fx && fx <= fd)) return make_float4(0.f, 0.f, 0.f, 0.f);
  fx = fx - 0.5f;
  float flr = floorf(fx);
  float a = fx - flr;
  int ix = iclamp((int)flr, 0, n - 1);
  int jx = iclamp((int)ceilf(fx), 0, n - 1);
  const float w = 1.f - a;
  const float *p = tf + 4 * ix, *q = tf + 4 * jx;
  return make_float4(p[0] * w + q[0] * a, p[1] * w + q[1] * a, p[2] * w + q[2] * a, p[3] * w + q[3] * a);
}

struct MarchState {
  f3 o, d;
  int nsteps, step_size;
  float4 v;
  f3 theta;
  float phi;
};

// Shade one resolved sample and composite it front to back
// (volume_shader.cu:52-70, plus the density scale on the TF alpha).  theta and
// phi are carried only for a film (the fetch pass exports them); without one
// they are never observed and THETA = false leaves them out.
// shade_sample gives the premultiplied colour of a sample (it does not depend
// on the ray's state), composite_ub adds it behind what the ray holds.
template <bool WEIGHTED = false>
__device__ __forceinline__ float4 shade_sample(const VolumeArgs &a, const float *tf, float s, int pgid, float w) {
  float4 ub = tf_sample(tf, a.tf_n, s);
  ub.w = ub.w * a.density;
  if (WEIGHTED) ub.w = ub.w * w;
  if (a.mode == HYD_MODE_PARTITION) {
    float l = length3(f3{ub.x, ub.y, ub.z});
    float r = a.rank, q = 1.f - r;
    ub = make_float4((1.f * q + 0.f * r) * l, (0.f * q + 0.f * r) * l, (0.f * q + 1.f * r) * l, ub.w);
  } else if (a.mode == HYD_MODE_PAGING) {
    ub = pgid >= a.lowest_blkcnt ? make_float4(0.f, 1.f, 0.f, ub.w) : make_float4(1.f, 0.f, 0.f, ub.w);
  }
  const float aw = ub.w;
  ub.x = ub.x * aw; ub.y = ub.y * aw; ub.z = ub.z * aw;
  return ub;
}

template <bool THETA = true>
__device__ __forceinline__ void composite_ub(float4 ub, MarchState &st) {
  if (THETA) {
    st.theta.x = st.theta.x + ub.x * st.phi;
    st.theta.y = st.theta.y + ub.y * st.phi;
    st.theta.z = st.theta.z + ub.z * st.phi;
    st.phi = st.phi * (1.f - ub.w);
  }
  const float t = 1.f - st.v.w;
  st.v.x = st.v.x + ub.x * t;
  st.v.y = st.v.y + ub.y * t;
  st.v.z = st.v.z + ub.z * t;
  st.v.w = st.v.w + ub.w * t;
}

template <bool WEIGHTED = false, bool THETA = true>
__device__ __forceinline__ void shade_composite_impl(const VolumeArgs &a, const float *tf, float s, int pgid,
                                                     MarchState &st, float w) {
  composite_ub<THETA>(shade_sample<WEIGHTED>(a, tf, s, pgid, w), st);
}

__device__ __forceinline__ void shade_composite(const VolumeArgs &a, const float *tf, float s, int pgid,
                                                MarchState &st) {
  shade_composite_impl<false>(a, tf, s, pgid, st, 1.f);
}
__device__ __forceinline__ void shade_composite(const VolumeArgs &a, const float *tf, float s, int pgid,
                                                MarchState &st, float w) {
  shade_composite_impl<true>(a, tf, s, pgid, st, w);
}

// The chebyshev and page-table entries of the brick the last sample fell in:
// consecutive samples inside one brick read them once (the page entry only
// for a non-empty brick).
// With the combined per-brick code (equal grid dims) this is one load.
struct BrickLookup {
  f3 ip = {-INFINITY, -INFINITY, -INFINITY};
  int cd = 0, gi = 0, pg = -1;
  __device__ __forceinline__ void update(const VolumeArgs &a, f3 p) {
    if (p.x == ip.x && p.y == ip.y && p.z == ip.z) return;
    ip = p;
    if (a.codes) {
      gi = grid_index(p, a.vdim);
      const int code = a.codes[gi];
      cd = code <= -2 ? -1 - code : 0;
      pg = code;
      return;
    }
    cd = a.cheb[grid_index(p, a.cdim)];
    if (cd == 0) {
      gi = grid_index(p, a.vdim);
      pg = a.vaddr[gi];
    }
  }
};

// Cell code of brick ip: page id, -1 not resident, or -1 - chebyshev distance.
__device__ __forceinline__ int cell_code(const VolumeArgs &a, f3 ip) {
  if (a.codes) return a.codes[grid_index(ip, a.vdim)];
  const int cd = a.cheb[grid_index(ip, a.cdim)];
  return cd ? -1 - cd : a.vaddr[grid_index(ip, a.vdim)];
}

// VolumeShaderKernel::main (volume_shader.cu:32-81), reference step schedule.
// The escalation tests compare float alpha with the double literals 0.93 and
// 0.85; for binary32 inputs `x > 0.93` equals `x >= 0.93f` (0.93f rounds up),
// likewise for 0.85.
template <int FMT, bool COUNT, bool TRACE>
__device__ void march_reference(const VolumeArgs &a, const float *tf, const float *lut, MarchState &st,
                                unsigned &count, hyd_trace_rec *trace, int max_trace, int &ntrace) {
  const float cdu = 1.f / comp_max(abs3(st.d));
  const f3 d = st.d;
  f3 o = st.o;
  int nsteps = st.nsteps;
  int step_size = st.step_size;
  VolTaps tc;                 // taps of the last sample
  int cpg = -1;               // page of the cached descriptor
  float4 pgd = make_float4(0.f, 0.f, 0.f, 0.f);
  BrickLookup bl;
  while (nsteps > 0) {
    const f3 ip = floor3(o);
    bl.update(a, ip);
    const int cd = bl.cd;
    if (cd) {
      // skip_nblock_steps (volume_shader.cu:5-14) with step * step_size
      const float sstep = a.step * float(step_size);
      float tn, tfar;
      intersect_box(o, d, ip, f3{ip.x + 1.f, ip.y + 1.f, ip.z + 1.f}, tn, tfar);
      const float di = ceilf((tfar + float(cd - 1) * cdu) / sstep);
      o = add(o, muls(muls(d, di), sstep));
      nsteps -= (int)di;
    } else {
      const int gi = bl.gi;
      const int pgid = bl.pg;
      if (pgid == -1) break;
      if (pgid != cpg) {
        pgd = a.page[pgid];
        cpg = pgid;
      }
      TapIdx taps;
      const float s = brick_sample_reuse<FMT>(a.pool, a.pad, pgid, pgd, sub(o, ip), lut, TRACE ? &taps : nullptr,
                                              tc);
      if (TRACE && ntrace < max_trace) {
        const int gx = gi % a.vdim[0], gy = (gi / a.vdim[0]) % a.vdim[1], gz = gi / (a.vdim[0] * a.vdim[1]);
        trace[ntrace++] = hyd_trace_rec{gx, gy, gz, pgid, taps.ix, taps.iy, taps.iz, taps.jx, taps.jy, taps.jz};
      }
      shade_composite(a, tf, s, pgid, st);
      if (COUNT) ++count;
      if (st.v.w >= 0.93f) {
        step_size = 16;
      } else if (st.v.w >= 0.85f) {
        step_size = 4;
      }
    }
    o = add(o, muls(muls(d, a.step), float(step_size)));
    nsteps -= step_size;
  }
  st.o = o;
  st.nsteps = nsteps;
  st.step_size = step_size;
}

// ---------------------------------------------------------------------------
// The reference schedule with the tap loads of up to U consecutive samples in
// flight together.  A ray is one long chain of dependent samples, and a frame
// holds only a few waves per SIMD (one lane per pixel), so issuing one
// sample's loads and waiting for them before the next leaves the memory
// latency exposed on every sample.  Inside one brick the next sample positions
// follow from the current step size alone (o += (d * step) * step_size, the
// same binary32 operations one at a time), so a group of U samples resolves
// its positions, tap offsets and weights first, issues the loads of every
// sample, and then composites the samples in order.  (Skipping the loads of a
// sample whose cell repeats its predecessor's saved memory work, but its
// branches and register copies cost more instructions than the loads: the
// 512^2 FP32 frame took 0.534 ms with that reuse and groups of 2, 0.509 /
// 0.499 / 0.607 ms without it and groups of 3 / 4 / 5.)  Escalation changes the step size; a group stops at the
// sample that escalates (the ones after it were placed with the old step and
// are dropped), so positions, taps, values and step budget are those of the
// one-sample-at-a-time march, bit for bit.  Samples are also cut at the
// brick's boundary and at the end of the step budget.
//
// Voxels are addressed by element offsets from the pool base: 32-bit when the
// pool is below 4 GiB (OFF = uint32_t, so the base stays in scalar registers),
// else 64-bit.

// Rays per wave below which a wave of the one-lane-per-ray march (FP32 and
// FP16 frames) hands its remaining rays to the cooperative march (kCoopLanes
// lanes per ray).  Measured when the paged 1024^3 u8 frame at 512^2 still ran
// one lane per ray: 0.587 ms without the handover; 0.543 / 0.519 / 0.550 /
// 0.565 ms with 2 / 4 / 8 / 16 lanes per ray.  The plain FP32 frame is
// unchanged by it: its last waves still hold more rays than the handover
// takes.
#ifndef HYD_COOP_LANES
#define HYD_COOP_LANES 4
#endif
constexpr int kCoopLanes = HYD_COOP_LANES;
constexpr int kCoopRays = 64 / kCoopLanes;

template <int FMT, int U, bool COUNT, bool THETA, class OFF>
__device__ void march_reference_grouped(const VolumeArgs &a, const float *tf, const float *lut, MarchState &st,
                                        unsigned &count, bool &handed) {
  const float cdu = 1.f / comp_max(abs3(st.d));
  const f3 d = st.d;
  const f3 invr = recip3(d);
  f3 o = st.o;
  int nsteps = st.nsteps;
  int step_size = st.step_size;
  const OFF pad3 = (OFF)a.pad * (OFF)a.pad * (OFF)a.pad;
  int cpg = -1;
  float4 pgd = make_float4(0.f, 0.f, 0.f, 0.f);
  BrickLookup bl;
  while (nsteps > 0) {
    // few rays left in this wave: the wave's lanes take them over together
    if (__popcll(__ballot(1)) <= kCoopRays) {
      handed = true;
      break;
    }
    const f3 ip = floor3(o);
    bl.update(a, ip);
    if (bl.cd) {
      // skip_nblock_steps (volume_shader.cu:5-14) with step * step_size
      const float sstep = a.step * float(step_size);
      float tn, tfar;
      intersect_box_inv(o, invr, ip, f3{ip.x + 1.f, ip.y + 1.f, ip.z + 1.f}, tn, tfar);
      const float di = ceilf((tfar + float(bl.cd - 1) * cdu) / sstep);
      o = add(o, muls(muls(d, di), sstep));
      nsteps -= (int)di;
      o = add(o, muls(muls(d, a.step), float(step_size)));
      nsteps -= step_size;
      continue;
    }
    const int pgid = bl.pg;
    if (pgid == -1) break;
    if (pgid != cpg) {
      pgd = a.page[pgid];
      cpg = pgid;
    }
    const OFF page_base = (OFF)pgid * pad3;
    // positions of the group: the same additions the one-sample march makes
    const f3 inc = muls(muls(d, a.step), float(step_size));
    f3 p[U];
    bool ok[U];
    p[0] = o;
    ok[0] = true;
#pragma unroll
    for (int j = 1; j < U; ++j) {
      p[j] = add(p[j - 1], inc);
      const f3 q = floor3(p[j]);
      ok[j] = ok[j - 1] && nsteps - j * step_size > 0 && q.x == ip.x && q.y == ip.y && q.z == ip.z;
    }
    // taps and loads of every sample of the group
    RefTap<OFF> t[U];
    RawTaps<FMT> raw[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      t[j] = ref_tap<OFF>(sub(p[j], ip), pgd, a.pad, page_base);
      if (ok[j]) ref_fetch<FMT, OFF>(a.pool, t[j], a.pad, raw[j]);
    }
    // composite in order; stop at the sample that escalates
    int last = 0, last_ss = step_size;
    f3 plast = p[0];
    bool live = true;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (ok[j] && live) {
        float cur[8];
        ref_values<FMT, OFF>(raw[j], t[j], a.pad, lut, cur);
        const float s = ref_interp<OFF>(cur, t[j]);
        shade_composite_impl<false, THETA>(a, tf, s, pgid, st, 1.f);
        if (COUNT) ++count;
        last = j;
        plast = p[j];
        if (st.v.w >= 0.93f) {   // v.w > 0.93 (double), see march_reference
          last_ss = 16;
        } else if (st.v.w >= 0.85f) {
          last_ss = 4;
        }
        if (last_ss != step_size) live = false;
      }
    }
    // samples 0 .. last-1 advanced by the group's step, the last by its own
    o = add(plast, muls(muls(d, a.step), float(last_ss)));
    nsteps -= last * step_size + last_ss;
    step_size = last_ss;
  }
  st.o = o;
  st.nsteps = nsteps;
  st.step_size = step_size;
}

// Calls f(std::integral_constant<int, j>) for j = 0 .. N-1, so that the body
// sees j as a constant (DPP controls must be).
template <int... I, class F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F &&f) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// Lane J of each group of G lanes, to every lane of the group.  Groups of 4
// are the hardware quads: one DPP quad_perm move (a VALU operation) instead
// of an LDS permute.
template <int G, int J> __device__ __forceinline__ float group_bcast(float v) {
  if constexpr (G == 4)
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), J * 0x55, 0xF, 0xF, false));
  else
    return __shfl(v, J, G);
}

// The reference march of one ray by G lanes that hold identical state
// (lane h of the ray's lane group): per group of G * UL consecutive samples
// of one brick, lane h resolves samples h, h + G, ... (taps, loads issued
// together, interpolation, shading: the colour does not depend on the ray's
// state) and every lane composites the group's colours in order, taken from
// their lanes, stopping where the step escalates.  Positions come from the
// same additions as the one-lane march and each value from the same taps, so
// the ray's samples, values and composite equal march_reference_grouped's bit
// for bit.  A lane's instruction stream per group holds UL samples' sampling
// and shading instead of G * UL, which is what a long ray's time is made of.
template <int FMT, bool COUNT, bool THETA, class OFF, int G, int UL = 1>
__device__ void march_reference_coop(const VolumeArgs &a, const float *tf, const float *lut, MarchState &st, int h,
                                     unsigned &count) {
  constexpr int N = G * UL;
  const float cdu = 1.f / comp_max(abs3(st.d));
  const f3 d = st.d;
  const f3 invr = recip3(d);
  f3 o = st.o;
  int nsteps = st.nsteps;
  int step_size = st.step_size;
  const OFF pad3 = (OFF)a.pad * (OFF)a.pad * (OFF)a.pad;
  int cpg = -1;
  float4 pgd = make_float4(0.f, 0.f, 0.f, 0.f);
  BrickLookup bl;
  while (nsteps > 0) {
    const f3 ip = floor3(o);
    bl.update(a, ip);
    if (bl.cd) {
      // skip_nblock_steps (volume_shader.cu:5-14) with step * step_size
      const float sstep = a.step * float(step_size);
      float tn, tfar;
      intersect_box_inv(o, invr, ip, f3{ip.x + 1.f, ip.y + 1.f, ip.z + 1.f}, tn, tfar);
      const float di = ceilf((tfar + float(bl.cd - 1) * cdu) / sstep);
      o = add(o, muls(muls(d, di), sstep));
      nsteps -= (int)di;
      o = add(o, muls(muls(d, a.step), float(step_size)));
      nsteps -= step_size;
      continue;
    }
    const int pgid = bl.pg;
    if (pgid == -1) break;
    if (pgid != cpg) {
      pgd = a.page[pgid];
      cpg = pgid;
    }
    const OFF page_base = (OFF)pgid * pad3;
    const f3 inc = muls(muls(d, a.step), float(step_size));
    f3 p[N];
    bool ok[N];
    p[0] = o;
    ok[0] = true;
#pragma unroll
    for (int j = 1; j < N; ++j) {
      p[j] = add(p[j - 1], inc);
      const f3 q = floor3(p[j]);
      ok[j] = ok[j - 1] && nsteps - j * step_size > 0 && q.x == ip.x && q.y == ip.y && q.z == ip.z;
    }
    // this lane's samples h + G * k: taps and loads first, then values and colours
    RefTap<OFF> t[UL];
    RawTaps<FMT> raw[UL];
    bool okh[UL];
#pragma unroll
    for (int k = 0; k < UL; ++k) {
      f3 ph = p[G * k];
      okh[k] = ok[G * k];
#pragma unroll
      for (int q = 1; q < G; ++q) {
        if (h == q) {
          ph = p[G * k + q];
          okh[k] = ok[G * k + q];
        }
      }
      t[k] = ref_tap<OFF>(sub(ph, ip), pgd, a.pad, page_base);
      if (okh[k]) ref_fetch<FMT, OFF>(a.pool, t[k], a.pad, raw[k]);
    }
    float4 ubh[UL];
#pragma unroll
    for (int k = 0; k < UL; ++k) {
      ubh[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (okh[k]) {
        float cur[8];
        ref_values<FMT, OFF>(raw[k], t[k], a.pad, lut, cur);
        ubh[k] = shade_sample<false>(a, tf, ref_interp<OFF>(cur, t[k]), pgid, 1.f);
      }
    }
    // composite in order; stop at the sample that escalates
    int last = 0, last_ss = step_size;
    f3 plast = p[0];
    bool live = true;
    static_for<N>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      const float4 &uk = ubh[j / G];
      const float4 ub = make_float4(group_bcast<G, j % G>(uk.x), group_bcast<G, j % G>(uk.y),
                                    group_bcast<G, j % G>(uk.z), group_bcast<G, j % G>(uk.w));
      if (ok[j] && live) {
        composite_ub<THETA>(ub, st);
        if (COUNT && h == 0) ++count;
        last = j;
        plast = p[j];
        if (st.v.w >= 0.93f) {   // v.w > 0.93 (double), see march_reference
          last_ss = 16;
        } else if (st.v.w >= 0.85f) {
          last_ss = 4;
        }
        if (last_ss != step_size) live = false;
      }
    });
    o = add(plast, muls(muls(d, a.step), float(last_ss)));
    nsteps -= last * step_size + last_ss;
    step_size = last_ss;
  }
  st.o = o;
  st.nsteps = nsteps;
  st.step_size = step_size;
}

// Lattice schedule: sample k covers w = min(1, L - k) steps (L the ray extent
// in steps, k < ceil(L)), sits at the covered segment's midpoint
// o0 + d * ((k + w/2) * step) and its opacity is scaled by w; empty regions are
// skipped to the next lattice point; early
// termination once v.w > ert.  Film: o = o0, nsteps slot = bits of L (float),
// step_size = next lattice index.
template <int FMT, bool COUNT, bool THETA = true>
__device__ void march_lattice(const VolumeArgs &a, const float *tf, const float *lut, MarchState &st,
                              unsigned &count) {
  const f3 o0 = st.o, d = st.d;
  const float lf = __int_as_float(st.nsteps);
  int k = st.step_size;
  const int kend = (int)ceilf(lf);
  const float cdu = 1.f / comp_max(abs3(d));
  VolTaps tc;                 // taps of the last sample
  int cpg = -1;               // page of the cached descriptor
  float4 pgd = make_float4(0.f, 0.f, 0.f, 0.f);
  BrickLookup bl;
  while (k < kend) {
    const float w = gmin(1.f, lf - float(k));       // coverage of sample k
    const float t = (float(k) + 0.5f * w) * a.step;  // midpoint of the covered segment
    const f3 o = add(o0, muls(d, t));
    const f3 ip = floor3(o);
    bl.update(a, ip);
    const int cd = bl.cd;
    if (cd) {
      float tn, tfar;
      intersect_box(o, d, ip, f3{ip.x + 1.f, ip.y + 1.f, ip.z + 1.f}, tn, tfar);
      const float t_exit = t + tfar + float(cd - 1) * cdu;
      const int kn = (int)ceilf(t_exit / a.step - 0.5f);
      k = kn > k + 1 ? kn : k + 1;
      continue;
    }
    const int pgid = bl.pg;
    if (pgid == -1) break;
    if (pgid != cpg) {
      pgd = a.page[pgid];
      cpg = pgid;
    }
    const float s = brick_sample_reuse<FMT>(a.pool, a.pad, pgid, pgd, sub(o, ip), lut, nullptr, tc);
    shade_composite_impl<true, THETA>(a, tf, s, pgid, st, w);
    if (COUNT) ++count;
    ++k;
    if (a.ert > 0.f && st.v.w > a.ert) { k = kend; break; }
  }
  st.step_size = k;
}

// ray_emit_shader_impl (shader.hpp:113-137) + VolumeShaderKernel::init.
__device__ __forceinline__ bool emit(const VolumeArgs &a, int x, int y, MarchState &st) {
  st.d = pixel_ray_dir(a.trans, a.ctg, a.ray_o, x, y, a.W, a.H);
  st.o = a.ray_o;
  st.v = make_float4(0.f, 0.f, 0.f, 0.f);
  st.theta = f3{0.f, 0.f, 0.f};
  st.phi = 1.f;
  float tnear, tfar;
  if (intersect_box(st.o, st.d, a.bmin, a.bmax, tnear, tfar)) {
    if (tnear < 0) tnear = 0;
    if (tfar < 0) tfar = 0;
    st.o = add(st.o, muls(st.d, tnear));
    if (a.schedule == HYD_SCHED_LATTICE) {
      st.nsteps = __float_as_int(gmin((tfar - tnear) / a.step, float(a.max_steps)));
      st.step_size = 0;
    } else {
      const int n = (int)((tfar - tnear) / a.step);
      st.nsteps = a.max_steps < n ? a.max_steps : n;
      st.step_size = 1;
    }
    return true;
  }
  st.nsteps = 0;   // also the bits of 0.f for the lattice schedule
  st.step_size = a.schedule == HYD_SCHED_LATTICE ? 0 : 1;
  return false;
}

__device__ __forceinline__ void store_pixel(const VolumeArgs &a, size_t idx, const MarchState &st) {
  if (a.film) {
    float4 *f = reinterpret_cast<float4 *>(a.film + idx);
    f[0] = make_float4(st.o.x, st.o.y, st.o.z, st.d.x);
    f[1] = make_float4(st.d.y, st.d.z, __int_as_float(st.nsteps), __int_as_float(st.step_size));
    f[2] = st.v;
    f[3] = make_float4(st.theta.x, st.theta.y, st.theta.z, st.phi);
  }
  if (a.rgba) a.rgba[idx] = st.v;
}

__device__ __forceinline__ void load_pixel(const VolumeArgs &a, size_t idx, MarchState &st) {
  const float4 *f = reinterpret_cast<const float4 *>(a.film + idx);
  float4 f0 = f[0], f1 = f[1];
  st.o = f3{f0.x, f0.y, f0.z};
  st.d = f3{f0.w, f1.x, f1.y};
  st.nsteps = __float_as_int(f1.z);
  st.step_size = __float_as_int(f1.w);
  st.v = f[2];
  float4 f3v = f[3];
  st.theta = f3{f3v.x, f3v.y, f3v.z};
  st.phi = f3v.w;
}

// Samples whose tap loads a reference-schedule group issues together, per
// brick format.  Measured on the 512^2 frames: FP32 (256^3 single brick) takes
// 0.499 ms with 4 (128 registers, 4 waves per SIMD hold the whole frame) and
// 0.509 / 0.607 / 0.603 ms with 3 / 5 / 6; u8 (1024^3 paged, 4 lanes per ray)
// 0.357, 0.354 and 0.350 ms with 4, 8 and 12.
#ifndef HYD_REF_GROUP_U8
#define HYD_REF_GROUP_U8 8
#endif
#ifndef HYD_REF_GROUP_F16
#define HYD_REF_GROUP_F16 4
#endif
#ifndef HYD_REF_GROUP_F32
#define HYD_REF_GROUP_F32 4
#endif
template <int FMT> constexpr int ref_group() {
  return FMT == HYD_FMT_U8 ? HYD_REF_GROUP_U8 : FMT == HYD_FMT_F16 ? HYD_REF_GROUP_F16 : HYD_REF_GROUP_F32;
}

// Lanes per ray of the reference-schedule frame, per brick format (a wave
// then covers an 8 x 8/LANES tile; LANES > 1 marches with
// march_reference_coop, ref_group / LANES samples per lane and group).  A
// frame whose time is the chain of its longest rays gains: the paged u8 frame
// takes 0.52 ms with 1 lane (groups of 8 on one lane, the last rays handed to
// 4 lanes each), 0.405 with 2, 0.354 with 4 and 0.42 with 8.  A frame bound by
// instruction issue loses, since every lane of a ray repeats the positions
// and the compositing: the plain FP32 frame takes 0.55 ms with 1 lane, 0.59
// with 2 and 0.69 with 4.
#ifndef HYD_REF_LANES_U8
#define HYD_REF_LANES_U8 4
#endif
#ifndef HYD_REF_LANES_F16
#define HYD_REF_LANES_F16 1
#endif
#ifndef HYD_REF_LANES_F32
#define HYD_REF_LANES_F32 1
#endif
template <int FMT> constexpr int ref_lanes() {
  return FMT == HYD_FMT_U8 ? HYD_REF_LANES_U8 : FMT == HYD_FMT_F16 ? HYD_REF_LANES_F16 : HYD_REF_LANES_F32;
}

// Render modes of volume_kernel: emit into an RGBA image only, emit into a film
// (and optionally RGBA), or resume a film.
enum { kEmitRgba = 0, kEmitFilm = 1, kResume = 2 };

template <int FMT, bool COUNT, int MODE, bool BIG, int LANES>
__global__ __launch_bounds__(64) void volume_kernel(VolumeArgs a) {
  constexpr bool RESUME = MODE == kResume;
  constexpr bool THETA = MODE != kEmitRgba;
  static_assert(LANES == 1 || (kTile % LANES == 0 && ref_group<FMT>() % LANES == 0), "lanes per ray");
  __shared__ float s_tf[4 * kMaxTf];
  __shared__ float s_lut[FMT == HYD_FMT_U8 ? 256 : 1];
  __shared__ unsigned long long s_count;
  for (int i = threadIdx.x; i < 4 * a.tf_n; i += blockDim.x) s_tf[i] = a.tf[i];
  if (FMT == HYD_FMT_U8)
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_lut[i] = a.u8_lut[i];
  if (COUNT && threadIdx.x == 0) s_count = 0;
  __syncthreads();

  using OFF = std::conditional_t<BIG, uint64_t, uint32_t>;
  const int ray = (int)threadIdx.x / LANES, h = (int)threadIdx.x % LANES;
  const int tx = rotated_tile_x();
  const int x = tx * kTile + (ray % kTile);
  const int y = blockIdx.y * (kTile / LANES) + (ray / kTile);
  unsigned count = 0;
  bool handed = false;   // this lane's ray went to the cooperative march
  MarchState st;
  size_t idx = 0;
  if (x < a.W && y < a.H) {
    idx = (size_t)y * a.W + x;
    bool live;
    if (RESUME) {
      load_pixel(a, idx, st);
      live = st.nsteps > 0;
    } else {
      live = emit(a, x, y, st);
    }
    if (live) {
      if (LANES > 1)   // reference schedule only (the launch sees to it)
        march_reference_coop<FMT, COUNT, THETA, OFF, LANES, ref_group<FMT>() / LANES>(a, s_tf, s_lut, st, h, count);
      else if (a.schedule == HYD_SCHED_LATTICE)
        march_lattice<FMT, COUNT, THETA>(a, s_tf, s_lut, st, count);
      else
        march_reference_grouped<FMT, ref_group<FMT>(), COUNT, THETA, OFF>(a, s_tf, s_lut, st, count, handed);
    }
    if (!handed && h == 0) store_pixel(a, idx, st);
  }
  // the rays handed over: kCoopLanes lanes each, with the state copied through LDS
  const unsigned long long hm = __ballot(handed);
  if (hm) {
    __shared__ MarchState s_st[kCoopRays];
    __shared__ size_t s_idx[kCoopRays];
    const int lane = (int)threadIdx.x;
    // at most kCoopRays lanes hand over: they all leave the march together, in
    // the iteration where the wave's marching lanes first number kCoopRays or fewer
    if (handed) {
      const int slot = __popcll(hm & ((1ull << lane) - 1ull));
      if (slot < kCoopRays) {
        s_st[slot] = st;
        s_idx[slot] = idx;
      }
    }
    __syncthreads();
    const int r = lane / kCoopLanes, h = lane % kCoopLanes;
    if (r < __popcll(hm)) {
      MarchState c = s_st[r];
      march_reference_coop<FMT, COUNT, THETA, OFF, kCoopLanes>(a, s_tf, s_lut, c, h, count);
      if (h == 0) store_pixel(a, s_idx[r], c);
    }
  }
  if (COUNT) {
    if (count) atomicAdd(&s_count, (unsigned long long)count);
    __syncthreads();
    if (threadIdx.x == 0 && s_count) atomicAdd(a.samples, s_count);
  }
}

template <int FMT>
__global__ __launch_bounds__(64) void volume_trace_kernel(VolumeArgs a, int x, int y, hyd_trace_rec *out,
                                                          int max_recs, int *n_out) {
  __shared__ float s_tf[4 * kMaxTf];
  __shared__ float s_lut[256];
  for (int i = threadIdx.x; i < 4 * a.tf_n; i += blockDim.x) s_tf[i] = a.tf[i];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) s_lut[i] = a.u8_lut ? a.u8_lut[i] : 0.f;
  __syncthreads();
  if (threadIdx.x != 0) return;
  MarchState st;
  unsigned count = 0;
  int nt = 0;
  if (emit(a, x, y, st)) march_reference<FMT, false, true>(a, s_tf, s_lut, st, count, out, max_recs, nt);
  *n_out = nt;
}

// ---------------------------------------------------------------------------
// Sort-last segments.  The lattice walk of a ray visits cells in order (every
// coordinate of o0 + d*t is monotone in t, so a cell is never revisited); the
// k-th visited cell owns slot k of the ray.  A pass composites, front to back
// from zero, the samples of every visited cell whose brick is resident in the
// current page table into its slot (each sample reads the brick of its own
// floor(p), exactly as the single march does); empty cells write a transparent
// slot; cells not resident now are left for a later pass.  The final fold
// composites the slots in order and stops at the first slot no pass filled
// (the reference stops a ray at its first non-resident brick).
//
// A cell's samples are a run of lattice indices [k, kend_cell) (positions are
// monotone in t, so a cell is entered once); the run's end is estimated from
// the cell's exit distance and then moved until it agrees with floor() of the
// sample positions, evaluated exactly as the walk evaluates them, so every
// pass sees the same cells and samples as the single march.  A cell not
// resident in this pass is passed over in one step; a resident one composites
// its run with the tap loads of up to U samples in flight together (every
// sample loads its taps: skipping a repeated cell's loads cost more in
// branches and register copies, 4.60 against 4.51 ms for the 3-batch set).
#ifndef HYD_SEG_GROUP
#define HYD_SEG_GROUP 4
#endif

// Sample k of the lattice walk: position o0 + d * ((k + w/2) * step), w = min(1, L - k).
__device__ __forceinline__ f3 lattice_pos(f3 o0, f3 d, float lf, float step, int k, float &w) {
  w = gmin(1.f, lf - float(k));
  const float t = (float(k) + 0.5f * w) * step;
  return add(o0, muls(d, t));
}

__device__ __forceinline__ bool same_cell(f3 p, f3 ip) {
  const f3 q = floor3(p);
  return q.x == ip.x && q.y == ip.y && q.z == ip.z;
}

// The walk's step at lattice index k.  An empty cell writes a transparent
// slot and a cell not resident in this pass leaves its slot; both advance k
// and the slot index and return false.  A resident cell returns true with
// its run [k, kc), its cell ip and its page id, k and ord unchanged.
__device__ __forceinline__ bool seg_walk_cell(const VolumeArgs &a, f3 o0, f3 d, f3 invr, float lf, int kend,
                                              float cdu, size_t P, size_t ray, int &k, int &ord, f3 &ip, int &kc,
                                              int &pgid) {
  float w;
  const f3 o = lattice_pos(o0, d, lf, a.step, k, w);
  const float t = (float(k) + 0.5f * w) * a.step;
  ip = floor3(o);
  const int code = cell_code(a, ip);
  const int cd = code <= -2 ? -1 - code : 0;
  float tn, tfar;
  intersect_box_inv(o, invr, ip, f3{ip.x + 1.f, ip.y + 1.f, ip.z + 1.f}, tn, tfar);
  if (cd) {   // empty cell: same skip as the single lattice march
    const float t_exit = t + tfar + float(cd - 1) * cdu;
    const int kn = (int)ceilf(t_exit / a.step - 0.5f);
    k = kn > k + 1 ? kn : k + 1;
    a.slots[(size_t)ord * P + ray] = make_float4(0.f, 0.f, 0.f, 0.f);
    ++ord;
    return false;
  }
  // the cell's run [k, kc): estimate from the exit distance, then agree with floor()
  kc = (int)ceilf((t + tfar) / a.step - 0.5f);
  kc = kc > k + 1 ? kc : k + 1;
  kc = kc < kend ? kc : kend;
  float wq;
  while (kc < kend && same_cell(lattice_pos(o0, d, lf, a.step, kc, wq), ip)) ++kc;
  while (kc - 1 > k && !same_cell(lattice_pos(o0, d, lf, a.step, kc - 1, wq), ip)) --kc;
  pgid = code;
  if (pgid < 0) {   // not resident in this pass: the slot is left for another
    ++ord;
    k = kc;
    return false;
  }
  return true;
}

// Samples [kk, min(kk + U, kc)) of a run in cell ip of page pgid, their tap
// loads in flight together, composited in order into seg.
template <int FMT, bool COUNT, int U, typename OFF>
__device__ __forceinline__ void seg_sample_group(const VolumeArgs &a, const float *s_tf, const float *s_lut, f3 o0,
                                                 f3 d, float lf, f3 ip, float4 pgd, OFF page_base, int pgid, int kk,
                                                 int kc, MarchState &seg, unsigned &count) {
  RefTap<OFF> tp[U];
  RawTaps<FMT> raw[U];
  bool ok[U];
  float wj[U];
#pragma unroll
  for (int j = 0; j < U; ++j) {
    ok[j] = kk + j < kc;
    const f3 pj = lattice_pos(o0, d, lf, a.step, ok[j] ? kk + j : kk, wj[j]);
    tp[j] = ref_tap<OFF>(sub(pj, ip), pgd, a.pad, page_base);
    if (ok[j]) ref_fetch<FMT, OFF>(a.pool, tp[j], a.pad, raw[j]);
  }
#pragma unroll
  for (int j = 0; j < U; ++j) {
    if (ok[j]) {
      float cur[8];
      ref_values<FMT, OFF>(raw[j], tp[j], a.pad, s_lut, cur);
      const float s = ref_interp<OFF>(cur, tp[j]);
      shade_composite_impl<true, false>(a, s_tf, s, pgid, seg, wj[j]);
      if (COUNT) ++count;
    }
  }
}

template <int FMT, bool COUNT, bool BIG, bool MULTI>
__global__ __launch_bounds__(64) void segment_kernel(VolumeArgs a, SegBatchTable tb) {
  using OFF = std::conditional_t<BIG, uint64_t, uint32_t>;
  if (MULTI) {   // this workgroup's residency batch
    a.pool = tb.pool[blockIdx.z];
    a.codes = tb.codes[blockIdx.z];
    a.page = tb.page[blockIdx.z];
  }
  constexpr int U = HYD_SEG_GROUP;
  __shared__ float s_tf[4 * kMaxTf];
  __shared__ float s_lut[FMT == HYD_FMT_U8 ? 256 : 1];
  __shared__ unsigned long long s_count;
  for (int i = threadIdx.x; i < 4 * a.tf_n; i += blockDim.x) s_tf[i] = a.tf[i];
  if (FMT == HYD_FMT_U8)
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_lut[i] = a.u8_lut[i];
  if (COUNT && threadIdx.x == 0) s_count = 0;
  __syncthreads();
  const int x = rotated_tile_x() * kTile + (threadIdx.x % kTile);
  const int y = blockIdx.y * kTile + (threadIdx.x / kTile);
  unsigned count = 0;
  if (x < a.W && y < a.H) {
    const size_t P = (size_t)a.W * a.H, ray = (size_t)y * a.W + x;
    MarchState st;
    int ord = 0;
    if (emit(a, x, y, st)) {
      const f3 o0 = st.o, d = st.d;
      const float lf = __int_as_float(st.nsteps);
      const int kend = (int)ceilf(lf);
      const float cdu = 1.f / comp_max(abs3(d));
      const f3 invr = recip3(d);
      const OFF pad3 = (OFF)a.pad * (OFF)a.pad * (OFF)a.pad;
      int k = 0, kc = 0, pgid = -1;
      f3 ip{0.f, 0.f, 0.f};
      MarchState seg;
      if constexpr (MULTI) {
        // the batches' waves fill the chip: the wave walks its rays' cells in
        // lock-step (the lanes in resident cells composite their runs while
        // the others wait), which issues the fewest instructions
        while (k < kend && ord < a.max_slots) {
          if (!seg_walk_cell(a, o0, d, invr, lf, kend, cdu, P, ray, k, ord, ip, kc, pgid)) continue;
          const float4 pgd = a.page[pgid];
          const OFF page_base = (OFF)pgid * pad3;
          seg.v = make_float4(0.f, 0.f, 0.f, 0.f);
          for (int kk = k; kk < kc; kk += U)
            seg_sample_group<FMT, COUNT, U, OFF>(a, s_tf, s_lut, o0, d, lf, ip, pgd, page_base, pgid, kk, kc, seg,
                                                 count);
          a.slots[(size_t)ord * P + ray] = seg.v;
          ++ord;
          k = kc;
        }
      } else {
        // one batch: the waves do not fill the chip and lanes idle in
        // lock-step cost time, so each lane composites U samples of its own
        // run per iteration; lanes whose run ended wait until a quarter of
        // the wave's busy lanes (or all of them) need to walk, then walk
        // together to their next resident run (7.47 -> 6.84 ms for three
        // one-batch passes of the 1024^3 frame; 5.16 against 4.58 ms in the
        // set launch, which is instruction-issue bound)
        bool in_run = false;
        int kk = 0;
        float4 pgd = make_float4(0.f, 0.f, 0.f, 0.f);
        OFF page_base = 0;
        while (true) {
          const bool want = !in_run && k < kend && ord < a.max_slots;
          const uint64_t m_want = __ballot(want), m_run = __ballot(in_run);
          if ((m_want | m_run) == 0) break;
          const bool walk_now = m_run == 0 || 4 * __popcll(m_want) >= __popcll(m_want | m_run);
          if (want && walk_now) {
            while (k < kend && ord < a.max_slots) {
              if (!seg_walk_cell(a, o0, d, invr, lf, kend, cdu, P, ray, k, ord, ip, kc, pgid)) continue;
              pgd = a.page[pgid];
              page_base = (OFF)pgid * pad3;
              seg.v = make_float4(0.f, 0.f, 0.f, 0.f);
              kk = k;
              in_run = true;
              break;
            }
          }
          if (!in_run) continue;
          seg_sample_group<FMT, COUNT, U, OFF>(a, s_tf, s_lut, o0, d, lf, ip, pgd, page_base, pgid, kk, kc, seg,
                                               count);
          kk += U;
          if (kk >= kc) {   // the run is complete: its slot, then the next cell
            a.slots[(size_t)ord * P + ray] = seg.v;
            ++ord;
            k = kc;
            in_run = false;
          }
        }
      }
    }
    a.nslots[ray] = ord;
  }
  if (COUNT) {
    if (count) atomicAdd(&s_count, (unsigned long long)count);
    __syncthreads();
    if (threadIdx.x == 0 && s_count) atomicAdd(a.samples, s_count);
  }
}

__global__ __launch_bounds__(256) void segment_clear_kernel(float4 *slots, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) slots[i] = make_float4(0.f, 0.f, 0.f, -1.f);   // alpha -1: not filled yet
}

__global__ __launch_bounds__(256) void segment_fold_kernel(const float4 *slots, const int32_t *nslots,
                                                           int max_slots, int P, float4 *rgba) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  const int n = min(nslots[i], max_slots);
  for (int ord = 0; ord < n; ++ord) {
    const float4 s = slots[(size_t)ord * P + i];
    if (s.w < 0.f) break;
    const float t = 1.f - v.w;
    v = make_float4(v.x + s.x * t, v.y + s.y * t, v.z + s.z * t, v.w + s.w * t);
  }
  rgba[i] = v;
}

template <int FMT>
hipError_t launch_segments_fmt(const VolumeArgs &a, hipStream_t s);

template <int FMT, int MODE, bool BIG>
void launch_mode(const VolumeArgs &a, hipStream_t s) {
  dim3 block(kTile * kTile);
  constexpr int L = ref_lanes<FMT>();
  if (L > 1 && a.schedule == HYD_SCHED_REFERENCE) {
    dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile / L - 1) / (kTile / L));
    if (a.samples) hipLaunchKernelGGL((volume_kernel<FMT, true, MODE, BIG, L>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((volume_kernel<FMT, false, MODE, BIG, L>), grid, block, 0, s, a);
  } else {
    dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile);
    if (a.samples) hipLaunchKernelGGL((volume_kernel<FMT, true, MODE, BIG, 1>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((volume_kernel<FMT, false, MODE, BIG, 1>), grid, block, 0, s, a);
  }
}

template <int FMT, bool BIG>
void launch_fmt_off(const VolumeArgs &a, bool resume, hipStream_t s) {
  if (resume) launch_mode<FMT, kResume, BIG>(a, s);
  else if (a.film) launch_mode<FMT, kEmitFilm, BIG>(a, s);
  else launch_mode<FMT, kEmitRgba, BIG>(a, s);
}

// HYD_VOLUME_OFFSETS64=1 selects the 64-bit offsets for any pool (they render the
// same image; the switch lets a small pool exercise that variant).
bool force_offsets64() {
  static const bool on = [] {
    const char *e = std::getenv("HYD_VOLUME_OFFSETS64");
    return e && e[0] == '1';
  }();
  return on;
}

template <int FMT, bool MULTI>
void launch_segments_multi(const VolumeArgs &a, const SegBatchTable &tb, hipStream_t s) {
  dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile, MULTI ? tb.n : 1);
  // pool_elems is the largest pool's when MULTI
  const uint64_t bytes = a.pool_elems * (uint64_t)sizeof(typename RawVox<FMT>::T);
  const bool big = bytes > (1ull << 32) || force_offsets64();
  if (a.samples) {
    if (big) hipLaunchKernelGGL((segment_kernel<FMT, true, true, MULTI>), grid, dim3(64), 0, s, a, tb);
    else hipLaunchKernelGGL((segment_kernel<FMT, true, false, MULTI>), grid, dim3(64), 0, s, a, tb);
  } else {
    if (big) hipLaunchKernelGGL((segment_kernel<FMT, false, true, MULTI>), grid, dim3(64), 0, s, a, tb);
    else hipLaunchKernelGGL((segment_kernel<FMT, false, false, MULTI>), grid, dim3(64), 0, s, a, tb);
  }
}

template <int FMT>
hipError_t launch_segments_fmt(const VolumeArgs &a, hipStream_t s) {
  SegBatchTable none{};
  launch_segments_multi<FMT, false>(a, none, s);
  return hipGetLastError();
}

template <int FMT>
hipError_t launch_fmt(const VolumeArgs &a, bool resume, hipStream_t s) {
  // 32-bit byte offsets while every byte of the pool is below 4 GiB
  const uint64_t bytes = a.pool_elems * (uint64_t)sizeof(typename RawVox<FMT>::T);
  // the x-pair loads need two voxels per row (hyd_upload_pool refuses smaller bricks)
  if (bytes <= (1ull << 32) && !force_offsets64()) launch_fmt_off<FMT, false>(a, resume, s);
  else launch_fmt_off<FMT, true>(a, resume, s);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_volume(const VolumeArgs &a, int fmt, bool resume, hipStream_t s) {
  switch (fmt) {
    case HYD_FMT_U8: return launch_fmt<HYD_FMT_U8>(a, resume, s);
    case HYD_FMT_F16: return launch_fmt<HYD_FMT_F16>(a, resume, s);
    default: return launch_fmt<HYD_FMT_F32>(a, resume, s);
  }
}

hipError_t launch_segments_clear(float4 *slots, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(segment_clear_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, slots, n);
  return hipGetLastError();
}

hipError_t launch_segments_pass(const VolumeArgs &a, int fmt, hipStream_t s) {
  switch (fmt) {
    case HYD_FMT_U8: return launch_segments_fmt<HYD_FMT_U8>(a, s);
    case HYD_FMT_F16: return launch_segments_fmt<HYD_FMT_F16>(a, s);
    default: return launch_segments_fmt<HYD_FMT_F32>(a, s);
  }
}

hipError_t launch_segments_batches(const VolumeArgs &a, const SegBatchTable &tb, int fmt, hipStream_t s) {
  switch (fmt) {
    case HYD_FMT_U8: launch_segments_multi<HYD_FMT_U8, true>(a, tb, s); break;
    case HYD_FMT_F16: launch_segments_multi<HYD_FMT_F16, true>(a, tb, s); break;
    default: launch_segments_multi<HYD_FMT_F32, true>(a, tb, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_segments_composite(const float4 *slots, const int32_t *nslots, int max_slots, int P, float4 *rgba,
                                     hipStream_t s) {
  hipLaunchKernelGGL(segment_fold_kernel, dim3((P + 255) / 256), dim3(256), 0, s, slots, nslots, max_slots, P, rgba);
  return hipGetLastError();
}

hipError_t launch_volume_trace(const VolumeArgs &a, int fmt, int x, int y, hyd_trace_rec *out_dev, int max_recs,
                               int *n_dev, hipStream_t s) {
  switch (fmt) {
    case HYD_FMT_U8:
      hipLaunchKernelGGL(volume_trace_kernel<HYD_FMT_U8>, dim3(1), dim3(64), 0, s, a, x, y, out_dev, max_recs, n_dev);
      break;
    case HYD_FMT_F16:
      hipLaunchKernelGGL(volume_trace_kernel<HYD_FMT_F16>, dim3(1), dim3(64), 0, s, a, x, y, out_dev, max_recs, n_dev);
      break;
    default:
      hipLaunchKernelGGL(volume_trace_kernel<HYD_FMT_F32>, dim3(1), dim3(64), 0, s, a, x, y, out_dev, max_recs, n_dev);
  }
  return hipGetLastError();
}

}  // namespace hyd
