// Blocks.Volume / Blocks.Solid and Isosurface kernels for gfx950.
// One lane per pixel, 8x8 pixel packets per 64-lane workgroup.
#include <hip/hip_fp16.h>

#include "hyd_internal.hpp"
#include "hyd_sampler.hpp"

#include <cstdlib>
#include <type_traits>

namespace hyd {
namespace {

constexpr int kTile = 8;

__device__ __forceinline__ int grid_index(f3 p, const int32_t dim[3]) {
  float fx = gclamp(p.x, 0.f, float(dim[0]));
  float fy = gclamp(p.y, 0.f, float(dim[1]));
  float fz = gclamp(p.z, 0.f, float(dim[2]));
  int ix = iclamp((int)floorf(fx), 0, dim[0] - 1);
  int iy = iclamp((int)floorf(fy), 0, dim[1] - 1);
  int iz = iclamp((int)floorf(fz), 0, dim[2] - 1);
  return iz * dim[0] * dim[1] + iy * dim[0] + ix;
}

// skip_nblock_steps (blocks_shader.cu:5-14, isosurface_shader.cu:5-14).
__device__ __forceinline__ int skip_nblock_steps(f3 &o, f3 d, f3 ip, int nblocks, float cdu, float step) {
  float tn, tfar;
  intersect_box(o, d, ip, f3{ip.x + 1.f, ip.y + 1.f, ip.z + 1.f}, tn, tfar);
  const float di = ceilf((tfar + float(nblocks - 1) * cdu) / step);
  o = add(o, muls(muls(d, di), step));
  return (int)di;
}

// skip_nblock_steps with the ray's reciprocal direction computed once (same result).
__device__ __forceinline__ int skip_nblock_steps_inv(f3 &o, f3 invr, f3 d, f3 ip, int nblocks, float cdu,
                                                     float step) {
  float tn, tfar;
  intersect_box_inv(o, invr, ip, f3{ip.x + 1.f, ip.y + 1.f, ip.z + 1.f}, tn, tfar);
  const float di = ceilf((tfar + float(nblocks - 1) * cdu) / step);
  o = add(o, muls(muls(d, di), step));
  return (int)di;
}

// VisBlocksShaderKernel::main, blocks_shader.cu:16-50; HYD_BLOCKS_PAGING is the
// Paging renderer's residency view, PagingShaderKernel::main (paging_shader.cu:24-61).
__global__ __launch_bounds__(64) void blocks_kernel(BlocksArgs a) {
  const int x = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int y = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (x >= a.W || y >= a.H) return;
  const f3 d = pixel_ray_dir(a.trans, a.ctg, a.ray_o, x, y, a.W, a.H);
  f3 o = a.ray_o;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  float tnear, tfar;
  if (intersect_box(o, d, a.bmin, a.bmax, tnear, tfar)) {
    if (tnear < 0) tnear = 0;
    if (tfar < 0) tfar = 0;
    o = add(o, muls(d, tnear));
    const int n = (int)((tfar - tnear) / a.step);
    int nsteps = a.max_steps < n ? a.max_steps : n;
    const float cdu = 1.f / comp_max(abs3(d));
    const float opacity_threshold = 0.95f;
    const f3 hi = f3{a.bmax.x - 1.f, a.bmax.y - 1.f, a.bmax.z - 1.f};
    const f3 ext = sub(a.bmax, a.bmin);
    while (nsteps > 0) {
      const f3 ip = floor3(o);
      const int gi = grid_index(ip, a.dim);
      const int cd = a.cheb[gi];
      if (cd) {
        nsteps -= skip_nblock_steps(o, d, ip, cd, cdu, a.step);
      } else if (a.mode == HYD_BLOCKS_PAGING) {
        // PagingShaderKernel::main (paging_shader.cu:34-59): one composite per
        // resident brick, green for a level-0 page, red for a lower-level one
        const int pgid = a.vaddr[grid_index(ip, a.vdim)];
        if (pgid == -1) break;   // page fault
        const float4 col = pgid >= a.lowest ? make_float4(0.f, .5f, 0.f, .5f) : make_float4(.2f, 0.f, 0.f, .2f);
        nsteps -= skip_nblock_steps(o, d, ip, 1, cdu, a.step);
        const float t = 1.f - v.w;
        v = make_float4(v.x + col.x * t, v.y + col.y * t, v.z + col.z * t, v.w + col.w * t);
        if (v.w > opacity_threshold) break;
      } else {
        const f3 c = {gclamp(ip.x, a.bmin.x, hi.x), gclamp(ip.y, a.bmin.y, hi.y), gclamp(ip.z, a.bmin.z, hi.z)};
        const f3 rip = {c.x / ext.x, c.y / ext.y, c.z / ext.z};
        if (a.mode == HYD_BLOCKS_VOLUME) {
          const float mean = a.mean[gi];
          const float4 col = make_float4(rip.x * a.density * mean, rip.y * a.density * mean,
                                         rip.z * a.density * mean, 1.f * a.density * mean);
          const float t = 1.f - v.w;
          v = make_float4(v.x + col.x * t, v.y + col.y * t, v.z + col.z * t, v.w + col.w * t);
          if (v.w > opacity_threshold) break;
        } else {
          v = make_float4(rip.x, rip.y, rip.z, 1.f);
          break;
        }
      }
      o = add(o, muls(d, a.step));
      nsteps -= 1;
    }
  }
  a.rgba[(size_t)y * a.W + x] = v;
}

template <int FMT> __device__ __forceinline__ float vox(const void *pool, size_t i, const float *lut);
template <> __device__ __forceinline__ float vox<HYD_FMT_F32>(const void *pool, size_t i, const float *) {
  return static_cast<const float *>(pool)[i];
}
template <> __device__ __forceinline__ float vox<HYD_FMT_F16>(const void *pool, size_t i, const float *) {
  return __half2float(static_cast<const __half *>(pool)[i]);
}
template <> __device__ __forceinline__ float vox<HYD_FMT_U8>(const void *pool, size_t i, const float *lut) {
  return lut[static_cast<const uint8_t *>(pool)[i]];
}

// BlockSampler::sample_3d (mapping + CpuSampler Wrap/Linear), as in the volume kernel.
template <int FMT>
__device__ float iso_sample(const IsoArgs &a, const float *lut, int pgid, f3 x) {
  const float4 pg = a.page[pgid];
  const int pad = a.pad;
  const float fd = float(pad);
  float u[3] = {x.x * pg.x + pg.y, x.y * pg.x + pg.z, x.z * pg.x + pg.w};
  float w[3];
  int i[3], j[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float fx = fd * u[c];
    fx = fx - fd * floorf(fx / fd);
    fx = fx - 0.5f;
    float flr = floorf(fx);
    w[c] = fx - flr;
    i[c] = iclamp((int)flr, 0, pad - 1);
    j[c] = iclamp((int)ceilf(fx), 0, pad - 1);
  }
  const size_t base = (size_t)pgid * pad * pad * pad;
  auto at = [&](int xx, int yy, int zz) { return vox<FMT>(a.pool, base + ((size_t)zz * pad + yy) * pad + xx, lut); };
  const float bx = 1.f - w[0], by = 1.f - w[1], bz = 1.f - w[2];
  float x00 = at(i[0], i[1], i[2]) * bx + at(j[0], i[1], i[2]) * w[0];
  float x10 = at(i[0], j[1], i[2]) * bx + at(j[0], j[1], i[2]) * w[0];
  float x01 = at(i[0], i[1], j[2]) * bx + at(j[0], i[1], j[2]) * w[0];
  float x11 = at(i[0], j[1], j[2]) * bx + at(j[0], j[1], j[2]) * w[0];
  float y0 = x00 * by + x10 * w[1];
  float y1 = x01 * by + x11 * w[1];
  return y0 * bz + y1 * w[2];
}

__device__ __forceinline__ float gsign(float x) { return float((0.f < x) - (x < 0.f)); }

// IsosurfaceShaderKernel::linear_to_srgb: double arithmetic on a float input
// (isosurface_shader.cu:25-32).
__device__ __forceinline__ float iso_srgb(float linear) {
  const double l = linear;
  if (l <= 0.0031308) return (float)(12.92 * l);
  return (float)((1.0 + 0.055) * pow(l, 1.0 / 2.4) - 0.055);
}

// IsosurfaceShaderKernel::main (isosurface_shader.cu:50-125) when
// corrected == 0; bracketing + bisection + oriented central differences when
// corrected == 1.  Two kernels: the march finds each ray's first crossing and
// leaves its state in the output pixel (v = march position and previous value,
// depth = value at the crossing, hit = kIsoPending); the hit kernel then
// refines and shades those pixels.  The march kernel carries none of the
// refinement's registers, so more of its waves fit on a SIMD to cover the
// latency of the tap loads.
constexpr int32_t kIsoPending = -1;

// The march of IsosurfaceShaderKernel::main up to the first crossing, with the
// tap loads of up to U consecutive samples in flight together (the step is
// fixed, so the next positions inside a brick follow from the same additions
// the one-step march makes).  The group is evaluated, then scanned in order for
// the first crossing; values, positions, counts and the pending pixel are those
// of the one-step march.
template <int FMT, int U, class OFF>
__device__ bool iso_march_grouped(const IsoArgs &a, const float *lut, f3 &o, f3 d, int nsteps, float cdu,
                                  hyd_iso_pixel &px, unsigned &count) {
  const f3 invr = recip3(d);
  const OFF pad3 = (OFF)a.pad * (OFF)a.pad * (OFF)a.pad;
  float prev_value = 0.f;
  bool have_prev = false;
  OFF c0 = ~(OFF)0, c7 = ~(OFF)0;   // taps of the last evaluated sample
  float cur[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  f3 cip = {-INFINITY, -INFINITY, -INFINITY};
  int ccd = 0, cpg = -1, dpg = -1;
  float4 pgd = make_float4(0.f, 0.f, 0.f, 0.f);
  while (nsteps > 0) {
    const f3 ip = floor3(o);
    if (ip.x != cip.x || ip.y != cip.y || ip.z != cip.z) {
      cip = ip;
      if (a.codes) {
        const int code = a.codes[grid_index(ip, a.vdim)];
        ccd = code <= -2 ? -1 - code : 0;
        cpg = code;
      } else {
        ccd = a.cheb[grid_index(ip, a.cdim)];
        if (ccd == 0) cpg = a.vaddr[grid_index(ip, a.vdim)];
      }
    }
    if (ccd) {
      nsteps -= skip_nblock_steps_inv(o, invr, d, ip, ccd, cdu, a.step);
      have_prev = false;
      o = add(o, muls(d, a.step));
      nsteps -= 1;
      continue;
    }
    const int pgid = cpg;
    if (pgid == -1) break;
    if (pgid != dpg) {
      pgd = a.page[pgid];
      dpg = pgid;
    }
    const OFF page_base = (OFF)pgid * pad3;
    const f3 inc = muls(d, a.step);
    f3 p[U];
    bool ok[U];
    p[0] = o;
    ok[0] = true;
#pragma unroll
    for (int j = 1; j < U; ++j) {
      p[j] = add(p[j - 1], inc);
      const f3 q = floor3(p[j]);
      ok[j] = ok[j - 1] && nsteps - j > 0 && q.x == ip.x && q.y == ip.y && q.z == ip.z;
    }
    RefTap<OFF> t[U];
    RawTaps<FMT> raw[U];
    bool fresh[U];
    OFF p0 = c0, p7 = c7;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      t[j] = ref_tap<OFF>(sub(p[j], ip), pgd, a.pad, page_base);
      const OFF i0 = t[j].r00 + (OFF)t[j].ix, i7 = t[j].r11 + (OFF)t[j].jx;
      fresh[j] = i0 != p0 || i7 != p7;
      p0 = i0;
      p7 = i7;
      if (ok[j] && fresh[j]) ref_fetch<FMT, OFF>(a.pool, t[j], a.pad, raw[j]);
    }
    int last = 0;
    f3 plast = p[0];
    bool live = true;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (ok[j] && live) {
        if (fresh[j]) ref_values<FMT, OFF>(raw[j], t[j], a.pad, lut, cur);
        const float value = ref_interp<OFF>(cur, t[j]);
        c0 = t[j].r00 + (OFF)t[j].ix;
        c7 = t[j].r11 + (OFF)t[j].jx;
        ++count;
        bool crossing = gsign(value - a.isovalue) != gsign(prev_value - a.isovalue);
        if (a.corrected) crossing = crossing && have_prev;
        if (crossing) {
          px.v[0] = p[j].x; px.v[1] = p[j].y; px.v[2] = p[j].z; px.v[3] = prev_value;
          px.depth = value;
          px.hit = kIsoPending;
          live = false;
        } else {
          if (a.corrected) { prev_value = value; have_prev = true; }
          last = j;
          plast = p[j];
        }
      }
    }
    if (!live) return true;
    o = add(plast, inc);
    nsteps -= last + 1;
  }
  return false;
}

#ifndef HYD_ISO_GROUP
#define HYD_ISO_GROUP 4
#endif

template <int FMT, bool BIG>
__global__ __launch_bounds__(64) void iso_march_grouped_kernel(IsoArgs a) {
  __shared__ float s_lut[FMT == HYD_FMT_U8 ? 256 : 1];
  if (FMT == HYD_FMT_U8)
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_lut[i] = a.u8_lut[i];
  __syncthreads();
  const int x = rotated_tile_x() * kTile + (threadIdx.x % kTile);
  const int y = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (x >= a.W || y >= a.H) return;
  unsigned count = 0;
  const f3 d = pixel_ray_dir(a.trans, a.ctg, a.ray_o, x, y, a.W, a.H);
  f3 o = a.ray_o;
  hyd_iso_pixel px;
  px.v[0] = px.v[1] = px.v[2] = px.v[3] = 0.f;
  px.depth = INFINITY;
  px.hit = 0;
  float tnear, tfar;
  if (intersect_box(o, d, a.bmin, a.bmax, tnear, tfar)) {
    if (tnear < 0) tnear = 0;
    if (tfar < 0) tfar = 0;
    o = add(o, muls(d, tnear));
    const int n = (int)((tfar - tnear) / a.step);
    const int nsteps = a.max_steps < n ? a.max_steps : n;
    const float cdu = 1.f / comp_max(abs3(d));
    iso_march_grouped<FMT, HYD_ISO_GROUP, std::conditional_t<BIG, uint64_t, uint32_t>>(a, s_lut, o, d, nsteps, cdu,
                                                                                        px, count);
  }
  a.out[(size_t)y * a.W + x] = px;
  if (a.samples && count) atomicAdd(a.samples, (unsigned long long)count);
}

// Refinement and shading of the pixels the march left pending: the crossing's
// march position o (its brick ip = floor(o) and page from the page table, as
// the march had them), the previous and the crossing value.
template <int FMT>
__global__ __launch_bounds__(64) void iso_hit_kernel(IsoArgs a) {
  __shared__ float s_lut[FMT == HYD_FMT_U8 ? 256 : 1];
  if (FMT == HYD_FMT_U8)
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_lut[i] = a.u8_lut[i];
  __syncthreads();
  const int x = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int y = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (x >= a.W || y >= a.H) return;
  hyd_iso_pixel *out = a.out + (size_t)y * a.W + x;
  const hyd_iso_pixel st = *out;
  if (st.hit != kIsoPending) return;
  const f3 d = pixel_ray_dir(a.trans, a.ctg, a.ray_o, x, y, a.W, a.H);
  const f3 origin = a.ray_o;
  const f3 o = f3{st.v[0], st.v[1], st.v[2]};
  const float prev_value = st.v[3], value = st.depth;
  const f3 ip = floor3(o);
  const int pgid = a.vaddr[grid_index(ip, a.vdim)];
  const f3 p = o;
  const float dt = a.step;
  float4 v = make_float4(a.surface_color.x, a.surface_color.y, a.surface_color.z, 1.f);
  f3 inter_p, nn;
  if (!a.corrected) {
    const float al = (a.isovalue - prev_value) / (value - prev_value);
    const float bl = 1.f - al;
    inter_p = f3{bl * (p.x - dt * d.x) + al * p.x, bl * (p.y - dt * d.y) + al * p.y,
                 bl * (p.z - dt * d.z) + al * p.z};
    const f3 q = sub(inter_p, ip);
    const float h = a.du * 2.f;
    nn = f3{iso_sample<FMT>(a, s_lut, pgid, f3{q.x - h, q.y - 0.f, q.z - 0.f}) -
                iso_sample<FMT>(a, s_lut, pgid, f3{q.x + h, q.y + 0.f, q.z + 0.f}),
            iso_sample<FMT>(a, s_lut, pgid, f3{q.x - 0.f, q.y - h, q.z - 0.f}) -
                iso_sample<FMT>(a, s_lut, pgid, f3{q.x + 0.f, q.y + h, q.z + 0.f}),
            iso_sample<FMT>(a, s_lut, pgid, f3{q.x - 0.f, q.y - 0.f, q.z - h}) -
                iso_sample<FMT>(a, s_lut, pgid, f3{q.x + 0.f, q.y + 0.f, q.z + h})};
  } else {
    f3 plo = p;
    f3 phi = f3{p.x - dt * d.x, p.y - dt * d.y, p.z - dt * d.z};
    const float sprev = gsign(prev_value - a.isovalue);
    for (int it = 0; it < a.bisect_steps; ++it) {
      const f3 mid = {(plo.x + phi.x) * 0.5f, (plo.y + phi.y) * 0.5f, (plo.z + phi.z) * 0.5f};
      const f3 mip = floor3(mid);
      const int mpg = a.vaddr[grid_index(mip, a.vdim)];
      const float mv = mpg == -1 ? value : iso_sample<FMT>(a, s_lut, mpg, sub(mid, mip));
      if (gsign(mv - a.isovalue) == sprev) phi = mid; else plo = mid;
    }
    inter_p = f3{(plo.x + phi.x) * 0.5f, (plo.y + phi.y) * 0.5f, (plo.z + phi.z) * 0.5f};
    f3 gip = floor3(inter_p);
    int gpg = a.vaddr[grid_index(gip, a.vdim)];
    if (gpg == -1) { gpg = pgid; gip = ip; }
    const f3 q = sub(inter_p, gip);
    const float h = a.du;
    const f3 g = {iso_sample<FMT>(a, s_lut, gpg, f3{q.x + h, q.y + 0.f, q.z + 0.f}) -
                      iso_sample<FMT>(a, s_lut, gpg, f3{q.x - h, q.y - 0.f, q.z - 0.f}),
                  iso_sample<FMT>(a, s_lut, gpg, f3{q.x + 0.f, q.y + h, q.z + 0.f}) -
                      iso_sample<FMT>(a, s_lut, gpg, f3{q.x - 0.f, q.y - h, q.z - 0.f}),
                  iso_sample<FMT>(a, s_lut, gpg, f3{q.x + 0.f, q.y + 0.f, q.z + h}) -
                      iso_sample<FMT>(a, s_lut, gpg, f3{q.x - 0.f, q.y - 0.f, q.z - h})};
    nn = muls(g, sprev);
  }
  float wp4[4], n4[4];
  mat4_vec4(a.to_world, inter_p.x, inter_p.y, inter_p.z, 1.f, wp4);
  mat4_vec4(a.to_world, nn.x, nn.y, nn.z, 0.f, n4);
  const f3 world_p = {wp4[0], wp4[1], wp4[2]};
  const f3 nrm = normalize3(f3{n4[0], n4[1], n4[2]});
  const f3 light_dir = normalize3(sub(a.light_pos, world_p));
  const f3 hv = a.corrected ? normalize3(add(light_dir, normalize3(sub(a.eye_pos, world_p))))
                            : normalize3(sub(light_dir, world_p));
  const float ambient = 0.2f;
  const float diffuse = .6f * gclamp(dot3(light_dir, nrm), 0.f, 1.f);
  const float specular = .2f * powf(gclamp(dot3(hv, nrm), 0.f, 1.f), 100.f);
  const float dist = length3(sub(world_p, a.eye_pos)) / 2.f;
  if (a.mode == HYD_ISO_COLOR) {
    const float sc = ambient + (diffuse + specular) / dist;
    v = make_float4(v.x * sc, v.y * sc, v.z * sc, v.w * sc);
    v = make_float4(iso_srgb(v.x), iso_srgb(v.y), iso_srgb(v.z), v.w);
  } else if (a.mode == HYD_ISO_POSITION) {
    v = make_float4(inter_p.x / a.bmax.x, inter_p.y / a.bmax.y, inter_p.z / a.bmax.z, 1.f);
  } else {
    v = make_float4(nrm.x, nrm.y, nrm.z, 1.f);
  }
  const float depth = a.corrected ? length3(sub(inter_p, origin)) : length3(sub(origin, o));
  hyd_iso_pixel px;
  px.v[0] = v.x; px.v[1] = v.y; px.v[2] = v.z; px.v[3] = v.w;
  px.depth = depth;
  px.hit = 1;
  *out = px;
}

}  // namespace

hipError_t launch_blocks(const BlocksArgs &a, hipStream_t s) {
  dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile);
  hipLaunchKernelGGL(blocks_kernel, grid, dim3(64), 0, s, a);
  return hipGetLastError();
}

namespace {
// HYD_VOLUME_OFFSETS64=1 selects the 64-bit offsets for any pool (same image).
bool iso_force_offsets64() {
  static const bool on = [] {
    const char *e = std::getenv("HYD_VOLUME_OFFSETS64");
    return e && e[0] == '1';
  }();
  return on;
}

template <int FMT>
void launch_iso_fmt(const IsoArgs &a, hipStream_t s) {
  dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile);
  const uint64_t bytes = a.pool_elems * (uint64_t)sizeof(typename RawVox<FMT>::T);
  if (bytes <= (1ull << 32) && !iso_force_offsets64())
    hipLaunchKernelGGL((iso_march_grouped_kernel<FMT, false>), grid, dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((iso_march_grouped_kernel<FMT, true>), grid, dim3(64), 0, s, a);
  hipLaunchKernelGGL(iso_hit_kernel<FMT>, grid, dim3(64), 0, s, a);
}
}  // namespace

hipError_t launch_iso(const IsoArgs &a, int fmt, hipStream_t s) {
  switch (fmt) {
    case HYD_FMT_U8: launch_iso_fmt<HYD_FMT_U8>(a, s); break;
    case HYD_FMT_F16: launch_iso_fmt<HYD_FMT_F16>(a, s); break;
    default: launch_iso_fmt<HYD_FMT_F32>(a, s);
  }
  return hipGetLastError();
}

}  // namespace hyd
