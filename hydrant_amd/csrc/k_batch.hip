// Batched persistent volume renderer for gfx950.
//
// Many views of several volumes (uneven resolutions) are rendered by ONE
// launch: a grid of persistent single-wave workgroups drains a queue of
// super-tiles (TW x 8 pixels of one view), each marched as 8x8 ray packets.
// The queue is split into 8 contiguous chunks, one per XCD (workgroups are
// dispatched round-robin over the XCDs, so workgroup i runs on XCD i % 8);
// a wave drains its own chunk first, so neighbouring tiles of a view - which
// touch the same bricks - share that XCD's L2, and then steals from the others.
//
// Schedule: the lattice schedule (sample k at the midpoint of its coverage,
// sample weighted by its coverage, empty regions skipped, early termination at
// ert).  The march is a brick-level DDA: the cell code (page / not resident /
// empty distance) is looked up once per brick, the number of lattice samples
// inside the brick comes from the ray's exit distance, and the samples inside
// a brick step a precomputed local frame q(k) = qo + qd*(k + 0.5) in padded
// voxel units, i.e. fx = (p - ip)*bs + (apron - 0.5) - the reference mapping
// pad*(k*x + b) - 0.5 with k = bs/pad, b = apron/pad, whose Wrap mod is the
// identity for an apron >= 1.  A sample that lands on a brick face may be read
// from either neighbour: the apron holds the neighbour's voxels, so the value
// is the same up to rounding.  UNROLL samples are resolved per inner
// iteration so their tap loads are in flight together; compositing stays
// strictly in order.  A lane whose sample falls in the same cell as its
// previous sample (about half of them at half-voxel steps) issues no tap loads
// and reuses that sample's taps.  FP16 / u8 / FP32 bricks, FP32 accumulation.
//
// Pool orientation: the pool is held in three layouts (x, y or z the fastest
// voxel axis, the other two following cyclically).  The vector-memory unit,
// not HBM, bounds this kernel, and it serves a gather at full rate only for
// quads (4 consecutive lanes) whose addresses share a cache line; a quad that
// spans lines is served one lane at a time.  The four rays of a quad sample
// points about a pixel apart along one screen direction, so each tile picks
// the lane order (quads along screen x or along screen y) and the layout whose
// fastest axis is best aligned with that direction at the tile: the quad's
// taps then lie along one voxel row.  The eight taps are put back in x-y-z
// order before the interpolation and a ray's work does not depend on its lane,
// so every choice gives bit-identical images.  FP16 and u8 pools are held as
// quads (each element also holds the voxel's neighbours along the layout's
// middle and slow axes): a sample's eight taps are then one 16-byte (FP16) or
// 8-byte (u8) load.  Row pairs (the neighbour along the middle axis only; two
// loads per sample) serve when the device cannot hold the quads; single-voxel
// FP16 layouts need four 2-byte-aligned dwords, which cost two accesses per
// lane each.
#include <hip/hip_fp16.h>

#include <type_traits>

#include "hyd_batch.hpp"

namespace hyd {
namespace {

constexpr int kMaxTf = 256;

// Stored voxel types.  The interpolated raw value is multiplied once by the
// volume's value_scale (1/255 for bricks that hold 8-bit counts).  Voxels are
// addressed by a 32-bit byte offset from the (uniform) pool base, so a load is
// one global_load with a scalar base, and the neighbour offsets of a brick
// edge known at compile time fold into the load's immediate offset.
template <int FMT> struct BVox;
template <> struct BVox<HYD_FMT_F32> { using T = float; };
template <> struct BVox<HYD_FMT_F16> { using T = __half; };
template <> struct BVox<HYD_FMT_U8> { using T = uint8_t; };

struct RayState {
  float ox, oy, oz, dx, dy, dz;
  float ix, iy, iz;   // 1 / d
  float lf;
  int k, kend;
  float cdu;
  float4 v;
};

// One resolved sample: tap base index and trilinear weights.
struct Tap {
  uint32_t off;
  float ax, ay, az;
};

// Pool layouts: PERM 0 stores (x, y, z) with x fastest, PERM 1 (y, z, x) with
// y fastest, PERM 2 (z, x, y) with z fastest; a voxel's index in a page is
// (a2 * pad + a1) * pad + a0 for the layout's axis order (a0, a1, a2).
template <int PERM>
__device__ __forceinline__ float layout_index(float flx, float fly, float flz, float fpad) {
  if (PERM == 1) return fmaf(fmaf(flx, fpad, flz), fpad, fly);
  if (PERM == 2) return fmaf(fmaf(fly, fpad, flx), fpad, flz);
  return fmaf(fmaf(flz, fpad, fly), fpad, flx);
}

// Position in the fetched taps (bits along a0, a1, a2) of the tap with x-y-z
// bits c = bx + 2 by + 4 bz.
template <int PERM>
__host__ __device__ constexpr int tap_slot(int c) {
  return PERM == 1 ? ((c >> 1) & 1) + 2 * (c >> 2) + 4 * (c & 1)
       : PERM == 2 ? (c >> 2) + 2 * (c & 1) + 4 * ((c >> 1) & 1)
                   : c;
}

// The tap base index is formed in float: exact while pad^3 < 2^24, one
// conversion instead of three plus two integer products.  The march only
// resolves samples inside the brick (local coordinates in [apron - 0.5,
// apron - 0.5 + bs] up to rounding, within [0, pad - 1) for an apron >= 1), so
// the coordinates need no clamp; the index is still bounded by lim (as
// unsigned, so a negative index is bounded too) to keep all eight taps inside
// the page whatever the position.
template <int PERM>
__device__ __forceinline__ Tap local_tap(float fx, float fy, float fz, float fpad, uint32_t lim, uint32_t base) {
  const float flx = floorf(fx), fly = floorf(fy), flz = floorf(fz);
  Tap t;
  t.ax = fx - flx;
  t.ay = fy - fly;
  t.az = fz - flz;
  t.off = base + min((uint32_t)(int32_t)layout_index<PERM>(flx, fly, flz, fpad), lim);
  return t;
}

// Fetched taps in the layout's order, as loaded: FP16 and u8 as four words
// each holding a pair of neighbours along the fastest axis (u8 in the low 16
// bits), FP32 as floats.  Values are unpacked / converted only in the interpolation,
// so a sample that reuses its predecessor's taps selects whole registers.
template <int FMT> struct TapVals { float v[8]; };
template <> struct TapVals<HYD_FMT_U8> { uint32_t v[4]; };
template <> struct TapVals<HYD_FMT_F16> { uint32_t v[4]; };

template <int FMT> constexpr int tap_regs() { return FMT == HYD_FMT_F32 ? 8 : 4; }

// The 8 taps: two planes along the slowest axis, each addressed once; PAD > 0
// is the padded brick edge as a compile-time constant (0: runtime pad).
template <int FMT, int PAD>
__device__ __forceinline__ TapVals<FMT> fetch(const char *pool, uint32_t off, uint32_t pad_rt) {
  using V = BVox<FMT>;
  const uint32_t pad = PAD > 0 ? (uint32_t)PAD : pad_rt;
  const uint32_t b0 = off * (uint32_t)sizeof(typename V::T);
  const uint32_t b1 = (off + pad * pad) * (uint32_t)sizeof(typename V::T);
  TapVals<FMT> r;
  if constexpr (FMT == HYD_FMT_F16) {
    // pairs (a0, a0 + 1) at 2-byte alignment: one unaligned dword each
    __builtin_memcpy(&r.v[0], pool + b0, 4);
    __builtin_memcpy(&r.v[1], pool + b0 + 2 * pad, 4);
    __builtin_memcpy(&r.v[2], pool + b1, 4);
    __builtin_memcpy(&r.v[3], pool + b1 + 2 * pad, 4);
  } else if constexpr (FMT == HYD_FMT_U8) {
    // pairs (a0, a0 + 1) at any alignment: one 2-byte load each, a register each
    uint16_t w[4];
    __builtin_memcpy(&w[0], pool + b0, 2);
    __builtin_memcpy(&w[1], pool + b0 + pad, 2);
    __builtin_memcpy(&w[2], pool + b1, 2);
    __builtin_memcpy(&w[3], pool + b1 + pad, 2);
#pragma unroll
    for (int i = 0; i < 4; ++i) r.v[i] = w[i];
  } else {
    const float *q0 = reinterpret_cast<const float *>(pool + b0), *q1 = reinterpret_cast<const float *>(pool + b1);
    r.v[0] = q0[0]; r.v[1] = q0[1]; r.v[2] = q0[pad]; r.v[3] = q0[pad + 1];
    r.v[4] = q1[0]; r.v[5] = q1[1]; r.v[6] = q1[pad]; r.v[7] = q1[pad + 1];
  }
  return r;
}

// FP16 row-pair layouts: element (a0, a1, a2) of layout PERM holds the two
// halves (v(a0, a1, a2), v(a0, a1 + 1, a2)) (a1 + 1 clamped at the page edge),
// so the four taps of one plane are the 8 bytes of elements a0 and a0 + 1: one
// 4-byte-aligned dwordx2 per plane, two loads per sample.  Taps: word
// b0 + 2 b2, half b1 of the layout bits (b0, b1, b2).
template <int PERM>
__host__ __device__ constexpr int pair_slot(int c) {
  return PERM == 1 ? 2 * (((c >> 1) & 1) + 2 * (c & 1)) + (c >> 2)
       : PERM == 2 ? 2 * ((c >> 2) + 2 * ((c >> 1) & 1)) + (c & 1)
                   : 2 * ((c & 1) + 2 * (c >> 2)) + ((c >> 1) & 1);
}

// u8 row-pair layouts: 2-byte elements (v(a0, a1, a2), v(a0, a1 + 1, a2)), so
// a plane's four taps are the 4 bytes of elements a0 and a0 + 1: one load per
// plane at 2-byte alignment.  Taps: byte 2 b0 + b1 of word b2.
template <int FMT, int PAD>
__device__ __forceinline__ TapVals<FMT> fetch_pairs(const char *pool, uint32_t off, uint32_t pad_rt) {
  const uint32_t pad = PAD > 0 ? (uint32_t)PAD : pad_rt;
  TapVals<FMT> r;
  if constexpr (FMT == HYD_FMT_F16) {
    __builtin_memcpy(&r.v[0], pool + off * 4u, 8);
    __builtin_memcpy(&r.v[2], pool + (off + pad * pad) * 4u, 8);
  } else {
    __builtin_memcpy(&r.v[0], pool + off * 2u, 4);
    __builtin_memcpy(&r.v[1], pool + (off + pad * pad) * 2u, 4);
    r.v[2] = r.v[3] = 0u;
  }
  return r;
}

// Quad layouts (FP16 and u8): element (a0, a1, a2) holds the four voxels
// v(a0, a1, a2), v(a0, a1 + 1, a2), v(a0, a1, a2 + 1), v(a0, a1 + 1, a2 + 1)
// (neighbours clamped at the page edge), so a sample's eight taps are elements
// a0 and a0 + 1: one 16-byte (FP16) or 8-byte (u8) load.  Tap 4 b0 + 2 b2 + b1
// of the layout bits (b0, b1, b2), counted in voxels.  A layout can pass
// 4 GiB: 64-bit byte offsets.
template <int PERM>
__host__ __device__ constexpr int quad_slot(int c) {
  return PERM == 1 ? 4 * ((c >> 1) & 1) + 2 * (c & 1) + (c >> 2)
       : PERM == 2 ? 4 * (c >> 2) + 2 * ((c >> 1) & 1) + (c & 1)
                   : 4 * (c & 1) + 2 * (c >> 2) + ((c >> 1) & 1);
}

template <int FMT>
__device__ __forceinline__ TapVals<FMT> fetch_quads(const char *pool, uint32_t off) {
  TapVals<FMT> r;
  if constexpr (FMT == HYD_FMT_F16) {
    __builtin_memcpy(&r.v[0], pool + (uint64_t)off * 8u, 16);
  } else {
    __builtin_memcpy(&r.v[0], pool + (uint64_t)off * 4u, 8);
    r.v[2] = r.v[3] = 0u;
  }
  return r;
}

// Tap with x-y-z bits c of layout PERM, as a float (u8 / FP32).
template <int PERM, int FMT, int PAIRS>
__device__ __forceinline__ float tap_f(const TapVals<FMT> &t, int c) {
  if constexpr (FMT == HYD_FMT_U8 && PAIRS) {
    const int sl = PAIRS == 2 ? quad_slot<PERM>(c) : pair_slot<PERM>(c);   // byte sl & 3 of word sl / 4
    return float((t.v[sl >> 2] >> (8 * (sl & 3))) & 0xffu);
  } else if constexpr (FMT == HYD_FMT_U8) {
    const int sl = tap_slot<PERM>(c);   // byte sl & 1 of pair sl / 2
    return float((t.v[sl >> 1] >> (8 * (sl & 1))) & 0xffu);
  } else {
    return t.v[tap_slot<PERM>(c)];
  }
}

// FP16 tap with x-y-z bits c of layout PERM (slot s: word s / 2, high half if s is odd).
template <int PERM, int PAIRS>
__device__ __forceinline__ __half tap_h(const TapVals<HYD_FMT_F16> &t, int c) {
  const int s = PAIRS == 2 ? quad_slot<PERM>(c) : PAIRS ? pair_slot<PERM>(c) : tap_slot<PERM>(c);
  const uint32_t w = t.v[s >> 1];
  return __ushort_as_half((unsigned short)((s & 1) ? (w >> 16) : (w & 0xffffu)));
}

// Trilinear x -> y -> z.
template <int PERM, int FMT, int PAIRS>
__device__ __forceinline__ float lerp8(const TapVals<FMT> &v, const Tap &t) {
  const auto tap = [&](int c) { return tap_f<PERM, FMT, PAIRS>(v, c); };
  const float x00 = fmaf(t.ax, tap(1) - tap(0), tap(0));
  const float x10 = fmaf(t.ax, tap(3) - tap(2), tap(2));
  const float x01 = fmaf(t.ax, tap(5) - tap(4), tap(4));
  const float x11 = fmaf(t.ax, tap(7) - tap(6), tap(6));
  const float y0 = fmaf(t.ay, x10 - x00, x00);
  const float y1 = fmaf(t.ay, x11 - x01, x01);
  return fmaf(t.az, y1 - y0, y0);
}

// FP16: x-lerps as (v0 - a v0) + a v1, both FMAs taking the half taps directly
// (v_fma_mix_f32), no separate conversions.
template <int PERM, int PAIRS>
__device__ __forceinline__ float lerp8_h(const TapVals<HYD_FMT_F16> &v, const Tap &t) {
  float x[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const __half2 h = __halves2half2(tap_h<PERM, PAIRS>(v, 2 * i), tap_h<PERM, PAIRS>(v, 2 * i + 1));
    const float h0 = __low2float(h), h1 = __high2float(h);
    x[i] = fmaf(t.ax, h1, fmaf(-t.ax, h0, h0));
  }
  const float y0 = fmaf(t.ay, x[1] - x[0], x[0]);
  const float y1 = fmaf(t.ay, x[3] - x[2], x[2]);
  return fmaf(t.az, y1 - y0, y0);
}

template <int PERM, int FMT, int PAIRS>
__device__ __forceinline__ float interpolate(const TapVals<FMT> &v, const Tap &t) {
  if constexpr (FMT == HYD_FMT_F16) return lerp8_h<PERM, PAIRS>(v, t);
  else return lerp8<PERM, FMT, PAIRS>(v, t);
}

// TF lookup on the edge-extended table in slope form: entry i holds
// p = ext[i] and dq = ext[i+1] - ext[i] with ext[0] = tf[0], ext[1..n] =
// tf[0..n-1], ext[n+1] = tf[n-1].  fx = n*s - 0.5 lies in [-0.5, n - 0.5] for
// s in [0, 1], so floor(fx) + 1 indexes the table without clamping and
// reproduces the reference's clamp-to-edge; p + t*dq is the lerp
// fma(t, q - p, p).  s outside [0, 1] is the Border zero.  A sample is
// composited only while the ray is live (alpha 0 leaves v unchanged), so the
// early-termination test needs no branch per sample.
struct TfSlope {
  float4 p, dq;
};

__device__ __forceinline__ void composite(const TfSlope *ext, float fsc, float smax, float raw, float w, bool live,
                                          float4 &v) {
  // s = raw * value_scale; fx = n s - 0.5 = (n value_scale) raw - 0.5; inside <=> 0 <= raw <= 1 / value_scale
  const float fx = fmaf(fsc, raw, -0.5f);
  const bool inside = (0.f <= raw) & (raw <= smax);
  const float flr = floorf(fx);
  const float t = fx - flr;
  const int i = inside ? (int)flr + 1 : 0;
  const TfSlope e = ext[i];
  // the alpha column of the table is premultiplied by the density
  const float alpha = ((inside & live) ? w : 0.f) * fmaf(t, e.dq.w, e.p.w);
  const float tr = (1.f - v.w) * alpha;
  v.x = fmaf(fmaf(t, e.dq.x, e.p.x), tr, v.x);
  v.y = fmaf(fmaf(t, e.dq.y, e.p.y), tr, v.y);
  v.z = fmaf(fmaf(t, e.dq.z, e.p.z), tr, v.z);
  v.w = fmaf(1.f - v.w, alpha, v.w);
}

// Cell code: >= 0 page id, -1 not resident, <= -2 empty with chebyshev -1 - code.
__device__ __forceinline__ int cell_code(const BatchVolDev &vol, const int32_t *cells, float ipx, float ipy,
                                         float ipz) {
  const int ix = min(max((int)ipx, 0), vol.grid[0] - 1);
  const int iy = min(max((int)ipy, 0), vol.grid[1] - 1);
  const int iz = min(max((int)ipz, 0), vol.grid[2] - 1);
  return cells[vol.cell_off + (iz * vol.grid[1] + iy) * vol.grid[0] + ix];
}

// Ray parameter (from o0) at which the ray leaves the unit box [ip, ip + 1].
// The exit face follows the sign of 1/d, so d = +0 / -0 give +inf slabs.
__device__ __forceinline__ float box_exit(const RayState &r, float ipx, float ipy, float ipz) {
  const float tx = ((r.ix > 0.f ? ipx + 1.f : ipx) - r.ox) * r.ix;
  const float ty = ((r.iy > 0.f ? ipy + 1.f : ipy) - r.oy) * r.iy;
  const float tz = ((r.iz > 0.f ? ipz + 1.f : ipz) - r.oz) * r.iz;
  return fminf(fminf(tx, ty), tz);
}

template <int FMT, int UNROLL, bool COUNT, int PAD, int PERM, int PAIRS>
__device__ void march(const BatchArgs &a, const BatchVolDev &vol, const TfSlope *ext, RayState &r,
                      unsigned &count) {
  const float step = vol.step, inv_step = 1.f / vol.step;
  const float fsc = float(a.tf_n) * vol.value_scale, smax = 1.f / vol.value_scale;
  const float ert = a.ert > 0.f ? a.ert : 3.f;   // opacity never exceeds 1: 3 disables termination
  const uint32_t pad = PAD > 0 ? (uint32_t)PAD : (uint32_t)a.pad;
  const float fpad = float(pad);
  const uint32_t lim = pad * pad * pad - pad * pad - pad - 2u;   // largest base index of a tap cell
  const char *pool = static_cast<const char *>(a.pools[PERM]);
  const float fbs = float(vol.bs), off = float(vol.apron) - 0.5f;
  // Tap reuse: a sample in its predecessor's cell takes that sample's taps
  // instead of loading them.  It pays for FP16 quads (16-byte loads keep the
  // data path busy); for u8 quads (8-byte loads) the registers and selects it
  // costs outweigh the loads it saves (55.3 ms with, 52.7 ms without, per
  // batch of the 96-view workload).
  constexpr bool kReuse = !(FMT == HYD_FMT_U8 && PAIRS == 2);
  uint32_t poff = 0xffffffffu;   // tap offset of the last resolved sample (none yet)
  TapVals<FMT> pv{};             // its taps
  // fetched taps (layout order) of the group's samples; kept across groups so
  // that a lane reusing taps leaves its registers untouched (stale values the
  // select below never picks) instead of initialising them
  TapVals<FMT> tv[UNROLL] = {};
  while (r.k < r.kend) {
    const float t0 = fmaf(0.5f, fminf(1.f, r.lf - float(r.k)), float(r.k)) * step;
    const float ipx = floorf(fmaf(r.dx, t0, r.ox));
    const float ipy = floorf(fmaf(r.dy, t0, r.oy));
    const float ipz = floorf(fmaf(r.dz, t0, r.oz));
    const int code = cell_code(vol, a.cells, ipx, ipy, ipz);
    const float t_exit = box_exit(r, ipx, ipy, ipz);
    if (code <= -2) {   // empty: jump to the first lattice point past the empty region
      const int kn = (int)ceilf(fmaf(t_exit + float(-1 - code - 1) * r.cdu, inv_step, -0.5f));
      r.k = kn > r.k + 1 ? kn : r.k + 1;
      continue;
    }
    if (code == -1) { r.k = r.kend; break; }   // page fault: the ray stops (reference semantics)
    // lattice samples with t < t_exit belong to this brick
    int kb = (int)ceilf(fmaf(t_exit, inv_step, -0.5f));
    kb = kb > r.k + 1 ? kb : r.k + 1;
    kb = kb < r.kend ? kb : r.kend;
    const uint32_t base = (uint32_t)code * a.page_elems;
    // local frame in padded-voxel units: q(k) = qo + qd * (k + 0.5)
    const float qox = fmaf(r.ox - ipx, fbs, off), qoy = fmaf(r.oy - ipy, fbs, off), qoz = fmaf(r.oz - ipz, fbs, off);
    const float qdx = r.dx * fbs * step, qdy = r.dy * fbs * step, qdz = r.dz * fbs * step;
    int k = r.k;
    bool done = false;
    // groups of UNROLL samples with all tap loads in flight together; slots
    // past kb repeat the brick's last sample with weight 0 (they composite
    // nothing; with reuse they take that sample's taps and issue no loads,
    // without it they load them again: every slot loading keeps no taps live
    // across groups, 86 VGPRs against 99 for loading only valid slots)
    const float klast = float(kb - 1);
    while (k < kb) {
      const float kf_base = float(k);   // sample indices stay exact in float (< 2^24)
      const int nvalid = min(UNROLL, kb - k);
      Tap tp[UNROLL];
      bool fresh[UNROLL];        // the sample's cell differs from the previous sample's
      float wv[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        const float kf0 = fminf(kf_base + float(u), klast);
        const float w = fminf(1.f, r.lf - kf0);
        wv[u] = u < nvalid ? w : 0.f;
        const float kf = fmaf(0.5f, w, kf0);
        tp[u] = local_tap<PERM>(fmaf(qdx, kf, qox), fmaf(qdy, kf, qoy), fmaf(qdz, kf, qoz), fpad, lim, base);
        fresh[u] = !kReuse || tp[u].off != (u == 0 ? poff : tp[u - 1].off);
        if (fresh[u]) {
          if constexpr (PAIRS == 2) tv[u] = fetch_quads<FMT>(pool, tp[u].off);
          else if constexpr (PAIRS) tv[u] = fetch_pairs<FMT, PAD>(pool, tp[u].off, pad);
          else tv[u] = fetch<FMT, PAD>(pool, tp[u].off, pad);
        }
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        // a sample in the previous sample's cell reuses its taps (same offset, same voxels)
        // the comparison is redone on an offset passed through an empty asm, so
        // that the compiler cannot fold this select into the fetch branch (as
        // "copy the previous taps, then load": the copy would wait for the
        // previous sample's loads)
        if constexpr (kReuse) {
          uint32_t o = tp[u].off;
          asm("" : "+v"(o));
          const bool keep = o == (u == 0 ? poff : tp[u - 1].off);
#pragma unroll
          for (int c = 0; c < tap_regs<FMT>(); ++c) tv[u].v[c] = keep ? (u == 0 ? pv.v[c] : tv[u - 1].v[c]) : tv[u].v[c];
        }
        const bool live = r.v.w <= ert;
        composite(ext, fsc, smax, interpolate<PERM, FMT, PAIRS>(tv[u], tp[u]), wv[u], live, r.v);
        if (COUNT) count += (live && u < nvalid) ? 1u : 0u;
      }
      poff = tp[UNROLL - 1].off;
      pv = tv[UNROLL - 1];
      k += nvalid;
      if (r.v.w > ert) { done = true; break; }
    }
    r.k = done ? r.kend : k;
  }
}

__device__ __forceinline__ unsigned claim(unsigned *counter) {
  unsigned t = 0;
  if (__lane_id() == 0) t = atomicAdd(counter, 1u);
  return __shfl(t, 0);
}

// GROUP 1: single-wave workgroups, each claiming TW x 8 tiles.  GROUP 4: four
// waves per workgroup claim a 16 x 16 block and take one 8 x 8 tile each, so
// the four rays packets that touch the same bricks share one CU's L1.
template <int FMT, int TW, int UNROLL, bool COUNT, int PAD, int GROUP, int PAIRS>
__global__ __launch_bounds__(64 * GROUP) void batch_kernel(BatchArgs a) {
  extern __shared__ TfSlope s_ext[];   // tf_n + 1 entries (dynamic LDS)
  __shared__ unsigned long long s_count;
  __shared__ unsigned s_claim;
  const float4 *tf4 = reinterpret_cast<const float4 *>(a.tf);
  for (int i = threadIdx.x; i < a.tf_n + 1; i += 64 * GROUP) {
    const int j0 = i == 0 ? 0 : i - 1, j1 = i > a.tf_n - 1 ? a.tf_n - 1 : i;
    const float4 p = tf4[j0], q = tf4[j1];
    const float pw = p.w * a.density, qw = q.w * a.density;
    s_ext[i].p = make_float4(p.x, p.y, p.z, pw);
    s_ext[i].dq = make_float4(q.x - p.x, q.y - p.y, q.z - p.z, qw - pw);
  }
  if (COUNT && threadIdx.x == 0) s_count = 0;
  __syncthreads();
  unsigned count = 0;
  const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
  const int home = blockIdx.x % kXcds;
  for (int c = 0; c < kXcds; ++c) {
    const int chunk = (home + c) % kXcds;
    const unsigned base = (unsigned)a.chunk_start[chunk];
    const unsigned len = (unsigned)(a.chunk_start[chunk + 1] - a.chunk_start[chunk]);
    while (true) {
      unsigned t;
      if (GROUP == 1) {
        t = claim(a.counters + chunk);
      } else {
        if (threadIdx.x == 0) s_claim = atomicAdd(a.counters + chunk, 1u);
        __syncthreads();
        t = s_claim;
        __syncthreads();   // every wave has read the claim before the next one is written
      }
      if (t >= len) break;
      const uint32_t desc = a.tiles[base + t];
      const int view = (int)(desc >> 16);
      const int tx = (int)(desc & 0xffu), ty = (int)((desc >> 8) & 0xffu);
      const BatchViewDev &vw = a.views[view];
      const BatchVolDev &vol = a.vols[vw.volume];
      const int x0 = GROUP == 1 ? tx * TW : tx * 16 + (wave & 1) * 8;
      const int y0 = GROUP == 1 ? ty * 8 : ty * 16 + (wave >> 1) * 8;
      const f3 ro = f3{vw.ray_o[0], vw.ray_o[1], vw.ray_o[2]};
      // lane order and layout: the screen direction (x or y) and volume axis of
      // best alignment, from the ray directions one pixel apart at the tile centre
      int perm, cols;
      {
        const int cx = x0 + (GROUP == 1 ? TW / 2 : 4), cy = y0 + 4;
        const f3 dc = pixel_ray_dir(vw.trans, vw.ctg, ro, cx, cy, vw.W, vw.H);
        const f3 dx = sub(pixel_ray_dir(vw.trans, vw.ctg, ro, cx + 1, cy, vw.W, vw.H), dc);
        const f3 dy = sub(pixel_ray_dir(vw.trans, vw.ctg, ro, cx, cy + 1, vw.W, vw.H), dc);
        const float nx = comp_max(abs3(dx)) * rsqrtf(dot3(dx, dx) + 1e-30f);
        const float ny = comp_max(abs3(dy)) * rsqrtf(dot3(dy, dy) + 1e-30f);
        cols = ny > nx ? 1 : 0;
        const f3 m = abs3(cols ? dy : dx);
        perm = !a.oriented ? 0 : (m.x >= m.y && m.x >= m.z) ? 0 : (m.y >= m.z ? 1 : 2);
        perm = __builtin_amdgcn_readfirstlane(perm);
        cols = __builtin_amdgcn_readfirstlane(cols);
      }
      const int lx = cols ? (lane >> 3) : (lane & 7), ly = cols ? (lane & 7) : (lane >> 3);
#pragma unroll 1
      for (int sub = 0; sub < (GROUP == 1 ? TW / 8 : 1); ++sub) {
        const int x = x0 + sub * 8 + lx;
        const int y = y0 + ly;
        if (x >= vw.W || y >= vw.H) continue;
        const f3 d = pixel_ray_dir(vw.trans, vw.ctg, ro, x, y, vw.W, vw.H);
        RayState r;
        r.v = make_float4(0.f, 0.f, 0.f, 0.f);
        float tnear, tfar;
        if (intersect_box(ro, d, f3{0.f, 0.f, 0.f}, f3{vol.bmax[0], vol.bmax[1], vol.bmax[2]}, tnear, tfar)) {
          if (tnear < 0) tnear = 0;
          if (tfar < 0) tfar = 0;
          r.ox = ro.x + d.x * tnear; r.oy = ro.y + d.y * tnear; r.oz = ro.z + d.z * tnear;
          r.dx = d.x; r.dy = d.y; r.dz = d.z;
          r.ix = 1.f / d.x; r.iy = 1.f / d.y; r.iz = 1.f / d.z;
          r.lf = fminf((tfar - tnear) / vol.step, float(a.max_steps));
          r.k = 0;
          r.kend = (int)ceilf(r.lf);
          r.cdu = 1.f / comp_max(abs3(d));
          if (perm == 0) march<FMT, UNROLL, COUNT, PAD, 0, PAIRS>(a, vol, s_ext, r, count);
          else if (perm == 1) march<FMT, UNROLL, COUNT, PAD, 1, PAIRS>(a, vol, s_ext, r, count);
          else march<FMT, UNROLL, COUNT, PAD, 2, PAIRS>(a, vol, s_ext, r, count);
        }
        a.out[vw.out_offset + (int64_t)y * vw.W + x] = r.v;
      }
    }
  }
  if (COUNT) {
    if (count) atomicAdd(&s_count, (unsigned long long)count);
    __syncthreads();
    if (threadIdx.x == 0 && s_count) atomicAdd(a.samples, s_count);
  }
}

// The oriented copies of an x-fastest pool, for pages first_page + blockIdx.x.
// A workgroup produces one plane of one layout (layout p of page P at dst +
// p * pad^3 * total_pages + P * pad^3) and, when `staged` (the source planes
// fit in LDS), first reads the source voxels it needs row by row into LDS, so
// that both its global reads and its writes run along rows:
//   layout 1 (y, z, x), plane a1 = z: source plane z (and z + 1 for pairs);
//   layout 2 (z, x, y), plane a2 = y: the source rows of every z at that y.
// Layout 0 (the x-fastest copy of the row-pair form) reads plane z directly.
// PAIRS: elements (v(a0, a1, a2), v(a0, min(a1 + 1, pad - 1), a2)) of twice the
// voxel's size (4 bytes for a 2-byte pool, 2 for a u8 pool); otherwise single
// voxels of type T.
template <typename T, bool PAIRS>
__global__ __launch_bounds__(256) void orient_kernel(const T *src, void *dst_v, int pad, int total_pages,
                                                     int first_page, int layout0, bool staged) {
  using Pair = typename std::conditional<sizeof(T) == 1, uint16_t, uint32_t>::type;
  using Out = typename std::conditional<PAIRS, Pair, T>::type;
  extern __shared__ unsigned char s_raw[];
  T *s_a = reinterpret_cast<T *>(s_raw), *s_b = s_a + pad * pad;
  const size_t page = (size_t)first_page + blockIdx.x;
  const size_t p3 = (size_t)pad * pad * pad;
  const int plane = blockIdx.y, p = layout0 + (int)blockIdx.z;
  const T *pg = src + page * p3;
  Out *out = static_cast<Out *>(dst_v) + (size_t)(p - (PAIRS ? 0 : 1)) * p3 * (size_t)total_pages + page * p3;
  const int n2 = pad * pad;
  const auto pack = [](T v0, T v1) -> Out {
    if constexpr (PAIRS) return (Out)((uint32_t)v0 | ((uint32_t)v1 << (8 * sizeof(T))));
    else return v0;
  };
  if (p == 0) {   // (x, y, z): plane z; the pair neighbour is y + 1
    const int z = plane;
    for (int r = threadIdx.x; r < n2; r += 256) {
      const int x = r % pad, y = r / pad, y1 = y + 1 < pad ? y + 1 : pad - 1;
      out[(size_t)z * n2 + r] = pack(pg[((size_t)z * pad + y) * pad + x], pg[((size_t)z * pad + y1) * pad + x]);
    }
  } else if (p == 1) {   // (y, z, x): plane a1 = z of every a2 = x; the pair neighbour is z + 1
    const int z = plane, z1 = z + 1 < pad ? z + 1 : pad - 1;
    if (staged) {
      for (int r = threadIdx.x; r < n2; r += 256) {
        s_a[r] = pg[(size_t)z * n2 + r];
        if (PAIRS) s_b[r] = pg[(size_t)z1 * n2 + r];
      }
      __syncthreads();
    }
    for (int r = threadIdx.x; r < n2; r += 256) {
      const int y = r % pad, x = r / pad;   // consecutive threads: consecutive y (a0)
      const T v0 = staged ? s_a[y * pad + x] : pg[((size_t)z * pad + y) * pad + x];
      const T v1 = !PAIRS ? v0 : staged ? s_b[y * pad + x] : pg[((size_t)z1 * pad + y) * pad + x];
      out[((size_t)x * pad + z) * pad + y] = pack(v0, v1);
    }
  } else {   // (z, x, y): plane a2 = y; the pair neighbour is x + 1
    const int y = plane;
    if (staged) {
      for (int r = threadIdx.x; r < n2; r += 256) {
        const int x = r % pad, z = r / pad;
        s_a[r] = pg[((size_t)z * pad + y) * pad + x];   // s_a[z][x]
      }
      __syncthreads();
    }
    for (int r = threadIdx.x; r < n2; r += 256) {
      const int z = r % pad, x = r / pad, x1 = x + 1 < pad ? x + 1 : pad - 1;   // consecutive z (a0)
      const T v0 = staged ? s_a[z * pad + x] : pg[((size_t)z * pad + y) * pad + x];
      const T v1 = !PAIRS ? v0 : staged ? s_a[z * pad + x1] : pg[((size_t)z * pad + y) * pad + x1];
      out[(size_t)y * n2 + r] = pack(v0, v1);
    }
  }
}

__global__ void reset_counters(unsigned *c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// 64^3 bricks with a one-voxel apron (pad 66) get the compile-time brick edge;
// FP16 and u8 pools read through quad or row-pair layouts when the host built them.
template <int FMT, int TW, int UNROLL, int GROUP, int PAIRS>
hipError_t launch5(const BatchArgs &a, int grid, hipStream_t s) {
  const size_t shmem = (size_t)(a.tf_n + 1) * sizeof(TfSlope);
  const dim3 blk(64 * GROUP);
  if (a.pad == 66) {
    if (a.samples)
      hipLaunchKernelGGL((batch_kernel<FMT, TW, UNROLL, true, 66, GROUP, PAIRS>), dim3(grid), blk, shmem, s, a);
    else
      hipLaunchKernelGGL((batch_kernel<FMT, TW, UNROLL, false, 66, GROUP, PAIRS>), dim3(grid), blk, shmem, s, a);
  } else {
    if (a.samples)
      hipLaunchKernelGGL((batch_kernel<FMT, TW, UNROLL, true, 0, GROUP, PAIRS>), dim3(grid), blk, shmem, s, a);
    else
      hipLaunchKernelGGL((batch_kernel<FMT, TW, UNROLL, false, 0, GROUP, PAIRS>), dim3(grid), blk, shmem, s, a);
  }
  return hipGetLastError();
}

template <int FMT, int TW, int UNROLL, int GROUP>
hipError_t launch4(const BatchArgs &a, int grid, hipStream_t s) {
  if constexpr (FMT == HYD_FMT_F16 || FMT == HYD_FMT_U8) {
    if (a.pairs == 2) return launch5<FMT, TW, UNROLL, GROUP, 2>(a, grid, s);
    if (a.pairs == 1) return launch5<FMT, TW, UNROLL, GROUP, 1>(a, grid, s);
  }
  return launch5<FMT, TW, UNROLL, GROUP, 0>(a, grid, s);
}

template <int FMT, int TW, int UNROLL>
hipError_t launch3(const BatchArgs &a, int group, int grid, hipStream_t s) {
  if (TW == 8 && group == 4) return launch4<FMT, 8, UNROLL, 4>(a, grid, s);
  return launch4<FMT, TW, UNROLL, 1>(a, grid, s);
}

template <int FMT, int TW>
hipError_t launch2(const BatchArgs &a, int unroll, int group, int grid, hipStream_t s) {
  switch (unroll) {
    case 4: return launch3<FMT, TW, 4>(a, group, grid, s);
    case 2: return launch3<FMT, TW, 2>(a, group, grid, s);
    default: return launch3<FMT, TW, 1>(a, group, grid, s);
  }
}

template <int FMT>
hipError_t launch1(const BatchArgs &a, int tw, int unroll, int group, int grid, hipStream_t s) {
  switch (tw) {
    case 32: return launch2<FMT, 32>(a, unroll, group, grid, s);
    case 16: return launch2<FMT, 16>(a, unroll, group, grid, s);
    default: return launch2<FMT, 8>(a, unroll, group, grid, s);
  }
}

}  // namespace

// Launches orient_kernel for layouts [layout0, layout0 + n_layouts); the LDS
// staging is used while two source planes fit in 64 KiB.
template <typename T, bool PAIRS>
hipError_t launch_orient(const void *src, void *dst, int pad, int total_pages, int first_page, int n_pages,
                         int layout0, int n_layouts, hipStream_t s) {
  if (n_pages <= 0) return hipSuccess;
  const size_t lds = 2 * (size_t)pad * pad * sizeof(T);
  const bool staged = lds <= 65536;
  hipLaunchKernelGGL((orient_kernel<T, PAIRS>), dim3(n_pages, pad, n_layouts), dim3(256), staged ? lds : 0, s,
                     static_cast<const T *>(src), dst, pad, total_pages, first_page, layout0, staged);
  return hipGetLastError();
}

hipError_t launch_orient_pool(int fmt, const void *src, void *dst, int pad, int total_pages, int first_page,
                              int n_pages, hipStream_t s) {
  switch (fmt) {
    case HYD_FMT_U8: return launch_orient<uint8_t, false>(src, dst, pad, total_pages, first_page, n_pages, 1, 2, s);
    case HYD_FMT_F16: return launch_orient<uint16_t, false>(src, dst, pad, total_pages, first_page, n_pages, 1, 2, s);
    default: return launch_orient<uint32_t, false>(src, dst, pad, total_pages, first_page, n_pages, 1, 2, s);
  }
}

hipError_t launch_orient_pairs(int fmt, const void *src, void *dst, int pad, int total_pages, int first_page,
                               int n_pages, hipStream_t s) {
  if (fmt == HYD_FMT_U8) return launch_orient<uint8_t, true>(src, dst, pad, total_pages, first_page, n_pages, 0, 3, s);
  return launch_orient<uint16_t, true>(src, dst, pad, total_pages, first_page, n_pages, 0, 3, s);
}

// The three quad layouts of an x-fastest pool of voxels T (FP16 bits or u8),
// for pages first_page + blockIdx.x: block (plane s = blockIdx.y, layout p =
// blockIdx.z) writes the pad^2 elements whose quad starts at s along one axis,
// one thread per element, consecutive threads writing consecutive elements.
// Their voxels lie in two source planes (z = s, s + 1 for layouts 0 and 1;
// rows y = s, s + 1 for layout 2), staged in LDS when `staged` (layout p at
// dst + p * 4 sizeof(T) pad^3 total_pages bytes).
template <typename T>
__global__ __launch_bounds__(256) void orient_quads_kernel(const T *src, void *dst_v, int pad, int total_pages,
                                                           int first_page, bool staged) {
  using Out = typename std::conditional<sizeof(T) == 1, uint32_t, uint64_t>::type;
  constexpr int bits = 8 * sizeof(T);
  extern __shared__ unsigned char s_raw[];
  T *s_a = reinterpret_cast<T *>(s_raw), *s_b = s_a + pad * pad;
  const size_t page = (size_t)first_page + blockIdx.x;
  const size_t p3 = (size_t)pad * pad * pad;
  const int s = blockIdx.y, s1 = s + 1 < pad ? s + 1 : pad - 1, p = blockIdx.z, n2 = pad * pad;
  const T *pg = src + page * p3;
  Out *out = static_cast<Out *>(dst_v) + (size_t)p * p3 * (size_t)total_pages + page * p3;
  // voxel (i, j) of source plane q (0: s, 1: s1): planes z = s / s1 indexed
  // [y][x] for layouts 0 and 1, rows y = s / s1 indexed [z][x] for layout 2
  const auto src_at = [&](int q, int i, int j) -> Out {
    const int t = q ? s1 : s;
    return (Out)(p < 2 ? pg[((size_t)t * pad + i) * pad + j] : pg[((size_t)i * pad + t) * pad + j]);
  };
  const auto at = [&](int q, int i, int j) -> Out {
    return staged ? (Out)(q ? s_b : s_a)[i * pad + j] : src_at(q, i, j);
  };
  if (staged) {
    for (int r = threadIdx.x; r < n2; r += 256) {
      s_a[r] = (T)src_at(0, r / pad, r % pad);
      s_b[r] = (T)src_at(1, r / pad, r % pad);
    }
    __syncthreads();
  }
  const auto quad = [](Out v00, Out v10, Out v01, Out v11) -> Out {
    return v00 | (v10 << bits) | (v01 << (2 * bits)) | (v11 << (3 * bits));
  };
  for (int r = threadIdx.x; r < n2; r += 256) {
    const int u = r % pad, w = r / pad, w1 = w + 1 < pad ? w + 1 : pad - 1;
    if (p == 0) {   // (x, y, z), z = s: element (x = u, y = w); quad over y, z
      out[(size_t)s * n2 + r] = quad(at(0, w, u), at(0, w1, u), at(1, w, u), at(1, w1, u));
    } else if (p == 1) {   // (y, z, x), z = s: element (y = u, x = w); quad over z, x
      out[((size_t)w * pad + s) * pad + u] = quad(at(0, u, w), at(1, u, w), at(0, u, w1), at(1, u, w1));
    } else {   // (z, x, y), y = s: element (z = u, x = w); quad over x, y
      out[(size_t)s * n2 + r] = quad(at(0, u, w), at(0, u, w1), at(1, u, w), at(1, u, w1));
    }
  }
}

hipError_t launch_orient_quads(int fmt, const void *src, void *dst, int pad, int total_pages, int first_page,
                               int n_pages, hipStream_t s) {
  if (n_pages <= 0) return hipSuccess;
  const dim3 grid(n_pages, pad, 3);
  const size_t e = fmt == HYD_FMT_U8 ? 1 : 2, lds = 2 * (size_t)pad * pad * e;
  const bool staged = lds <= 65536;
  if (fmt == HYD_FMT_U8) {
    hipLaunchKernelGGL(orient_quads_kernel<uint8_t>, grid, dim3(256), staged ? lds : 0, s,
                       static_cast<const uint8_t *>(src), dst, pad, total_pages, first_page, staged);
  } else if (fmt == HYD_FMT_F16) {
    hipLaunchKernelGGL(orient_quads_kernel<uint16_t>, grid, dim3(256), staged ? lds : 0, s,
                       static_cast<const uint16_t *>(src), dst, pad, total_pages, first_page, staged);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &a, int fmt, int tile_w, int unroll, int group, int grid, hipStream_t s) {
  hipLaunchKernelGGL(reset_counters, dim3(1), dim3(kXcds), 0, s, a.counters, kXcds);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  switch (fmt) {
    case HYD_FMT_U8: return launch1<HYD_FMT_U8>(a, tile_w, unroll, group, grid, s);
    case HYD_FMT_F16: return launch1<HYD_FMT_F16>(a, tile_w, unroll, group, grid, s);
    default: return launch1<HYD_FMT_F32>(a, tile_w, unroll, group, grid, s);
  }
}

}  // namespace hyd
