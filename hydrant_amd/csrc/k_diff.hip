// Differentiable render: forward with per-ray checkpoints and the adjoint
// (gradients with respect to the transfer-function table and the voxels).
//
// Model (one FP32 brick of nx*ny*nz voxels spanning the unit box, K samples per
// ray at the midpoints of K equal sub-intervals of the clipped chord):
//   s_k = trilinear(volume, n*p_k - 0.5) with clamped taps,
//   (c, a) = TF(s_k) (Border + Linear lookup, transfer_fn.hpp:14-28),
//   alpha_k = 1 - exp(-sigma * max(a, 0) * dt),
//   C = sum T_k alpha_k c_k, T_{k+1} = T_k (1 - alpha_k), out = (C, 1 - T_K).
// Rays come from the reference camera model (cuda_shading.cu:17-29).
//
// Checkpoints: the forward stores T at k = 0, ckpt, 2*ckpt, ... and the final
// T (ckpt_every = 64 for 512 steps: 9 floats per ray).  The sign bit of a
// segment's checkpoint flags a segment containing a sample with alpha > 0.5.
// Adjoint: segments are processed back to front; each sample is recomputed
// (re-sampled) and T_k = T_{k+1} / (1 - alpha_k) is recovered within the
// segment (stable while alpha <= 0.5) and re-anchored at every checkpoint; a
// flagged segment recomputes T_k forward from its checkpoint instead.
// With the suffix colour S and suffix transmittance P (back to front):
//   dL/dalpha_k = T_k (g . (c_k - S_{k+1}) + gA * P_{k+1}),  dL/dc_k = T_k alpha_k g.
// TF gradients accumulate in an LDS copy per workgroup (flushed with global
// atomics); voxel gradients use hardware float atomics.
#include "hyd_diff.hpp"

namespace hyd {
namespace {

constexpr int kTile = 8;
constexpr int kMaxTf = 256;

struct Smp {
  float s;
};

// Volume taps use unsigned 32-bit byte offsets from the (wave-uniform) base, so
// each load is one global_load with a scalar base and no 64-bit address
// arithmetic (the volume is limited to 2^31 - 1 bytes).
struct VolumeRef {
  const char *base;
};

__device__ __forceinline__ VolumeRef volume_ref(const DiffArgs &a) {
  return VolumeRef{reinterpret_cast<const char *>(a.vol)};
}

__device__ __forceinline__ float vol_at(const VolumeRef &v, int byte_offset) {
  return *reinterpret_cast<const float *>(v.base + (unsigned)byte_offset);
}

__device__ __forceinline__ void sample(const DiffArgs &a, const VolumeRef &vr, float px, float py, float pz,
                                       Smp &sm) {
  const float fx = float(a.nx) * px - 0.5f, fy = float(a.ny) * py - 0.5f, fz = float(a.nz) * pz - 0.5f;
  const float flx = floorf(fx), fly = floorf(fy), flz = floorf(fz);
  const float gx = fx - flx, gy = fy - fly, gz = fz - flz;
  const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
  const int sx = a.nx * 4, sxy = a.nx * a.ny * 4;   // byte strides
  const int x0 = iclamp(ix, 0, a.nx - 1) * 4, x1 = iclamp(ix + 1, 0, a.nx - 1) * 4;
  const int y0 = iclamp(iy, 0, a.ny - 1) * sx, y1 = iclamp(iy + 1, 0, a.ny - 1) * sx;
  const int z0 = iclamp(iz, 0, a.nz - 1) * sxy, z1 = iclamp(iz + 1, 0, a.nz - 1) * sxy;
  const int r00 = z0 + y0, r01 = z0 + y1, r10 = z1 + y0, r11 = z1 + y1;
  const int id[8] = {r00 + x0, r00 + x1, r01 + x0, r01 + x1, r10 + x0, r10 + x1, r11 + x0, r11 + x1};
  const float hx = 1.f - gx, hy = 1.f - gy, hz = 1.f - gz;
  const float w[8] = {hx * hy * hz, gx * hy * hz, hx * gy * hz, gx * gy * hz,
                      hx * hy * gz, gx * hy * gz, hx * gy * gz, gx * gy * gz};
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) s = fmaf(w[q], vol_at(vr, id[q]), s);
  sm.s = s;
}

// Taps of the last sample of sample_reuse: its first and last tap offsets
// (which fix all eight) and their values.
struct TapCache {
  int id0 = -1, id7 = -1;
  float v[8];
};

// sample() for consecutive samples of one ray: a sample whose eight taps are
// the previous sample's (same first and last offset) loads nothing.  The
// weights and the sum are the same operations, so the value is bit-identical.
__device__ __forceinline__ void sample_reuse(const DiffArgs &a, const VolumeRef &vr, float px, float py, float pz,
                                             Smp &sm, TapCache &tc) {
  const float fx = float(a.nx) * px - 0.5f, fy = float(a.ny) * py - 0.5f, fz = float(a.nz) * pz - 0.5f;
  const float flx = floorf(fx), fly = floorf(fy), flz = floorf(fz);
  const float gx = fx - flx, gy = fy - fly, gz = fz - flz;
  const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
  const int sx = a.nx * 4, sxy = a.nx * a.ny * 4;   // byte strides
  const int x0 = iclamp(ix, 0, a.nx - 1) * 4, x1 = iclamp(ix + 1, 0, a.nx - 1) * 4;
  const int y0 = iclamp(iy, 0, a.ny - 1) * sx, y1 = iclamp(iy + 1, 0, a.ny - 1) * sx;
  const int z0 = iclamp(iz, 0, a.nz - 1) * sxy, z1 = iclamp(iz + 1, 0, a.nz - 1) * sxy;
  const int r00 = z0 + y0, r01 = z0 + y1, r10 = z1 + y0, r11 = z1 + y1;
  const int id[8] = {r00 + x0, r00 + x1, r01 + x0, r01 + x1, r10 + x0, r10 + x1, r11 + x0, r11 + x1};
  if (id[0] != tc.id0 || id[7] != tc.id7) {
#pragma unroll
    for (int q = 0; q < 8; ++q) tc.v[q] = vol_at(vr, id[q]);
    tc.id0 = id[0];
    tc.id7 = id[7];
  }
  const float hx = 1.f - gx, hy = 1.f - gy, hz = 1.f - gz;
  const float w[8] = {hx * hy * hz, gx * hy * hz, hx * gy * hz, gx * gy * hz,
                      hx * hy * gz, gx * hy * gz, hx * gy * gz, gx * gy * gz};
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) s = fmaf(w[q], tc.v[q], s);
  sm.s = s;
}

struct Shade {
  float c[3], a, alpha, g;
  int j0, j1;
  bool inside;
};

__device__ __forceinline__ void shade(const float4 *tf, int n, float sigma, float dt, float s, Shade &sh) {
  sh.inside = (0.f <= s) & (s <= 1.f);
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  sh.j0 = sh.j1 = 0;
  sh.g = 0.f;
  if (sh.inside) {
    const float gx = float(n) * s - 0.5f;
    const float fl = floorf(gx);
    sh.g = gx - fl;
    const int j = (int)fl;
    sh.j0 = iclamp(j, 0, n - 1);
    sh.j1 = iclamp(j + 1, 0, n - 1);
    const float4 p = tf[sh.j0], q = tf[sh.j1];
    const float h = 1.f - sh.g;
    v = make_float4(p.x * h + q.x * sh.g, p.y * h + q.y * sh.g, p.z * h + q.z * sh.g, p.w * h + q.w * sh.g);
  }
  sh.c[0] = v.x; sh.c[1] = v.y; sh.c[2] = v.z;
  sh.a = v.w;
  sh.alpha = 1.f - expf(-sigma * fmaxf(v.w, 0.f) * dt);
}

struct RaySetup {
  float ox, oy, oz, dx, dy, dz, tnear, dt;
};

__device__ __forceinline__ bool setup(const DiffArgs &a, const DiffView &v, int x, int y, RaySetup &r) {
  const f3 ro = f3{v.ray_o[0], v.ray_o[1], v.ray_o[2]};
  const f3 d = pixel_ray_dir(v.trans, v.ctg, ro, x, y, a.W, a.H);
  float tn, tf;
  if (!intersect_box(ro, d, f3{0.f, 0.f, 0.f}, f3{1.f, 1.f, 1.f}, tn, tf)) return false;
  if (tn < 0) tn = 0;
  if (tf < 0) tf = 0;
  r.ox = ro.x; r.oy = ro.y; r.oz = ro.z;
  r.dx = d.x; r.dy = d.y; r.dz = d.z;
  r.tnear = tn;
  r.dt = (tf - tn) / float(a.K);
  return true;
}

__device__ __forceinline__ void position(const RaySetup &r, int k, float &px, float &py, float &pz) {
  const float t = r.tnear + (float(k) + 0.5f) * r.dt;
  px = r.ox + r.dx * t;
  py = r.oy + r.dy * t;
  pz = r.oz + r.dz * t;
}

template <bool COUNT>
__global__ __launch_bounds__(64) void diff_forward_kernel(DiffArgs a) {
  __shared__ float4 s_tf[kMaxTf];
  __shared__ unsigned long long s_count;
  for (int i = threadIdx.x; i < a.tf_n; i += 64) s_tf[i] = reinterpret_cast<const float4 *>(a.tf)[i];
  if (COUNT && threadIdx.x == 0) s_count = 0;
  __syncthreads();
  const int view = blockIdx.z;
  const int x = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int y = blockIdx.y * kTile + (threadIdx.x / kTile);
  unsigned count = 0;
  if (x < a.W && y < a.H) {
    const size_t ray = ((size_t)view * a.H + y) * a.W + x;
    float *ck = a.ckpts + ray * a.nck;
    RaySetup r;
    const VolumeRef vr = volume_ref(a);
    float C0 = 0.f, C1 = 0.f, C2 = 0.f, T = 1.f;
    const int nseg = a.nck - 1;
    if (setup(a, a.views[view], x, y, r)) {
      TapCache tcf;   // consecutive samples in one cell share their taps
      for (int seg = 0; seg < nseg; ++seg) {
        const float Tstart = T;
        float amax = 0.f;
        const int k0 = seg * a.ckpt, k1 = min(a.K, k0 + a.ckpt);
        for (int k = k0; k < k1; ++k) {
          float px, py, pz;
          position(r, k, px, py, pz);
          Smp sm;
          sample_reuse(a, vr, px, py, pz, sm, tcf);
          Shade sh;
          shade(s_tf, a.tf_n, a.sigma, r.dt, sm.s, sh);
          const float ta = T * sh.alpha;
          C0 = fmaf(ta, sh.c[0], C0);
          C1 = fmaf(ta, sh.c[1], C1);
          C2 = fmaf(ta, sh.c[2], C2);
          T = T * (1.f - sh.alpha);
          amax = fmaxf(amax, sh.alpha);
        }
        if (COUNT) count += (unsigned)(k1 - k0);
        ck[seg] = amax > 0.5f ? -Tstart : Tstart;   // sign bit: recompute this segment exactly
      }
      ck[nseg] = T;
    } else {
      for (int i = 0; i < a.nck; ++i) ck[i] = 1.f;
    }
    a.rgba[ray] = make_float4(C0, C1, C2, 1.f - T);
  }
  if (COUNT) {
    if (count) atomicAdd(&s_count, (unsigned long long)count);
    __syncthreads();
    if (threadIdx.x == 0 && s_count) atomicAdd(a.samples, s_count);
  }
}

constexpr int kBoundParts = kDiffBoundParts;
constexpr int kBrick = 8;          // gradient bricks of the scatter pass

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// One fixed-point term: v * 2^e rounded to an int32 (|v| 2^e < 2^30 by the
// choice of e), sign-extended for the int64 accumulators.
__device__ __forceinline__ unsigned long long fixed_term(float v, float scale) {
  return (unsigned long long)(long long)(int)rintf(v * scale);
}

// A term already scaled (|v| < 2^30) to fixed point by truncation: one
// conversion; the bias (under one unit of 2^-30 of the bound per term) is far
// below the FP32 rounding of the final sums.
__device__ __forceinline__ unsigned long long fixed_trunc(float v) {
  return (unsigned long long)(long long)(int)v;
}

// Scale 2^e with b 2^e < 2^30 (frexp: b < 2^E), less the headroom, in [2^-100, 2^100].
__device__ __forceinline__ int fixed_exponent(float b, int headroom) {
  int E = 0;
  (void)frexpf(b, &E);
  return min(100, max(-100, 30 - E - headroom));
}

// Prepass: partial maxima of the bound b >= |h d_ch| of one TF-gradient term,
//   |d_ch| = T alpha |g_ch| <= |g_ch|              (colour channels),
//   |d_3| <= |dalpha| sigma dt, |dalpha| <= 2 max|tf_rgb| |g|_1 + |g_A|
// (|c|, |S| <= max|tf_rgb|, T, P <= 1, dt <= sqrt(3)/K for the unit box).  Workgroup 0
// also resets the work queue and clears grad_tf (both outputs are overwritten).
__global__ __launch_bounds__(256) void diff_bound_kernel(DiffArgs a) {
  __shared__ float s_red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (blockIdx.x == 0) {
    if (tid == 0) *a.tile_counter = 0;
    for (int i = tid; i < 4 * a.tf_n; i += 256) a.grad_tf[i] = 0.f;
  }
  float mc = 0.f;
  for (int i = tid; i < a.tf_n; i += 256) {
    const float4 t = reinterpret_cast<const float4 *>(a.tf)[i];
    mc = fmaxf(mc, fmaxf(fabsf(t.x), fmaxf(fabsf(t.y), fabsf(t.z))));
  }
  mc = wave_max(mc);
  if (lane == 0) s_red[wave] = mc;
  __syncthreads();
  mc = fmaxf(fmaxf(s_red[0], s_red[1]), fmaxf(s_red[2], s_red[3]));
  __syncthreads();
  const float dtmax = 1.7322f / float(a.K);
  const long long n_rays = (long long)a.n_views * a.W * a.H;
  float b = 0.f;
  for (long long r = (long long)blockIdx.x * 256 + tid; r < n_rays; r += (long long)gridDim.x * 256) {
    const float4 g = a.grad_rgba[r];
    const float ginf = fmaxf(fabsf(g.x), fmaxf(fabsf(g.y), fabsf(g.z)));
    const float g1 = fabsf(g.x) + fabsf(g.y) + fabsf(g.z);
    b = fmaxf(b, fmaxf(ginf, (2.f * mc * g1 + fabsf(g.w)) * a.sigma * dtmax));
  }
  b = wave_max(b);
  if (lane == 0) s_red[wave] = b;
  __syncthreads();
  if (tid == 0) a.bound[blockIdx.x] = fmaxf(fmaxf(s_red[0], s_red[1]), fmaxf(s_red[2], s_red[3])) * 1.0625f;
}

constexpr int kRun = 8;      // samples per register run (TF terms, scatter corners): 3 bits of headroom

__device__ __forceinline__ void flush_tf(unsigned long long *s_gtf, int j0, int j1, const float *T0, const float *T1,
                                         float scale) {
#pragma unroll
  for (int ch = 0; ch < 4; ++ch) {
    atomicAdd(&s_gtf[ch * kMaxTf + j0], fixed_term(T0[ch], scale));
    atomicAdd(&s_gtf[ch * kMaxTf + j1], fixed_term(T1[ch], scale));
  }
}

constexpr int kStage = 16;   // samples per ray staged in LDS before a store
constexpr int kStagePitch = kStage + 1;

// Back-to-front adjoint of one 8x8 ray tile; the wave runs the sample loop in
// lockstep (inactive rays masked).  The TF table is read from LDS: the
// kernel is bound by the vector-memory data path (tap gathers), and TF reads
// from global memory would add two scattered 16 B lane-loads per sample.
// Consecutive samples in one cell reuse their taps.  Writes dL/ds of every sample of the
// active rays (dL/ds layout [view][y][x][k]: staged 16 samples at a time in
// LDS so each ray's run leaves as one 64 B store) and every ray's max |dL/ds|
// (0 for inactive rays); adds the TF-gradient terms to the fixed-point LDS
// accumulators.
__device__ void adjoint_tile(const DiffArgs &a, const float4 *tf, unsigned long long *s_gtf, float *stage,
                             float scale, int view, int tx, int ty, int lane) {
  const int x = tx * kTile + (lane % kTile), y = ty * kTile + (lane / kTile);
  const bool in_image = x < a.W && y < a.H;
  const size_t ray = in_image ? ((size_t)view * a.H + y) * a.W + x : 0;
  float g[3] = {0.f, 0.f, 0.f}, gA = 0.f;
  if (in_image) {
    const float4 gr = a.grad_rgba[ray];
    g[0] = gr.x; g[1] = gr.y; g[2] = gr.z; gA = gr.w;
  }
  RaySetup r{};
  const VolumeRef vr = volume_ref(a);
  const bool active =
      in_image && (g[0] != 0.f || g[1] != 0.f || g[2] != 0.f || gA != 0.f) && setup(a, a.views[view], x, y, r);
  const unsigned long long amask = __ballot(active);
  if (active) {   // the scatter pass reuses the ray setup
    a.rays[2 * ray] = make_float4(r.dx, r.dy, r.dz, r.tnear);
    a.rays[2 * ray + 1] = make_float4(r.dt, 0.f, 0.f, 0.f);
  }
  float dmax = 0.f;
  // TF-gradient run: terms of consecutive samples on the same pair of TF entries
  // are summed in registers (at most kRun of them) before they reach LDS
  int tj0 = -1, tj1 = -1, tn = 0;
  float T0[4] = {0.f, 0.f, 0.f, 0.f}, T1[4] = {0.f, 0.f, 0.f, 0.f};
  if (amask != 0ull) {
    const float *ck = a.ckpts + ray * a.nck;
    float S[3] = {0.f, 0.f, 0.f}, P = 1.f;
    TapCache tc;
    const int nseg = a.nck - 1;
    float Tnext = active ? fabsf(ck[nseg]) : 1.f;
    for (int seg = nseg - 1; seg >= 0; --seg) {
      const int k0 = seg * a.ckpt, k1 = min(a.K, k0 + a.ckpt);
      const float cks = active ? ck[seg] : 1.f;
      const float Tstart = fabsf(cks);
      const bool exact = signbit(cks);
      for (int k = k1 - 1; k >= k0; --k) {
        float ds = 0.f;
        if (active) {
          float px, py, pz;
          position(r, k, px, py, pz);
          Smp sm;
          sample_reuse(a, vr, px, py, pz, sm, tc);
          Shade sh;
          shade(tf, a.tf_n, a.sigma, r.dt, sm.s, sh);
          float Tk;
          if (!exact) {
            Tk = Tnext * __builtin_amdgcn_rcpf(1.f - sh.alpha);   // alpha <= 0.5 here
          } else {
            Tk = Tstart;   // recompute T_k forward from the checkpoint
            for (int j = k0; j < k; ++j) {
              float qx, qy, qz;
              position(r, j, qx, qy, qz);
              Smp sj;
              sample(a, vr, qx, qy, qz, sj);
              Shade hj;
              shade(tf, a.tf_n, a.sigma, r.dt, sj.s, hj);
              Tk = Tk * (1.f - hj.alpha);
            }
          }
          // dL/dalpha_k = T_k (g . (c_k - S_{k+1}) + gA P_{k+1}),  dL/dc_k = T_k alpha_k g
          const float dalpha =
              Tk * (g[0] * (sh.c[0] - S[0]) + g[1] * (sh.c[1] - S[1]) + g[2] * (sh.c[2] - S[2]) + gA * P);
          const float ta = Tk * sh.alpha;
          float d[4];
          d[0] = ta * g[0];
          d[1] = ta * g[1];
          d[2] = ta * g[2];
          d[3] = sh.a > 0.f ? dalpha * a.sigma * r.dt * (1.f - sh.alpha) : 0.f;
          if (sh.inside) {
            const float h = 1.f - sh.g;
            const float4 p = tf[sh.j0], q = tf[sh.j1];
            ds = float(a.tf_n) * ((q.x - p.x) * d[0] + (q.y - p.y) * d[1] + (q.z - p.z) * d[2] + (q.w - p.w) * d[3]);
            if (sh.j0 != tj0 || sh.j1 != tj1 || tn == kRun) {
              if (tn > 0) flush_tf(s_gtf, tj0, tj1, T0, T1, scale);
              tj0 = sh.j0;
              tj1 = sh.j1;
              tn = 0;
#pragma unroll
              for (int ch = 0; ch < 4; ++ch) T0[ch] = T1[ch] = 0.f;
            }
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
              T0[ch] += h * d[ch];
              T1[ch] += sh.g * d[ch];
            }
            ++tn;
          }
          dmax = fmaxf(dmax, fabsf(ds));
          S[0] = sh.alpha * sh.c[0] + (1.f - sh.alpha) * S[0];
          S[1] = sh.alpha * sh.c[1] + (1.f - sh.alpha) * S[1];
          S[2] = sh.alpha * sh.c[2] + (1.f - sh.alpha) * S[2];
          P = (1.f - sh.alpha) * P;
          Tnext = Tk;
        }
        stage[lane * kStagePitch + (k % kStage)] = ds;
        if (k % kStage == 0) {
          // samples k .. k+15 of all 64 rays are staged: store 16 rays x 4 at a time,
          // 16 lanes per ray, one 64 B run each
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          const int kk = lane % kStage, kg = k + kk;
#pragma unroll 1
          for (int i = 0; i < 64 / (64 / kStage); ++i) {
            const int rr = i * (64 / kStage) + lane / kStage;
            if (kg < a.K && ((amask >> rr) & 1ull)) {
              const int xr = tx * kTile + (rr % kTile), yr = ty * kTile + (rr / kTile);
              a.ds[(((size_t)view * a.H + yr) * a.W + xr) * a.K + kg] = stage[rr * kStagePitch + kk];
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
      }
      if (active) Tnext = Tstart;   // re-anchor at the checkpoint
    }
    if (tn > 0) flush_tf(s_gtf, tj0, tj1, T0, T1, scale);
  }
  if (in_image) a.dsmax[ray] = dmax;
}

// Ray-order adjoint: persistent workgroups of four waves; each wave claims
// 8x8 ray tiles from the work queue.  TF gradients: register runs, then int64
// LDS accumulators with one scale per launch (from the prepass bound), folded
// once per workgroup into FP32 global atomics.
__global__ __launch_bounds__(256) void diff_adjoint_kernel(DiffArgs a) {
  __shared__ unsigned long long s_gtf[4 * kMaxTf];   // [channel][entry]
  __shared__ float s_b[kBoundParts];
  __shared__ float s_stage[4][64 * kStagePitch];
  __shared__ float4 s_tfl[kMaxTf];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < 4 * kMaxTf; i += 256) s_gtf[i] = 0ull;
  for (int i = tid; i < kBoundParts; i += 256) s_b[i] = a.bound[i];
  for (int i = tid; i < a.tf_n; i += 256) s_tfl[i] = reinterpret_cast<const float4 *>(a.tf)[i];
  __syncthreads();
  float b = 0.f;
  for (int i = lane; i < kBoundParts; i += 64) b = fmaxf(b, s_b[i]);
  b = wave_max(b);
  const int e = fixed_exponent(b, a.headroom + 3);   // a run sums <= kRun = 8 terms
  const float scale = ldexpf(1.f, e);
  const int tiles_x = (a.W + kTile - 1) / kTile, tiles_y = (a.H + kTile - 1) / kTile;
  const int per_view = tiles_x * tiles_y, total = per_view * a.n_views;
  for (;;) {
    int t = 0;
    if (lane == 0) t = atomicAdd(a.tile_counter, 1);
    t = __shfl(t, 0);
    if (t >= total) break;
    const int view = t / per_view, rem = t - view * per_view;
    const int ty = rem / tiles_x, tx = rem - ty * tiles_x;
    adjoint_tile(a, s_tfl, s_gtf, s_stage[wave], scale, view, tx, ty, lane);
  }
  __syncthreads();
  const double unscale = ldexp(1.0, -e);
  for (int i = tid; i < 4 * kMaxTf; i += 256) {
    const int ch = i / kMaxTf, j = i - ch * kMaxTf;
    const long long v = (long long)s_gtf[i];
    if (j < a.tf_n && v != 0) unsafeAtomicAdd(&a.grad_tf[4 * j + ch], (float)((double)v * unscale));
  }
}

// Conservative slab test (the hardware reciprocal's error is far inside the
// support box's margin).
__device__ __forceinline__ void slab(float o, float d, float lo, float hi, float &ta, float &tb, bool &hit) {
  if (d != 0.f) {
    const float rd = __builtin_amdgcn_rcpf(d);
    const float t1 = (lo - o) * rd, t2 = (hi - o) * rd;
    ta = fmaxf(ta, fminf(t1, t2));
    tb = fminf(tb, fmaxf(t1, t2));
  } else if (o < lo || o > hi) {
    hit = false;
  }
}

// Scatter of dL/ds into the voxel gradient, one wave per 8^3 brick.  Every
// sample belongs to the brick that holds its trilinear cell floor(n p - 0.5)
// (cell -1 of the clamped low border goes to the first brick); its corners
// (unclamped) lie in the brick or one voxel beyond it on either side, so a
// brick accumulates a kBlk^3 = 10^3 block indexed by the corner coordinates
// relative to the brick, plus one: no border clamping here, and the eight
// corner slots of a sample sit at fixed offsets from the first (0, 1, 10, 11,
// 100, ...).  Per view: the pixels in the projection of the box holding the
// brick's cells (plus a margin); the rays that reach the box go to an LDS
// list with their sample range inside it (conservative); the list's samples
// are split into 64 contiguous runs, one per lane; each candidate sample's
// position and cell are recomputed with the forward's operations and samples
// of other bricks' cells dropped.  Each corner term goes to an int64 LDS
// accumulator (one scale per (brick, view) from the rays' max |dL/ds|),
// folded per view into FP32 sums held in registers and stored once as the
// brick's block in global scratch; diff_combine_kernel then adds, per voxel,
// the block entries that map to it (own brick, the high halo of the lower
// neighbour, and the border slots that clamp onto it) in a fixed order.
// The views are split into groups (one workgroup per brick and group, each
// with its own block) so that the launch has several waves per SIMD slot.
// No float atomics: the voxel gradient is deterministic.
constexpr int kList = 64;    // rays per scatter list
constexpr int kBlk = kBrick + 2;                    // corner coordinates -1 .. kBrick, plus one
constexpr int kBlk3 = kBlk * kBlk * kBlk;
static_assert(kBlk3 == kDiffBlock, "block size shared with the host");
constexpr int kBlkPerLane = (kBlk3 + 63) / 64;

struct ScatterList {
  float dx[kList], dy[kList], dz[kList], tn[kList], dt[kList];
  int k0[kList], off[kList], ray[kList];
};

struct BrickBox {
  int X0, Y0, Z0;        // first voxel
  int cx0, cy0, cz0;     // first owned cell (-1 on the low border)
  int cx1, cy1, cz1;     // last owned cell
};

// Scatter the samples of the first n rays of the list (total samples `total`);
// lane l takes the contiguous run [l per, (l+1) per) and prefetches the next
// sample's dL/ds while it works on the current one.
__device__ void scatter_list(const DiffArgs &a, const DiffView &v, const ScatterList &L, int n, int total,
                             float scale, unsigned long long *s_acc, int lane, const BrickBox &bb) {
  const int per = (total + 63) / 64;
  const int start = lane * per, end = min(total, start + per);
  int j = 0;   // the last ray whose first sample is <= start (every listed ray has samples)
  for (int step = kList / 2; step > 0; step >>= 1)
    if (j + step < n && L.off[j + step] <= start) j += step;
  RaySetup q;
  q.ox = v.ray_o[0]; q.oy = v.ray_o[1]; q.oz = v.ray_o[2];
  q.dx = L.dx[j]; q.dy = L.dy[j]; q.dz = L.dz[j]; q.tnear = L.tn[j]; q.dt = L.dt[j];
  int next = (j + 1 < n) ? L.off[j + 1] : total;
  long long base = (long long)L.ray[j] * a.K + L.k0[j] - L.off[j];   // ds index of list sample `it`: base + it
  int kb = L.k0[j] - L.off[j];                                         // its sample index: kb + it
  // one step: item it with dL/ds `cur`; issues the load of item it + 1 and
  // returns it.  Two steps per loop iteration, alternating two registers, so a
  // load is consumed a whole step after it is issued; the load is issued
  // unconditionally (past the run's end it rereads the run's last item) so that
  // no branch forces an early wait for it.
  const auto step = [&](int it, float cur) -> float {
    const bool act = it < end;
    if (act && it == next) {   // next ray of the list
      ++j;
      q.dx = L.dx[j]; q.dy = L.dy[j]; q.dz = L.dz[j]; q.tnear = L.tn[j]; q.dt = L.dt[j];
      next = (j + 1 < n) ? L.off[j + 1] : total;
      base = (long long)L.ray[j] * a.K + L.k0[j] - L.off[j];
      kb = L.k0[j] - L.off[j];
    }
    const int in = min(it + 1, end - 1);   // a listed item (end >= 1)
    const long long nb = (in == next) ? (long long)L.ray[j + 1] * a.K + L.k0[j + 1] - L.off[j + 1] : base;
    const float nxt = a.ds[nb + in];
    float px, py, pz;
    position(q, kb + it, px, py, pz);
    const float fx = float(a.nx) * px - 0.5f, fy = float(a.ny) * py - 0.5f, fz = float(a.nz) * pz - 0.5f;
    const float flx = floorf(fx), fly = floorf(fy), flz = floorf(fz);
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    if (act && ix >= bb.cx0 && ix <= bb.cx1 && iy >= bb.cy0 && iy <= bb.cy1 && iz >= bb.cz0 && iz <= bb.cz1 &&
        cur != 0.f) {
      const float gx = fx - flx, gy = fy - fly, gz = fz - flz;
      const float vs = cur * scale;
      const float hx = 1.f - gx, hy = 1.f - gy;
      const float wz0 = (1.f - gz) * vs, wz1 = gz * vs;   // z weights times the scaled dL/ds
      const float w00 = hx * hy, w10 = gx * hy, w01 = hx * gy, w11 = gx * gy;
      // corner slots relative to the brick, 0 .. kBlk - 1 on each axis (24-bit products)
      const unsigned sxl = (unsigned)(ix - bb.X0 + 1), syl = (unsigned)(iy - bb.Y0 + 1),
                     szl = (unsigned)(iz - bb.Z0 + 1);
      unsigned long long *p = s_acc + (__umul24(__umul24(szl, kBlk) + syl, kBlk) + sxl);
      atomicAdd(p, fixed_trunc(w00 * wz0));
      atomicAdd(p + 1, fixed_trunc(w10 * wz0));
      atomicAdd(p + kBlk, fixed_trunc(w01 * wz0));
      atomicAdd(p + kBlk + 1, fixed_trunc(w11 * wz0));
      atomicAdd(p + kBlk * kBlk, fixed_trunc(w00 * wz1));
      atomicAdd(p + kBlk * kBlk + 1, fixed_trunc(w10 * wz1));
      atomicAdd(p + kBlk * kBlk + kBlk, fixed_trunc(w01 * wz1));
      atomicAdd(p + kBlk * kBlk + kBlk + 1, fixed_trunc(w11 * wz1));
    }
    return nxt;
  };
  float dA = a.ds[base + min(start, end - 1)], dB;
  for (int i = 0; i < per; i += 2) {
    dB = step(start + i, dA);
    dA = step(start + i + 1, dB);
  }
}

__global__ __launch_bounds__(64) void diff_scatter_kernel(DiffArgs a) {
  __shared__ unsigned long long s_acc[kBlk3];
  __shared__ ScatterList L;
  const int lane = threadIdx.x;
  const int nbx = (a.nx + kBrick - 1) / kBrick, nby = (a.ny + kBrick - 1) / kBrick,
            nbz = (a.nz + kBrick - 1) / kBrick;
  const int bi = blockIdx.x % (nbx * nby * nbz), grp = blockIdx.x / (nbx * nby * nbz);
  const int bx = bi % nbx, by = (bi / nbx) % nby, bz = bi / (nbx * nby);
  // views of this workgroup: a contiguous share of the launch's views
  const int v_begin = (int)((long long)a.n_views * grp / a.view_groups);
  const int v_end = (int)((long long)a.n_views * (grp + 1) / a.view_groups);
  BrickBox bb;
  bb.X0 = bx * kBrick; bb.Y0 = by * kBrick; bb.Z0 = bz * kBrick;
  bb.cx0 = bb.X0 == 0 ? -1 : bb.X0; bb.cy0 = bb.Y0 == 0 ? -1 : bb.Y0; bb.cz0 = bb.Z0 == 0 ? -1 : bb.Z0;
  bb.cx1 = min(bb.X0 + kBrick - 1, a.nx - 1); bb.cy1 = min(bb.Y0 + kBrick - 1, a.ny - 1);
  bb.cz1 = min(bb.Z0 + kBrick - 1, a.nz - 1);
  float *blk = a.blocks + ((size_t)grp * nbx * nby * nbz + bi) * kBlk3;
  for (int i = lane; i < kBlk3; i += 64) s_acc[i] = 0ull;
  float acc[kBlkPerLane];   // the block's FP32 sums over this group's views; entry lane + 64 i
#pragma unroll
  for (int i = 0; i < kBlkPerLane; ++i) acc[i] = 0.f;
  // the box of the owned cells: floor(n p - 0.5) in [c0, c1]  <=>  p in [(c0 + 0.5) / n, (c1 + 1.5) / n)
  const float delta = 1e-4f;
  const float lo[3] = {(float(bb.cx0) + 0.5f) / float(a.nx) - delta, (float(bb.cy0) + 0.5f) / float(a.ny) - delta,
                       (float(bb.cz0) + 0.5f) / float(a.nz) - delta};
  const float hi[3] = {(float(bb.cx1) + 1.5f) / float(a.nx) + delta, (float(bb.cy1) + 1.5f) / float(a.ny) + delta,
                       (float(bb.cz1) + 1.5f) / float(a.nz) + delta};
  __syncthreads();
  for (int view = v_begin; view < v_end; ++view) {
    const DiffView &v = a.views[view];
    // pixel bounding box of the projected box (whole image if a corner is not in front)
    int px0 = 0, px1 = a.W - 1, py0 = 0, py1 = a.H - 1;
    {
      bool front = true;
      float umin = 3e38f, umax = -3e38f, vmin = 3e38f, vmax = -3e38f;
      for (int c = 0; c < 8 && front; ++c) {
        const float X = (c & 1) ? hi[0] : lo[0], Y = (c & 2) ? hi[1] : lo[1], Z = (c & 4) ? hi[2] : lo[2];
        const float qx = v.inv[0] * X + v.inv[1] * Y + v.inv[2] * Z + v.inv[3];
        const float qy = v.inv[4] * X + v.inv[5] * Y + v.inv[6] * Z + v.inv[7];
        const float qz = v.inv[8] * X + v.inv[9] * Y + v.inv[10] * Z + v.inv[11];
        if (!(qz < -1e-4f)) {
          front = false;
        } else {
          const float sc = v.ctg / (-qz) * float(a.H) * 0.5f;
          const float u = float(a.W) * 0.5f + qx * sc, w = float(a.H) * 0.5f - qy * sc;
          umin = fminf(umin, u); umax = fmaxf(umax, u);
          vmin = fminf(vmin, w); vmax = fmaxf(vmax, w);
        }
      }
      if (front) {
        px0 = max(px0, (int)floorf(fmaxf(umin, -4.f)) - 2);
        px1 = min(px1, (int)ceilf(fminf(umax, float(a.W) + 4.f)) + 2);
        py0 = max(py0, (int)floorf(fmaxf(vmin, -4.f)) - 2);
        py1 = min(py1, (int)ceilf(fminf(vmax, float(a.H) + 4.f)) + 2);
      }
    }
    const int bw = px1 - px0 + 1, bh = py1 - py0 + 1;
    if (bw <= 0 || bh <= 0) continue;
    const int nR = bw * bh;
    float B = 0.f;
    for (int r = lane; r < nR; r += 64) {
      const int y = py0 + r / bw, x = px0 + r % bw;
      B = fmaxf(B, a.dsmax[((size_t)view * a.H + y) * a.W + x]);
    }
    B = wave_max(B);
    if (!(B > 0.f)) continue;
    // terms below 2^30 each; terms per accumulator in this view <= 8 nR K, so the
    // int64 sums hold while nR K <= 2^29
    const double items = (double)nR * (double)a.K;
    const int hv = items > 536870912.0 ? (int)ceil(log2(items / 536870912.0)) : 0;
    const int e = fixed_exponent(B, hv);
    const float scale = ldexpf(1.f, e);
    int n = 0, total = 0;   // rays and samples in the list (wave-uniform)
    for (int base = 0; base < nR; base += 64) {
      const int rr = base + lane;
      int cnt = 0, k0 = 0, ray = 0;
      RaySetup rs{};
      if (rr < nR) {
        const int y = py0 + rr / bw, x = px0 + rr % bw;
        ray = (view * a.H + y) * a.W + x;
        if (a.dsmax[ray] > 0.f) {   // an active ray: its setup was stored by the ray-order pass
          const float4 r0 = a.rays[2 * (size_t)ray], r1 = a.rays[2 * (size_t)ray + 1];
          rs.ox = v.ray_o[0]; rs.oy = v.ray_o[1]; rs.oz = v.ray_o[2];
          rs.dx = r0.x; rs.dy = r0.y; rs.dz = r0.z; rs.tnear = r0.w; rs.dt = r1.x;
          float ta = -3e38f, tb = 3e38f;
          bool hit = true;
          slab(rs.ox, rs.dx, lo[0], hi[0], ta, tb, hit);
          slab(rs.oy, rs.dy, lo[1], hi[1], ta, tb, hit);
          slab(rs.oz, rs.dz, lo[2], hi[2], ta, tb, hit);
          if (hit && ta <= tb) {
            if (rs.dt > 0.f) {
              const float rdt = __builtin_amdgcn_rcpf(rs.dt);
              const float fa = (ta - rs.tnear) * rdt - 0.5f, fb = (tb - rs.tnear) * rdt - 0.5f;
              const float ka = fminf(fmaxf(ceilf(fa) - 1.f, 0.f), float(a.K));
              const float kb = fminf(fmaxf(floorf(fb) + 1.f, -1.f), float(a.K - 1));
              k0 = (int)ka;
              cnt = max(0, (int)kb - k0 + 1);
            } else {
              cnt = a.K;
            }
          }
        }
      }
      const unsigned long long m = __ballot(cnt > 0);
      const int n_new = __popcll(m);
      if (n_new == 0) continue;
      if (n + n_new > kList) {   // list full: scatter it first
        __syncthreads();
        scatter_list(a, v, L, n, total, scale, s_acc, lane, bb);
        __syncthreads();
        n = 0;
        total = 0;
      }
      const int incl = wave_incl_scan(cnt, lane);
      if (cnt > 0) {
        const int pos = n + __popcll(m & ((1ull << lane) - 1ull));
        L.dx[pos] = rs.dx; L.dy[pos] = rs.dy; L.dz[pos] = rs.dz; L.tn[pos] = rs.tnear; L.dt[pos] = rs.dt;
        L.k0[pos] = k0; L.off[pos] = total + incl - cnt; L.ray[pos] = ray;
      }
      n += n_new;
      total += __shfl(incl, 63);
    }
    if (n > 0) {
      __syncthreads();
      scatter_list(a, v, L, n, total, scale, s_acc, lane, bb);
    }
    __syncthreads();
    const double unscale = ldexp(1.0, -e);
#pragma unroll
    for (int i = 0; i < kBlkPerLane; ++i) {
      const int idx = lane + 64 * i;
      if (idx < kBlk3) {
        const long long val = (long long)s_acc[idx];
        if (val != 0) {
          acc[i] += (float)((double)val * unscale);
          s_acc[idx] = 0ull;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < kBlkPerLane; ++i) {
    const int idx = lane + 64 * i;
    if (idx < kBlk3) blk[idx] = acc[i];
  }
}

// Block entries that map to voxel coordinate v on one axis of extent n: the
// own brick's slot v - X0 + 1; slot kBrick + 1 of the brick below when v is
// its brick's first voxel (the lower neighbour's high halo); slot 0 of the
// first brick for v = 0 (corner -1, clamped) and slot n - X0 + 1 of the last
// brick for v = n - 1 (corner n, clamped).  Returns the count (<= 3).
__device__ __forceinline__ int axis_sources(int v, int n, int *db, int *slot) {
  const int b = v / kBrick, l = v - b * kBrick;
  int c = 0;
  db[c] = 0; slot[c] = l + 1; ++c;
  if (l == 0 && b > 0) { db[c] = -1; slot[c] = kBrick + 1; ++c; }
  if (v == 0) { db[c] = 0; slot[c] = 0; ++c; }
  if (v == n - 1) { db[c] = 0; slot[c] = n - b * kBrick + 1; ++c; }
  return c;
}

// grad_vol[v] = over the view groups, every block entry that maps to v
// (axis_sources on each axis), summed in a fixed order.
__global__ __launch_bounds__(256) void diff_combine_kernel(DiffArgs a) {
  const size_t nvox = (size_t)a.nx * a.ny * a.nz;
  const int nbx = (a.nx + kBrick - 1) / kBrick, nby = (a.ny + kBrick - 1) / kBrick;
  const size_t nb = (size_t)nbx * nby * ((a.nz + kBrick - 1) / kBrick);
  for (size_t vi = (size_t)blockIdx.x * 256 + threadIdx.x; vi < nvox; vi += (size_t)gridDim.x * 256) {
    const int X = (int)(vi % a.nx), Y = (int)((vi / a.nx) % a.ny), Z = (int)(vi / ((size_t)a.nx * a.ny));
    const int bx = X / kBrick, by = Y / kBrick, bz = Z / kBrick;
    int dbx[3], sx[3], dby[3], sy[3], dbz[3], sz[3];
    const int cx = axis_sources(X, a.nx, dbx, sx), cy = axis_sources(Y, a.ny, dby, sy),
              cz = axis_sources(Z, a.nz, dbz, sz);
    float g = 0.f;
    for (int grp = 0; grp < a.view_groups; ++grp) {
      for (int k = 0; k < cz; ++k)
        for (int j = 0; j < cy; ++j)
          for (int i = 0; i < cx; ++i) {
            const size_t b = grp * nb + ((size_t)(bz + dbz[k]) * nby + (by + dby[j])) * nbx + (bx + dbx[i]);
            g += a.blocks[b * kBlk3 + (sz[k] * kBlk + sy[j]) * kBlk + sx[i]];
          }
    }
    a.grad_vol[vi] = g;
  }
}

}  // namespace

hipError_t launch_diff_forward(const DiffArgs &a, hipStream_t s) {
  dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile, a.n_views);
  if (a.samples) hipLaunchKernelGGL(diff_forward_kernel<true>, grid, dim3(64), 0, s, a);
  else hipLaunchKernelGGL(diff_forward_kernel<false>, grid, dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_diff_backward(const DiffArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(diff_bound_kernel, dim3(kBoundParts), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(diff_adjoint_kernel, dim3(a.n_workers), dim3(256), 0, s, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int nb = ((a.nx + kBrick - 1) / kBrick) * ((a.ny + kBrick - 1) / kBrick) * ((a.nz + kBrick - 1) / kBrick);
  hipLaunchKernelGGL(diff_scatter_kernel, dim3(nb * a.view_groups), dim3(64), 0, s, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const size_t nvox = (size_t)a.nx * a.ny * a.nz;
  const int cg = (int)std::min<size_t>((nvox + 255) / 256, 4096);
  hipLaunchKernelGGL(diff_combine_kernel, dim3(cg), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace hyd
