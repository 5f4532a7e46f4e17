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
// TF gradients: int64 fixed-point sums (LDS per workgroup, then the launch's
// global sums), converted to FP32 once.  Voxel gradients: dL/ds of every
// sample is stored; a brick-order pass adds the corner terms of each brick's
// samples in LDS, one wave per brick block, and a final pass sums the blocks
// per voxel in a fixed order.  Both gradients are bit-identical from run to run.
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

// Two x-adjacent voxels from one 8-byte load at 4-byte alignment.
__device__ __forceinline__ float2 vol_pair(const VolumeRef &v, unsigned byte_offset) {
  float2 r;
  __builtin_memcpy(&r, v.base + byte_offset, sizeof(r));
  return r;
}

// The taps of a sample's cell: the four x-pairs of rows (y0, z0), (y1, z0),
// (y0, z1), (y1, z1), loaded at x = clamp(ix, 0, nx - 2) (nx >= 2), and the
// cell they belong to.
struct TapCache {
  int ix = -0x40000000, iy = 0, iz = 0;
  float2 p[4];
};

// s = trilinear(volume, n p - 0.5) with clamped taps.  Rows clamp through
// their offsets; along x the pair at clamp(ix, 0, nx - 2) holds both taps, and
// a cell below 0 (at or above nx - 1) takes the pair's first (second) voxel
// alone (x weight 0 or 1).  Linear interpolations along x, then y, then z.
// With REUSE a sample in the previous sample's cell loads nothing.
template <bool REUSE>
__device__ __forceinline__ void sample_taps(const DiffArgs &a, const VolumeRef &vr, float px, float py, float pz,
                                            Smp &sm, TapCache &tc) {
  const float fx = float(a.nx) * px - 0.5f, fy = float(a.ny) * py - 0.5f, fz = float(a.nz) * pz - 0.5f;
  const float flx = floorf(fx), fly = floorf(fy), flz = floorf(fz);
  const float gx = fx - flx, gy = fy - fly, gz = fz - flz;
  const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
  if (!REUSE || ix != tc.ix || iy != tc.iy || iz != tc.iz) {
    const unsigned sx = (unsigned)a.nx * 4u, sxy = (unsigned)a.nx * (unsigned)a.ny * 4u;   // byte strides
    const unsigned xm = (unsigned)iclamp(ix, 0, a.nx - 2) * 4u;
    const unsigned cy0 = (unsigned)iclamp(iy, 0, a.ny - 1), cy1 = (unsigned)iclamp(iy + 1, 0, a.ny - 1);
    const unsigned cz0 = (unsigned)iclamp(iz, 0, a.nz - 1), cz1 = (unsigned)iclamp(iz + 1, 0, a.nz - 1);
    const unsigned y0 = __umul24(cy0, sx) + xm, y1 = __umul24(cy1, sx) + xm;   // 4 nx < 2^24 (host check)
    const unsigned z0 = cz0 * sxy, z1 = cz1 == cz0 ? z0 : z0 + sxy;
    tc.p[0] = vol_pair(vr, z0 + y0);
    tc.p[1] = vol_pair(vr, z0 + y1);
    tc.p[2] = vol_pair(vr, z1 + y0);
    tc.p[3] = vol_pair(vr, z1 + y1);
    tc.ix = ix;
    tc.iy = iy;
    tc.iz = iz;
  }
  const float wx = ix < 0 ? 0.f : (ix >= a.nx - 1 ? 1.f : gx);
  const float c00 = fmaf(wx, tc.p[0].y - tc.p[0].x, tc.p[0].x), c10 = fmaf(wx, tc.p[1].y - tc.p[1].x, tc.p[1].x);
  const float c01 = fmaf(wx, tc.p[2].y - tc.p[2].x, tc.p[2].x), c11 = fmaf(wx, tc.p[3].y - tc.p[3].x, tc.p[3].x);
  const float c0 = fmaf(gy, c10 - c00, c00), c1 = fmaf(gy, c11 - c01, c01);
  sm.s = fmaf(gz, c1 - c0, c0);
}

__device__ __forceinline__ void sample(const DiffArgs &a, const VolumeRef &vr, float px, float py, float pz, Smp &sm) {
  TapCache tc;
  sample_taps<false>(a, vr, px, py, pz, sm, tc);
}

__device__ __forceinline__ void sample_reuse(const DiffArgs &a, const VolumeRef &vr, float px, float py, float pz,
                                             Smp &sm, TapCache &tc) {
  sample_taps<true>(a, vr, px, py, pz, sm, tc);
}

struct Shade {
  float c[3], a, alpha, g;
  int j0, j1;
  bool inside;
};

// (c, a) = TF(s) (Border + Linear lookup) and alpha = 1 - exp(-sigma a dt),
// computed as 1 - 2^(nl a) with nl = -sigma dt log2(e), the ray's constant.
__device__ __forceinline__ void shade(const float4 *tf, int n, float nl, float s, Shade &sh) {
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
    v = make_float4(fmaf(sh.g, q.x - p.x, p.x), fmaf(sh.g, q.y - p.y, p.y), fmaf(sh.g, q.z - p.z, p.z),
                    fmaf(sh.g, q.w - p.w, p.w));
  }
  sh.c[0] = v.x; sh.c[1] = v.y; sh.c[2] = v.z;
  sh.a = v.w;
  sh.alpha = 1.f - __builtin_amdgcn_exp2f(nl * fmaxf(v.w, 0.f));
}

// nl = -sigma dt log2(e), so that alpha = 1 - 2^(nl a)
__device__ __forceinline__ float opacity_const(float sigma, float dt) { return -(sigma * dt) * 1.44269504f; }

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
  const int x = rotated_tile_x() * kTile + (threadIdx.x % kTile);
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
      const float nl = opacity_const(a.sigma, r.dt);
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
          shade(s_tf, a.tf_n, nl, sm.s, sh);
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

// Scale 2^e with b 2^e < 2^30 (frexp: b < 2^E), less the headroom, in [2^-100, 2^100].
__device__ __forceinline__ int fixed_exponent(float b, int headroom) {
  int E = 0;
  (void)frexpf(b, &E);
  return min(100, max(-100, 30 - E - headroom));
}

// A term already scaled (|v| < 2^30) to fixed point, rounded to nearest (an
// unbiased error of at most half a unit per term), sign-extended for the int64
// accumulators.
__device__ __forceinline__ unsigned long long fixed_round(float v) {
  return (unsigned long long)(long long)(int)rintf(v);
}

// The TF-gradient scale exponent of the launch, from the prepass's partial
// bounds (every wave that calls this gets the same value).
__device__ __forceinline__ int tf_exponent(const DiffArgs &a, int lane) {
  float b = 0.f;
  for (int i = lane; i < kDiffBoundParts; i += 64) b = fmaxf(b, a.bound[i]);
  b = wave_max(b);
  return fixed_exponent(b, a.headroom + 3);   // a run sums <= kRun = 8 terms
}

// Prepass: partial maxima of the bound b >= |h d_ch| of one TF-gradient term,
//   |d_ch| = T alpha |g_ch| <= |g_ch|              (colour channels),
//   |d_3| <= |dalpha| sigma dt, |dalpha| <= 2 max|tf_rgb| |g|_1 + |g_A|
// (|c|, |S| <= max|tf_rgb|, T, P <= 1, dt <= sqrt(3)/K for the unit box).  Workgroup 0
// also resets the work queue and clears the TF-gradient sums.
__global__ __launch_bounds__(256) void diff_bound_kernel(DiffArgs a) {
  __shared__ float s_red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (blockIdx.x == 0) {
    if (tid == 0) *a.tile_counter = 0;
    for (int i = tid; i < 4 * kMaxTf; i += 256) a.gtf_acc[i] = 0ull;
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

// The four channel sums of one TF entry into the fixed-point LDS accumulators.
__device__ __forceinline__ void flush_tf_entry(unsigned long long *s_gtf, int j, const float *T, float scale) {
#pragma unroll
  for (int ch = 0; ch < 4; ++ch) atomicAdd(&s_gtf[ch * kMaxTf + j], fixed_term(T[ch], scale));
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
  const float nl = opacity_const(a.sigma, r.dt);
  const unsigned long long amask = __ballot(active);
  if (active) {   // the scatter pass reuses the ray setup
    a.rays[2 * ray] = make_float4(r.dx, r.dy, r.dz, r.tnear);
    a.rays[2 * ray + 1] = make_float4(r.dt, 0.f, 0.f, 0.f);
  }
  float dmax = 0.f;
  // TF-gradient run: terms of consecutive samples on the same pair of TF entries
  // (or on pairs one entry apart, for the entry they share) are summed in
  // registers (at most kRun samples since the last full flush) before they
  // reach LDS
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
          shade(tf, a.tf_n, nl, sm.s, sh);
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
              shade(tf, a.tf_n, nl, sj.s, hj);
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
              // a move by one entry keeps the sums of the entry both pairs share
              const bool up = tn < kRun && sh.j0 == tj1;     // old second entry = new first
              const bool down = tn < kRun && sh.j1 == tj0;   // old first entry = new second
              if (tn > 0) {
                if (!down) flush_tf_entry(s_gtf, tj0, T0, scale);
                if (!up) flush_tf_entry(s_gtf, tj1, T1, scale);
              }
#pragma unroll
              for (int ch = 0; ch < 4; ++ch) {
                const float n0 = up ? T1[ch] : 0.f, n1 = down ? T0[ch] : 0.f;
                T0[ch] = n0;
                T1[ch] = n1;
              }
              if (!up && !down) tn = 0;
              tj0 = sh.j0;
              tj1 = sh.j1;
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
          // samples k .. k+15 of all 64 rays are staged: with K a multiple of 4,
          // 16 rays at a time, 4 lanes per ray storing 16 B each (one 64 B run
          // per ray); else 4 rays at a time, 16 lanes per ray, 4 B each
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
          if ((a.K & 3) == 0) {
            const int kk = 4 * (lane & 3), kg = k + kk;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int rr = 16 * i + (lane >> 2);
              if (kg < a.K && ((amask >> rr) & 1ull)) {
                const int xr = tx * kTile + (rr % kTile), yr = ty * kTile + (rr / kTile);
                const float *st = stage + rr * kStagePitch + kk;
                *reinterpret_cast<float4 *>(a.ds + (((size_t)view * a.H + yr) * a.W + xr) * a.K + kg) =
                    make_float4(st[0], st[1], st[2], st[3]);
              }
            }
          } else {
            const int kk = lane % kStage, kg = k + kk;
#pragma unroll 1
            for (int i = 0; i < 64 / (64 / kStage); ++i) {
              const int rr = i * (64 / kStage) + lane / kStage;
              if (kg < a.K && ((amask >> rr) & 1ull)) {
                const int xr = tx * kTile + (rr % kTile), yr = ty * kTile + (rr / kTile);
                a.ds[(((size_t)view * a.H + yr) * a.W + xr) * a.K + kg] = stage[rr * kStagePitch + kk];
              }
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          __builtin_amdgcn_wave_barrier();
        }
      }
      if (active) Tnext = Tstart;   // re-anchor at the checkpoint
    }
    if (tn > 0) {
      flush_tf_entry(s_gtf, tj0, T0, scale);
      flush_tf_entry(s_gtf, tj1, T1, scale);
    }
  }
  if (in_image) a.dsmax[ray] = dmax;
}

// Ray-order adjoint: persistent workgroups of four waves; each wave claims
// 8x8 ray tiles from the work queue.  TF gradients: register runs, then int64
// LDS accumulators with one scale per launch (from the prepass bound), added
// once per workgroup to the launch's int64 sums (integer additions: the sums
// do not depend on which workgroup took which tile, or in what order).
__global__ __launch_bounds__(256) void diff_adjoint_kernel(DiffArgs a) {
  __shared__ unsigned long long s_gtf[4 * kMaxTf];   // [channel][entry]
  __shared__ float s_stage[4][64 * kStagePitch];
  __shared__ float4 s_tfl[kMaxTf];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < 4 * kMaxTf; i += 256) s_gtf[i] = 0ull;
  for (int i = tid; i < a.tf_n; i += 256) s_tfl[i] = reinterpret_cast<const float4 *>(a.tf)[i];
  const int e = tf_exponent(a, lane);
  __syncthreads();
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
  for (int i = tid; i < 4 * kMaxTf; i += 256) {
    const unsigned long long v = s_gtf[i];
    if (v != 0ull) atomicAdd(&a.gtf_acc[i], v);
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
// are cut into segments of 8, one per lane; each candidate sample's position
// and cell are recomputed with the forward's operations and samples of other
// bricks' cells dropped.  Each corner term is added to the brick's int64
// fixed-point block in LDS (one scale per view, from the largest |dL/ds| of
// the rays reaching the box), folded per view into FP32 sums in registers
// and stored once, after the group's views, in global scratch;
// diff_combine_kernel then adds, per voxel, the block entries that
// map to it (own brick, the high halo of the lower neighbour, and the border
// slots that clamp onto it) in a fixed order.  The views are split into groups
// (one workgroup per brick and group, each with its own block) so that the
// launch has several waves per SIMD slot.  The voxel gradient is
// deterministic: integer sums in LDS, no global atomics.
constexpr int kList = 64;    // rays per scatter list
constexpr int kBlk = kBrick + 2;                    // corner coordinates -1 .. kBrick, plus one
constexpr int kBlk3 = kBlk * kBlk * kBlk;
static_assert(kBlk3 == kDiffBlock, "block size shared with the host");
constexpr int kBlkPerLane = (kBlk3 + 63) / 64;

// Rays of a scatter list: setup, candidate samples k0 .. k1, first segment.
struct ScatterList {
  float dx[kList], dy[kList], dz[kList], tn[kList], dt[kList];
  int k0[kList], k1[kList], off[kList], ray[kList];
};

struct BrickBox {
  int X0, Y0, Z0;        // first voxel
  int cx0, cy0, cz0;     // first owned cell (-1 on the low border)
  int cx1, cy1, cz1;     // last owned cell
};

// A lane's segment has at most 8 samples, so a register sum holds at most 8
// terms; the view's scale leaves 3 bits of headroom for that, so a sum stays
// below 2^30.  Measured for 32 views at 256^2 of a 128^3 volume (K = 512):
// lanes on single samples (neighbouring lanes adding to the same slots)
// 12.7 ms, runs of 4 samples 5.2 ms, of 8 4.9 ms, of 16 5.8 ms; keeping the
// sums of corners shared with the next cell was slower at every length.
constexpr int kScatterBits = 3;

// First corner slot of cell (x, y, z) in the brick's block (24-bit products).
__device__ __forceinline__ unsigned corner_slot(int x, int y, int z, const BrickBox &bb) {
  const unsigned sxl = (unsigned)(x - bb.X0 + 1), syl = (unsigned)(y - bb.Y0 + 1), szl = (unsigned)(z - bb.Z0 + 1);
  return __umul24(szl, (unsigned)(kBlk * kBlk)) + __umul24(syl, (unsigned)kBlk) + sxl;
}

// The eight corner sums of one cell into the brick's int64 block.
__device__ __forceinline__ void flush_cell(unsigned long long *p, const float *A) {
  constexpr int off[8] = {0, 1, kBlk, kBlk + 1, kBlk * kBlk, kBlk * kBlk + 1, kBlk * kBlk + kBlk,
                          kBlk * kBlk + kBlk + 1};
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (A[q] != 0.f) atomicAdd(p + off[q], fixed_round(A[q]));
}

// Scatter the first n rays of the list, whose candidate samples form nseg
// segments (aligned runs of 8 sample indices, clipped to the ray's candidate
// range), in rounds of 64 segments: lane l takes segment l of the round, so a
// lane's samples are consecutive samples of one ray (which mostly share cells)
// and its dL/ds is one 32-byte run (two 16-byte loads when K is a multiple of
// 8).  Each owned sample's corner terms (dL/ds times the view's scale, times
// the weight) are summed in registers per cell and added in fixed point (int32,
// rounded, sign-extended) to the brick's int64 block in LDS.
__device__ void scatter_list(const DiffArgs &a, const DiffView &v, const ScatterList &L, int n, int nseg,
                             float scale, unsigned long long *s_acc, int lane, const BrickBox &bb) {
  const float fnx = float(a.nx), fny = float(a.ny), fnz = float(a.nz);
  const bool vec = (a.K & 7) == 0;
  for (int g0 = 0; g0 < nseg; g0 += 64) {
    const int g = g0 + lane;
    if (g >= nseg) continue;
    int j = 0;   // the ray of segment g: the last ray whose first segment is <= g
    for (int step = kList / 2; step > 0; step >>= 1)
      if (j + step < n && L.off[j + step] <= g) j += step;
    const int c8 = 8 * ((L.k0[j] >> 3) + (g - L.off[j]));
    const int ks = max(L.k0[j], c8), ke = min(L.k1[j], c8 + 7);
    const float *p = a.ds + (size_t)L.ray[j] * a.K + c8;
    float dv[8];
    if (vec) {
      const float4 u0 = reinterpret_cast<const float4 *>(p)[0], u1 = reinterpret_cast<const float4 *>(p)[1];
      dv[0] = u0.x; dv[1] = u0.y; dv[2] = u0.z; dv[3] = u0.w;
      dv[4] = u1.x; dv[5] = u1.y; dv[6] = u1.z; dv[7] = u1.w;
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u) dv[u] = (c8 + u >= ks && c8 + u <= ke) ? p[u] : 0.f;
    }
    RaySetup q;
    q.ox = v.ray_o[0]; q.oy = v.ray_o[1]; q.oz = v.ray_o[2];
    q.dx = L.dx[j]; q.dy = L.dy[j]; q.dz = L.dz[j]; q.tnear = L.tn[j]; q.dt = L.dt[j];
    float A[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int cx = 0, cy = 0, cz = 0;
    bool held = false;   // A holds sums of cell (cx, cy, cz)
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = c8 + u;
      const float d = dv[u];
      if (k < ks || k > ke || d == 0.f) continue;
      float px, py, pz;
      position(q, k, px, py, pz);
      const float fx = fnx * px - 0.5f, fy = fny * py - 0.5f, fz = fnz * pz - 0.5f;
      const float flx = floorf(fx), fly = floorf(fy), flz = floorf(fz);
      const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
      if (ix >= bb.cx0 && ix <= bb.cx1 && iy >= bb.cy0 && iy <= bb.cy1 && iz >= bb.cz0 && iz <= bb.cz1) {
        if (held && (ix != cx || iy != cy || iz != cz)) {
          flush_cell(s_acc + corner_slot(cx, cy, cz, bb), A);
#pragma unroll
          for (int q8 = 0; q8 < 8; ++q8) A[q8] = 0.f;
        }
        held = true;
        cx = ix; cy = iy; cz = iz;
        const float gx = fx - flx, gy = fy - fly, gz = fz - flz;
        const float hx = 1.f - gx, hy = 1.f - gy;
        const float vs = d * scale;
        const float wz0 = (1.f - gz) * vs, wz1 = gz * vs;   // z weights times the scaled dL/ds
        const float w00 = hx * hy, w10 = gx * hy, w01 = hx * gy, w11 = gx * gy;
        A[0] = fmaf(w00, wz0, A[0]); A[1] = fmaf(w10, wz0, A[1]);
        A[2] = fmaf(w01, wz0, A[2]); A[3] = fmaf(w11, wz0, A[3]);
        A[4] = fmaf(w00, wz1, A[4]); A[5] = fmaf(w10, wz1, A[5]);
        A[6] = fmaf(w01, wz1, A[6]); A[7] = fmaf(w11, wz1, A[7]);
      }
    }
    if (held) flush_cell(s_acc + corner_slot(cx, cy, cz, bb), A);
  }
}

// One wave per brick workgroup: measured 4.95 ms for 32 views at 256^2 of a
// 128^3 volume (K = 512) against 5.2 and 7.1 ms with two and four waves
// sharing the block (more waves contend for the same LDS slots).
// Orders the wave's LDS list writes and reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One workgroup (one wave) per (brick, view group).  Per view: one scale from
// the largest |dL/ds| of the rays reaching the brick's box, the view's terms
// in the int64 block, then the block folded into FP32 sums held in registers
// (fixed order over the views: deterministic) and cleared.
__global__ __launch_bounds__(64) void diff_scatter_kernel(DiffArgs a) {
  __shared__ unsigned long long s_acc[kBlk3];
  __shared__ ScatterList L;
  const int lane = threadIdx.x;
  const int nbx = (a.nx + kBrick - 1) / kBrick, nby = (a.ny + kBrick - 1) / kBrick,
            nbz = (a.nz + kBrick - 1) / kBrick;
  // brick order: each z slice of the brick grid is shifted by (5, 3) * z bricks
  // in x and y, so the workgroups a CU receives (256 apart in dispatch order,
  // otherwise one z column of bricks) spread over the volume's central and
  // outer bricks
  const int bl = blockIdx.x % (nbx * nby * nbz), grp = blockIdx.x / (nbx * nby * nbz);
  const int bz = bl / (nbx * nby);
  const int bx = (bl % nbx + 5 * bz) % nbx, by = ((bl / nbx) % nby + 3 * bz) % nby;
  const int bi = (bz * nby + by) * nbx + bx;
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
    // the view's scale: the largest |dL/ds| of the box's rays times 2^e below
    // 2^30, less 3 bits for the register sums, less the bits for every listed
    // sample adding to one slot beyond 2^33 terms
    float B = 0.f;
    for (int r = lane; r < nR; r += 64) B = fmaxf(B, a.dsmax[((size_t)view * a.H + py0 + r / bw) * a.W + px0 + r % bw]);
    B = wave_max(B);
    if (!(B > 0.f)) continue;
    const unsigned long long items = (unsigned long long)nR * (unsigned long long)a.K;
    const int hv = items > (1ull << 33) ? 64 - __clzll(items - 1ull) - 33 : 0;
    const int e = fixed_exponent(B, hv + kScatterBits);
    const float scale = ldexpf(1.f, e);
    int n = 0, total = 0;   // rays and segments in the list (wave-uniform)
    for (int base = 0; base < nR; base += 64) {
      const int rr = base + lane;
      int cnt = 0, k0 = 0, k1 = -1, ray = 0;
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
              // the box's margin (1e-4) is far above the rounding of ta, tb and
              // of these quotients, so the range holds every owned sample
              const float ka = fminf(fmaxf(ceilf(fa), 0.f), float(a.K));
              const float kb = fminf(fmaxf(floorf(fb), -1.f), float(a.K - 1));
              k0 = (int)ka;
              k1 = (int)kb;
            } else {
              k0 = 0;
              k1 = a.K - 1;
            }
            cnt = k1 >= k0 ? (k1 >> 3) - (k0 >> 3) + 1 : 0;   // segments
          }
        }
      }
      const unsigned long long m = __ballot(cnt > 0);
      const int n_new = __popcll(m);
      if (n_new == 0) continue;
      if (n + n_new > kList) {   // list full: scatter it first
        wave_sync();
        scatter_list(a, v, L, n, total, scale, s_acc, lane, bb);
        wave_sync();
        n = 0;
        total = 0;
      }
      const int incl = wave_incl_scan(cnt, lane);
      if (cnt > 0) {
        const int pos = n + __popcll(m & ((1ull << lane) - 1ull));
        L.dx[pos] = rs.dx; L.dy[pos] = rs.dy; L.dz[pos] = rs.dz; L.tn[pos] = rs.tnear; L.dt[pos] = rs.dt;
        L.k0[pos] = k0; L.k1[pos] = k1; L.off[pos] = total + incl - cnt; L.ray[pos] = ray;
      }
      n += n_new;
      total += __shfl(incl, 63);
    }
    if (n > 0) {
      wave_sync();
      scatter_list(a, v, L, n, total, scale, s_acc, lane, bb);
    }
    wave_sync();
    const float unscale = ldexpf(1.f, -e);
#pragma unroll
    for (int i = 0; i < kBlkPerLane; ++i) {
      const int idx = lane + 64 * i;
      if (idx < kBlk3) {
        const long long val = (long long)s_acc[idx];
        if (val != 0) {
          acc[i] += (float)val * unscale;
          s_acc[idx] = 0ull;
        }
      }
    }
    wave_sync();
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
// (axis_sources on each axis), summed in a fixed order.  Workgroup 0 also
// writes grad_tf from the int64 sums: one rounding (int64 -> float) and an
// exact power-of-two scale.
__global__ __launch_bounds__(256) void diff_combine_kernel(DiffArgs a) {
  if (blockIdx.x == 0) {
    const int e = tf_exponent(a, threadIdx.x & 63);
    const float unscale = ldexpf(1.f, -e);
    for (int i = threadIdx.x; i < 4 * a.tf_n; i += 256) {
      const int j = i >> 2, ch = i & 3;
      a.grad_tf[i] = (float)(long long)a.gtf_acc[ch * kMaxTf + j] * unscale;
    }
  }
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
