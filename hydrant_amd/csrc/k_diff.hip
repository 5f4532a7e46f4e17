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
  int idx[8];
  float w[8];
};

__device__ __forceinline__ void sample(const DiffArgs &a, float px, float py, float pz, Smp &sm, bool want_taps) {
  const float fx = float(a.nx) * px - 0.5f, fy = float(a.ny) * py - 0.5f, fz = float(a.nz) * pz - 0.5f;
  const float flx = floorf(fx), fly = floorf(fy), flz = floorf(fz);
  const float gx = fx - flx, gy = fy - fly, gz = fz - flz;
  const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
  const int x0 = iclamp(ix, 0, a.nx - 1), x1 = iclamp(ix + 1, 0, a.nx - 1);
  const int y0 = iclamp(iy, 0, a.ny - 1), y1 = iclamp(iy + 1, 0, a.ny - 1);
  const int z0 = iclamp(iz, 0, a.nz - 1), z1 = iclamp(iz + 1, 0, a.nz - 1);
  const int r00 = (z0 * a.ny + y0) * a.nx, r01 = (z0 * a.ny + y1) * a.nx;
  const int r10 = (z1 * a.ny + y0) * a.nx, r11 = (z1 * a.ny + y1) * a.nx;
  const int id[8] = {r00 + x0, r00 + x1, r01 + x0, r01 + x1, r10 + x0, r10 + x1, r11 + x0, r11 + x1};
  const float hx = 1.f - gx, hy = 1.f - gy, hz = 1.f - gz;
  const float w[8] = {hx * hy * hz, gx * hy * hz, hx * gy * hz, gx * gy * hz,
                      hx * hy * gz, gx * hy * gz, hx * gy * gz, gx * gy * gz};
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) s = fmaf(w[q], a.vol[id[q]], s);
  sm.s = s;
  if (want_taps) {
#pragma unroll
    for (int q = 0; q < 8; ++q) { sm.idx[q] = id[q]; sm.w[q] = w[q]; }
  }
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
    float C0 = 0.f, C1 = 0.f, C2 = 0.f, T = 1.f;
    const int nseg = a.nck - 1;
    if (setup(a, a.views[view], x, y, r)) {
      for (int seg = 0; seg < nseg; ++seg) {
        const float Tstart = T;
        float amax = 0.f;
        const int k0 = seg * a.ckpt, k1 = min(a.K, k0 + a.ckpt);
        for (int k = k0; k < k1; ++k) {
          float px, py, pz;
          position(r, k, px, py, pz);
          Smp sm;
          sample(a, px, py, pz, sm, false);
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

__device__ __forceinline__ void accumulate(const DiffArgs &a, float *s_gtf, const float4 *s_tf, const RaySetup &r,
                                           const Smp &sm, const Shade &sh, float Tk, const float g[3], float gA,
                                           float S[3], float &P) {
  const float dalpha = Tk * (g[0] * (sh.c[0] - S[0]) + g[1] * (sh.c[1] - S[1]) + g[2] * (sh.c[2] - S[2]) + gA * P);
  float d[4];
  const float ta = Tk * sh.alpha;
  d[0] = ta * g[0];
  d[1] = ta * g[1];
  d[2] = ta * g[2];
  d[3] = sh.a > 0.f ? dalpha * a.sigma * r.dt * (1.f - sh.alpha) : 0.f;
  if (sh.inside) {
    const float h = 1.f - sh.g;
    const float4 p = s_tf[sh.j0], q = s_tf[sh.j1];
    const float ds = float(a.tf_n) * ((q.x - p.x) * d[0] + (q.y - p.y) * d[1] + (q.z - p.z) * d[2] + (q.w - p.w) * d[3]);
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      atomicAdd(&s_gtf[4 * sh.j0 + ch], h * d[ch]);
      atomicAdd(&s_gtf[4 * sh.j1 + ch], sh.g * d[ch]);
    }
    if (ds != 0.f) {
#pragma unroll
      for (int q2 = 0; q2 < 8; ++q2) unsafeAtomicAdd(&a.grad_vol[sm.idx[q2]], sm.w[q2] * ds);
    }
  }
  S[0] = sh.alpha * sh.c[0] + (1.f - sh.alpha) * S[0];
  S[1] = sh.alpha * sh.c[1] + (1.f - sh.alpha) * S[1];
  S[2] = sh.alpha * sh.c[2] + (1.f - sh.alpha) * S[2];
  P = (1.f - sh.alpha) * P;
}

__global__ __launch_bounds__(64) void diff_backward_kernel(DiffArgs a) {
  __shared__ float4 s_tf[kMaxTf];
  __shared__ float s_gtf[4 * kMaxTf];
  for (int i = threadIdx.x; i < a.tf_n; i += 64) s_tf[i] = reinterpret_cast<const float4 *>(a.tf)[i];
  for (int i = threadIdx.x; i < 4 * a.tf_n; i += 64) s_gtf[i] = 0.f;
  __syncthreads();
  const int view = blockIdx.z;
  const int x = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int y = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (x < a.W && y < a.H) {
    const size_t ray = ((size_t)view * a.H + y) * a.W + x;
    const float *ck = a.ckpts + ray * a.nck;
    const float4 gr = a.grad_rgba[ray];
    const float g[3] = {gr.x, gr.y, gr.z};
    const float gA = gr.w;
    RaySetup r;
    if ((g[0] != 0.f || g[1] != 0.f || g[2] != 0.f || gA != 0.f) && setup(a, a.views[view], x, y, r)) {
      float S[3] = {0.f, 0.f, 0.f}, P = 1.f;
      const int nseg = a.nck - 1;
      float Tnext = fabsf(ck[nseg]);
      for (int seg = nseg - 1; seg >= 0; --seg) {
        const int k0 = seg * a.ckpt, k1 = min(a.K, k0 + a.ckpt);
        const float cks = ck[seg];
        const float Tstart = fabsf(cks);
        const bool exact = signbit(cks);
        for (int k = k1 - 1; k >= k0; --k) {
          float px, py, pz;
          position(r, k, px, py, pz);
          Smp sm;
          sample(a, px, py, pz, sm, true);
          Shade sh;
          shade(s_tf, a.tf_n, a.sigma, r.dt, sm.s, sh);
          float Tk;
          if (!exact) {
            Tk = Tnext / (1.f - sh.alpha);
          } else {
            Tk = Tstart;   // recompute T_k forward from the checkpoint
            for (int j = k0; j < k; ++j) {
              float qx, qy, qz;
              position(r, j, qx, qy, qz);
              Smp sj;
              sample(a, qx, qy, qz, sj, false);
              Shade hj;
              shade(s_tf, a.tf_n, a.sigma, r.dt, sj.s, hj);
              Tk = Tk * (1.f - hj.alpha);
            }
          }
          accumulate(a, s_gtf, s_tf, r, sm, sh, Tk, g, gA, S, P);
          Tnext = Tk;
        }
        Tnext = Tstart;   // re-anchor at the checkpoint
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * a.tf_n; i += 64)
    if (s_gtf[i] != 0.f) unsafeAtomicAdd(&a.grad_tf[i], s_gtf[i]);
}

}  // namespace

hipError_t launch_diff_forward(const DiffArgs &a, hipStream_t s) {
  dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile, a.n_views);
  if (a.samples) hipLaunchKernelGGL(diff_forward_kernel<true>, grid, dim3(64), 0, s, a);
  else hipLaunchKernelGGL(diff_forward_kernel<false>, grid, dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_diff_backward(const DiffArgs &a, hipStream_t s) {
  dim3 grid((a.W + kTile - 1) / kTile, (a.H + kTile - 1) / kTile, a.n_views);
  hipLaunchKernelGGL(diff_backward_kernel, grid, dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace hyd
