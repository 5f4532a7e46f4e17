// Oracle of the differentiable render (TEST INFRASTRUCTURE ONLY; see oracle.hpp).
//
// Forward model per ray (build-defined; the reference has no differentiable path):
//   rays from the reference camera (Raycaster ray generation, cpu_shading.cc:16-26),
//   clipped to the unit bbox of a single brick of nx*ny*nz voxels (Ray::intersect,
//   glm_math.hpp:58-72), K samples at the midpoints t_k = tnear + (k + 0.5) dt,
//   dt = (tfar - tnear) / K;
//   s_k = trilinear value at fx = n*p - 0.5 with clamped taps (CpuSampler
//   filter_linear, cpu_sampler.hpp:88-104, single brick k = 1, b = 0);
//   (c, a) = TF(s_k) with the reference's Border + Linear lookup (transfer_fn.hpp:14-28);
//   alpha_k = 1 - exp(-sigma * max(a, 0) * dt);
//   C = sum_k T_k alpha_k c_k, T_{k+1} = T_k (1 - alpha_k), output (C, 1 - T_K).
// The adjoint is the exact derivative of this model, evaluated back to front with
// the suffix colour S_{k+1} and suffix transmittance P_{k+1}:
//   dL/dalpha_k = T_k (g . (c_k - S_{k+1}) + gA P_{k+1}),  dL/dc_k = T_k alpha_k g.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.hpp"

using namespace orc;

extern "C" {

typedef struct {
  double trans[16];
  double ctg_fovy_2;
  double ray_o[3];
} orc_view_g;

typedef struct {
  int W, H;
  int K;
  double sigma;
  int dims[3];
  const float *vol;   // dims[0]*dims[1]*dims[2], x fastest
  const float *tf;    // tf_n * 4
  int tf_n;
} orc_diff_args;
}

namespace {

template <class T> View<T> view_of(const orc_view_g *v) {
  View<T> r;
  for (int c = 0; c < 4; ++c)
    r.trans.c[c] = V4<T>{T(v->trans[4 * c]), T(v->trans[4 * c + 1]), T(v->trans[4 * c + 2]), T(v->trans[4 * c + 3])};
  r.ctg_fovy_2 = T(v->ctg_fovy_2);
  r.ray_o = V3<T>{T(v->ray_o[0]), T(v->ray_o[1]), T(v->ray_o[2])};
  return r;
}

template <class T> struct Sample {
  int idx[8];
  T w[8];
  T s;
  int j0, j1;
  T g;
  bool inside;
  T c[3], a, alpha;
};

template <class T> void sample_at(const orc_diff_args *a, V3<T> p, T dt, Sample<T> &sm) {
  const int n[3] = {a->dims[0], a->dims[1], a->dims[2]};
  const T pc[3] = {p.x, p.y, p.z};
  int i0[3], i1[3];
  T f[3];
  for (int c = 0; c < 3; ++c) {
    const T fx = T(n[c]) * pc[c] - T(0.5f);
    const T fl = std::floor(fx);
    f[c] = fx - fl;
    const int i = int(fl);
    i0[c] = gclamp(i, 0, n[c] - 1);
    i1[c] = gclamp(i + 1, 0, n[c] - 1);
  }
  T s = 0;
  for (int q = 0; q < 8; ++q) {
    const int bx = q & 1, by = (q >> 1) & 1, bz = (q >> 2) & 1;
    const int ix = bx ? i1[0] : i0[0], iy = by ? i1[1] : i0[1], iz = bz ? i1[2] : i0[2];
    const T w = (bx ? f[0] : T(1) - f[0]) * (by ? f[1] : T(1) - f[1]) * (bz ? f[2] : T(1) - f[2]);
    sm.idx[q] = (iz * n[1] + iy) * n[0] + ix;
    sm.w[q] = w;
    s += w * T(a->vol[sm.idx[q]]);
  }
  sm.s = s;
  const int N = a->tf_n;
  sm.inside = T(0) <= s && s <= T(1);
  T rgba[4] = {0, 0, 0, 0};
  if (sm.inside) {
    const T gx = T(N) * s - T(0.5f);
    const T fl = std::floor(gx);
    sm.g = gx - fl;
    const int j = int(fl);
    sm.j0 = gclamp(j, 0, N - 1);
    sm.j1 = gclamp(j + 1, 0, N - 1);
    for (int ch = 0; ch < 4; ++ch)
      rgba[ch] = T(a->tf[4 * sm.j0 + ch]) * (T(1) - sm.g) + T(a->tf[4 * sm.j1 + ch]) * sm.g;
  } else {
    sm.g = 0; sm.j0 = sm.j1 = 0;
  }
  sm.c[0] = rgba[0]; sm.c[1] = rgba[1]; sm.c[2] = rgba[2];
  sm.a = rgba[3];
  sm.alpha = T(1) - std::exp(-T(a->sigma) * std::max(sm.a, T(0)) * dt);
}

template <class T> bool ray_setup(const orc_diff_args *a, const View<T> &v, int x, int y, Ray<T> &r, T &tnear, T &dt) {
  r = gen_ray(v, x, y, a->W, a->H);
  T tn, tf;
  if (!r.intersect(Box<T>{V3<T>{0, 0, 0}, V3<T>{1, 1, 1}}, tn, tf)) return false;
  if (tn < 0) tn = 0;
  if (tf < 0) tf = 0;
  tnear = tn;
  dt = (tf - tn) / T(a->K);
  return true;
}

template <class T> void forward_ray(const orc_diff_args *a, const View<T> &v, int x, int y, double out[4]) {
  Ray<T> r;
  T tnear, dt;
  out[0] = out[1] = out[2] = out[3] = 0;
  if (!ray_setup(a, v, x, y, r, tnear, dt)) return;
  T C[3] = {0, 0, 0}, Tr = 1;
  Sample<T> sm;
  for (int k = 0; k < a->K; ++k) {
    const T t = tnear + (T(k) + T(0.5f)) * dt;
    sample_at(a, r.o + r.d * t, dt, sm);
    for (int ch = 0; ch < 3; ++ch) C[ch] += Tr * sm.alpha * sm.c[ch];
    Tr *= T(1) - sm.alpha;
  }
  out[0] = C[0]; out[1] = C[1]; out[2] = C[2]; out[3] = T(1) - Tr;
}

template <class T>
void grad_ray(const orc_diff_args *a, const View<T> &v, int x, int y, const double g4[4], double *gvol, double *gtf) {
  Ray<T> r;
  T tnear, dt;
  if (!ray_setup(a, v, x, y, r, tnear, dt)) return;
  std::vector<Sample<T>> s(a->K);
  std::vector<T> Tk(a->K);
  T Tr = 1;
  for (int k = 0; k < a->K; ++k) {
    const T t = tnear + (T(k) + T(0.5f)) * dt;
    sample_at(a, r.o + r.d * t, dt, s[k]);
    Tk[k] = Tr;
    Tr *= T(1) - s[k].alpha;
  }
  const T g[3] = {T(g4[0]), T(g4[1]), T(g4[2])}, gA = T(g4[3]);
  T S[3] = {0, 0, 0}, P = 1;
  const int N = a->tf_n;
  for (int k = a->K - 1; k >= 0; --k) {
    const Sample<T> &sm = s[k];
    const T dalpha = Tk[k] * (g[0] * (sm.c[0] - S[0]) + g[1] * (sm.c[1] - S[1]) + g[2] * (sm.c[2] - S[2]) + gA * P);
    T d[4];
    for (int ch = 0; ch < 3; ++ch) d[ch] = Tk[k] * sm.alpha * g[ch];
    d[3] = sm.a > T(0) ? dalpha * T(a->sigma) * dt * (T(1) - sm.alpha) : T(0);
    if (sm.inside) {
      T ds = 0;
      for (int ch = 0; ch < 4; ++ch) {
        gtf[4 * sm.j0 + ch] += double((T(1) - sm.g) * d[ch]);
        gtf[4 * sm.j1 + ch] += double(sm.g * d[ch]);
        ds += (T(a->tf[4 * sm.j1 + ch]) - T(a->tf[4 * sm.j0 + ch])) * d[ch];
      }
      ds *= T(N);
      for (int q = 0; q < 8; ++q) gvol[sm.idx[q]] += double(sm.w[q] * ds);
    }
    for (int ch = 0; ch < 3; ++ch) S[ch] = sm.alpha * sm.c[ch] + (T(1) - sm.alpha) * S[ch];
    P = (T(1) - sm.alpha) * P;
  }
}

}  // namespace

extern "C" {

// RGBA of the listed pixels (xy: n pairs).
int orc_diff_render(int prec, int nthreads, const orc_view_g *view, const orc_diff_args *a, int n, const int *xy,
                    double *rgba) {
  int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
  auto run = [&](auto tag) {
    using T = decltype(tag);
    View<T> v = view_of<T>(view);
    row_interleave((n + 63) / 64, nt, [&](int r) {
      for (int i = r * 64; i < std::min(n, r * 64 + 64); ++i) forward_ray<T>(a, v, xy[2 * i], xy[2 * i + 1], rgba + 4 * i);
    });
  };
  if (prec == 32) run(float{}); else run(double{});
  return 0;
}

// Adds dL/dvolume and dL/dTF for the listed pixels given dL/dRGBA (n x 4).
int orc_diff_grad(int prec, const orc_view_g *view, const orc_diff_args *a, int n, const int *xy,
                  const double *grad_rgba, double *grad_vol, double *grad_tf) {
  auto run = [&](auto tag) {
    using T = decltype(tag);
    View<T> v = view_of<T>(view);
    for (int i = 0; i < n; ++i) grad_ray<T>(a, v, xy[2 * i], xy[2 * i + 1], grad_rgba + 4 * i, grad_vol, grad_tf);
  };
  if (prec == 32) run(float{}); else run(double{});
  return 0;
}

}  // extern "C"
