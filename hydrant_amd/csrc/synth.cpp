// Host-side dataset preparation: deterministic volume generators, brick
// packing with apron, thumbnails and seeded page tables (hydrant_dataset.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hydrant_dataset.h"
#include "hyd_internal.hpp"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

int n_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (const char *e = std::getenv("OMP_NUM_THREADS")) {
    int v = std::atoi(e);
    if (v > 0) h = (unsigned)v;
  }
  return h ? (int)std::min(h, 64u) : 1;
}

template <class F> void parallel_for(int n, F &&fn) {
  int nt = std::min(n_threads(), std::max(n, 1));
  if (nt <= 1) { for (int i = 0; i < n; ++i) fn(i); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] { for (int i = t; i < n; i += nt) fn(i); });
  for (auto &x : th) x.join();
}

uint8_t quant_u8(double v) {
  v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  return (uint8_t)std::floor(255.0 * v + 0.5);
}

// Improved noise (Perlin, "Improving Noise", SIGGRAPH 2002): quintic fade,
// 12 gradient directions selected from the low 4 hash bits.
struct Perlin {
  int p[512];
  explicit Perlin(uint64_t seed) {
    SplitMix64 rng{seed};
    int perm[256];
    for (int i = 0; i < 256; ++i) perm[i] = i;
    for (int i = 255; i > 0; --i) {
      int j = (int)(rng.next() % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
  }
  static float fade(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }
  static float lerp(float t, float a, float b) { return a + t * (b - a); }
  static float grad(int hash, float x, float y, float z) {
    int h = hash & 15;
    float u = h < 8 ? x : y;
    float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
  }
  float noise(float x, float y, float z) const {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx; y -= fy; z -= fz;
    float u = fade(x), v = fade(y), w = fade(z);
    int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))));
  }
};

float read_norm(const void *vol, int fmt, size_t i) {
  switch (fmt) {
    case HYD_FMT_U8: return float(static_cast<const uint8_t *>(vol)[i]) / 255.f;
    case HYD_FMT_F16: return hyd_f16_to_f32(static_cast<const uint16_t *>(vol)[i]);
    default: return static_cast<const float *>(vol)[i];
  }
}

}  // namespace

extern "C" {

uint16_t hyd_f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));  // NaN / inf
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520 -> inf
  if (ax < 0x38800000u) {
    // subnormal half (or zero): value = ax_float * 2^24 rounded to an integer
    if (ax < 0x33000000u) return (uint16_t)sign;  // < 2^-25 rounds to 0
    const uint32_t e = ax >> 23;
    const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126 - e;  // 0x38800000 has e = 113 -> shift 13
    uint32_t r = m >> (shift);
    const uint32_t rem = m & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (r & 1u))) ++r;
    return (uint16_t)(sign | r);
  }
  uint32_t r = ((ax - 0x38000000u) >> 13);  // rebias exponent 127 -> 15
  const uint32_t rem = ax & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return (uint16_t)(sign | r);
}

float hyd_f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  const uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {  // subnormal: m * 2^-24
      float v = float(m) * 5.9604644775390625e-08f;
      std::memcpy(&x, &v, 4);
      x |= sign;
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

int32_t hyd_gen_blobs_u8(int32_t nx, int32_t ny, int32_t nz, int32_t n_blobs, uint64_t seed, double sig_lo,
                         double sig_hi, double sigma_scale, double amp_lo, double amp_hi, uint8_t *out) {
  HYD_REQUIRE(out && nx > 0 && ny > 0 && nz > 0 && n_blobs >= 0, HYD_ERR_ARG, "hyd_gen_blobs_u8: invalid arguments");
  SplitMix64 rng{seed};
  struct Blob { double c[3], s, a; };
  std::vector<Blob> blobs(n_blobs);
  for (auto &b : blobs) {
    b.c[0] = rng.uniform(0, nx);
    b.c[1] = rng.uniform(0, ny);
    b.c[2] = rng.uniform(0, nz);
    b.s = rng.uniform(sig_lo, sig_hi) * sigma_scale;
    b.a = rng.uniform(amp_lo, amp_hi);
  }
  const int dims[3] = {nx, ny, nz};
  std::vector<double> g[3];
  for (int ax = 0; ax < 3; ++ax) {
    g[ax].resize((size_t)n_blobs * dims[ax]);
    for (int b = 0; b < n_blobs; ++b)
      for (int i = 0; i < dims[ax]; ++i) {
        const double dx = (i + 0.5) - blobs[b].c[ax];
        g[ax][(size_t)b * dims[ax] + i] = std::exp(-(dx * dx) / (2.0 * blobs[b].s * blobs[b].s));
      }
  }
  parallel_for(nz, [&](int z) {
    std::vector<double> acc((size_t)nx * ny, 0.0);
    for (int b = 0; b < n_blobs; ++b) {
      const double tz = blobs[b].a * g[2][(size_t)b * nz + z];
      const double *gx = &g[0][(size_t)b * nx];
      for (int y = 0; y < ny; ++y) {
        const double tzy = tz * g[1][(size_t)b * ny + y];
        double *row = &acc[(size_t)y * nx];
        for (int x = 0; x < nx; ++x) row[x] += tzy * gx[x];
      }
    }
    uint8_t *dst = out + (size_t)z * nx * ny;
    for (size_t i = 0; i < (size_t)nx * ny; ++i) dst[i] = quant_u8(acc[i]);
  });
  return HYD_OK;
}

int32_t hyd_gen_perlin_u8(int32_t n, int32_t octaves, float persistence, float base_period, float gain,
                          uint64_t seed, uint8_t *out) {
  HYD_REQUIRE(out && n > 0 && octaves > 0 && base_period > 0.f, HYD_ERR_ARG, "hyd_gen_perlin_u8: invalid arguments");
  const Perlin pn(seed);
  std::vector<float> amp(octaves), freq(octaves);
  float a = 1.f, f = 1.f / base_period;
  for (int o = 0; o < octaves; ++o) { amp[o] = a; freq[o] = f; a *= persistence; f *= 2.f; }
  parallel_for(n, [&](int z) {
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        float sum = 0.f;
        for (int o = 0; o < octaves; ++o)
          sum += amp[o] * pn.noise((x + 0.5f) * freq[o], (y + 0.5f) * freq[o], (z + 0.5f) * freq[o]);
        out[((size_t)z * n + y) * n + x] = quant_u8(0.5 + (double)(gain * sum));
      }
  });
  return HYD_OK;
}

int32_t hyd_gen_sdf_spheres_f16(int32_t n, int32_t n_spheres, float r_lo, float r_hi, float k, uint64_t seed,
                                uint16_t *out_half) {
  HYD_REQUIRE(out_half && n > 0 && n_spheres > 0 && k > 0.f, HYD_ERR_ARG, "hyd_gen_sdf_spheres_f16: invalid arguments");
  SplitMix64 rng{seed};
  std::vector<float> cx(n_spheres), cy(n_spheres), cz(n_spheres), r(n_spheres);
  for (int i = 0; i < n_spheres; ++i) {
    cx[i] = (float)rng.uniform(0, n);
    cy[i] = (float)rng.uniform(0, n);
    cz[i] = (float)rng.uniform(0, n);
    r[i] = (float)rng.uniform(r_lo, r_hi);
  }
  const float quarter_k = k * 0.25f;
  parallel_for(n, [&](int z) {
    std::vector<float> acc(n);
    for (int y = 0; y < n; ++y) {
      const float pz = z + 0.5f, py = y + 0.5f;
      for (int s = 0; s < n_spheres; ++s) {
        const float dz = pz - cz[s], dy = py - cy[s];
        const float dzy = dz * dz + dy * dy;
        for (int x = 0; x < n; ++x) {
          const float dx = (x + 0.5f) - cx[s];
          const float d = std::sqrt(dx * dx + dzy) - r[s];
          if (s == 0) {
            acc[x] = d;
          } else {
            const float a = acc[x];
            const float h = std::max(k - std::fabs(a - d), 0.f) / k;
            acc[x] = std::min(a, d) - h * h * quarter_k;
          }
        }
      }
      uint16_t *dst = out_half + ((size_t)z * n + y) * n;
      for (int x = 0; x < n; ++x) dst[x] = hyd_f32_to_f16(acc[x]);
    }
  });
  return HYD_OK;
}

int32_t hyd_pack_bricks(const void *vol, int32_t fmt_in, const int32_t dims[3], int32_t bs, int32_t apron,
                        int32_t fmt_out, int32_t normalize, const int32_t *bricks, int32_t n_bricks, void *out) {
  const bool raw_counts = fmt_in == HYD_FMT_U8 && !normalize;
  HYD_REQUIRE(vol && dims && out && bs > 0 && apron >= 0, HYD_ERR_ARG, "hyd_pack_bricks: invalid arguments");
  HYD_REQUIRE(fmt_in >= HYD_FMT_U8 && fmt_in <= HYD_FMT_F32 && fmt_out >= HYD_FMT_U8 && fmt_out <= HYD_FMT_F32,
              HYD_ERR_ARG, "hyd_pack_bricks: unknown format");
  HYD_REQUIRE(fmt_out != HYD_FMT_U8 || fmt_in == HYD_FMT_U8, HYD_ERR_ARG, "u8 bricks need a u8 volume");
  const int g[3] = {(dims[0] + bs - 1) / bs, (dims[1] + bs - 1) / bs, (dims[2] + bs - 1) / bs};
  const int total = g[0] * g[1] * g[2];
  // NULL stands for the bricks 0 .. n_bricks - 1, so an empty list packs nothing
  HYD_REQUIRE(n_bricks >= 0 && (bricks || n_bricks <= total), HYD_ERR_RANGE,
              "hyd_pack_bricks: %d bricks requested from a grid of %d", n_bricks, total);
  const int count = n_bricks;
  if (bricks)
    for (int i = 0; i < count; ++i)
      HYD_REQUIRE(bricks[i] >= 0 && bricks[i] < total, HYD_ERR_RANGE, "brick index %d out of range", bricks[i]);
  const int pad = bs + 2 * apron;
  const size_t pv = (size_t)pad * pad * pad;
  parallel_for(count, [&](int p) {
    const int b = bricks ? bricks[p] : p;
    const int bx = b % g[0], by = (b / g[0]) % g[1], bz = b / (g[0] * g[1]);
    for (int k = 0; k < pad; ++k) {
      const int z = std::min(std::max(bz * bs + k - apron, 0), dims[2] - 1);
      for (int j = 0; j < pad; ++j) {
        const int y = std::min(std::max(by * bs + j - apron, 0), dims[1] - 1);
        for (int i = 0; i < pad; ++i) {
          const int x = std::min(std::max(bx * bs + i - apron, 0), dims[0] - 1);
          const size_t src = ((size_t)z * dims[1] + y) * dims[0] + x;
          const size_t dst = (size_t)p * pv + ((size_t)k * pad + j) * pad + i;
          const float value = raw_counts ? float(static_cast<const uint8_t *>(vol)[src]) : 0.f;
          if (fmt_out == HYD_FMT_U8) {
            static_cast<uint8_t *>(out)[dst] = static_cast<const uint8_t *>(vol)[src];
          } else if (fmt_out == HYD_FMT_F16) {
            static_cast<uint16_t *>(out)[dst] =
                fmt_in == HYD_FMT_F16 ? static_cast<const uint16_t *>(vol)[src]
                                      : hyd_f32_to_f16(raw_counts ? value : read_norm(vol, fmt_in, src));
          } else {
            static_cast<float *>(out)[dst] = raw_counts ? value : read_norm(vol, fmt_in, src);
          }
        }
      }
    }
  });
  return HYD_OK;
}

int32_t hyd_brick_stats(const void *vol, int32_t fmt_in, const int32_t dims[3], int32_t bs, int32_t apron,
                        float *bmin, float *bmax, float *bmean) {
  HYD_REQUIRE(vol && dims && bs > 0 && apron >= 0, HYD_ERR_ARG, "hyd_brick_stats: invalid arguments");
  const int g[3] = {(dims[0] + bs - 1) / bs, (dims[1] + bs - 1) / bs, (dims[2] + bs - 1) / bs};
  const int total = g[0] * g[1] * g[2];
  parallel_for(total, [&](int b) {
    const int bx = b % g[0], by = (b / g[0]) % g[1], bz = b / (g[0] * g[1]);
    float lo = INFINITY, hi = -INFINITY;
    double sum = 0.0;
    size_t cnt = 0;
    for (int k = -apron; k < bs + apron; ++k) {
      const int zz = bz * bs + k;
      const int z = std::min(std::max(zz, 0), dims[2] - 1);
      for (int j = -apron; j < bs + apron; ++j) {
        const int yy = by * bs + j;
        const int y = std::min(std::max(yy, 0), dims[1] - 1);
        for (int i = -apron; i < bs + apron; ++i) {
          const int xx = bx * bs + i;
          const int x = std::min(std::max(xx, 0), dims[0] - 1);
          const float v = read_norm(vol, fmt_in, ((size_t)z * dims[1] + y) * dims[0] + x);
          lo = std::min(lo, v);
          hi = std::max(hi, v);
          const bool interior = i >= 0 && i < bs && j >= 0 && j < bs && k >= 0 && k < bs && xx < dims[0] &&
                                yy < dims[1] && zz < dims[2];
          if (interior) { sum += v; ++cnt; }
        }
      }
    }
    if (bmin) bmin[b] = lo;
    if (bmax) bmax[b] = hi;
    if (bmean) bmean[b] = cnt ? float(sum / double(cnt)) : 0.f;
  });
  return HYD_OK;
}

int32_t hyd_tf_alpha_zero_on(const float *rgba, int32_t n, float lo, float hi) {
  if (!rgba || n < 1) return 0;
  // entries that the linear lookup at fx = n*s - 0.5 can touch for s in [lo, hi]
  const float flo = std::floor(float(n) * lo - 0.5f), fhi = std::ceil(float(n) * hi - 0.5f);
  const int i0 = std::min(std::max((int)flo, 0), n - 1);
  const int i1 = std::min(std::max((int)fhi, 0), n - 1);
  for (int i = i0; i <= i1; ++i)
    if (rgba[4 * i + 3] != 0.f) return 0;
  return 1;
}

int32_t hyd_chebyshev(const int32_t dims[3], const uint8_t *empty, int32_t *cheb) {
  HYD_REQUIRE(dims && empty && cheb && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, HYD_ERR_ARG,
              "hyd_chebyshev: invalid arguments");
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  const int total = nx * ny * nz;
  std::vector<int> filled;
  for (int i = 0; i < total; ++i)
    if (!empty[i]) filled.push_back(i);
  const int none = std::max(nx, std::max(ny, nz));
  parallel_for(nz, [&](int z) {
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int i = (z * ny + y) * nx + x;
        if (!empty[i]) { cheb[i] = 0; continue; }
        int best = none;
        for (int f : filled) {
          const int fx = f % nx, fy = (f / nx) % ny, fz = f / (nx * ny);
          const int d = std::max(std::abs(fx - x), std::max(std::abs(fy - y), std::abs(fz - z)));
          best = std::min(best, d);
        }
        cheb[i] = best;
      }
  });
  return HYD_OK;
}

int32_t hyd_seeded_page_table(const int32_t dims[3], const uint8_t *empty, int32_t n_resident, uint64_t seed,
                              int32_t *vaddr, int32_t *order_out, int32_t *n_nonempty_out) {
  HYD_REQUIRE(dims && empty && vaddr && n_resident >= 0, HYD_ERR_ARG, "hyd_seeded_page_table: invalid arguments");
  const int total = dims[0] * dims[1] * dims[2];
  std::vector<int32_t> order;
  for (int i = 0; i < total; ++i)
    if (!empty[i]) order.push_back(i);
  SplitMix64 rng{seed};
  for (int i = (int)order.size() - 1; i > 0; --i) {
    const int j = (int)(rng.next() % (uint64_t)(i + 1));
    std::swap(order[i], order[j]);
  }
  for (int i = 0; i < total; ++i) vaddr[i] = -1;
  const int nr = std::min<int>(n_resident, (int)order.size());
  for (int p = 0; p < nr; ++p) vaddr[order[p]] = p;
  if (order_out) {
    for (int i = 0; i < total; ++i) order_out[i] = i < (int)order.size() ? order[i] : -1;
  }
  if (n_nonempty_out) *n_nonempty_out = (int32_t)order.size();
  return HYD_OK;
}

}  // extern "C"
