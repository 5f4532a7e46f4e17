// Random-grid fuzz of the frustum culler and the residency bookkeeping
// (hydrant_amd/csrc/residency.cpp) under AddressSanitizer and UBSan: culled
// indices are non-empty bricks in increasing order, slots stay inside
// [first_slot, first_slot + max_pages) and unique, out-of-grid commits are
// refused.  Exit status names the first broken invariant (0 = all held).
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <random>
#include <set>
#include <vector>
#include "hydrant_residency.h"
#include "hyd_internal.hpp"
namespace hyd { void set_error(const char *fmt, ...) { (void)fmt; } }
int main() {
  std::mt19937 rng(99);
  long culled = 0, commits = 0;
  for (int it = 0; it < 1000; ++it) {
    int32_t dims[3] = {1 + (int)(rng() % 12), 1 + (int)(rng() % 12), 1 + (int)(rng() % 12)};
    const int n = dims[0] * dims[1] * dims[2];
    std::vector<int32_t> cheb(n);
    for (auto &c : cheb) c = rng() % 3 == 0 ? 1 : 0;
    hyd_view v{};
    // random rigid transform: rotation from random angles, origin around the grid
    float a = (rng() % 6283) / 1000.f, b = (rng() % 3141) / 1000.f;
    float fx = std::cos(a) * std::sin(b), fy = std::cos(b), fz = std::sin(a) * std::sin(b);   // forward
    float rx = fz, ry = 0, rz = -fx; float rl = std::sqrt(rx * rx + rz * rz) + 1e-6f; rx /= rl; rz /= rl;
    float ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
    float o[3] = {dims[0] * 0.5f - 2.f * dims[0] * fx, dims[1] * 0.5f - 2.f * dims[1] * fy, dims[2] * 0.5f - 2.f * dims[2] * fz};
    float m[16] = {rx, ry, rz, 0, ux, uy, uz, 0, -fx, -fy, -fz, 0, o[0], o[1], o[2], 1};
    for (int i = 0; i < 16; ++i) v.trans[i] = m[i];
    v.ctg_fovy_2 = 0.5f + (rng() % 300) / 100.f;
    v.ray_o[0] = o[0]; v.ray_o[1] = o[1]; v.ray_o[2] = o[2];
    std::vector<int32_t> out(n);
    int64_t nout = -1;
    int32_t lo[3] = {0, 0, 0};
    int64_t limit = (int64_t)(rng() % (n + 3)) - 1;
    if (hyd_cull_bricks(&v, dims, cheb.data(), lo, dims, limit, out.data(), &nout) != 0) return 1;
    if (nout < 0 || nout > n) return 2;
    for (int64_t i = 0; i < nout; ++i) if (out[i] < 0 || out[i] >= n || cheb[out[i]] != 0 || (i && out[i] <= out[i - 1])) return 3;
    culled += nout;
    const int32_t max_pages = rng() % (n + 1), first = rng() % 5;
    hyd_residency *r = nullptr;
    if (hyd_residency_create(dims, first, max_pages, nullptr, &r) != 0) return 4;
    for (int frame = 0; frame < 3; ++frame) {
      std::vector<int32_t> miss(n);
      int64_t nm = -1;
      if (hyd_residency_update(r, &v, cheb.data(), miss.data(), &nm) != 0) return 5;
      if (nm < 0 || nm > max_pages) return 6;
      for (int64_t i = 0; i < nm; ++i) {
        int32_t slot;
        if (hyd_residency_commit(r, miss[i], &slot) != 0) return 7;
        ++commits;
        if (slot != -1 && (slot < first || slot >= first + max_pages)) return 8;
      }
      int32_t bad;
      if (hyd_residency_commit(r, n + 5, &bad) == 0) return 9;   // out of range must be refused
      std::vector<int32_t> va(n);
      hyd_residency_vaddr(r, va.data());
      std::set<int32_t> used;
      for (int32_t s : va) if (s >= 0 && !used.insert(s).second) return 10;   // slots unique
      v.trans[12] += 0.7f; v.ray_o[0] += 0.7f;                                // move the camera
    }
    hyd_residency_destroy(r);
  }
  std::printf("culled %ld commits %ld\n", culled, commits);
  return 0;
}
