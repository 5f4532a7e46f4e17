// Random-shape fuzz of the dataset preparation (hydrant_amd/csrc/synth.cpp)
// under AddressSanitizer and UBSan: brick packing and statistics over volumes
// whose sizes are not multiples of the brick edge (clamp-to-edge aprons), the
// chebyshev thumbnail against brute force, and the seeded page table.  Every
// buffer is allocated at exactly the size the header documents.  Exit status
// names the first broken invariant (0 = all held).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "hydrant_dataset.h"
#include "hydrant_hip.h"
#include "hyd_internal.hpp"
namespace hyd { void set_error(const char *fmt, ...) { (void)fmt; } }
int main() {
  std::mt19937 rng(7);
  long bricks_total = 0;
  for (int it = 0; it < 150; ++it) {
    const int32_t dims[3] = {1 + (int)(rng() % 40), 1 + (int)(rng() % 40), 1 + (int)(rng() % 40)};
    const int32_t bs = 1 + rng() % 16, apron = rng() % 3;
    const int32_t grid[3] = {(dims[0] + bs - 1) / bs, (dims[1] + bs - 1) / bs, (dims[2] + bs - 1) / bs};
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2], nb = (size_t)grid[0] * grid[1] * grid[2];
    const size_t pad = bs + 2 * apron, pv = pad * pad * pad;
    const int fin = rng() % 3 == 0 ? HYD_FMT_U8 : (rng() % 2 ? HYD_FMT_F16 : HYD_FMT_F32);
    const size_t esz_in = fin == HYD_FMT_U8 ? 1 : fin == HYD_FMT_F16 ? 2 : 4;
    std::vector<uint8_t> vol(nvox * esz_in);
    for (size_t i = 0; i < nvox; ++i) {
      if (fin == HYD_FMT_U8) vol[i] = rng() % 256;
      else if (fin == HYD_FMT_F16) { uint16_t h = hyd_f32_to_f16((rng() % 1000) / 999.f); std::memcpy(&vol[2 * i], &h, 2); }
      else { float f = (rng() % 1000) / 999.f; std::memcpy(&vol[4 * i], &f, 4); }
    }
    const int fout = fin == HYD_FMT_U8 && rng() % 2 ? HYD_FMT_U8 : (rng() % 2 ? HYD_FMT_F16 : HYD_FMT_F32);
    const size_t esz_out = fout == HYD_FMT_U8 ? 1 : fout == HYD_FMT_F16 ? 2 : 4;
    std::vector<int32_t> list;
    const bool all = rng() % 2;
    if (!all) for (size_t b = 0; b < nb; ++b) if (rng() % 2) list.push_back((int32_t)b);
    const size_t npk = all ? nb : list.size();
    std::vector<uint8_t> out(std::max<size_t>(1, npk * pv * esz_out));
    if (hyd_pack_bricks(vol.data(), fin, dims, bs, apron, fout, rng() % 2, all ? nullptr : list.data(),
                        (int32_t)npk, out.data()) != 0) return 1;
    int32_t too_many = (int32_t)nb + 1;   // NULL list longer than the grid: refused
    if (hyd_pack_bricks(vol.data(), fin, dims, bs, apron, fout, 0, nullptr, too_many, out.data()) == 0) return 9;
    std::vector<float> bmin(nb), bmax(nb), bmean(nb);
    if (hyd_brick_stats(vol.data(), fin, dims, bs, apron, bmin.data(), bmax.data(), bmean.data()) != 0) return 2;
    for (size_t b = 0; b < nb; ++b) if (!(bmin[b] <= bmean[b] + 1e-6f && bmean[b] <= bmax[b] + 1e-6f)) return 3;
    std::vector<uint8_t> empty(nb);
    for (auto &e : empty) e = rng() % 4 == 0 ? 0 : 1;
    std::vector<int32_t> cheb(nb);
    if (hyd_chebyshev(grid, empty.data(), cheb.data()) != 0) return 4;
    const int dmax = std::max(grid[0], std::max(grid[1], grid[2]));
    for (int z = 0; z < (nb <= 4096 ? grid[2] : 0); ++z)   // brute force on the smaller grids
      for (int y = 0; y < grid[1]; ++y)
        for (int x = 0; x < grid[0]; ++x) {
          int best = dmax;
          for (int zz = 0; zz < grid[2]; ++zz)
            for (int yy = 0; yy < grid[1]; ++yy)
              for (int xx = 0; xx < grid[0]; ++xx)
                if (!empty[((size_t)zz * grid[1] + yy) * grid[0] + xx])
                  best = std::min(best, std::max(std::abs(x - xx), std::max(std::abs(y - yy), std::abs(z - zz))));
          if (cheb[((size_t)z * grid[1] + y) * grid[0] + x] != best) return 5;
        }
    int32_t n_ne = -1;
    std::vector<int32_t> vaddr(nb), order(nb);
    const int32_t n_res = rng() % (nb + 1);
    if (hyd_seeded_page_table(grid, empty.data(), n_res, rng(), vaddr.data(), order.data(), &n_ne) != 0) return 6;
    int32_t resident = 0;
    for (size_t b = 0; b < nb; ++b) {
      if (vaddr[b] >= 0 && empty[b]) return 7;
      resident += vaddr[b] >= 0;
    }
    if (resident != std::min(n_res, n_ne)) return 8;
    bricks_total += (long)nb;
  }
  std::printf("bricks %ld\n", bricks_total);
  return 0;
}
