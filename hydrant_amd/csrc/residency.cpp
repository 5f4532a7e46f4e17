// View-driven brick residency: frustum culling of the non-empty bricks and the
// present / missing / redundant bookkeeping that decides which pool slot a
// decoded brick goes to.  Host code; the caller writes the bricks into the
// pool and uploads the page table.
#include <algorithm>
#include <cstring>
#include <iterator>
#include <vector>

#include "hydrant_residency.h"
#include "hyd_internal.hpp"

namespace hyd {
namespace {

// Frustum of a view (octree_culler.hpp:190-213): four side planes through the
// camera position, normals from the screen-rect corner directions; the
// corner directions themselves stay in camera space, as in the reference,
// where they are used as ray directions from the world-space origin.
struct Frustum {
  f3 norm[4], border[4], orig;

  bool contains(f3 p) const {   // octree_culler.hpp:20-26
    const f3 q = sub(p, orig);
    return dot3(norm[0], q) > 0.f && dot3(norm[1], q) > 0.f && dot3(norm[2], q) > 0.f && dot3(norm[3], q) > 0.f;
  }

  // octree_culler.hpp:28-50: a corner inside, or a border ray through the box
  bool contains_fast(f3 bmin, f3 bmax) const {
    const f3 vp[2] = {bmin, bmax};
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k)
          if (contains(f3{vp[i].x, vp[j].y, vp[k].z})) return true;
    for (int i = 0; i < 4; ++i) {
      float tnear, tfar;
      if (intersect_box(orig, normalize3(border[i]), bmin, bmax, tnear, tfar) && tfar > 0.f) return true;
    }
    return false;
  }
};

Frustum view_frustum(const hyd_view &v) {
  const float c = v.ctg_fovy_2;
  Frustum f;
  f.border[0] = f3{1.f, 1.f, -c};
  f.border[1] = f3{1.f, -1.f, -c};
  f.border[2] = f3{-1.f, -1.f, -c};
  f.border[3] = f3{-1.f, 1.f, -c};
  float o[4];
  mat4_vec4(v.trans, 0.f, 0.f, 0.f, 1.f, o);
  f.orig = f3{o[0], o[1], o[2]};
  for (int i = 0; i < 4; ++i) {
    const f3 n = cross3(f.border[i], f.border[(i + 1) & 3]);
    float w[4];
    mat4_vec4(v.trans, n.x, n.y, n.z, 0.f, w);
    f.norm[i] = f3{w[0], w[1], w[2]};
  }
  return f;
}

// squared distance of brick idx's centre to p (distance2(orig, vec3(idx) + .5))
float centre_dist2(f3 p, int x, int y, int z) {
  const f3 d = sub(p, f3{float(x) + .5f, float(y) + .5f, float(z) + .5f});
  return dot3(d, d);
}

// The culled bricks, nearest `limit` (ties by index), in increasing index order.
void cull(const hyd_view &view, const int32_t dims[3], const int32_t *cheb, const int32_t lo[3],
          const int32_t hi[3], int64_t limit, std::vector<int32_t> &out, f3 &orig) {
  const Frustum f = view_frustum(view);
  orig = f.orig;
  struct Cand { float d; int32_t idx; };
  std::vector<Cand> c;
  for (int z = lo[2]; z < hi[2]; ++z)
    for (int y = lo[1]; y < hi[1]; ++y)
      for (int x = lo[0]; x < hi[0]; ++x) {
        const int32_t i = (int32_t)(((int64_t)z * dims[1] + y) * dims[0] + x);
        if (cheb[i] != 0) continue;
        const f3 bmin{float(x), float(y), float(z)};
        if (f.contains_fast(bmin, add(bmin, f3{1.f, 1.f, 1.f}))) c.push_back({centre_dist2(f.orig, x, y, z), i});
      }
  const auto nearer = [](const Cand &a, const Cand &b) { return a.d < b.d || (a.d == b.d && a.idx < b.idx); };
  const size_t n = (size_t)std::min<int64_t>(std::max<int64_t>(limit, 0), (int64_t)c.size());
  if (n < c.size()) std::nth_element(c.begin(), c.begin() + n, c.end(), nearer);
  c.resize(n);
  out.resize(n);
  for (size_t k = 0; k < n; ++k) out[k] = c[k].idx;
  std::sort(out.begin(), out.end());
}

int32_t check_grid(const int32_t dims[3]) {
  HYD_REQUIRE(dims && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, HYD_ERR_ARG, "brick grid dims must be > 0");
  HYD_REQUIRE((int64_t)dims[0] * dims[1] * dims[2] < (1ll << 31), HYD_ERR_RANGE, "brick grid too large");
  return HYD_OK;
}

}  // namespace
}  // namespace hyd

using namespace hyd;

struct hyd_residency {
  int32_t dims[3];
  int32_t first_slot = 0, max_pages = 0, n_slots = 0;
  std::vector<int32_t> basic;       // fallback page-table entry per brick
  std::vector<int32_t> vaddr;       // current page table
  std::vector<int32_t> present;     // resident bricks, increasing index
  std::vector<int32_t> redundant;   // present but no longer required (increasing; taken from the back)
};

extern "C" {

int32_t hyd_cull_bricks(const hyd_view *view, const int32_t dims[3], const int32_t *chebyshev,
                        const int32_t bbox_min[3], const int32_t bbox_max[3], int64_t limit, int32_t *out_idx,
                        int64_t *out_n) {
  int32_t st = check_grid(dims);
  if (st) return st;
  HYD_REQUIRE(view && chebyshev && bbox_min && bbox_max && out_idx && out_n, HYD_ERR_ARG, "null argument");
  for (int a = 0; a < 3; ++a)
    HYD_REQUIRE(0 <= bbox_min[a] && bbox_min[a] <= bbox_max[a] && bbox_max[a] <= dims[a], HYD_ERR_RANGE,
                "cull box [%d, %d) on axis %d outside the grid of %d", bbox_min[a], bbox_max[a], a, dims[a]);
  std::vector<int32_t> out;
  f3 orig;
  cull(*view, dims, chebyshev, bbox_min, bbox_max, limit, out, orig);
  if (!out.empty()) std::memcpy(out_idx, out.data(), out.size() * sizeof(int32_t));
  *out_n = (int64_t)out.size();
  return HYD_OK;
}

int32_t hyd_residency_create(const int32_t dims[3], int32_t first_slot, int32_t max_pages,
                             const int32_t *basic_vaddr, hyd_residency **out) {
  int32_t st = check_grid(dims);
  if (st) return st;
  HYD_REQUIRE(out && first_slot >= 0 && max_pages >= 0 && (int64_t)first_slot + max_pages < (1ll << 31),
              HYD_ERR_ARG, "invalid residency arguments");
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  hyd_residency *r = new hyd_residency();
  std::memcpy(r->dims, dims, sizeof(r->dims));
  r->first_slot = first_slot;
  r->max_pages = max_pages;
  if (basic_vaddr) r->basic.assign(basic_vaddr, basic_vaddr + n);
  else r->basic.assign(n, -1);
  r->vaddr = r->basic;
  *out = r;
  return HYD_OK;
}

int32_t hyd_residency_destroy(hyd_residency *res) {
  delete res;
  return HYD_OK;
}

int32_t hyd_residency_update(hyd_residency *res, const hyd_view *view, const int32_t *chebyshev,
                             int32_t *out_missing, int64_t *n_missing) {
  HYD_REQUIRE(res && view && chebyshev && out_missing && n_missing, HYD_ERR_ARG, "null argument");
  const int32_t lo[3] = {0, 0, 0};
  std::vector<int32_t> required;
  f3 orig;
  cull(*view, res->dims, chebyshev, lo, res->dims, res->max_pages, required, orig);
  std::vector<int32_t> missing;
  std::set_difference(required.begin(), required.end(), res->present.begin(), res->present.end(),
                      std::back_inserter(missing));
  res->redundant.clear();
  std::set_difference(res->present.begin(), res->present.end(), required.begin(), required.end(),
                      std::back_inserter(res->redundant));
  const auto dist = [&](int32_t i) {
    const int x = i % res->dims[0], y = (i / res->dims[0]) % res->dims[1], z = i / (res->dims[0] * res->dims[1]);
    return centre_dist2(orig, x, y, z);
  };
  std::sort(missing.begin(), missing.end(), [&](int32_t a, int32_t b) {
    const float da = dist(a), db = dist(b);
    return da < db || (da == db && a < b);
  });
  if (!missing.empty()) std::memcpy(out_missing, missing.data(), missing.size() * sizeof(int32_t));
  *n_missing = (int64_t)missing.size();
  return HYD_OK;
}

int32_t hyd_residency_commit(hyd_residency *res, int32_t idx, int32_t *slot) {
  HYD_REQUIRE(res && slot, HYD_ERR_ARG, "null argument");
  HYD_REQUIRE(idx >= 0 && (size_t)idx < res->vaddr.size(), HYD_ERR_RANGE, "brick %d outside the grid", idx);
  *slot = -1;
  auto at = std::lower_bound(res->present.begin(), res->present.end(), idx);
  if (at != res->present.end() && *at == idx) return HYD_OK;   // already present: abandoned
  int32_t s;
  if (res->n_slots < res->max_pages) {
    s = res->first_slot + res->n_slots++;
  } else if (!res->redundant.empty()) {
    // swap out the last redundant brick; the reference's std::set<vol::Idx> is
    // assumed to order bricks by linear index (z, then y, then x)
    const int32_t out = res->redundant.back();
    res->redundant.pop_back();
    res->present.erase(std::lower_bound(res->present.begin(), res->present.end(), out));
    s = res->vaddr[out];
    res->vaddr[out] = res->basic[out];
  } else {
    return HYD_OK;   // no slot can be freed: artifact
  }
  res->vaddr[idx] = s;
  res->present.insert(std::lower_bound(res->present.begin(), res->present.end(), idx), idx);
  *slot = s;
  return HYD_OK;
}

int32_t hyd_residency_vaddr(const hyd_residency *res, int32_t *out) {
  HYD_REQUIRE(res && out, HYD_ERR_ARG, "null argument");
  std::memcpy(out, res->vaddr.data(), res->vaddr.size() * sizeof(int32_t));
  return HYD_OK;
}

int32_t hyd_residency_stats(const hyd_residency *res, int64_t *n_present, int32_t *n_slots) {
  HYD_REQUIRE(res && n_present && n_slots, HYD_ERR_ARG, "null argument");
  *n_present = (int64_t)res->present.size();
  *n_slots = res->n_slots;
  return HYD_OK;
}

}  // extern "C"
