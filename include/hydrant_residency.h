/*
 * hydrant_residency.h - view-driven brick residency (host side of the render path).
 *
 * The reference keeps a bounded set of decoded bricks resident for the
 * current camera (RtBlockPagingServerImpl::update, src/slave/paging/
 * rt_block_paging.cc:296-326): it culls the non-empty bricks against the view
 * frustum (OctreeCuller::cull, include/hydrant/octree_culler.hpp:121-161),
 * keeps the nearest max_block_count of them, and diffs them against the
 * present set; a decoded brick then takes a free pool slot or the slot of a
 * brick no longer required, whose page-table entry falls back to its basic
 * (lowest-level) entry (rt_block_paging.cc:249-291).
 *
 * Here the decode thread is the caller: hyd_residency_update returns the
 * missing bricks nearest first; for each one the caller calls
 * hyd_residency_commit, which returns the pool slot to write the brick into
 * (hyd_write_pool_pages, hydrant_hip.h) or -1 when the brick is skipped,
 * and keeps the page table (hyd_residency_vaddr, for hyd_set_page_table).
 * Host-only, deterministic; every call returns HYD_OK (0) or an HYD_ERR_*
 * code of hydrant_hip.h with hyd_last_error() set.
 */
#ifndef HYDRANT_RESIDENCY_H
#define HYDRANT_RESIDENCY_H

#include <stdint.h>

#include "hydrant_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* OctreeCuller::cull (octree_culler.hpp:121-161) with the default screen rect
 * [-1, 1]^2 (ScreenRect, :88-92): the bricks of [bbox_min, bbox_max) whose
 * chebyshev entry is 0 and whose unit box the frustum of `view` contains
 * (Frustum::contains_fast, :28-50; Frustum from Camera, :190-213, with the
 * reference's border rays taken in camera space), the `limit` nearest to the
 * camera by squared centre distance (ties by brick index), as linear brick
 * indices (x fastest) in increasing order.  out_idx holds room for the grid's
 * brick count; *out_n receives the number written.  chebyshev: dims[0] *
 * dims[1] * dims[2] int32, x fastest. */
int32_t hyd_cull_bricks(const hyd_view *view, const int32_t dims[3], const int32_t *chebyshev,
                        const int32_t bbox_min[3], const int32_t bbox_max[3], int64_t limit, int32_t *out_idx,
                        int64_t *out_n);

typedef struct hyd_residency hyd_residency;

/* A residency manager over a brick grid of dims with max_pages pool slots,
 * numbered first_slot, first_slot + 1, ... (the reference numbers them after
 * the lowest-level blocks, rt_block_paging.cc:256-260).  basic_vaddr
 * (optional, dims-sized, x fastest) is the page-table entry each brick falls
 * back to when it is not resident: the reference's lowest sample level
 * (rt_block_paging.cc:215-219); NULL means -1 (not resident) for all. */
int32_t hyd_residency_create(const int32_t dims[3], int32_t first_slot, int32_t max_pages,
                             const int32_t *basic_vaddr, hyd_residency **out);
int32_t hyd_residency_destroy(hyd_residency *res);

/* RtBlockPagingServerImpl::update (rt_block_paging.cc:296-326): cull the grid
 * for `view` with limit max_pages, then missing = required - present (written
 * to out_missing nearest first, ties by brick index) and redundant = present
 * - required (kept for hyd_residency_commit).  out_missing holds room for
 * max_pages entries. */
int32_t hyd_residency_update(hyd_residency *res, const hyd_view *view, const int32_t *chebyshev,
                             int32_t *out_missing, int64_t *n_missing);

/* The decode-completion step of the reference pipeline (rt_block_paging.cc:
 * 249-291) for brick `idx`: *slot receives the pool slot its voxels go to, a
 * new slot while the pool has room, else the slot of the last redundant brick
 * (whose page-table entry reverts to its basic entry), or -1 when the brick is
 * already present ("abandoned") or no slot can be freed ("artifact").  "Last"
 * assumes the reference's std::set<vol::Idx> orders bricks by linear index
 * (z, then y, then x): varch's Idx ordering is not available to check. */
int32_t hyd_residency_commit(hyd_residency *res, int32_t idx, int32_t *slot);

/* The current page table (dims-sized, x fastest): slot of each resident brick,
 * basic entry elsewhere. */
int32_t hyd_residency_vaddr(const hyd_residency *res, int32_t *out);
/* Present bricks (increasing index) and slots in use. */
int32_t hyd_residency_stats(const hyd_residency *res, int64_t *n_present, int32_t *n_slots);

#ifdef __cplusplus
}
#endif

#endif
