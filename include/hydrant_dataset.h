/*
 * hydrant_dataset.h - host-side dataset preparation for the render path.
 *
 * The reference reads bricked datasets produced by the varch tools
 * (build_dataset.sh:79-124: archiver with 64^3 bricks + padding, chebyshev and
 * mean thumbnails) and decodes H.264 bricks at render time
 * (lossless_block_paging.cc:137-174).  Decode is outside this build: these
 * functions synthesise volumes deterministically, pack them into padded
 * bricks and derive the thumbnails and page tables the renderer consumes.
 *
 * Randomness: SplitMix64 (Steele et al. 2014) -> 53-bit uniform doubles,
 * drawn in the order documented per generator.  All functions are host-only,
 * multi-threaded, deterministic and return HYD_OK (0) or an HYD_ERR_* code.
 */
#ifndef HYDRANT_DATASET_H
#define HYDRANT_DATASET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sum of n seeded isotropic Gaussian blobs, quantised to u8.
 * Draw order per blob: cx ~ U[0,nx), cy ~ U[0,ny), cz ~ U[0,nz),
 * sigma ~ U[sig_lo, sig_hi) * sigma_scale, amp ~ U[amp_lo, amp_hi).
 * f = sum amp * exp(-dx^2/2s^2) exp(-dy^2/2s^2) exp(-dz^2/2s^2) at voxel
 * centres (i + 0.5), in double; out = floor(255 * min(1, f) + 0.5). */
int32_t hyd_gen_blobs_u8(int32_t nx, int32_t ny, int32_t nz, int32_t n_blobs, uint64_t seed, double sig_lo,
                         double sig_hi, double sigma_scale, double amp_lo, double amp_hi, uint8_t *out);

/* Improved Perlin noise (Perlin 2002) with a seeded permutation of 0..255
 * (Fisher-Yates, j = next() % (i + 1)), `octaves` octaves, amplitude
 * persistence^o, frequency 2^o / base_period at voxel centres; binary32.
 * v = clamp(0.5 + gain * sum, 0, 1), out = floor(255 * v + 0.5). */
int32_t hyd_gen_perlin_u8(int32_t n, int32_t octaves, float persistence, float base_period, float gain,
                          uint64_t seed, uint8_t *out);

/* Signed distance of n seeded spheres (centre ~ U[0,n)^3, radius ~ U[r_lo, r_hi)),
 * folded in order with the polynomial smooth minimum
 * smin(a, b) = min(a, b) - h*h*k/4, h = max(k - |a - b|, 0) / k; binary32,
 * in voxel units at voxel centres, stored as IEEE half (round to nearest even). */
int32_t hyd_gen_sdf_spheres_f16(int32_t n, int32_t n_spheres, float r_lo, float r_hi, float k, uint64_t seed,
                                uint16_t *out_half);

/* IEEE binary32 -> binary16 (round to nearest even) and back. */
uint16_t hyd_f32_to_f16(float f);
float hyd_f16_to_f32(uint16_t h);

/* Pack a volume (x fastest) into padded bricks of (bs + 2*apron)^3, brick
 * after brick in grid order (x fastest), apron voxels taken from the
 * neighbours with clamp-to-edge at the volume border.
 *   fmt_in:  HYD_FMT_U8 (uint8) or HYD_FMT_F16 (uint16 half bits) or HYD_FMT_F32
 *   fmt_out: HYD_FMT_U8 (only from u8), HYD_FMT_F16 or HYD_FMT_F32.
 * normalize: for u8 input and f16/f32 output, 1 stores float(v) / 255 (the
 *            reference's NormalizedFloat read), 0 stores the count v itself,
 *            which f16 holds exactly (renderers then scale by 1/255).
 * bricks: list of n_bricks linear brick indices to pack (page order); NULL
 * stands for the bricks 0 .. n_bricks - 1 (n_bricks = the grid's brick count
 * packs every brick).  out receives n_bricks pages; n_bricks = 0 packs nothing. */
int32_t hyd_pack_bricks(const void *vol, int32_t fmt_in, const int32_t dims[3], int32_t bs, int32_t apron,
                        int32_t fmt_out, int32_t normalize, const int32_t *bricks, int32_t n_bricks, void *out);

/* Per-brick min / max / mean of normalised values.  min/max cover the padded
 * brick (every voxel a sample in the brick can touch); mean covers the bs^3
 * interior (the mean thumbnail of Blocks.Volume). */
int32_t hyd_brick_stats(const void *vol, int32_t fmt_in, const int32_t dims[3], int32_t bs, int32_t apron,
                        float *bmin, float *bmax, float *bmean);

/* 1 if the TF alpha is identically zero for every s in [lo, hi] (the TF is the
 * piecewise-linear lookup of transfer_fn.hpp:14-28), else 0. */
int32_t hyd_tf_alpha_zero_on(const float *rgba, int32_t n, float lo, float hi);

/* Chebyshev thumbnail from per-brick empty flags: 0 for a non-empty brick,
 * otherwise the L-inf distance (in bricks) to the nearest non-empty brick
 * (max(dims) if there is none). */
int32_t hyd_chebyshev(const int32_t dims[3], const uint8_t *empty, int32_t *cheb);

/* Seeded page table: the non-empty bricks are shuffled (Fisher-Yates with the
 * seed) and the first n_resident get page ids 0..n_resident-1 in that order;
 * every other brick maps to -1.  order_out (optional, n_bricks entries)
 * receives the full shuffled order of non-empty bricks followed by -1 fill;
 * *n_nonempty_out the number of non-empty bricks. */
int32_t hyd_seeded_page_table(const int32_t dims[3], const uint8_t *empty, int32_t n_resident, uint64_t seed,
                              int32_t *vaddr, int32_t *order_out, int32_t *n_nonempty_out);

#ifdef __cplusplus
}
#endif

#endif /* HYDRANT_DATASET_H */
