/*
 * hydrant_hip.h - C ABI of the MI355X (gfx950) render hot path.
 *
 * This is the thin boundary that replaces the reference's shader-device
 * backend (ShaderRegistry devices[Cpu|Cuda], include/hydrant/core/shader.hpp:95-109,
 * 304-396) and its Raycaster passes (include/hydrant/core/raycaster.hpp:32-123).
 * Everything is plain C: opaque context handle, plain pointers and sizes.
 *
 * Conventions
 *  - Every call returns int32 status (HYD_OK = 0).  On failure a message is
 *    kept in a thread-local buffer readable through hyd_last_error().
 *  - Host pointers stay caller-owned and are only read during the call.
 *    Pool, tables and transfer function are copied into context-owned device
 *    memory.  Arguments named *_dev are device pointers owned by the caller
 *    (hipMalloc'd or a torch tensor's data_ptr()).
 *  - One context per device; a context is not re-entrant.  All work is issued
 *    on the context's stream (hyd_set_stream), asynchronously unless stated.
 *  - Floating point follows the reference CPU backend operation by operation
 *    (FP32, no contraction), so "reference" schedule images match it to the bit
 *    up to the order of float reductions, which the render path does not have.
 */
#ifndef HYDRANT_HIP_H
#define HYDRANT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hyd_ctx hyd_ctx;

enum {
  HYD_OK = 0,
  HYD_ERR_ARG = 1,    /* invalid argument (null, size, range)            */
  HYD_ERR_HIP = 2,    /* HIP runtime error (message carries hipGetErrorString) */
  HYD_ERR_STATE = 3,  /* call order: e.g. render before pool/TF were set   */
  HYD_ERR_RANGE = 4   /* input exceeds a documented limit                  */
};

/* Brick storage formats.  u8 is read as float(v)/255 (Texture3D<u8>
 * NormalizedFloat, include/hydrant/bridge/texture_3d.hpp:74-95). */
enum { HYD_FMT_U8 = 0, HYD_FMT_F16 = 1, HYD_FMT_F32 = 2 };

/* VolumeRenderMode, ui/volume.schema.hpp:11-14 */
enum { HYD_MODE_DEFAULT = 0, HYD_MODE_PARTITION = 1, HYD_MODE_PAGING = 2 };

/* Step schedules.  REFERENCE = volume_shader.cu:32-81 (step escalation x4 at
 * alpha > 0.85, x16 at > 0.93, no early termination).  LATTICE = sample k at
 * entry + d*(k*step), no escalation, early termination when alpha > ert
 * (brick-order independent; used by sort-last segments and the batched form). */
enum { HYD_SCHED_REFERENCE = 0, HYD_SCHED_LATTICE = 1 };

/* BlocksRenderMode, ui/blocks.schema.hpp:11-12 */
/* Blocks.Volume, Blocks.Solid, and the Paging renderer's residency view
 * (green: a page at or above the page table's lowest_blkcnt, red: below it). */
enum { HYD_BLOCKS_VOLUME = 0, HYD_BLOCKS_SOLID = 1, HYD_BLOCKS_PAGING = 2 };

/* IsosurfaceRenderMode, ui/isosurface.schema.hpp:11-12 */
enum { HYD_ISO_COLOR = 0, HYD_ISO_NORMAL = 1, HYD_ISO_POSITION = 2 };

/* ViewArgs, include/hydrant/core/shader.hpp:63-68 */
typedef struct {
  float trans[16];  /* column-major: iet * inverse(lookAt(...)) */
  float ctg_fovy_2; /* 1 / tan(fovy / 2) */
  float ray_o[3];   /* camera position in object (block) space */
} hyd_view;

/* Camera, include/hydrant/core/scene.hpp:32-82 */
typedef struct {
  float position[3];
  float target[3];
  float up[3];
  float fovy;       /* degrees (CameraPerspective.fovy, default 60) */
} hyd_camera;

/* Volume shader fields (IShaderTypeErased shader.hpp:16-22 + VolumeShader
 * volume_shader.hpp:24-31) plus the density scale on the TF alpha
 * (1.0 reproduces the reference, which has no density in Volume mode). */
typedef struct {
  float bbox_min[3];
  float bbox_max[3];
  float step;       /* BasicRenderer::update: du / 4 / sample_rate */
  float du;         /* 1 / block_size */
  int32_t max_steps;
  int32_t mode;     /* HYD_MODE_* */
  float rank;       /* Partition mode colour parameter */
  float density;
  int32_t schedule; /* HYD_SCHED_* */
  float ert;        /* LATTICE early-termination alpha; <= 0 disables */
} hyd_volume_params;

/* VolumePixel film record, 64 bytes (volume_shader.hpp:11-15): the resumable
 * per-ray state.  In LATTICE schedule o is the clipped entry point, nsteps
 * holds the bits of the float ray extent in steps and step_size the next
 * lattice index. */
typedef struct {
  float o[3];
  float d[3];
  int32_t nsteps;
  int32_t step_size;
  float v[4];
  float theta[3];
  float phi;
} hyd_volume_pixel;

/* VolumeFetchPixel, volume_shader.hpp:17-22 (32 bytes) */
typedef struct {
  float theta[3];
  float phi;
  float val[4];
} hyd_fetch_pixel;

/* One record of the per-sample trace: brick cell, page id and the 6 filter
 * tap indices (CpuSampler::filter_linear, cpu_sampler.hpp:88-104). */
typedef struct {
  int32_t ix, iy, iz, pgid;
  int32_t tix, tiy, tiz, tjx, tjy, tjz;
} hyd_trace_rec;

typedef struct {
  float bbox_min[3];
  float bbox_max[3];
  float step;
  int32_t max_steps;
  int32_t mode;     /* HYD_BLOCKS_* */
  float density;
} hyd_blocks_params;

typedef struct {
  float bbox_min[3];
  float bbox_max[3];
  float step;
  float du;
  int32_t max_steps;
  int32_t mode;          /* HYD_ISO_* */
  float to_world[16];    /* inverse(iet), column-major (isosurface.cc:108) */
  float eye_pos[3];
  float light_pos[3];
  float surface_color[3];
  float isovalue;
  int32_t corrected;     /* 0: reference crossing test; 1: bracket + bisection */
  int32_t bisect_steps;  /* bisection iterations when corrected (8) */
} hyd_iso_params;

/* Isosurface output per pixel: rgba, depth (block units, +inf on miss), hit */
typedef struct {
  float v[4];
  float depth;
  int32_t hit;
} hyd_iso_pixel;

/* IsosurfaceFetchPixel, isosurface_shader.hpp:17-21 (8 bytes: uchar3, padding,
 * float depth). */
typedef struct {
  uint8_t val[3];
  uint8_t pad;
  float depth;
} hyd_iso_fetch_pixel;

/* ---------------------------------------------------------------- core -- */
const char *hyd_last_error(void);
int32_t hyd_version(void);

/* IRenderer device selection (basic_renderer.hpp:43-55): one context per device. */
int32_t hyd_create(int32_t device, hyd_ctx **out);
int32_t hyd_destroy(hyd_ctx *ctx);
/* Issue all work on this HIP stream (a hipStream_t; NULL selects the device's
 * null stream).  A new context uses a private non-blocking stream. */
int32_t hyd_set_stream(hyd_ctx *ctx, void *stream);
int32_t hyd_synchronize(hyd_ctx *ctx);

/* ------------------------------------------------------ camera (host) -- */
/* Camera::update_params(CameraOrbit), scene.hpp:68-81 (+ perspective fovy). */
int32_t hyd_orbit_camera(const float arm[3], const float center[3], float fovy, hyd_camera *out);
/* Raycaster::fill_ray_emit_args, raycaster.hpp:126-135, with Exhibit::get_iet
 * (scene.hpp:20-29) for an exhibit of the given centre and size. */
/* A camera whose position equals its target, whose up is parallel to the view
 * direction, or whose fovy lies outside (0, 180) is refused with HYD_ERR_ARG
 * (the reference renders NaN rays for it). */
int32_t hyd_make_view(const hyd_camera *cam, const float ex_center[3], const float ex_size[3], hyd_view *out);
/* Exhibit::get_iet and glm inverse (for the isosurface to_world). */
int32_t hyd_exhibit_iet(const float ex_center[3], const float ex_size[3], float out[16]);
int32_t hyd_mat4_inverse(const float m[16], float out[16]);

/* --------------------------------------------------- resident state ---- */
/* Brick pool: n_pages padded bricks of pad_edge^3 voxels (2 <= pad_edge <= 1024),
 * x fastest, brick after brick (LosslessBlockPagingServer block_storage,
 * src/slave/paging/lossless_block_paging.cc:15-83).  Copied to the device. */
int32_t hyd_upload_pool(hyd_ctx *ctx, int32_t fmt, int32_t pad_edge, int32_t n_pages, const void *host);
/* Same, but the pool already lives on the device; the context keeps the pointer
 * (not a copy), which must stay valid while the context renders from it.
 * The batched renderer reads the pool through oriented copies that it derives
 * at its first launch after an upload or bind: after changing a bound pool's
 * contents in place, call this again before the next batched launch.  The
 * copies take 12x the pool's bytes for FP16 and u8 pools (three quad
 * layouts), or 6x (three row-pair layouts) when the device cannot hold those,
 * else 2x (y- and z-fastest layouts); with less free memory the renderer
 * reads the pool alone.  Every choice renders the same image. */
int32_t hyd_bind_pool_device(hyd_ctx *ctx, int32_t fmt, int32_t pad_edge, int32_t n_pages, const void *pool_dev);
/* Overwrite pages [first_page, first_page + n_pages) of the current pool with
 * n_pages host bricks of its format and edge: the reference's block_storage
 * upload when a decoded brick takes a slot (rt_block_paging.cc:277-280).
 * Synchronous.  The batched renderer's oriented copies of those pages are
 * updated in place, so captured batch graphs stay valid. */
int32_t hyd_write_pool_pages(hyd_ctx *ctx, int32_t first_page, int32_t n_pages, const void *host);
/* Page table (BlockPaging, block_paging.hpp:38-44): vaddr grid of dims[0]*dims[1]*dims[2]
 * int32 (x fastest, -1 = not resident) and per-page BlockSamplerMapping {k, b[3]}
 * (block_paging.hpp:11-22). */
int32_t hyd_set_page_table(hyd_ctx *ctx, const int32_t dims[3], const int32_t *vaddr, int32_t n_pages,
                           const float *page_k, const float *page_b, int32_t lowest_blkcnt);
/* Empty-space grid (chebyshev thumbnail, volume.cc:47-51): 0 = non-empty brick,
 * k > 0 = L-inf brick distance to the nearest non-empty brick. */
int32_t hyd_set_empty_grid(hyd_ctx *ctx, const int32_t dims[3], const int32_t *chebyshev);
/* Per-brick mean thumbnail for Blocks.Volume (blocks.cc:44-50). */
int32_t hyd_set_mean_grid(hyd_ctx *ctx, const int32_t dims[3], const float *mean);
/* Transfer function table: n RGBA float entries (TransferFn, transfer_fn.hpp:14-28). */
int32_t hyd_set_tf(hyd_ctx *ctx, int32_t n, const float *rgba);

/* ----------------------------------------------------------- passes ---- */
/* Ray-emit pass with the march fused (ray_emit_kernel_impl + VolumeShaderKernel::main,
 * cuda_shading.cu:14-37, volume_shader.cu:32-81).  film_dev (hyd_volume_pixel
 * per pixel) and rgba_dev (float4 per pixel) may each be NULL.  samples_dev, if
 * not NULL, is a device uint64 to which the number of composited samples is
 * added (slower counting variant). */
int32_t hyd_render_volume(hyd_ctx *ctx, const hyd_volume_params *p, const hyd_view *view, int32_t width,
                          int32_t height, hyd_volume_pixel *film_dev, float *rgba_dev,
                          unsigned long long *samples_dev);
/* Ray-march pass: resume every film pixel after the page table changed
 * (ray_march_kernel_impl, cuda_shading.cu:41-56; shader.hpp:146-157). */
int32_t hyd_resume_volume(hyd_ctx *ctx, const hyd_volume_params *p, int32_t width, int32_t height,
                          hyd_volume_pixel *film_dev, float *rgba_dev, unsigned long long *samples_dev);
/* Pixel pass (StdVec4Pixel::write_to, pixel_template.hpp:11-20): RGBA float
 * -> RGB8, clear colour where alpha == 0. */
int32_t hyd_pixel_pass(hyd_ctx *ctx, const float *rgba_dev, int32_t width, int32_t height,
                       const uint8_t clear_color[3], uint8_t *rgb_dev);
/* Fetch pass (VolumeShaderKernel::fetch, volume_shader.cu:23-30). */
int32_t hyd_fetch_pass(hyd_ctx *ctx, const hyd_volume_pixel *film_dev, int32_t width, int32_t height,
                       hyd_fetch_pixel *out_dev);
/* In-GPU sort-last composite of n partial images (hyd_fetch_pixel, n images
 * of width*height back to back) in z_order (host array of n indices), then
 * sRGB + saturate (volume.cc:267-292).  rgba_dev / rgb_dev may be NULL. */
int32_t hyd_composite(hyd_ctx *ctx, int32_t n, const hyd_fetch_pixel *partials_dev, const int32_t *z_order,
                      int32_t width, int32_t height, float *rgba_dev, uint8_t *rgb_dev);
/* hyd_composite without any device allocation (no counterpart in the reference):
 * for callers that cannot allocate, such as code under HIP graph capture or a
 * serving loop that composites every frame.  The z order is staged in a fixed
 * device array of the library (one per device, resolved by hyd_create); n is at
 * most 4096 (HYD_ERR_RANGE beyond).  Every call on one device uses that same
 * array, so calls that overlap on a device (other streams, other contexts,
 * other threads) race.  Same results as hyd_composite.  Synchronous. */
int32_t hyd_composite_noalloc(hyd_ctx *ctx, int32_t n, const hyd_fetch_pixel *partials_dev, const int32_t *z_order,
                              int32_t width, int32_t height, float *rgba_dev, uint8_t *rgb_dev);
/* The bytes hyd_composite writes: t[0] = -inf and, for k = 1..255, t[k] = the
 * least float whose saturate(linear_to_srgb(x)) byte (volume.cc:175-181,284-288,
 * binary32 on the host) is k or more; the composite writes the number of t[k]
 * at or below each channel.  Host only. */
int32_t hyd_srgb8_thresholds(float out[256]);
/* Isosurface fetch pass (IsosurfaceShaderKernel::fetch, isosurface_shader.cu:42-49):
 * bytes saturate(v.rgb) and the depth of every pixel. */
int32_t hyd_iso_fetch_pass(hyd_ctx *ctx, const hyd_iso_pixel *px_dev, int32_t width, int32_t height,
                           hyd_iso_fetch_pixel *out_dev);
/* In-GPU isosurface sort-last composite (isosurface.cc:272-289) of n partial
 * images (n images of width*height back to back) in z_order (host array of n
 * indices): starting from image z_order[0], each later image replaces a pixel
 * whose depth is strictly greater.  rgb_dev (3 bytes per pixel) and depth_dev
 * (float per pixel) may each be NULL, not both.  Synchronous. */
int32_t hyd_composite_iso(hyd_ctx *ctx, int32_t n, const hyd_iso_fetch_pixel *partials_dev, const int32_t *z_order,
                          int32_t width, int32_t height, uint8_t *rgb_dev, float *depth_dev);
/* Sort-last rendering over residency batches.  The reference renders a volume
 * larger than the pool in batches by pausing rays at non-resident bricks and
 * resuming them (volume.cc:95-134), which needs the bricks in ray order; here
 * each ray's per-brick partial (premultiplied RGBA of the samples inside one
 * brick, LATTICE schedule, no early termination) is written to slot k of the
 * ray for its k-th visited brick, in any brick order, and the partials are
 * composited front to back on the GPU (the sort-last "over" of volume.cc:267-292
 * applied per ray).  slots_dev holds max_slots * width * height float4
 * (slot-major); max_slots >= sum of the brick grid dims.  nslots_dev holds one
 * int32 per pixel. */
int32_t hyd_segments_begin(hyd_ctx *ctx, int32_t width, int32_t height, int32_t max_slots, float *slots_dev);
/* One pass with the context's current pool and page table: fills the slots of
 * every visited brick resident now.  Repeat for every residency batch. */
int32_t hyd_segments_pass(hyd_ctx *ctx, const hyd_volume_params *p, const hyd_view *view, int32_t width,
                          int32_t height, int32_t max_slots, float *slots_dev, int32_t *nslots_dev,
                          unsigned long long *samples_dev);
/* Residency batches kept in device memory together (no counterpart in the
 * reference, whose batches are rendered one after another, or by different
 * ranks): a set renders every batch's segments in one launch.  A batch is a
 * device pool (hyd_bind_pool_device's layout) and its page table
 * (hyd_set_page_table's arguments, host memory, over the brick grid of the
 * context's empty-space grid); every batch has the same format and page edge.
 * The set copies the page tables and the empty-space grid at creation (a later
 * hyd_set_empty_grid does not change it) and holds the pools' pointers, which
 * must stay valid while the set is used. */
typedef struct {
  int32_t fmt;            /* HYD_FMT_* of the pool */
  int32_t pad_edge;       /* page edge in voxels */
  int32_t n_pages;
  const void *pool_dev;
  const int32_t *vaddr;   /* host, empty-grid dims: page id or -1 */
  const float *page_k;    /* host, n_pages */
  const float *page_b;    /* host, 3 * n_pages */
} hyd_segment_batch;
typedef struct hyd_segment_set hyd_segment_set;
/* n_batches in [1, 16]. */
int32_t hyd_segment_set_create(hyd_ctx *ctx, int32_t n_batches, const hyd_segment_batch *batches,
                               hyd_segment_set **out);
int32_t hyd_segment_set_destroy(hyd_segment_set *set);
/* hyd_segments_pass for every batch of the set, in one launch (the batches'
 * bricks are disjoint, so their slots are too): the same slots as one pass per
 * batch with that batch bound. */
int32_t hyd_segments_pass_set(hyd_ctx *ctx, const hyd_segment_set *set, const hyd_volume_params *p,
                              const hyd_view *view, int32_t width, int32_t height, int32_t max_slots,
                              float *slots_dev, int32_t *nslots_dev, unsigned long long *samples_dev);
/* Fold the slots in ray order into RGBA; a ray stops at its first slot that no
 * pass filled (a brick never resident), as the reference stops at a page fault. */
int32_t hyd_segments_composite(hyd_ctx *ctx, int32_t width, int32_t height, int32_t max_slots,
                               const float *slots_dev, const int32_t *nslots_dev, float *rgba_dev);
/* Per-sample trace of one pixel (debug; synchronous; FP32 reference schedule):
 * writes up to max_recs records to host memory, *n_out = count. */
int32_t hyd_trace_volume(hyd_ctx *ctx, const hyd_volume_params *p, const hyd_view *view, int32_t width,
                         int32_t height, int32_t x, int32_t y, hyd_trace_rec *out, int32_t max_recs,
                         int32_t *n_out);

/* Blocks.Volume / Blocks.Solid (VisBlocksShaderKernel::main, blocks_shader.cu:16-50) and the
 * residency view (PagingShaderKernel::main, paging_shader.cu:24-61: the page table of
 * hyd_set_page_table over the empty-space grid; density is not used). */
int32_t hyd_render_blocks(hyd_ctx *ctx, const hyd_blocks_params *p, const hyd_view *view, int32_t width,
                          int32_t height, float *rgba_dev);
/* Isosurface (IsosurfaceShaderKernel::main, isosurface_shader.cu:50-125). */
int32_t hyd_render_iso(hyd_ctx *ctx, const hyd_iso_params *p, const hyd_view *view, int32_t width,
                       int32_t height, hyd_iso_pixel *out_dev, unsigned long long *samples_dev);

/* --------------------------------------------------- batched form ------ */
/* The reference renders one view per offline_render call (volume.cc:95-134),
 * re-launching per pass.  The batched form renders many views of several
 * volumes in ONE persistent launch over the shared brick pool, with the
 * LATTICE schedule (no step escalation; FP32 accumulation over u8/f16/f32
 * bricks).  It has no counterpart in the reference. */

/* One volume of the set.  Page ids in vaddr index the shared pool. */
typedef struct {
  int32_t grid[3];           /* brick grid (x, y, z) */
  int32_t bs;                /* block size (brick edge without apron) */
  int32_t apron;             /* apron voxels per side, >= 1 (bs + 2*apron = pool pad edge) */
  float bbox_max[3];         /* raw dims / bs, block units (bbox_min = 0) */
  float step;                /* lattice step, block units */
  float value_scale;         /* multiplies the interpolated brick value: 1/255 for bricks that hold
                                8-bit counts (u8, or f16 packed with normalize = 0), else 1 */
  const int32_t *vaddr;      /* host: grid cells, page id or -1 */
  const int32_t *chebyshev;  /* host: grid cells, 0 = non-empty */
} hyd_volume_desc;

typedef struct {
  hyd_view view;
  int32_t width;
  int32_t height;
  int32_t volume;            /* index into the volume set */
  int32_t reserved;
  int64_t out_offset;        /* first pixel of this view in the RGBA output */
} hyd_batch_view;

/* Launch shape (results do not depend on it). */
typedef struct {
  int32_t tile_w;            /* ray-packet super-tile: 8, 16 or 32 pixels wide, 8 high */
  int32_t unroll;            /* samples resolved per inner iteration: 1, 2 or 4 */
  int32_t waves_per_cu;      /* persistent waves per CU: 1..32 */
  int32_t group;             /* waves per workgroup: 0 or 1, or 4 sharing a 16x16 block of
                              * 8x8 tiles on one CU (tile_w 8 only) */
} hyd_batch_config;

typedef struct hyd_batch hyd_batch;

/* Define the volume set (copied to the device; cells are re-encoded). */
int32_t hyd_set_volumes(hyd_ctx *ctx, int32_t n, const hyd_volume_desc *vols);
/* Build a plan for n_views views (<= 65535, <= 256 tiles per axis). */
int32_t hyd_batch_create(hyd_ctx *ctx, const hyd_batch_view *views, int32_t n_views, hyd_batch **out);
int32_t hyd_batch_destroy(hyd_batch *batch);
/* Pixels the output buffer must hold (max over views of out_offset + w*h). */
int32_t hyd_batch_output_pixels(const hyd_batch *batch, int64_t *n_out);
/* One persistent launch (a counter-reset kernel + the march); RGBA float out.
 * samples_dev, if not NULL, accumulates the composited-sample count.  The pool
 * must be at most 4 GiB with a padded brick edge of at most 256 (voxels are
 * addressed by 32-bit byte offsets): HYD_ERR_RANGE otherwise. */
int32_t hyd_batch_render(hyd_ctx *ctx, hyd_batch *batch, const hyd_batch_config *cfg, float density, float ert,
                         float *rgba_dev, unsigned long long *samples_dev);
/* Record that launch into a HIP graph owned by the batch (no memset nodes). */
int32_t hyd_batch_capture(hyd_ctx *ctx, hyd_batch *batch, const hyd_batch_config *cfg, float density, float ert,
                          float *rgba_dev);
/* Replay the captured graph on the context's stream.  The graph holds the
 * device pointers of the pool, transfer function and volume set it was
 * captured with: after hyd_upload_pool, hyd_bind_pool_device, hyd_set_tf or
 * hyd_set_volumes, replay returns HYD_ERR_STATE until the batch is captured
 * again. */
int32_t hyd_batch_replay(hyd_ctx *ctx, hyd_batch *batch);
/* Time every launch shape (tile_w x unroll x waves_per_cu) with `reps`
 * launches each and return the fastest; writes the output buffer. */
int32_t hyd_batch_autotune(hyd_ctx *ctx, hyd_batch *batch, float density, float ert, float *rgba_dev, int32_t reps,
                           hyd_batch_config *best, float *best_ms);
/* Also tune the step schedule: the early-termination threshold, over
 * {0.99, 0.995, 0.999, 0.9995, 0.9999, off}.  Terminating once alpha exceeds ert
 * drops at most 1 - ert of any channel, so only thresholds with
 * 1 - ert <= max_error / 2 are timed (the other half of the budget is left to
 * storage and accumulation rounding); returns the fastest launch shape and
 * threshold (0 = no termination). */
int32_t hyd_batch_autotune_schedule(hyd_ctx *ctx, hyd_batch *batch, float density, float max_error, float *rgba_dev,
                                    int32_t reps, hyd_batch_config *best, float *best_ert, float *best_ms);

/* -------------------------------------------- differentiable form ----- */
/* A differentiable variant of the Volume march (no counterpart in the
 * reference): one FP32 brick of dims voxels spanning the unit box, n_steps
 * samples per ray at the midpoints of equal sub-intervals of the clipped
 * chord, the reference's trilinear and TF lookups, opacity
 * alpha = 1 - exp(-sigma * max(a, 0) * dt), front-to-back compositing.  The
 * forward keeps transmittance checkpoints every ckpt_every samples; the adjoint
 * recomputes each segment from them and returns dL/dTF and dL/dvolume. */
typedef struct {
  int32_t dims[3];        /* voxels (x, y, z); x in [2, 2^22) */
  int32_t n_steps;        /* samples per ray (512) */
  int32_t ckpt_every;     /* checkpoint interval (64) */
  float sigma;            /* extinction scale */
} hyd_diff_params;

/* Floats per ray of the checkpoint buffer (ceil(n_steps / ckpt_every) + 1). */
int32_t hyd_diff_ckpt_floats(const hyd_diff_params *p, int32_t *per_ray);
/* Forward: rgba_dev gets n_views * height * width float4, ckpt_dev
 * n_views * height * width * hyd_diff_ckpt_floats floats.  volume_dev and
 * tf_dev (tf_n RGBA entries, <= 256) are device FP32 arrays. */
int32_t hyd_diff_forward(hyd_ctx *ctx, const hyd_diff_params *p, const hyd_view *views, int32_t n_views,
                         int32_t width, int32_t height, const float *volume_dev, const float *tf_dev, int32_t tf_n,
                         float *rgba_dev, float *ckpt_dev, unsigned long long *samples_dev);
/* Adjoint: writes dL/dvolume to grad_volume_dev and dL/dTF to grad_tf_dev
 * (both overwritten) given dL/dRGBA (grad_rgba_dev).  The context keeps a
 * scratch buffer of n_views * height * width * (n_steps + 8) floats (dL/ds of
 * every sample, ray setups), grown on demand; n_views * height * width must be
 * < 2^31.  Both gradients are deterministic (integer accumulation). */
int32_t hyd_diff_backward(hyd_ctx *ctx, const hyd_diff_params *p, const hyd_view *views, int32_t n_views,
                          int32_t width, int32_t height, const float *volume_dev, const float *tf_dev, int32_t tf_n,
                          const float *ckpt_dev, const float *grad_rgba_dev, float *grad_volume_dev,
                          float *grad_tf_dev);

#ifdef __cplusplus
}
#endif

#endif /* HYDRANT_HIP_H */
