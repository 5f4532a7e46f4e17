/*
 * hydrant_renderer.h - the JSON-configured renderer interface over the C ABI.
 *
 * Mirrors the reference's renderer plugin API (include/hydrant/core/renderer.hpp:30-89):
 *   RendererFactory(dataset).create(RendererConfig)  -> hyd_renderer_create
 *   IRenderer::update(json params)                    -> hyd_renderer_update
 *   IRenderer::offline_render(Camera) -> RGB8 image    -> hyd_renderer_offline_render
 *   RendererFactory::list_candidates()                -> hyd_renderer_candidates
 * with the renderers "Volume", "Blocks", "Isosurface" and "Paging" registered by
 * name as REGISTER_RENDERER does (volume.cc:315, blocks.cc:74, isosurface.cc:313,
 * paging.cc:193).  "Paging" renders nothing offline, as the reference's
 * offline_render_ctxed (paging.cc:51-56): every pixel is the clear colour.
 *
 * Configuration JSON follows the reference schemas:
 *   RendererConfig  {"resolution": [w, h], "renderer": "Volume", "params": {...}}
 *                   (core/renderer.schema.hpp:8-13); a whole config file
 *                   {"camera": ..., "render": RendererConfig} is accepted too.
 *   params          BasicRendererParams (basic_renderer.schema.hpp:10-16):
 *                   device, sample_rate (1), max_steps (4000000), clear_color;
 *                   VolumeRendererParams (ui/volume.schema.hpp:16-21): mode
 *                   Default|Partition|Paging ("Volume" = Default, as the README
 *                   writes it), transfer_fn {values: flat or nested RGBA},
 *                   mem_limit_mb (2048), and density (opacity scale, 1 = the
 *                   reference, which has none);
 *                   BlocksRendererParams (ui/blocks.schema.hpp:14-18): mode
 *                   Volume|Solid, density (1e-2);
 *                   IsosurfaceRendererParams (ui/isosurface.schema.hpp:16-22): mode
 *                   Color|Normal|Position, surface_color, isovalue (0.5),
 *                   mem_limit_mb; plus "corrected" (0 = the reference's crossing
 *                   test, 1 = bracket + "bisect_steps" bisections);
 *                   PagingRendererConfig (paging.schema.hpp:9-13): mem_limit_mb.
 *   camera          CameraConfig (core/scene.schema.hpp:10-35): "orbit" {arm,
 *                   center} or "ptu" {position, target, up}, "perspective" {fovy}.
 * device: "Cuda" or "Hip" render on this context's GPU; "Cpu" (the reference's
 * multi-thread backend) is refused - this library has no CPU path.
 *
 * Errors: every call returns HYD_OK (0) or an HYD_ERR_* code and sets
 * hyd_last_error() (the reference aborts with LOG(FATAL) instead).
 */
#ifndef HYDRANT_RENDERER_H
#define HYDRANT_RENDERER_H

#include <stdint.h>

#include "hydrant_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hyd_renderer hyd_renderer;

/* The dataset: raw voxels in host memory, x fastest (the reference reads a
 * varch package of decoded bricks, renderer.hpp:17-21 and
 * lossless_block_paging.cc:137-174; decode is outside this library).  The
 * renderer copies what it needs during create; the pointer need not outlive
 * the call. */
typedef struct {
  int32_t dims[3];       /* voxels (x, y, z) */
  int32_t fmt;           /* HYD_FMT_U8 (uint8), HYD_FMT_F16 (half bits), HYD_FMT_F32 */
  const void *voxels;
  int32_t block_size;    /* brick edge (the package's block size, e.g. 64) */
  int32_t apron;         /* padding voxels per side (e.g. 1) */
} hyd_dataset;

/* A dataset package on disk (the reference's RendererFactory(dataset_path),
 * core/renderer.hpp:53-59, reading package_meta.json and the level-0
 * thumbnails its renderers load, volume.cc:47-51, blocks.cc:35-44,
 * isosurface.cc:50-53):
 *   package_meta.json  {"block_size": bs, "padding": p, "sample_levels": [{
 *                        "raw": {x, y, z}, "dim": {x, y, z},
 *                        "thumbnails": {"chebyshev": file, "mean": file},
 *                        "voxels": {"path": file, "format": "u8"|"f16"|"f32"}}]}
 *   voxels             raw level-0 voxels, x fastest (the reference's package
 *                      holds H.264 brick archives that it decodes; decode is
 *                      outside this library)
 *   chebyshev          dim.x*dim.y*dim.z little-endian int32, x fastest: 0 for a
 *                      brick with data, else the L-inf distance in bricks to the
 *                      nearest such brick
 *   mean               the same count of little-endian float32 (Blocks mode)
 * File names resolve against the package directory.  dim must equal
 * ceil(raw / block_size); every file must hold exactly the bytes named. */
typedef struct {
  int32_t raw[3];        /* voxels (x, y, z) */
  int32_t dim[3];        /* bricks (x, y, z) */
  int32_t block_size;
  int32_t padding;
  int32_t fmt;           /* HYD_FMT_* of the voxel file */
  int32_t has_mean;      /* 1 if the package has a "mean" thumbnail */
} hyd_package_info;

/* Read and validate package_meta.json and the thumbnails (not the voxels). */
int32_t hyd_package_read_info(const char *dataset_path, hyd_package_info *out);
/* Copy the chebyshev (and, when mean is not NULL, mean) thumbnails; n must be
 * dim.x*dim.y*dim.z. */
int32_t hyd_package_read_thumbnails(const char *dataset_path, int32_t *chebyshev, float *mean, int64_t n);
/* RendererFactory(dataset_path).create + IRenderer::init: Volume, Blocks and
 * Isosurface skip the bricks the chebyshev thumbnail marks empty (Blocks shades
 * with the mean thumbnail), as the reference's renderers do, instead of deriving
 * emptiness from the transfer function or iso value. */
int32_t hyd_renderer_create_package(hyd_ctx *ctx, const char *dataset_path, const char *config_json,
                                    hyd_renderer **out);

/* Registered renderer names, comma separated, into buf (cap bytes, NUL
 * terminated); *len receives the full length. */
int32_t hyd_renderer_candidates(char *buf, int32_t cap, int32_t *len);
/* RendererFactory::create + IRenderer::init. */
int32_t hyd_renderer_create(hyd_ctx *ctx, const hyd_dataset *dataset, const char *config_json, hyd_renderer **out);
/* IRenderer::update with a params object. */
int32_t hyd_renderer_update(hyd_renderer *r, const char *params_json);
int32_t hyd_renderer_resolution(const hyd_renderer *r, int32_t *width, int32_t *height);
/* IRenderer::offline_render: RGB8, width*height*3 bytes into host memory (row 0 first). */
int32_t hyd_renderer_offline_render(hyd_renderer *r, const char *camera_json, uint8_t *rgb_host);
int32_t hyd_renderer_destroy(hyd_renderer *r);

/* Parse a renderer configuration without creating a renderer (no GPU use): the
 * renderer name into name (cap bytes, NUL terminated) and the resolution;
 * fails on malformed JSON or a missing / unregistered renderer name. */
int32_t hyd_renderer_parse_config(const char *config_json, char *name, int32_t cap, int32_t *width,
                                  int32_t *height);
/* Parse a CameraConfig (or a whole config holding "camera") into a hyd_camera
 * (no GPU use): orbit via Camera::update_params (scene.hpp:68-81) or ptu. */
int32_t hyd_renderer_parse_camera(const char *camera_json, hyd_camera *out);

#ifdef __cplusplus
}
#endif

#endif /* HYDRANT_RENDERER_H */
