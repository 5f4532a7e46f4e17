// Film passes: pixel (RGBA float -> RGB8), fetch (film -> partial image) and
// the in-GPU sort-last composite of partial images.
#include "hyd_internal.hpp"

namespace hyd {
namespace {

constexpr int kBlock = 256;

// StdVec4Pixel::write_to (pixel_template.hpp:11-20): premultiplied colour
// saturated to bytes (clamp(v * 255, 0, 255), truncation), clear colour when
// alpha == 0.
__global__ __launch_bounds__(kBlock) void pixel_kernel(const float4 *rgba, int n, uchar3 clear, uint8_t *rgb) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 v = rgba[i];
  uint8_t r, g, b;
  if (v.w != 0.f) {
    r = (uint8_t)(int)gclamp(v.x * 255.f, 0.f, 255.f);
    g = (uint8_t)(int)gclamp(v.y * 255.f, 0.f, 255.f);
    b = (uint8_t)(int)gclamp(v.z * 255.f, 0.f, 255.f);
  } else {
    r = clear.x; g = clear.y; b = clear.z;
  }
  rgb[3 * i] = r; rgb[3 * i + 1] = g; rgb[3 * i + 2] = b;
}

// The same write_to over pixel records of stride_floats floats whose first four
// are the RGBA value (IsosurfacePixel derives from StdVec4Pixel,
// isosurface_shader.hpp:11-15).
__global__ __launch_bounds__(kBlock) void pixel_strided_kernel(const float *v, int stride, int n, uchar3 clear,
                                                               uint8_t *rgb) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float *p = v + (size_t)i * stride;
  uint8_t r, g, b;
  if (p[3] != 0.f) {
    r = (uint8_t)(int)gclamp(p[0] * 255.f, 0.f, 255.f);
    g = (uint8_t)(int)gclamp(p[1] * 255.f, 0.f, 255.f);
    b = (uint8_t)(int)gclamp(p[2] * 255.f, 0.f, 255.f);
  } else {
    r = clear.x; g = clear.y; b = clear.z;
  }
  rgb[3 * i] = r; rgb[3 * i + 1] = g; rgb[3 * i + 2] = b;
}

// VolumeShaderKernel::fetch (volume_shader.cu:23-30).
__global__ __launch_bounds__(kBlock) void fetch_kernel(const hyd_volume_pixel *film, int n, hyd_fetch_pixel *out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 *f = reinterpret_cast<const float4 *>(film + i);
  float4 *o = reinterpret_cast<float4 *>(out + i);
  o[0] = f[3];   // theta, phi
  o[1] = f[2];   // val
}

// saturate(linear_to_srgb(x)) of volume.cc:175-181,284-288 as a byte: the number
// of thresholds t[1..255] at or below x, where t[k] is the least float whose
// byte is k or more (computed on the host with the same binary32 sRGB the
// reference evaluates there, so the bytes are the host's).  Byte 0 for x below
// t[1] and for NaN.
__device__ __forceinline__ uint8_t srgb8(const float *t, float x) {
  int k = 0;
#pragma unroll
  for (int step = 128; step > 0; step >>= 1)
    if (k + step <= 255 && t[k + step] <= x) k += step;
  return (uint8_t)k;
}

// Front-to-back "over" of n partial images in z order, then sRGB + saturate
// (volume.cc:267-292).  One lane per pixel; partials are image-major.
__global__ __launch_bounds__(kBlock) void composite_kernel(int n, const hyd_fetch_pixel *partials,
                                                           const int32_t *z_order, int P, float4 *rgba,
                                                           uint8_t *rgb, const float *srgb_t) {
  __shared__ float s_t[256];
  s_t[threadIdx.x] = srgb_t[threadIdx.x];
  __syncthreads();
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= P) return;
  const float4 *base = reinterpret_cast<const float4 *>(partials);
  float4 f = base[((size_t)z_order[0] * P + i) * 2 + 1];
  for (int k = 1; k < n; ++k) {
    const size_t j = ((size_t)z_order[k] * P + i) * 2;
    const float4 bt = base[j];      // theta, phi
    const float4 bv = base[j + 1];  // val
    const float vx = f.x + bv.x - f.w * bt.x;
    const float vy = f.y + bv.y - f.w * bt.y;
    const float vz = f.z + bv.z - f.w * bt.z;
    const float an = bt.w * f.w + bv.w;
    f = make_float4(vx, vy, vz, an);
  }
  if (rgba) rgba[i] = f;
  if (rgb) {
    rgb[3 * i] = srgb8(s_t, f.x);
    rgb[3 * i + 1] = srgb8(s_t, f.y);
    rgb[3 * i + 2] = srgb8(s_t, f.z);
  }
}

// IsosurfaceShaderKernel::fetch (isosurface_shader.cu:42-49): the pixel's
// colour saturated to bytes (clamp(v * 255, 0, 255), truncation) and its depth.
__global__ __launch_bounds__(kBlock) void iso_fetch_kernel(const hyd_iso_pixel *px, int n, hyd_iso_fetch_pixel *out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const hyd_iso_pixel p = px[i];
  hyd_iso_fetch_pixel o;
  o.val[0] = (uint8_t)(int)gclamp(p.v[0] * 255.f, 0.f, 255.f);
  o.val[1] = (uint8_t)(int)gclamp(p.v[1] * 255.f, 0.f, 255.f);
  o.val[2] = (uint8_t)(int)gclamp(p.v[2] * 255.f, 0.f, 255.f);
  o.pad = 0;
  o.depth = p.depth;
  out[i] = o;
}

// Isosurface sort-last composite (isosurface.cc:272-289): the partial image of
// z_order[0] is the front; each later one in z order replaces a front pixel
// whose depth is strictly greater.  Writes the bytes and, if asked, the depth.
__global__ __launch_bounds__(kBlock) void iso_composite_kernel(int n, const hyd_iso_fetch_pixel *partials,
                                                               const int32_t *z_order, int P, uint8_t *rgb,
                                                               float *depth) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= P) return;
  hyd_iso_fetch_pixel front = partials[(size_t)z_order[0] * P + i];
  for (int k = 1; k < n; ++k) {
    const hyd_iso_fetch_pixel back = partials[(size_t)z_order[k] * P + i];
    if (back.depth < front.depth) front = back;
  }
  if (rgb) {
    rgb[3 * i] = front.val[0];
    rgb[3 * i + 1] = front.val[1];
    rgb[3 * i + 2] = front.val[2];
  }
  if (depth) depth[i] = front.depth;
}

}  // namespace

hipError_t launch_pixel_pass(const float4 *rgba, int W, int H, uchar3 clear, uint8_t *rgb, hipStream_t s) {
  const int n = W * H;
  hipLaunchKernelGGL(pixel_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, rgba, n, clear, rgb);
  return hipGetLastError();
}

hipError_t launch_pixel_pass_strided(const float *v, int stride_floats, int W, int H, uchar3 clear, uint8_t *rgb,
                                     hipStream_t s) {
  const int n = W * H;
  hipLaunchKernelGGL(pixel_strided_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, v, stride_floats, n,
                     clear, rgb);
  return hipGetLastError();
}

hipError_t launch_fetch_pass(const hyd_volume_pixel *film, int W, int H, hyd_fetch_pixel *out, hipStream_t s) {
  const int n = W * H;
  hipLaunchKernelGGL(fetch_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, film, n, out);
  return hipGetLastError();
}

hipError_t launch_composite(int n, const hyd_fetch_pixel *partials, const int32_t *z_order_dev, int W, int H,
                            float4 *rgba, uint8_t *rgb, const float *srgb_t, hipStream_t s) {
  const int P = W * H;
  hipLaunchKernelGGL(composite_kernel, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, partials,
                     z_order_dev, P, rgba, rgb, srgb_t);
  return hipGetLastError();
}

hipError_t launch_iso_fetch(const hyd_iso_pixel *px, int W, int H, hyd_iso_fetch_pixel *out, hipStream_t s) {
  const int n = W * H;
  hipLaunchKernelGGL(iso_fetch_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, px, n, out);
  return hipGetLastError();
}

hipError_t launch_iso_composite(int n, const hyd_iso_fetch_pixel *partials, const int32_t *z_order_dev, int W, int H,
                                uint8_t *rgb, float *depth, hipStream_t s) {
  const int P = W * H;
  hipLaunchKernelGGL(iso_composite_kernel, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, partials,
                     z_order_dev, P, rgb, depth);
  return hipGetLastError();
}

}  // namespace hyd
