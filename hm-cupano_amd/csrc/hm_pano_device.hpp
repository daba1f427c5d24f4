// hm_pano_device.hpp -- per-pixel remap helpers shared by the pano kernels.
#pragma once

#include "hm_device.hpp"
#include "hm_stitch.h"

namespace hmdev {

// Remap launches: 64 x 4 threads, each thread kPx pixels of one canvas row,
// kBlock.x apart, so a wavefront's loads stay coalesced while every lane
// keeps kPx independent map loads (then gathers) in flight per image.
inline const dim3 kBlock(64, 4, 1);
constexpr int kPx = 4;

template <int C>
__device__ __forceinline__ uint32_t load_src_px(const Surf& s, int b, int x, int y) {
  const uint8_t* p = row_ptr(s, b, y) + (size_t)x * C;
  if constexpr (C == 4) return *reinterpret_cast<const uint32_t*>(p);
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16);
}

template <int C>
__device__ __forceinline__ void store_px(uint8_t* p, uint32_t v) {
  if constexpr (C == 4) {
    *reinterpret_cast<uint32_t*>(p) = v;
  } else {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
  }
}

// BatchedRemapKernelExOffset pixel (src/cuda/cudaRemap.cu:95-122), default
// pixel all-zero, no_unmapped_write = false: in range -> source pixel, else
// zero with alpha 0 for the 65535 sentinel and 255 otherwise.
template <int C>
__device__ __forceinline__ uint32_t remap_px(const Surf& s, int b, uint32_t m) {
  const uint32_t sx = m & 0xffffu, sy = m >> 16;
  if (sx < s.w && sy < s.h) return load_src_px<C>(s, b, (int)sx, (int)sy);
  if constexpr (C == 4) return sx == (uint32_t)kUnmapped ? 0u : 0xff000000u;
  return 0u;
}

// BatchedRemapKernelExOffsetWithDestMap pixel (cudaRemap.cu:235-244):
// out-of-range or unmapped -> the all-zero default pixel.
template <int C>
__device__ __forceinline__ uint32_t remap_px_dm(const Surf& s, int b, uint32_t m) {
  const uint32_t sx = m & 0xffffu, sy = m >> 16;
  if (sx < s.w && sy < s.h) return load_src_px<C>(s, b, (int)sx, (int)sy);
  return 0u;
}

inline Surf to_surf(const hmSurface& s) {
  return Surf{static_cast<uint8_t*>(s.d_ptr), s.width, s.height, s.pitch};
}

}  // namespace hmdev
