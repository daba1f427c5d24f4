// hm_device.hpp -- device-side building blocks for the gfx950 stitch kernels.
//
// Pixel arithmetic follows the hm-cupano CUDA sources exactly (order of
// operations, casts, alpha rules); the whole library is compiled with
// -ffp-contract=off so no multiply-add is fused.  Index arithmetic is 64-bit
// throughout (the reference's 32-bit products overflow on batched canvases).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hmdev {

constexpr int kUnmapped = 65535;  // src/cuda/cudaRemap.cu:13

// Pitched, batch-major image: base + (b*h + y)*pitch + x*sizeof(px)
// (src/cuda/cudaTypes.h:195-230, with 64-bit offsets).
struct Surf {
  uint8_t* ptr;
  uint32_t w, h, pitch;
};

__device__ __forceinline__ uint8_t* row_ptr(const Surf& s, int b, int y) {
  return s.ptr + ((size_t)b * s.h + (size_t)y) * (size_t)s.pitch;
}

// ---- the workgroup a kernel body runs as ----------------------------------
// Kernel bodies take the workgroup's index (x: tile, z: batch entry) rather
// than reading blockIdx, so that the same body runs as its own launch or as a
// work item of a batched pass (hm_batch.hip).  Every body runs on 256 threads;
// flat_tid() is a thread's index among them for either launch shape (256 x 1
// or 64 x 4).
struct Blk {
  int x, z;
};
__device__ __forceinline__ Blk hw_blk() { return Blk{(int)blockIdx.x, (int)blockIdx.z}; }

// A kernel's single by-value argument, read in place in the kernel-argument
// segment (constant memory: scalar loads).  Handing the parameter itself to a
// body by reference can make the compiler copy all of it to scratch memory
// once the body indexes its arrays with run-time (uniform) indices.
template <typename A>
__device__ __forceinline__ const A& kernel_arg() {
  return *(const A*)(const __attribute__((address_space(4))) A*)__builtin_amdgcn_kernarg_segment_ptr();
}
constexpr int kWgThreads = 256;
__device__ __forceinline__ int flat_tid() { return (int)(threadIdx.x + blockDim.x * threadIdx.y); }

// A uniform value pinned in a scalar register.  A per-lane choice between two
// pinned values stays a select; between two argument fields read in place, the
// compiler turns it into a per-lane load of the chosen field.  Pointers come
// back as global-memory pointers (global, not flat, accesses).
__device__ __forceinline__ int pin_sgpr(int v) {
  asm volatile("" : "+s"(v));
  return v;
}
__device__ __forceinline__ uint32_t pin_sgpr(uint32_t v) {
  asm volatile("" : "+s"(v));
  return v;
}
template <typename T>
__device__ __forceinline__ T* pin_sgpr(T* p) {
  uint64_t v = (uint64_t)p;
  asm volatile("" : "+s"(v));
  return (T*)(__attribute__((address_space(1))) T*)v;
}
__device__ __forceinline__ Surf pin_sgpr(const Surf& s) {
  return Surf{pin_sgpr(s.ptr), pin_sgpr(s.w), pin_sgpr(s.h), pin_sgpr(s.pitch)};
}

// ---- XCD-aware tile order ------------------------------------------------
// MI355X dispatches consecutive workgroup ids round-robin over its 8 XCDs,
// each with its own L2.  Renumbering so that XCD k runs the k-th contiguous
// eighth of the row-major tile list keeps vertically adjacent tiles (which
// share source lines and pyramid taps) in one L2.  A bijection on [0, n).
constexpr int kNumXcd = 8;
constexpr bool kXcdSwizzle = false;  // measured: no gain for these kernels
__device__ __forceinline__ int xcd_tile(int bid, int n) {
  if (!kXcdSwizzle) return bid;
  const int per = n / kNumXcd, rem = n % kNumXcd;
  const int xcd = bid % kNumXcd, idx = bid / kNumXcd;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Pixel of this thread in a 2-D tiling launched as a 1-D grid of
// (tiles_x * tiles_y) workgroups of blockDim.x * blockDim.y threads.
struct TileXY {
  int x, y;
};
__device__ __forceinline__ TileXY tile_pixel(int tiles_x, int tiles_y) {
  const int t = xcd_tile((int)blockIdx.x, tiles_x * tiles_y);
  const int tx = t % tiles_x, ty = t / tiles_x;
  return TileXY{tx * (int)blockDim.x + (int)threadIdx.x, ty * (int)blockDim.y + (int)threadIdx.y};
}

// ---- pixel load/store (C = 3 or 4), values carried in a float4 ----------
template <int C>
__device__ __forceinline__ float4 ld_px(const float* base, size_t pix) {
  if constexpr (C == 4) {
    return *reinterpret_cast<const float4*>(base + pix * 4);
  } else {
    const float* p = base + pix * 3;
    return make_float4(p[0], p[1], p[2], 0.0f);
  }
}
template <int C>
__device__ __forceinline__ float4 ld_px(const uint8_t* base, size_t pix) {
  if constexpr (C == 4) {
    const uchar4 u = *reinterpret_cast<const uchar4*>(base + pix * 4);
    return make_float4((float)u.x, (float)u.y, (float)u.z, (float)u.w);
  } else {
    const uint8_t* p = base + pix * 3;
    return make_float4((float)p[0], (float)p[1], (float)p[2], 0.0f);
  }
}
template <int C>
__device__ __forceinline__ void st_px(float* base, size_t pix, float4 v) {
  if constexpr (C == 4) {
    *reinterpret_cast<float4*>(base + pix * 4) = v;
  } else {
    float* p = base + pix * 3;
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
  }
}

// round_to_uchar: clamp to [0,255] then truncate (src/cuda/cudaUtils.cuh:18-28).
__device__ __forceinline__ uint8_t to_u8(float x) {
  if (x <= 0.0f) return 0;
  if (x >= 255.0f) return 255;
  return (uint8_t)x;
}

// float -> uint8 pixel store with perform_cast (cudaUtils.cuh:118-127, 190-193).
template <int C>
__device__ __forceinline__ void st_px_u8(uint8_t* p, float4 v) {
  if constexpr (C == 4) {
    *reinterpret_cast<uchar4*>(p) = make_uchar4(to_u8(v.x), to_u8(v.y), to_u8(v.z), to_u8(v.w));
  } else {
    p[0] = to_u8(v.x);
    p[1] = to_u8(v.y);
    p[2] = to_u8(v.z);
  }
}

// ---- the top-left aligned upsample taps (src/cuda/cudaBlend.cu:383-390,
// :613-623).  For pyramid levels built with (w+1)/2 the reconstruct kernel's
// clamp of the base tap never changes it, so one form serves both.
struct Taps {
  int x0, x1, y0, y1;
  float dx, dy;
};
__device__ __forceinline__ Taps make_taps(int x, int y, int lw, int lh) {
  Taps t;
  const float gx = (float)x / 2.0f, gy = (float)y / 2.0f;
  int gxi = (int)floorf(gx), gyi = (int)floorf(gy);
  gxi = max(0, min(gxi, lw - 1));
  gyi = max(0, min(gyi, lh - 1));
  t.x0 = gxi;
  t.y0 = gyi;
  t.x1 = min(gxi + 1, lw - 1);
  t.y1 = min(gyi + 1, lh - 1);
  t.dx = gx - (float)gxi;
  t.dy = gy - (float)gyi;
  return t;
}

// Correctly rounded x / d for the divisors of the pyramid arithmetic.  A
// power-of-two divisor (sum of bilinear weights 0.25, 0.5 or 1; a 2x2 tap
// count of 1, 2 or 4) divides exactly by multiplying with its exact
// reciprocal, so only the other divisors (0.75, 3) take the IEEE division;
// the result is bit-identical to x / d either way.
__device__ __forceinline__ float div_exact(float x, float d) {
  if (d == 1.0f) return x;
  if (d == 0.5f) return x * 2.0f;
  if (d == 0.25f) return x * 4.0f;
  if (d == 2.0f) return x * 0.5f;
  if (d == 4.0f) return x * 0.25f;
  return x / d;
}

// Alpha-aware bilinear (cudaBlend.cu:416-444, :649-666): taps 00,10,01,11;
// with C == 4 a tap whose alpha is 0 is skipped; value = sumV/sumW or 0.
template <int C>
__device__ __forceinline__ float4 up_alpha(const float* low, size_t row0, size_t row1, const Taps& t) {
  const float4 p00 = ld_px<C>(low, row0 + t.x0);
  const float4 p10 = ld_px<C>(low, row0 + t.x1);
  const float4 p01 = ld_px<C>(low, row1 + t.x0);
  const float4 p11 = ld_px<C>(low, row1 + t.x1);
  const float w00 = (1.0f - t.dx) * (1.0f - t.dy);
  const float w10 = t.dx * (1.0f - t.dy);
  const float w01 = (1.0f - t.dx) * t.dy;
  const float w11 = t.dx * t.dy;
  float sw = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
  if (C != 4 || p00.w != 0.0f) { sw += w00; sx += p00.x * w00; sy += p00.y * w00; sz += p00.z * w00; }
  if (C != 4 || p10.w != 0.0f) { sw += w10; sx += p10.x * w10; sy += p10.y * w10; sz += p10.z * w10; }
  if (C != 4 || p01.w != 0.0f) { sw += w01; sx += p01.x * w01; sy += p01.y * w01; sz += p01.z * w01; }
  if (C != 4 || p11.w != 0.0f) { sw += w11; sx += p11.x * w11; sy += p11.y * w11; sz += p11.z * w11; }
  if (sw > 0.0f) return make_float4(div_exact(sx, sw), div_exact(sy, sw), div_exact(sz, sw), 0.0f);
  return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Plain two-stage bilinear of the 3-channel Laplacian (cudaBlend.cu:446-449).
__device__ __forceinline__ float lerp2(float v00, float v10, float v01, float v11, float dx, float dy) {
  const float v0 = v00 * (1.0f - dx) + v10 * dx;
  const float v1 = v01 * (1.0f - dx) + v11 * dx;
  return v0 * (1.0f - dy) + v1 * dy;
}
__device__ __forceinline__ float4 up_plain3(const float* low, size_t row0, size_t row1, const Taps& t) {
  const float4 p00 = ld_px<3>(low, row0 + t.x0);
  const float4 p10 = ld_px<3>(low, row0 + t.x1);
  const float4 p01 = ld_px<3>(low, row1 + t.x0);
  const float4 p11 = ld_px<3>(low, row1 + t.x1);
  return make_float4(lerp2(p00.x, p10.x, p01.x, p11.x, t.dx, t.dy), lerp2(p00.y, p10.y, p01.y, p11.y, t.dx, t.dy),
                     lerp2(p00.z, p10.z, p01.z, p11.z, t.dx, t.dy), 0.0f);
}

// Upsample used by the Laplacian: alpha-aware for RGBA, plain for RGB.
template <int C>
__device__ __forceinline__ float4 up_lap(const float* low, size_t row0, size_t row1, const Taps& t) {
  if constexpr (C == 4)
    return up_alpha<4>(low, row0, row1, t);
  else
    return up_plain3(low, row0, row1, t);
}

}  // namespace hmdev
