// hm_remap.hip -- kernel-tier remap / ROI-copy entry points (C-ABI).
//
// Stand-alone launchers with the semantics of src/cuda/cudaRemap.cu and
// src/cuda/cudaMakeFull.cu, for callers that drive the stages themselves.
// The pano classes do not use them: they run the fused kernels of
// hm_pano2.hip / hm_panoN.hip.  Pixels travel as float4 in registers (uint8
// to float is exact); stores convert back with the reference's perform_cast.
#include <hip/hip_fp16.h>

#include <type_traits>

#include "hm_internal.hpp"
#include "hm_stitch.h"

using namespace hmdev;

namespace {

const dim3 kBlock(64, 4, 1);

// Element types of pixel components.
enum Elem { kU8 = 0, kF32 = 1, kF16 = 2 };

Surf to_surf(const hmSurface& s) { return Surf{static_cast<uint8_t*>(s.d_ptr), s.width, s.height, s.pitch}; }

bool valid_type(int t) { return t >= HM_PIX_U8C3 && t <= HM_PIX_F16C4; }
int channels_of(int t) {
  switch (t) {
    case HM_PIX_U8C1: case HM_PIX_F32C1: case HM_PIX_F16C1: return 1;
    case HM_PIX_U8C3: case HM_PIX_F32C3: case HM_PIX_F16C3: return 3;
    default: return 4;
  }
}
int elem_of(int t) {
  switch (t) {
    case HM_PIX_F32C1: case HM_PIX_F32C3: case HM_PIX_F32C4: return kF32;
    case HM_PIX_F16C1: case HM_PIX_F16C3: case HM_PIX_F16C4: return kF16;
    default: return kU8;
  }
}
size_t px_bytes(int t) { return (size_t)channels_of(t) * (elem_of(t) == kF32 ? 4 : (elem_of(t) == kF16 ? 2 : 1)); }

// One pixel as float4 (every component type widens exactly); 3 channels read
// alpha 255 (the 3 -> 4 cast), 1 channel only x.
template <int C, int E>
__device__ __forceinline__ float4 load_px(const Surf& s, int b, int x, int y) {
  const uint8_t* row = row_ptr(s, b, y);
  float v[4] = {0.0f, 0.0f, 0.0f, 255.0f};
#pragma unroll
  for (int k = 0; k < C; ++k) {
    if constexpr (E == kF32)
      v[k] = reinterpret_cast<const float*>(row)[(size_t)x * C + k];
    else if constexpr (E == kF16)
      v[k] = __half2float(reinterpret_cast<const __half*>(row)[(size_t)x * C + k]);
    else
      v[k] = (float)row[(size_t)x * C + k];
  }
  return make_float4(v[0], v[1], v[2], v[3]);
}

// perform_cast into the destination type (src/cuda/cudaUtils.cuh:80-213):
// float stores are exact, half stores round to nearest even, uint8 stores
// truncate (round_to_uchar) when the source was float or half and copy when
// it was uint8.
template <int C, int E>
__device__ __forceinline__ void store_px(const Surf& s, int b, int x, int y, float4 v, bool src_u8) {
  uint8_t* row = row_ptr(s, b, y);
  const float c[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < C; ++k) {
    if constexpr (E == kF32)
      reinterpret_cast<float*>(row)[(size_t)x * C + k] = c[k];
    else if constexpr (E == kF16)
      reinterpret_cast<__half*>(row)[(size_t)x * C + k] = __float2half(c[k]);
    else
      row[(size_t)x * C + k] = src_u8 ? (uint8_t)c[k] : to_u8(c[k]);
  }
}

// BatchedRemapKernelExOffset / ...ROI (src/cuda/cudaRemap.cu:61-198).
template <int C, int EI, int EO>
__global__ void __launch_bounds__(256)
    remap_offset_kernel(Surf src, Surf dst, const uint16_t* mx, const uint16_t* my, int remap_w, int remap_h, int off_x,
                        int off_y, int roi_x, int roi_y, int roi_w, int roi_h, int no_unmapped_write) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int b = blockIdx.z;
  if (x >= roi_w || y >= roi_h) return;
  const int mxp = roi_x + x, myp = roi_y + y;
  if (mxp < 0 || mxp >= remap_w || myp < 0 || myp >= remap_h) return;
  const int dx = off_x + mxp, dy = off_y + myp;
  if (dx < 0 || dx >= (int)dst.w || dy < 0 || dy >= (int)dst.h) return;
  const size_t mi = (size_t)myp * remap_w + mxp;
  const int sx = mx[mi], sy = my[mi];
  if (C == 4 && no_unmapped_write && sx == kUnmapped) return;
  if (sx < (int)src.w && sy < (int)src.h) {
    store_px<C, EO>(dst, b, dx, dy, load_px<C, EI>(src, b, sx, sy), EI == kU8);
  } else if (!no_unmapped_write || sx != kUnmapped) {
    const float a = (C == 4) ? (sx == kUnmapped ? 0.0f : 255.0f) : 0.0f;
    store_px<C, EO>(dst, b, dx, dy, make_float4(0.0f, 0.0f, 0.0f, a), true);
  }
}

// BatchedRemapKernelExOffsetWithDestMap (cudaRemap.cu:204-246).
template <int C, int E>
__global__ void __launch_bounds__(256)
    remap_dest_map_kernel(Surf src, Surf dst, const uint16_t* mx, const uint16_t* my, int image_index,
                          const uint8_t* dest_map, int remap_w, int remap_h, int off_x, int off_y) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int b = blockIdx.z;
  if (x >= remap_w || y >= remap_h) return;
  const int dx = off_x + x, dy = off_y + y;
  if (dx < 0 || dx >= (int)dst.w || dy < 0 || dy >= (int)dst.h) return;
  if (dest_map[(size_t)dy * dst.w + dx] != image_index) return;
  const size_t mi = (size_t)y * remap_w + x;
  const int sx = mx[mi], sy = my[mi];
  float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (sx < (int)src.w && sy < (int)src.h) v = load_px<C, E>(src, b, sx, sy);
  store_px<C, E>(dst, b, dx, dy, v, E == kU8);
}

struct HardSeamN {
  Surf src[kMaxImages];
  const uint16_t* mx[kMaxImages];
  const uint16_t* my[kMaxImages];
  int ox[kMaxImages], oy[kMaxImages], w[kMaxImages], h[kMaxImages];
  int n;
};

// BatchedRemapHardSeamKernelN (cudaRemap.cu:252-294).
template <int C, int E>
__global__ void __launch_bounds__(256) hard_seam_n_kernel(HardSeamN a, const uint8_t* dest_map, Surf dst) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int b = blockIdx.z;
  if (x >= (int)dst.w || y >= (int)dst.h) return;
  const int i = dest_map[(size_t)y * dst.w + x];
  if (i < 0 || i >= a.n) return;
  const int rx = x - a.ox[i], ry = y - a.oy[i];
  if (rx < 0 || ry < 0 || rx >= a.w[i] || ry >= a.h[i]) return;
  const size_t mi = (size_t)ry * a.w[i] + rx;
  const int sx = a.mx[i][mi], sy = a.my[i][mi];
  if (sx == kUnmapped) return;
  if (sx < (int)a.src[i].w && sy < (int)a.src[i].h)
    store_px<C, E>(dst, b, x, y, load_px<C, E>(a.src[i], b, sx, sy), E == kU8);
}

// copyRoiKernelBatched (src/cuda/cudaMakeFull.cu:40-69).
template <int CI, int EI, int CO, int EO>
__global__ void __launch_bounds__(256) copy_roi_kernel(Surf src, int region_w, int region_h, int sx0, int sy0,
                                                       Surf dst, int ox, int oy) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int b = blockIdx.z;
  if (x >= region_w || y >= region_h) return;
  const int sx = sx0 + x, sy = sy0 + y;
  if (!(sx < (int)src.w && sy < (int)src.h)) return;
  const int dx = ox + x, dy = oy + y;
  if (!(dx < (int)dst.w && dy < (int)dst.h)) return;
  store_px<CO, EO>(dst, b, dx, dy, load_px<CI, EI>(src, b, sx, sy), EI == kU8);
}

// Calls f(integral_constant C, integral_constant E) for a pixel type.
template <typename F>
hipError_t with_type(int t, F&& f) {
  using std::integral_constant;
  switch (t) {
    case HM_PIX_U8C1: return f(integral_constant<int, 1>{}, integral_constant<int, kU8>{});
    case HM_PIX_U8C3: return f(integral_constant<int, 3>{}, integral_constant<int, kU8>{});
    case HM_PIX_U8C4: return f(integral_constant<int, 4>{}, integral_constant<int, kU8>{});
    case HM_PIX_F32C1: return f(integral_constant<int, 1>{}, integral_constant<int, kF32>{});
    case HM_PIX_F32C3: return f(integral_constant<int, 3>{}, integral_constant<int, kF32>{});
    case HM_PIX_F32C4: return f(integral_constant<int, 4>{}, integral_constant<int, kF32>{});
    case HM_PIX_F16C1: return f(integral_constant<int, 1>{}, integral_constant<int, kF16>{});
    case HM_PIX_F16C3: return f(integral_constant<int, 3>{}, integral_constant<int, kF16>{});
    case HM_PIX_F16C4: return f(integral_constant<int, 4>{}, integral_constant<int, kF16>{});
    default: return hipErrorInvalidValue;
  }
}

bool surf_ok(const hmSurface& s, int type) {
  return s.d_ptr && s.pitch >= s.width * px_bytes(type);
}

}  // namespace

extern "C" int hmRemapExOffsetRoi(hmSurface src, hmPixelType src_type, hmSurface dest, hmPixelType dest_type,
                                  const uint16_t* d_map_x, const uint16_t* d_map_y, int batch_size, int remap_w,
                                  int remap_h, int offset_x, int offset_y, int roi_x, int roi_y, int roi_w, int roi_h,
                                  int no_unmapped_write, hipStream_t stream) {
  // Launcher validation of cudaRemap.cu:364-369.
  if (roi_w <= 0 || roi_h <= 0) return hipSuccess;
  if (roi_x < 0 || roi_y < 0) return hipErrorInvalidValue;
  if (roi_x + roi_w > remap_w || roi_y + roi_h > remap_h) return hipErrorInvalidValue;
  if (!valid_type(src_type) || !valid_type(dest_type) || channels_of(src_type) != channels_of(dest_type))
    return hipErrorInvalidValue;
  // dest: the source type, or float for a uint8 source (the instantiated pairs)
  if (src_type != dest_type && !(elem_of(src_type) == kU8 && elem_of(dest_type) == kF32)) return hipErrorInvalidValue;
  if (!surf_ok(src, src_type) || !surf_ok(dest, dest_type) || !d_map_x || !d_map_y || batch_size < 1)
    return hipErrorInvalidValue;
  const Surf s = to_surf(src), d = to_surf(dest);
  const dim3 g = hm::grid_for(roi_w, roi_h, batch_size, kBlock);
  const int eo = elem_of(dest_type);
  return with_type(src_type, [&](auto Cv, auto EIv) -> hipError_t {
    constexpr int C = decltype(Cv)::value, EI = decltype(EIv)::value;
    if (eo == kF32 && EI != kF32)
      hipLaunchKernelGGL((remap_offset_kernel<C, EI, kF32>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, remap_w,
                         remap_h, offset_x, offset_y, roi_x, roi_y, roi_w, roi_h, no_unmapped_write);
    else
      hipLaunchKernelGGL((remap_offset_kernel<C, EI, EI>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, remap_w,
                         remap_h, offset_x, offset_y, roi_x, roi_y, roi_w, roi_h, no_unmapped_write);
    return hipGetLastError();
  });
}

extern "C" int hmRemapExOffset(hmSurface src, hmPixelType src_type, hmSurface dest, hmPixelType dest_type,
                               const uint16_t* d_map_x, const uint16_t* d_map_y, int batch_size, int remap_w,
                               int remap_h, int offset_x, int offset_y, int no_unmapped_write, hipStream_t stream) {
  if (remap_w <= 0 || remap_h <= 0) return hipSuccess;
  return hmRemapExOffsetRoi(src, src_type, dest, dest_type, d_map_x, d_map_y, batch_size, remap_w, remap_h, offset_x,
                            offset_y, 0, 0, remap_w, remap_h, no_unmapped_write, stream);
}

extern "C" int hmRemapExOffsetWithDestMap(hmSurface src, hmSurface dest, hmPixelType type, const uint16_t* d_map_x,
                                          const uint16_t* d_map_y, int this_image_index,
                                          const uint8_t* d_dest_image_map, int batch_size, int remap_w, int remap_h,
                                          int offset_x, int offset_y, hipStream_t stream) {
  if (!valid_type(type) || !surf_ok(src, type) || !surf_ok(dest, type) || !d_map_x || !d_map_y ||
      !d_dest_image_map || batch_size < 1)
    return hipErrorInvalidValue;
  if (remap_w <= 0 || remap_h <= 0) return hipSuccess;
  const Surf s = to_surf(src), d = to_surf(dest);
  const dim3 g = hm::grid_for(remap_w, remap_h, batch_size, kBlock);
  return with_type(type, [&](auto Cv, auto Ev) -> hipError_t {
    constexpr int C = decltype(Cv)::value, E = decltype(Ev)::value;
    hipLaunchKernelGGL((remap_dest_map_kernel<C, E>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, this_image_index,
                       d_dest_image_map, remap_w, remap_h, offset_x, offset_y);
    return hipGetLastError();
  });
}

extern "C" int hmRemapHardSeamN(int n_images, const hmSurface* inputs, const uint16_t* const* d_map_x,
                                const uint16_t* const* d_map_y, const int* offsets_xy, const int* sizes_wh,
                                const uint8_t* d_dest_image_map, hmSurface dest, hmPixelType type, int batch_size,
                                hipStream_t stream) {
  if (n_images < 1 || n_images > kMaxImages || !inputs || !d_map_x || !d_map_y || !offsets_xy || !sizes_wh ||
      !d_dest_image_map || !valid_type(type) || !surf_ok(dest, type) || batch_size < 1)
    return hipErrorInvalidValue;
  HardSeamN a{};
  a.n = n_images;
  for (int i = 0; i < n_images; ++i) {
    if (!surf_ok(inputs[i], type) || !d_map_x[i] || !d_map_y[i]) return hipErrorInvalidValue;
    a.src[i] = to_surf(inputs[i]);
    a.mx[i] = d_map_x[i];
    a.my[i] = d_map_y[i];
    a.ox[i] = offsets_xy[2 * i];
    a.oy[i] = offsets_xy[2 * i + 1];
    a.w[i] = sizes_wh[2 * i];
    a.h[i] = sizes_wh[2 * i + 1];
  }
  const Surf d = to_surf(dest);
  const dim3 g = hm::grid_for((int)dest.width, (int)dest.height, batch_size, kBlock);
  return with_type(type, [&](auto Cv, auto Ev) -> hipError_t {
    constexpr int C = decltype(Cv)::value, E = decltype(Ev)::value;
    hipLaunchKernelGGL((hard_seam_n_kernel<C, E>), g, kBlock, 0, stream, a, d_dest_image_map, d);
    return hipGetLastError();
  });
}

extern "C" int hmCopyRoi(hmSurface src, hmPixelType src_type, int region_w, int region_h, int src_roi_x, int src_roi_y,
                         hmSurface dest, hmPixelType dest_type, int offset_x, int offset_y, int batch_size,
                         hipStream_t stream) {
  if (!valid_type(src_type) || !valid_type(dest_type) || !surf_ok(src, src_type) || !surf_ok(dest, dest_type) ||
      batch_size < 1)
    return hipErrorInvalidValue;
  if (region_w <= 0 || region_h <= 0) return hipSuccess;
  // 1-channel types copy to 1 channel; 3 and 4 channels convert (alpha 255 / dropped).
  if ((channels_of(src_type) == 1) != (channels_of(dest_type) == 1)) return hipErrorInvalidValue;
  const Surf s = to_surf(src), d = to_surf(dest);
  const dim3 g = hm::grid_for(region_w, region_h, batch_size, kBlock);
  return with_type(src_type, [&](auto CIv, auto EIv) -> hipError_t {
    constexpr int CI = decltype(CIv)::value, EI = decltype(EIv)::value;
    return with_type(dest_type, [&](auto COv, auto EOv) -> hipError_t {
      constexpr int CO = decltype(COv)::value, EO = decltype(EOv)::value;
      if constexpr ((CI == 1) == (CO == 1)) {
        hipLaunchKernelGGL((copy_roi_kernel<CI, EI, CO, EO>), g, kBlock, 0, stream, s, region_w, region_h, src_roi_x,
                           src_roi_y, d, offset_x, offset_y);
        return hipGetLastError();
      }
      return hipErrorInvalidValue;
    });
  });
}

extern "C" int hmMakeFull(hmSurface src, hmPixelType src_type, int region_w, int region_h, int src_roi_x,
                          int src_roi_y, int dest_offset_x, int dest_offset_y, int batch_size, hmSurface dest,
                          hmPixelType dest_type, hipStream_t stream) {
  if (!valid_type(dest_type) || !surf_ok(dest, dest_type) || batch_size < 1) return hipErrorInvalidValue;
  // cudaMakeFull.cu:200: memset of the whole destination batch.
  HM_RETURN_IF_ERROR(hipMemsetAsync(dest.d_ptr, 0, (size_t)dest.pitch * dest.height * batch_size, stream));
  return hmCopyRoi(src, src_type, region_w, region_h, src_roi_x, src_roi_y, dest, dest_type, dest_offset_x,
                   dest_offset_y, batch_size, stream);
}
