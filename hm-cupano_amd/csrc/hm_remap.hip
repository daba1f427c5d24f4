// hm_remap.hip -- kernel-tier remap / ROI-copy entry points (C-ABI).
//
// Stand-alone launchers with the semantics of src/cuda/cudaRemap.cu and
// src/cuda/cudaMakeFull.cu, for callers that drive the stages themselves.
// The pano classes do not use them: they run the fused kernels of
// hm_pano2.hip / hm_panoN.hip.  Pixels travel as float4 in registers (uint8
// to float is exact); stores convert back with the reference's perform_cast.
#include "hm_internal.hpp"
#include "hm_stitch.h"

using namespace hmdev;

namespace {

const dim3 kBlock(64, 4, 1);

Surf to_surf(const hmSurface& s) { return Surf{static_cast<uint8_t*>(s.d_ptr), s.width, s.height, s.pitch}; }

int channels_of(int t) { return (t == HM_PIX_U8C3 || t == HM_PIX_F32C3) ? 3 : 4; }
bool is_float(int t) { return t == HM_PIX_F32C3 || t == HM_PIX_F32C4; }
bool valid_type(int t) { return t >= HM_PIX_U8C3 && t <= HM_PIX_F32C4; }
size_t px_bytes(int t) { return (size_t)channels_of(t) * (is_float(t) ? 4 : 1); }

template <int C, bool F>
__device__ __forceinline__ float4 load_px(const Surf& s, int b, int x, int y) {
  const uint8_t* row = row_ptr(s, b, y);
  if constexpr (F) {
    const float* p = reinterpret_cast<const float*>(row) + (size_t)x * C;
    return make_float4(p[0], p[1], p[2], C == 4 ? p[3] : 255.0f);
  } else {
    const uint8_t* p = row + (size_t)x * C;
    return make_float4((float)p[0], (float)p[1], (float)p[2], C == 4 ? (float)p[3] : 255.0f);
  }
}

// perform_cast into the destination type: float stores are exact copies,
// uint8 stores truncate (round_to_uchar) when the source was float.
template <int C, bool F>
__device__ __forceinline__ void store_px(const Surf& s, int b, int x, int y, float4 v, bool src_float) {
  uint8_t* row = row_ptr(s, b, y);
  if constexpr (F) {
    float* p = reinterpret_cast<float*>(row) + (size_t)x * C;
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
    if (C == 4) p[3] = v.w;
  } else {
    uint8_t* p = row + (size_t)x * C;
    if (src_float) {
      p[0] = to_u8(v.x);
      p[1] = to_u8(v.y);
      p[2] = to_u8(v.z);
      if (C == 4) p[3] = to_u8(v.w);
    } else {
      p[0] = (uint8_t)v.x;
      p[1] = (uint8_t)v.y;
      p[2] = (uint8_t)v.z;
      if (C == 4) p[3] = (uint8_t)v.w;
    }
  }
}

// BatchedRemapKernelExOffset / ...ROI (src/cuda/cudaRemap.cu:61-198).
template <int C, bool FI, bool FO>
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
    store_px<C, FO>(dst, b, dx, dy, load_px<C, FI>(src, b, sx, sy), FI);
  } else if (!no_unmapped_write || sx != kUnmapped) {
    const float a = (C == 4) ? (sx == kUnmapped ? 0.0f : 255.0f) : 0.0f;
    store_px<C, FO>(dst, b, dx, dy, make_float4(0.0f, 0.0f, 0.0f, a), false);
  }
}

// BatchedRemapKernelExOffsetWithDestMap (cudaRemap.cu:204-246).
template <int C, bool F>
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
  if (sx < (int)src.w && sy < (int)src.h) v = load_px<C, F>(src, b, sx, sy);
  store_px<C, F>(dst, b, dx, dy, v, false);
}

struct HardSeamN {
  Surf src[kMaxImages];
  const uint16_t* mx[kMaxImages];
  const uint16_t* my[kMaxImages];
  int ox[kMaxImages], oy[kMaxImages], w[kMaxImages], h[kMaxImages];
  int n;
};

// BatchedRemapHardSeamKernelN (cudaRemap.cu:252-294).
template <int C, bool F>
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
  if (sx < (int)a.src[i].w && sy < (int)a.src[i].h) store_px<C, F>(dst, b, x, y, load_px<C, F>(a.src[i], b, sx, sy), false);
}

// copyRoiKernelBatched (src/cuda/cudaMakeFull.cu:40-69).
template <int CI, bool FI, int CO, bool FO>
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
  store_px<CO, FO>(dst, b, dx, dy, load_px<CI, FI>(src, b, sx, sy), FI);
}

template <int CI, bool FI>
hipError_t copy_dispatch_out(int dt, const Surf& s, int rw, int rh, int sx, int sy, const Surf& d, int ox, int oy,
                             int batch, hipStream_t st) {
  const dim3 g = hm::grid_for(rw, rh, batch, kBlock);
  switch (dt) {
    case HM_PIX_U8C3: hipLaunchKernelGGL((copy_roi_kernel<CI, FI, 3, false>), g, kBlock, 0, st, s, rw, rh, sx, sy, d, ox, oy); break;
    case HM_PIX_U8C4: hipLaunchKernelGGL((copy_roi_kernel<CI, FI, 4, false>), g, kBlock, 0, st, s, rw, rh, sx, sy, d, ox, oy); break;
    case HM_PIX_F32C3: hipLaunchKernelGGL((copy_roi_kernel<CI, FI, 3, true>), g, kBlock, 0, st, s, rw, rh, sx, sy, d, ox, oy); break;
    case HM_PIX_F32C4: hipLaunchKernelGGL((copy_roi_kernel<CI, FI, 4, true>), g, kBlock, 0, st, s, rw, rh, sx, sy, d, ox, oy); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
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
  if (is_float(src_type) && !is_float(dest_type)) return hipErrorInvalidValue;  // not instantiated upstream
  if (!surf_ok(src, src_type) || !surf_ok(dest, dest_type) || !d_map_x || !d_map_y || batch_size < 1)
    return hipErrorInvalidValue;
  const Surf s = to_surf(src), d = to_surf(dest);
  const dim3 g = hm::grid_for(roi_w, roi_h, batch_size, kBlock);
  const int c = channels_of(src_type);
  const bool fi = is_float(src_type), fo = is_float(dest_type);
#define HM_REMAP_LAUNCH(C, FI, FO)                                                                                  \
  hipLaunchKernelGGL((remap_offset_kernel<C, FI, FO>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, remap_w,      \
                     remap_h, offset_x, offset_y, roi_x, roi_y, roi_w, roi_h, no_unmapped_write)
  if (c == 4) {
    if (fi) HM_REMAP_LAUNCH(4, true, true);
    else if (fo) HM_REMAP_LAUNCH(4, false, true);
    else HM_REMAP_LAUNCH(4, false, false);
  } else {
    if (fi) HM_REMAP_LAUNCH(3, true, true);
    else if (fo) HM_REMAP_LAUNCH(3, false, true);
    else HM_REMAP_LAUNCH(3, false, false);
  }
#undef HM_REMAP_LAUNCH
  return hipGetLastError();
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
  switch (type) {
    case HM_PIX_U8C3: hipLaunchKernelGGL((remap_dest_map_kernel<3, false>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, this_image_index, d_dest_image_map, remap_w, remap_h, offset_x, offset_y); break;
    case HM_PIX_U8C4: hipLaunchKernelGGL((remap_dest_map_kernel<4, false>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, this_image_index, d_dest_image_map, remap_w, remap_h, offset_x, offset_y); break;
    case HM_PIX_F32C3: hipLaunchKernelGGL((remap_dest_map_kernel<3, true>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, this_image_index, d_dest_image_map, remap_w, remap_h, offset_x, offset_y); break;
    default: hipLaunchKernelGGL((remap_dest_map_kernel<4, true>), g, kBlock, 0, stream, s, d, d_map_x, d_map_y, this_image_index, d_dest_image_map, remap_w, remap_h, offset_x, offset_y); break;
  }
  return hipGetLastError();
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
  switch (type) {
    case HM_PIX_U8C3: hipLaunchKernelGGL((hard_seam_n_kernel<3, false>), g, kBlock, 0, stream, a, d_dest_image_map, d); break;
    case HM_PIX_U8C4: hipLaunchKernelGGL((hard_seam_n_kernel<4, false>), g, kBlock, 0, stream, a, d_dest_image_map, d); break;
    case HM_PIX_F32C3: hipLaunchKernelGGL((hard_seam_n_kernel<3, true>), g, kBlock, 0, stream, a, d_dest_image_map, d); break;
    default: hipLaunchKernelGGL((hard_seam_n_kernel<4, true>), g, kBlock, 0, stream, a, d_dest_image_map, d); break;
  }
  return hipGetLastError();
}

extern "C" int hmCopyRoi(hmSurface src, hmPixelType src_type, int region_w, int region_h, int src_roi_x, int src_roi_y,
                         hmSurface dest, hmPixelType dest_type, int offset_x, int offset_y, int batch_size,
                         hipStream_t stream) {
  if (!valid_type(src_type) || !valid_type(dest_type) || !surf_ok(src, src_type) || !surf_ok(dest, dest_type) ||
      batch_size < 1)
    return hipErrorInvalidValue;
  if (region_w <= 0 || region_h <= 0) return hipSuccess;
  const Surf s = to_surf(src), d = to_surf(dest);
  switch (src_type) {
    case HM_PIX_U8C3: return copy_dispatch_out<3, false>(dest_type, s, region_w, region_h, src_roi_x, src_roi_y, d, offset_x, offset_y, batch_size, stream);
    case HM_PIX_U8C4: return copy_dispatch_out<4, false>(dest_type, s, region_w, region_h, src_roi_x, src_roi_y, d, offset_x, offset_y, batch_size, stream);
    case HM_PIX_F32C3: return copy_dispatch_out<3, true>(dest_type, s, region_w, region_h, src_roi_x, src_roi_y, d, offset_x, offset_y, batch_size, stream);
    default: return copy_dispatch_out<4, true>(dest_type, s, region_w, region_h, src_roi_x, src_roi_y, d, offset_x, offset_y, batch_size, stream);
  }
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
