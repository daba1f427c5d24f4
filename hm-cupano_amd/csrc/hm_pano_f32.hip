// hm_pano_f32.hip -- the float pipelines of the pano classes
// (CudaStitchPano / CudaStitchPanoN / CudaStitchPano3 <float4|float3, ...>).
//
// The float pipeline takes float images and writes a float canvas.  Its
// stages run in the reference's order, one launch each, on pitched float
// images:
//   - the canvas and the float full buffers are cleared;
//   - the remaps write the canvas or the full buffers;
//   - the ROI copies move pixels between them;
//   - the blend runs the class's pyramid on the float full buffers.
// Levels are the uint8 pipeline's kernels with float level-0 inputs.
// Clears are kernels, not memsets, so captured graphs replay them safely.
// The maps are the classes' packed tables (x | y << 16).
#include "hm_pano.hpp"
#include "hm_pano_device.hpp"

namespace {

using namespace hmdev;

const dim3 kF32Block(64, 4, 1);

__device__ __forceinline__ float* f32_px(const Surf& s, int b, int x, int y, int c) {
  return reinterpret_cast<float*>(row_ptr(s, b, y)) + (size_t)x * c;
}

template <int C>
__device__ __forceinline__ void copy_px(float* d, const float* p) {
#pragma unroll
  for (int k = 0; k < C; ++k) d[k] = p[k];
}

__global__ void __launch_bounds__(256) f32_zero_kernel(Surf s, int row_floats, int rows) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= row_floats || y >= rows) return;
  reinterpret_cast<float*>(s.ptr + (size_t)y * s.pitch)[x] = 0.0f;
}

// BatchedRemapKernelExOffset(ROI) (src/cuda/cudaRemap.cu:61-198) with the
// all-zero default pixel and no_unmapped_write = false (mode kRemapOffset);
// BatchedRemapKernelExOffsetWithDestMap (:204-246: only where the canvas-wide
// label equals image_index, unmapped -> all zero; mode kRemapDestMap);
// FusedRemapToFullKernel3 (src/cuda/cudaFusedKernels3.cu:25-119: mapped,
// in-range pixels only; mode kRemapMapped).
enum F32RemapMode { kRemapOffset = 0, kRemapDestMap = 1, kRemapMapped = 2 };

template <int C, int MODE>
__global__ void __launch_bounds__(256)
    f32_remap_kernel(Surf src, Surf dst, const uint32_t* map, int mw, int mh, int off_x, int off_y, int roi_x,
                     int roi_y, int roi_w, int roi_h, int image_index, const uint8_t* labels) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int b = blockIdx.z;
  if (x >= roi_w || y >= roi_h) return;
  const int mx = roi_x + x, my = roi_y + y;
  if (mx < 0 || mx >= mw || my < 0 || my >= mh) return;
  const int dx = off_x + mx, dy = off_y + my;
  if (dx < 0 || dx >= (int)dst.w || dy < 0 || dy >= (int)dst.h) return;
  if (MODE == kRemapDestMap && labels[(size_t)dy * dst.w + dx] != image_index) return;
  const uint32_t m = map[(size_t)my * mw + mx];
  const int sx = (int)(m & 0xffffu), sy = (int)(m >> 16);
  float* d = f32_px(dst, b, dx, dy, C);
  if (sx < (int)src.w && sy < (int)src.h) {
    copy_px<C>(d, f32_px(src, b, sx, sy, C));
  } else if (MODE != kRemapMapped) {
#pragma unroll
    for (int k = 0; k < C; ++k) d[k] = 0.0f;
    if (C == 4 && MODE == kRemapOffset) d[3] = sx == kUnmapped ? 0.0f : 255.0f;
  }
}

struct F32HardN {
  Surf src[kMaxImages];
  const uint32_t* map[kMaxImages];
  int px[kMaxImages], py[kMaxImages], mw[kMaxImages], mh[kMaxImages];
  int n;
};

// BatchedRemapHardSeamKernelN (src/cuda/cudaRemap.cu:252-294) and
// FusedRemapHardSeam3 (cudaFusedKernels3.cu:124-217): the labelled image's
// pixel where its map is in range; sentinel, out-of-range and labels >= n
// leave the canvas alone.
template <int C>
__global__ void __launch_bounds__(256) f32_hard_n_kernel(F32HardN a, const uint8_t* labels, int labels_w, Surf dst) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int b = blockIdx.z;
  if (x >= (int)dst.w || y >= (int)dst.h) return;
  const int i = labels[(size_t)y * labels_w + x];
  if (i >= a.n) return;
  const int rx = x - a.px[i], ry = y - a.py[i];
  if (rx < 0 || ry < 0 || rx >= a.mw[i] || ry >= a.mh[i]) return;
  const uint32_t m = a.map[i][(size_t)ry * a.mw[i] + rx];
  const int sx = (int)(m & 0xffffu), sy = (int)(m >> 16);
  if (sx == kUnmapped) return;
  if (sx < (int)a.src[i].w && sy < (int)a.src[i].h) copy_px<C>(f32_px(dst, b, x, y, C), f32_px(a.src[i], b, sx, sy, C));
}

// copyRoiKernelBatched (src/cuda/cudaMakeFull.cu:40-69), float to float.
template <int C>
__global__ void __launch_bounds__(256) f32_copy_kernel(Surf src, int region_w, int region_h, int sx0, int sy0, Surf dst,
                                                       int ox, int oy) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int b = blockIdx.z;
  if (x >= region_w || y >= region_h) return;
  const int sx = sx0 + x, sy = sy0 + y;
  if (!(sx < (int)src.w && sy < (int)src.h)) return;
  const int dx = ox + x, dy = oy + y;
  if (!(dx < (int)dst.w && dy < (int)dst.h)) return;
  copy_px<C>(f32_px(dst, b, dx, dy, C), f32_px(src, b, sx, sy, C));
}

}  // namespace

namespace hm {

bool surface_ok_f32(const hmSurface& s, int channels) {
  return s.d_ptr && s.width > 0 && s.height > 0 && s.pitch >= s.width * (uint32_t)channels * 4u &&
         !(reinterpret_cast<uintptr_t>(s.d_ptr) & 3u) && !(s.pitch & 3u);
}

hmdev::Surf f32_tight(float* p, int w, int h, int channels) {
  return hmdev::Surf{reinterpret_cast<uint8_t*>(p), (uint32_t)w, (uint32_t)h, (uint32_t)((size_t)w * channels * 4)};
}

hipError_t f32_zero(const hmdev::Surf& s, int channels, int batch, hipStream_t st) {
  const int row_floats = (int)s.w * channels, rows = (int)s.h * batch;
  if (row_floats <= 0 || rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(f32_zero_kernel, grid_for(row_floats, rows, 1, kF32Block), kF32Block, 0, st, s, row_floats, rows);
  return hipGetLastError();
}

hipError_t f32_remap(int mode, int channels, const hmdev::Surf& src, const uint32_t* map, int mw, int mh, int off_x,
                     int off_y, int roi_x, int roi_y, int roi_w, int roi_h, const hmdev::Surf& dst, int image_index,
                     const uint8_t* labels, int batch, hipStream_t st) {
  if (roi_w <= 0 || roi_h <= 0) return hipSuccess;
  const dim3 g = grid_for(roi_w, roi_h, batch, kF32Block);
#define HM_F32_REMAP(C, M)                                                                                         \
  hipLaunchKernelGGL((f32_remap_kernel<C, M>), g, kF32Block, 0, st, src, dst, map, mw, mh, off_x, off_y, roi_x,    \
                     roi_y, roi_w, roi_h, image_index, labels)
  if (channels == 4) {
    if (mode == kRemapOffset) HM_F32_REMAP(4, kRemapOffset);
    else if (mode == kRemapDestMap) HM_F32_REMAP(4, kRemapDestMap);
    else HM_F32_REMAP(4, kRemapMapped);
  } else {
    if (mode == kRemapOffset) HM_F32_REMAP(3, kRemapOffset);
    else if (mode == kRemapDestMap) HM_F32_REMAP(3, kRemapDestMap);
    else HM_F32_REMAP(3, kRemapMapped);
  }
#undef HM_F32_REMAP
  return hipGetLastError();
}

hipError_t f32_hard_n(int n, int channels, const hmSurface* inputs, const DevBuf* maps, const int* px, const int* py,
                      const int* mw, const int* mh, const uint8_t* labels, int labels_w, const hmdev::Surf& dst, int batch,
                      hipStream_t st) {
  F32HardN a{};
  a.n = n;
  for (int i = 0; i < n; ++i) {
    a.src[i] = hmdev::to_surf(inputs[i]);
    a.map[i] = maps[i].as<uint32_t>();
    a.px[i] = px[i];
    a.py[i] = py[i];
    a.mw[i] = mw[i];
    a.mh[i] = mh[i];
  }
  const dim3 g = grid_for((int)dst.w, (int)dst.h, batch, kF32Block);
  if (channels == 4)
    hipLaunchKernelGGL((f32_hard_n_kernel<4>), g, kF32Block, 0, st, a, labels, labels_w, dst);
  else
    hipLaunchKernelGGL((f32_hard_n_kernel<3>), g, kF32Block, 0, st, a, labels, labels_w, dst);
  return hipGetLastError();
}

hipError_t f32_copy(int channels, const hmdev::Surf& src, int region_w, int region_h, int sx, int sy, const hmdev::Surf& dst, int ox,
                    int oy, int batch, hipStream_t st) {
  if (region_w <= 0 || region_h <= 0) return hipSuccess;
  const dim3 g = grid_for(region_w, region_h, batch, kF32Block);
  if (channels == 4)
    hipLaunchKernelGGL((f32_copy_kernel<4>), g, kF32Block, 0, st, src, region_w, region_h, sx, sy, dst, ox, oy);
  else
    hipLaunchKernelGGL((f32_copy_kernel<3>), g, kF32Block, 0, st, src, region_w, region_h, sx, sy, dst, ox, oy);
  return hipGetLastError();
}

}  // namespace hm
