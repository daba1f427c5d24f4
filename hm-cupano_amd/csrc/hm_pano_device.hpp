// hm_pano_device.hpp -- per-pixel remap helpers shared by the pano kernels.
#pragma once

#include "hm_batch.hpp"
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

// The same two pixels without a branch around the gather: a pixel that is
// not `take`n, or lies outside the source, loads source pixel (0, 0) (always
// a valid address) and the result is selected afterwards.  A branch per
// gather makes the compiler wait for each load before the next one; this form
// keeps every gather of a thread in flight at once.
template <int C>
__device__ __forceinline__ uint32_t remap_px_sel(const Surf& s, int b, uint32_t m, bool take) {
  const uint32_t sx = m & 0xffffu, sy = m >> 16;
  const bool inr = take && sx < s.w && sy < s.h;
  const uint32_t v = load_src_px<C>(s, b, inr ? (int)sx : 0, inr ? (int)sy : 0);
  uint32_t fb = 0u;
  if constexpr (C == 4) fb = (take && sx != (uint32_t)kUnmapped) ? 0xff000000u : 0u;
  return inr ? v : fb;
}

template <int C>
__device__ __forceinline__ uint32_t remap_px_dm_sel(const Surf& s, int b, uint32_t m, bool take) {
  const uint32_t sx = m & 0xffffu, sy = m >> 16;
  const bool inr = take && sx < s.w && sy < s.h;
  const uint32_t v = load_src_px<C>(s, b, inr ? (int)sx : 0, inr ? (int)sy : 0);
  return inr ? v : 0u;
}

inline Surf to_surf(const hmSurface& s) {
  return Surf{static_cast<uint8_t*>(s.d_ptr), s.width, s.height, s.pitch};
}

// ---------------------------------------------------------------------------
// Multi-image (N- and 3-image) remap kernels
// ---------------------------------------------------------------------------
template <int N>
struct MultiArgs {
  Surf src[N];
  const uint32_t* map[N];  // packed x | y << 16
  int mw[N], mh[N], px[N], py[N];
  int x0, x1, y0, y1;      // canvas region of this launch, [x0, x1) x [y0, y1)
  int tiles_x;
  // blend inputs: uint8 bw x bh images whose origin is canvas (x0, y0),
  // each written only on its valid region z0[i] (blend-input coordinates)
  uint8_t* g0[N];
  int bw, bh;
  ImgRect z0[N];
  // canvas
  uint8_t* canvas;
  uint32_t cpitch;
  int cw, ch;
  const uint8_t* seam;                     // canvas-sized labels
  int skip_x0, skip_x1, skip_y0, skip_y1;  // left to the blend write-back
  int mode;                                // kCanvasLabel, kCanvasLast or kCanvasClear
  int invalid_zero;    // label mode: a labelled pixel whose map is unmapped / out of range gets 0, else is left
  int uncovered_zero;  // a pixel no image writes gets 0, else is left
};

// Level-0 blend input i = image i remapped at its position inside its box,
// and the cleared buffer's zero outside.  DM = false: the default pixel of
// BatchedRemapKernelExOffset(ROI) (cudaRemap.cu:95-122 / :129-198: alpha 0
// for the sentinel, 255 for other out-of-range entries); DM = true: zero for
// both (FusedRemapToFullKernel3, cudaFusedKernels3.cu:25-119, writes only
// mapped in-range pixels into the cleared buffer).
//
// A workgroup covers 64 * kPx x 4 canvas pixels; only the images whose valid
// region meets that tile are written there.  They are taken up to
// kRemapSlots at a time, the slot count S a template parameter, so that the
// map loads of every slot and pixel are issued together, then every gather:
// two memory round trips per group of images.  Map loads are branch-free
// (a pixel off the map loads the row's first entry, ignored).
constexpr int kRemapSlots = 4;

template <int C, int N, bool DM, int S>
__device__ __forceinline__ void multi_inputs_slots(const MultiArgs<N>& a, const int* idx, int X0, int Y, int b) {
  const int v = Y - a.y0;
  uint32_t m[S][kPx];
  bool in[S][kPx], keep[S][kPx];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = idx[s];
    const ImgRect z = a.z0[i];
    const int ry = Y - a.py[i], mw = a.mw[i];
    const bool row_keep = v >= z.y0 && v <= z.y1;
    const bool row_ok = row_keep && ry >= 0 && ry < a.mh[i];
    const uint32_t* mrow = a.map[i] + (size_t)(row_ok ? ry : 0) * mw;
#pragma unroll
    for (int j = 0; j < kPx; ++j) {
      const int X = X0 + j * (int)blockDim.x;
      const int u = X - a.x0;
      keep[s][j] = row_keep && X < a.x1 && u >= z.x0 && u <= z.x1;
      const int rx = X - a.px[i];
      in[s][j] = keep[s][j] && row_ok && rx >= 0 && rx < mw;
      m[s][j] = mrow[in[s][j] ? rx : 0];
    }
  }
  uint32_t r[S][kPx];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const Surf src = a.src[idx[s]];
#pragma unroll
    for (int j = 0; j < kPx; ++j)
      r[s][j] = DM ? remap_px_dm_sel<C>(src, b, m[s][j], in[s][j]) : remap_px_sel<C>(src, b, m[s][j], in[s][j]);
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    uint8_t* grow = a.g0[idx[s]] + ((size_t)b * a.bh + v) * a.bw * C;
#pragma unroll
    for (int j = 0; j < kPx; ++j) {
      const int X = X0 + j * (int)blockDim.x;
      if (keep[s][j]) store_px<C>(grow + (size_t)(X - a.x0) * C, r[s][j]);
    }
  }
}

// Runs on 64 x 4 threads (kBlock).
template <int C, int N, bool DM>
__device__ __forceinline__ void multi_inputs_body(const MultiArgs<N>& a, Blk blk) {
  const int t = blk.x;
  const int tx0 = a.x0 + (t % a.tiles_x) * (int)blockDim.x * kPx, ty0 = a.y0 + (t / a.tiles_x) * (int)blockDim.y;
  const int X0 = tx0 + (int)threadIdx.x;
  const int Y = ty0 + (int)threadIdx.y;
  const int b = blk.z;
  if (Y >= a.y1) return;
  // the images whose valid region meets the tile (uniform)
  const int u0 = tx0 - a.x0, u1 = u0 + (int)blockDim.x * kPx - 1;
  const int v0 = ty0 - a.y0, v1 = v0 + (int)blockDim.y - 1;
  unsigned act = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const ImgRect& z = a.z0[i];
    if (z.x0 <= u1 && z.x1 >= u0 && z.y0 <= v1 && z.y1 >= v0) act |= 1u << i;
  }
  while (act) {
    int idx[kRemapSlots];
    const int n = min(__builtin_popcount(act), kRemapSlots);
#pragma unroll
    for (int s = 0; s < kRemapSlots; ++s) {
      idx[s] = act ? __builtin_ctz(act) : 0;
      act &= act - 1u;
    }
    switch (n) {  // uniform
      case 1: multi_inputs_slots<C, N, DM, 1>(a, idx, X0, Y, b); break;
      case 2: multi_inputs_slots<C, N, DM, 2>(a, idx, X0, Y, b); break;
      case 3: multi_inputs_slots<C, N, DM, 3>(a, idx, X0, Y, b); break;
      default: multi_inputs_slots<C, N, DM, 4>(a, idx, X0, Y, b); break;
    }
  }
}

template <int C, int N, bool DM>
__global__ void __launch_bounds__(256) multi_inputs_kernel(MultiArgs<N> a) {
  (void)a;
  multi_inputs_body<C, N, DM>(kernel_arg<MultiArgs<N>>(), hw_blk());
}

// Canvas writers (pixels inside the skip rectangle are never written):
//  kCanvasLabel: the seam label selects the image (batched_remap_hard_seam_kernel_n,
//    cudaRemap.cu:252-294; BatchedRemapKernelExOffsetWithDestMap, :204-246;
//    FusedRemapHardSeam3, cudaFusedKernels3.cu:124-217): its source pixel
//    where the map is in range, else zero or nothing (invalid_zero);
//  kCanvasLast: remaps of every image in order onto the canvas
//    (BatchedRemapKernelExOffset): the last image whose box covers the pixel,
//    with its default pixel (alpha 0 / 255) where unmapped / out of range;
//  kCanvasClear: no image.
// Pixels no image writes get zero (the reference's canvas clear) or are left
// (uncovered_zero).
enum CanvasMode { kCanvasLabel = 0, kCanvasLast = 1, kCanvasClear = 2 };

// One image's remap geometry, copied to LDS once per workgroup so that each
// pixel reads its labelled image's with a per-lane LDS load.
struct CanvasGeo {
  const uint32_t* map;
  uint8_t* src;
  int px, py, mw, mh;
  uint32_t sw, sh, spitch, pad;
};
constexpr int kCanvasGeoFloat4 = kMaxImages * (int)(sizeof(CanvasGeo) / 16);
static_assert(sizeof(CanvasGeo) % 16 == 0, "CanvasGeo is carved from float4 LDS");

// Label mode: per pixel only the labelled image, in three rounds of loads
// each issued in full before the next (labels, the labelled images' map
// entries, their source pixels); a load for a pixel that takes nothing goes
// to a valid address and is discarded.
template <int C, int N>
__device__ __forceinline__ void multi_canvas_label(const MultiArgs<N>& a, Blk blk, float4* lds) {
  CanvasGeo* geo = reinterpret_cast<CanvasGeo*>(lds);
  const int tid = flat_tid();
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (tid == i)
      geo[i] = CanvasGeo{a.map[i], a.src[i].ptr, a.px[i], a.py[i], a.mw[i], a.mh[i], a.src[i].w, a.src[i].h,
                         a.src[i].pitch, 0u};
  __syncthreads();
  const int t = blk.x;
  const int X0 = a.x0 + (t % a.tiles_x) * (int)blockDim.x * kPx + (int)threadIdx.x;
  const int Y = a.y0 + (t / a.tiles_x) * (int)blockDim.y + (int)threadIdx.y;
  const int b = blk.z;
  if (Y >= a.y1) return;
  const bool skip_row = Y >= a.skip_y0 && Y < a.skip_y1;
  const uint8_t* srow = a.seam + (size_t)Y * a.cw;
  int lab[kPx];
  bool act[kPx];
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    const int X = X0 + j * (int)blockDim.x;
    act[j] = X < a.x1 && !(skip_row && X >= a.skip_x0 && X < a.skip_x1);
    lab[j] = srow[act[j] ? X : a.x0];
  }
  uint32_t m[kPx];
  bool hit[kPx];
  int img[kPx];
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    const int X = X0 + j * (int)blockDim.x;
    img[j] = lab[j] < N ? lab[j] : 0;
    const CanvasGeo& g = geo[img[j]];
    const int rx = X - g.px, ry = Y - g.py;
    hit[j] = act[j] && lab[j] < N && rx >= 0 && ry >= 0 && rx < g.mw && ry < g.mh;
    m[j] = g.map[hit[j] ? (size_t)ry * g.mw + rx : 0];
  }
  uint32_t v[kPx];
  bool ok[kPx];
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    const CanvasGeo& g = geo[img[j]];
    const uint32_t sx = m[j] & 0xffffu, sy = m[j] >> 16;
    ok[j] = sx < g.sw && sy < g.sh;
    const Surf src{g.src, g.sw, g.sh, g.spitch};
    v[j] = remap_px_dm_sel<C>(src, b, m[j], hit[j]);
  }
  uint8_t* crow = a.canvas + ((size_t)b * a.ch + Y) * a.cpitch;
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    const bool wr = hit[j] ? (ok[j] || a.invalid_zero != 0) : (act[j] && a.uncovered_zero != 0);
    if (wr) store_px<C>(crow + (size_t)(X0 + j * (int)blockDim.x) * C, v[j]);
  }
}

// Runs on 64 x 4 threads (kBlock).  lds: kCanvasGeoFloat4 float4.
template <int C, int N>
__device__ __forceinline__ void multi_canvas_body(const MultiArgs<N>& a, Blk blk, float4* lds) {
  if (a.mode == kCanvasLabel && a.seam) {  // uniform
    multi_canvas_label<C, N>(a, blk, lds);
    return;
  }
  const int t = blk.x;
  const int X0 = a.x0 + (t % a.tiles_x) * (int)blockDim.x * kPx + (int)threadIdx.x;
  const int Y = a.y0 + (t / a.tiles_x) * (int)blockDim.y + (int)threadIdx.y;
  const int b = blk.z;
  if (Y >= a.y1) return;
  const bool skip_row = Y >= a.skip_y0 && Y < a.skip_y1;
  const uint8_t* srow = a.seam ? a.seam + (size_t)Y * a.cw : nullptr;
  int lab[kPx];
  bool act[kPx];
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    const int X = X0 + j * (int)blockDim.x;
    act[j] = X < a.x1 && !(skip_row && X >= a.skip_x0 && X < a.skip_x1);
    lab[j] = (act[j] && a.mode == kCanvasLabel) ? (int)srow[X] : -1;
  }
  uint32_t v[kPx] = {};
  bool wr[kPx];
#pragma unroll
  for (int j = 0; j < kPx; ++j) wr[j] = act[j] && a.uncovered_zero;
  if (a.mode != kCanvasClear) {
    // One branch per image with compile-time kernel-argument indices; a
    // wavefront runs only the branches its lanes take.
#pragma unroll
    for (int i = 0; i < N; ++i) {
      uint32_t m[kPx];
      bool hit[kPx];
#pragma unroll
      for (int j = 0; j < kPx; ++j) {
        const int rx = X0 + j * (int)blockDim.x - a.px[i], ry = Y - a.py[i];
        const bool in_box = act[j] && rx >= 0 && ry >= 0 && rx < a.mw[i] && ry < a.mh[i];
        hit[j] = in_box && (a.mode == kCanvasLast || lab[j] == i);
        m[j] = hit[j] ? a.map[i][(size_t)ry * a.mw[i] + rx] : 0u;
      }
      bool any = false;
#pragma unroll
      for (int j = 0; j < kPx; ++j) any = any || hit[j];
      if (!any) continue;
      // every gather first (branch-free, pixel (0, 0) for a miss), then the selects
      uint32_t g[kPx];
#pragma unroll
      for (int j = 0; j < kPx; ++j) g[j] = remap_px_dm_sel<C>(a.src[i], b, m[j], hit[j]);
#pragma unroll
      for (int j = 0; j < kPx; ++j) {
        if (!hit[j]) continue;
        const uint32_t sx = m[j] & 0xffffu, sy = m[j] >> 16;
        const bool ok = sx < a.src[i].w && sy < a.src[i].h;
        if (a.mode == kCanvasLast) {
          // remap_px: out of range -> alpha 0 for the sentinel, 255 otherwise (4 channels)
          uint32_t fb = 0u;
          if constexpr (C == 4) fb = sx == (uint32_t)kUnmapped ? 0u : 0xff000000u;
          v[j] = ok ? g[j] : fb;
          wr[j] = true;
        } else {
          v[j] = ok ? g[j] : 0u;
          wr[j] = ok || a.invalid_zero;
        }
      }
    }
  }
  uint8_t* crow = a.canvas + ((size_t)b * a.ch + Y) * a.cpitch;
#pragma unroll
  for (int j = 0; j < kPx; ++j)
    if (wr[j]) store_px<C>(crow + (size_t)(X0 + j * (int)blockDim.x) * C, v[j]);
}

template <int C, int N>
__global__ void __launch_bounds__(256) multi_canvas_kernel(MultiArgs<N> a) {
  (void)a;
  __shared__ float4 lds[kCanvasGeoFloat4];
  multi_canvas_body<C, N>(kernel_arg<MultiArgs<N>>(), hw_blk(), lds);
}

// ---------------------------------------------------------------------------
// Host side of the multi-image kernels
// ---------------------------------------------------------------------------
// Per-pano geometry (constant after set-up).
struct MultiGeom {
  const uint32_t* map[kMaxImages] = {};
  int mw[kMaxImages] = {}, mh[kMaxImages] = {}, px[kMaxImages] = {}, py[kMaxImages] = {};
  uint8_t* g0[kMaxImages] = {};
  int bw = 0, bh = 0;
  ImgRect z0[kMaxImages] = {};    // valid region of each blend input (blend-input coordinates)
  const uint8_t* seam = nullptr;  // canvas-sized labels
  int cw = 0, ch = 0;
};
enum MultiKernel { kMultiInputs = 0, kMultiInputsDM = 1, kMultiCanvas = 2 };
// One launch over the canvas region [x0, x1) x [y0, y1).
struct MultiRegion {
  int kernel = kMultiInputs;
  int x0 = 0, x1 = 0, y0 = 0, y1 = 0;
  int skip_x = 0, skip_y = 0, skip_w = 0, skip_h = 0;
  int mode = kCanvasLabel, invalid_zero = 1, uncovered_zero = 1;
};

template <int N>
hipError_t launch_multi(const MultiGeom& g, const MultiRegion& r, int channels, const hmSurface* inputs,
                        const hmSurface& canvas, int batch, hipStream_t s) {
  if (r.x1 <= r.x0 || r.y1 <= r.y0) return hipSuccess;
  MultiArgs<N> a{};
  for (int i = 0; i < N; ++i) {
    a.src[i] = to_surf(inputs[i]);
    a.map[i] = g.map[i];
    a.mw[i] = g.mw[i];
    a.mh[i] = g.mh[i];
    a.px[i] = g.px[i];
    a.py[i] = g.py[i];
    a.g0[i] = g.g0[i];
    a.z0[i] = g.z0[i];
  }
  a.x0 = r.x0;
  a.x1 = r.x1;
  a.y0 = r.y0;
  a.y1 = r.y1;
  const int tile_w = (int)kBlock.x * kPx;
  a.tiles_x = (r.x1 - r.x0 + tile_w - 1) / tile_w;
  const int tiles_y = (r.y1 - r.y0 + (int)kBlock.y - 1) / (int)kBlock.y;
  a.bw = g.bw;
  a.bh = g.bh;
  a.canvas = static_cast<uint8_t*>(canvas.d_ptr);
  a.cpitch = canvas.pitch;
  a.cw = g.cw;
  a.ch = g.ch;
  a.seam = g.seam;
  a.skip_x0 = r.skip_x;
  a.skip_x1 = r.skip_x + r.skip_w;
  a.skip_y0 = r.skip_y;
  a.skip_y1 = r.skip_y + r.skip_h;
  a.mode = r.mode;
  a.invalid_zero = r.invalid_zero;
  a.uncovered_zero = r.uncovered_zero;
  const dim3 grid((unsigned)(a.tiles_x * tiles_y), 1u, (unsigned)batch);
  const bool c4 = channels == 4;
  switch (r.kernel) {
    case kMultiInputs:
      if (hm::g_recorder)
        return c4 ? hm::g_recorder->add(hm::op_multi_inputs<4, N, false>(), grid, kBlock, s, a, "multi_inputs")
                  : hm::g_recorder->add(hm::op_multi_inputs<3, N, false>(), grid, kBlock, s, a, "multi_inputs");
      if (c4)
        hipLaunchKernelGGL((multi_inputs_kernel<4, N, false>), grid, kBlock, 0, s, a);
      else
        hipLaunchKernelGGL((multi_inputs_kernel<3, N, false>), grid, kBlock, 0, s, a);
      break;
    case kMultiInputsDM:
      if (hm::g_recorder)
        return c4 ? hm::g_recorder->add(hm::op_multi_inputs<4, N, true>(), grid, kBlock, s, a, "multi_inputs")
                  : hm::g_recorder->add(hm::op_multi_inputs<3, N, true>(), grid, kBlock, s, a, "multi_inputs");
      if (c4)
        hipLaunchKernelGGL((multi_inputs_kernel<4, N, true>), grid, kBlock, 0, s, a);
      else
        hipLaunchKernelGGL((multi_inputs_kernel<3, N, true>), grid, kBlock, 0, s, a);
      break;
    default:
      if (hm::g_recorder)
        return c4 ? hm::g_recorder->add(hm::op_multi_canvas<4, N>(), grid, kBlock, s, a, "multi_canvas")
                  : hm::g_recorder->add(hm::op_multi_canvas<3, N>(), grid, kBlock, s, a, "multi_canvas");
      if (c4)
        hipLaunchKernelGGL((multi_canvas_kernel<4, N>), grid, kBlock, 0, s, a);
      else
        hipLaunchKernelGGL((multi_canvas_kernel<3, N>), grid, kBlock, 0, s, a);
      break;
  }
  return hipGetLastError();
}

}  // namespace hmdev
