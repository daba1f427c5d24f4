// hm_pano2.hip -- CudaStitchPano (2-image) on gfx950.
//
// Reference: src/pano/cudaPano.inl (ctor :19-94, process_impl :167-418),
// src/pano/canvasManager.cpp (:16-101).  The reference runs, per frame,
// memset(canvas, full1, full2) -> remap1 -> make_full1 -> remap2 -> make_full2
// -> 34-launch Laplacian blend -> copy-back (about 44 stream operations).
// Here one kernel produces, from the two remaps, both the canvas outside the
// blend region and the level-0 blend inputs (kept as exact uint8), the
// pyramid runs one down launch and one fused up launch per level, and the
// level-0 up launch writes the truncated result straight into the canvas.
// The hard seam (levels == 0) is one launch.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "hm_pano.hpp"

using namespace hmdev;

namespace {

constexpr int kOverlapPad = 128;  // CanvasManager overlap_pad default (canvasManager.h:65)
const dim3 kBlock(64, 4, 1);

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

struct Pano2L0 {
  Surf src[2];
  const uint32_t* map[2];
  int mw[2], mh[2], px[2], py[2];
  uint8_t* canvas;
  uint32_t cpitch;
  int cw, ch;
  int write_canvas, skip_x0, skip_x1;  // canvas columns [skip_x0, skip_x1) are left to the blend write-back
  uint8_t* g0[2];
  int bw, bh;
  int d[2], ulo[2], uhi[2];  // G_i[u] <- canvas column u + d[i] for u in [ulo[i], uhi[i])
  int g2_canvas;             // 1: G2 takes the canvas after both remaps; 0: image 2 alone
  int tiles_x, tiles_y;
};

// Threads that remap canvas rows handle kPx pixels kBlock.x apart, so a
// wavefront's loads stay coalesced while every lane keeps several
// independent map loads (then gathers) in flight.
constexpr int kPx = 4;

// ---------------------------------------------------------------------------
// Remap + Gaussian pyramid in one launch.  Workgroups [0, roi_tiles) each take
// a 32x32 tile of the blend region: they build both level-0 images from the
// two remaps (the make_full copies of cudaPano.inl:225-236, 306-317), store
// them as uint8, and reduce them to up to five coarser levels in registers
// and LDS.  The remaining workgroups remap 256x4 canvas strips outside the
// blend columns (the final canvas there, cudaPano.inl:210-303).
// ---------------------------------------------------------------------------
struct Pano2Fused {
  Pano2L0 l0;
  const LevelDesc* lv;  // lv[0] = level 1 ...
  int n_down;           // Gaussian levels produced here (0..5)
  int roi_tiles_x, roi_tiles_y;
  int strips_x0, strips_x1, strips_y;  // canvas strips per row left / right of the blend columns
};

// Canvas value after remap 1 (c1) at (X, Y), and whether X, Y lies in box i.
template <int C>
__device__ __forceinline__ bool in_box(const Pano2L0& a, int i, int X, int Y) {
  const int rx = X - a.px[i], ry = Y - a.py[i];
  return rx >= 0 && ry >= 0 && rx < a.mw[i] && ry < a.mh[i];
}
template <int C>
__device__ __forceinline__ uint32_t remap_at(const Pano2L0& a, int i, int b, int X, int Y) {
  return remap_px<C>(a.src[i], b, a.map[i][(size_t)(Y - a.py[i]) * a.mw[i] + (X - a.px[i])]);
}

template <int C>
__global__ void __launch_bounds__(256) pano2_fused_kernel(Pano2Fused f) {
  __shared__ float4 s1[2][16 * 16];
  __shared__ float4 s2[2][8 * 8];
  const Pano2L0& a = f.l0;
  const int b = blockIdx.z;
  const int tid = threadIdx.x;
  const int n_roi = f.roi_tiles_x * f.roi_tiles_y;
  if ((int)blockIdx.x >= n_roi) {
    // ---- canvas strip outside the blend columns
    const int s = (int)blockIdx.x - n_roi;
    const int per_row = f.strips_x0 + f.strips_x1;
    const int sy = s / per_row, sx = s % per_row;
    const int xbase = sx < f.strips_x0 ? sx * 64 * kPx : a.skip_x1 + (sx - f.strips_x0) * 64 * kPx;
    const int xend = sx < f.strips_x0 ? a.skip_x0 : a.cw;
    const int Y = sy * 4 + (tid >> 6);
    if (Y >= a.ch) return;
    const int X0 = xbase + (tid & 63);
    uint32_t m[2][kPx];
    bool valid[2][kPx];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int ry = Y - a.py[i];
      const bool row_ok = ry >= 0 && ry < a.mh[i];
      const uint32_t* mrow = a.map[i] + (size_t)(row_ok ? ry : 0) * a.mw[i];
#pragma unroll
      for (int j = 0; j < kPx; ++j) {
        const int X = X0 + 64 * j;
        const int rx = X - a.px[i];
        valid[i][j] = row_ok && rx >= 0 && rx < a.mw[i] && X < xend;
        m[i][j] = valid[i][j] ? mrow[rx] : 0u;
      }
    }
    uint8_t* crow = a.canvas + ((size_t)b * a.ch + Y) * a.cpitch;
#pragma unroll
    for (int j = 0; j < kPx; ++j) {
      const int X = X0 + 64 * j;
      if (X >= xend) continue;
      uint32_t v = 0u;
      if (valid[1][j])
        v = remap_px<C>(a.src[1], b, m[1][j]);
      else if (valid[0][j])
        v = remap_px<C>(a.src[0], b, m[0][j]);
      store_px<C>(crow + (size_t)X * C, v);
    }
    return;
  }
  // ---- blend-region tile: level 0 of both images, then the Gaussian levels
  const int t = xcd_tile((int)blockIdx.x, n_roi);
  const int tx = t % f.roi_tiles_x, ty = t / f.roi_tiles_x;
  const int lx = tid & 15, ly = tid >> 4;
  const int u0 = tx * 32 + 2 * lx, y0 = ty * 32 + 2 * ly;
  // Phase 1: every map word this thread needs, issued back to back.
  //   slot 0: image 1 at X1 (G1 and, when X2 == X1, G2's fallback)
  //   slot 1: image 2 at X2 (G2)
  //   slot 2: image 1 at X2 (G2's fallback when X2 != X1)
  uint32_t mw[3][4];
  bool use[3][4];
  bool inr[4], take1[4], take2[4];
  const bool same_x = a.d[0] == a.d[1];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int u = u0 + (e & 1), y = y0 + (e >> 1);
    inr[e] = u < a.bw && y < a.bh;
    const int X1 = u + a.d[0], X2 = u + a.d[1];
    take1[e] = inr[e] && u >= a.ulo[0] && u < a.uhi[0] && X1 >= 0 && X1 < a.cw && y < a.ch;
    take2[e] = inr[e] && u >= a.ulo[1] && u < a.uhi[1] && X2 >= 0 && X2 < a.cw && y < a.ch;
    use[0][e] = take1[e] && in_box<C>(a, 0, X1, y);
    use[1][e] = take2[e] && in_box<C>(a, 1, X2, y);
    use[2][e] = take2[e] && !use[1][e] && a.g2_canvas && !(same_x && take1[e]) && in_box<C>(a, 0, X2, y);
    mw[0][e] = use[0][e] ? a.map[0][(size_t)(y - a.py[0]) * a.mw[0] + (X1 - a.px[0])] : 0u;
    mw[1][e] = use[1][e] ? a.map[1][(size_t)(y - a.py[1]) * a.mw[1] + (X2 - a.px[1])] : 0u;
    mw[2][e] = use[2][e] ? a.map[0][(size_t)(y - a.py[0]) * a.mw[0] + (X2 - a.px[0])] : 0u;
  }
  // Phase 2: the source gathers.
  uint32_t rv[3][4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    rv[0][e] = use[0][e] ? remap_px<C>(a.src[0], b, mw[0][e]) : 0u;
    rv[1][e] = use[1][e] ? remap_px<C>(a.src[1], b, mw[1][e]) : 0u;
    rv[2][e] = use[2][e] ? remap_px<C>(a.src[0], b, mw[2][e]) : 0u;
  }
  // Phase 3: the make_full values (cudaPano.inl:225-236, 306-317).
  uint32_t g[2][4];  // [image][tap (0,0),(1,0),(0,1),(1,1)]
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    g[0][e] = rv[0][e];  // canvas after remap 1 (0 outside image 1's box)
    uint32_t v2 = 0u;
    if (use[1][e])
      v2 = rv[1][e];
    else if (a.g2_canvas && take2[e])
      v2 = (same_x && take1[e]) ? rv[0][e] : rv[2][e];  // image 1 shows through outside image 2's box
    g[1][e] = take2[e] ? v2 : 0u;
  }
  // level-0 stores (uint8, exact)
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (!inr[e]) continue;
    const size_t o = ((size_t)b * a.bh + (y0 + (e >> 1))) * a.bw + (u0 + (e & 1));
    store_px<C>(a.g0[0] + o * C, g[0][e]);
    store_px<C>(a.g0[1] + o * C, g[1][e]);
  }
  if (f.n_down <= 0) return;
  // level 1 from the 2x2 block in registers
  {
    const LevelDesc& d = f.lv[0];
    const int x = tx * 16 + lx, y = ty * 16 + ly;
    if (x < d.w && y < d.h) {
      const size_t o = ((size_t)b * d.h + y) * d.w + x;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        float4 p[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t v = g[i][e];
          p[e] = make_float4((float)(v & 0xffu), (float)((v >> 8) & 0xffu), (float)((v >> 16) & 0xffu),
                             (float)(v >> 24));
        }
        const float4 r = down_px<C>(p, inr);
        s1[i][ly * 16 + lx] = r;
        st_px<C>(d.g_out[i], o, r);
      }
    }
  }
  // levels 2 .. n_down from LDS
  int edge = 16;
  for (int k = 1; k < f.n_down; ++k) {
    __syncthreads();
    const LevelDesc& din = f.lv[k - 1];
    const LevelDesc& d = f.lv[k];
    const int oedge = edge >> 1;
    float4 r[2];
    bool active = false;
    int qx = 0, qy = 0;
    if (tid < oedge * oedge) {
      qx = tid % oedge;
      qy = tid / oedge;
      const int x = tx * oedge + qx, y = ty * oedge + qy;
      if (x < d.w && y < d.h) {
        active = true;
        const bool has_x1 = 2 * x + 1 < din.w, has_y1 = 2 * y + 1 < din.h;
        const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
        const int sx = 2 * qx, sy = 2 * qy;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const float4* src = k == 1 ? s1[i] : s2[i];
          float4 p[4];
          p[0] = src[sy * edge + sx];
          if (has_x1) p[1] = src[sy * edge + sx + 1];
          if (has_y1) p[2] = src[(sy + 1) * edge + sx];
          if (has_x1 && has_y1) p[3] = src[(sy + 1) * edge + sx + 1];
          r[i] = down_px<C>(p, in_range);
          st_px<C>(d.g_out[i], ((size_t)b * d.h + y) * d.w + x, r[i]);
        }
      }
    }
    __syncthreads();
    if (active) {
      s2[0][qy * oedge + qx] = r[0];
      s2[1][qy * oedge + qx] = r[1];
    }
    edge = oedge;
  }
}

struct Pano2Hard {
  Surf src[2];
  const uint32_t* map[2];
  int mw[2], mh[2], px[2], py[2];
  uint8_t* canvas;
  uint32_t cpitch;
  int cw, ch;
  const uint8_t* seam;  // canvas-sized, 1 selects image 1, 0 image 2
  int tiles_x, tiles_y;
};

// Hard seam: the two dest-map remaps of cudaPano.inl:262-276 and :345-359
// (image 1 where seam == 1, image 2 where seam == 0) plus the canvas memset,
// as one pass; kPx pixels per thread as above.
template <int C>
__global__ void __launch_bounds__(256) pano2_hard_kernel(Pano2Hard a) {
  const int t = xcd_tile((int)blockIdx.x, a.tiles_x * a.tiles_y);
  const int X0 = (t % a.tiles_x) * (int)blockDim.x * kPx + (int)threadIdx.x;
  const int Y = (t / a.tiles_x) * (int)blockDim.y + (int)threadIdx.y;
  const int b = blockIdx.z;
  if (Y >= a.ch) return;
  const uint8_t* srow = a.seam + (size_t)Y * a.cw;
  uint32_t m[kPx];
  int sel[kPx];  // 0: image 1, 1: image 2, -1: nothing
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    const int X = X0 + j * (int)blockDim.x;
    sel[j] = -1;
    m[j] = 0u;
    if (X >= a.cw) continue;
    const uint8_t sv = srow[X];
    if (sv > 1) continue;
    // Scalar selects: a dynamic index into the kernel-argument arrays would
    // force them into scratch memory.
    const bool first = sv == 1;
    const int rx = X - (first ? a.px[0] : a.px[1]);
    const int ry = Y - (first ? a.py[0] : a.py[1]);
    const int mw = first ? a.mw[0] : a.mw[1];
    const int mh = first ? a.mh[0] : a.mh[1];
    if (rx >= 0 && ry >= 0 && rx < mw && ry < mh) {
      sel[j] = first ? 0 : 1;
      m[j] = (first ? a.map[0] : a.map[1])[(size_t)ry * mw + rx];
    }
  }
  uint32_t v[kPx];
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    v[j] = 0u;
    if (sel[j] >= 0) v[j] = remap_px_dm<C>(sel[j] == 0 ? a.src[0] : a.src[1], b, m[j]);
  }
  uint8_t* crow = a.canvas + ((size_t)b * a.ch + Y) * a.cpitch;
#pragma unroll
  for (int j = 0; j < kPx; ++j) {
    const int X = X0 + j * (int)blockDim.x;
    if (X < a.cw) store_px<C>(crow + (size_t)X * C, v[j]);
  }
}

Surf to_surf(const hmSurface& s) { return Surf{static_cast<uint8_t*>(s.d_ptr), s.width, s.height, s.pitch}; }

}  // namespace

// ===========================================================================
// shared helpers
// ===========================================================================
namespace hm {

hipError_t upload_packed_map(const uint16_t* map_x, const uint16_t* map_y, int w, int h, DevBuf& dst) {
  if (!map_x || !map_y || w < 1 || h < 1) return hipErrorInvalidValue;
  std::vector<uint32_t> packed((size_t)w * h);
  for (size_t i = 0; i < packed.size(); ++i) packed[i] = (uint32_t)map_x[i] | ((uint32_t)map_y[i] << 16);
  HM_RETURN_IF_ERROR(dst.alloc(packed.size() * sizeof(uint32_t)));
  return hipMemcpy(dst.p, packed.data(), packed.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
}

bool pad_replicate(const uint8_t* src, int sw, int sh, int cw, int ch, std::vector<uint8_t>& out) {
  if (!src || sw < 1 || sh < 1 || sw > cw || sh > ch) return false;
  out.resize((size_t)cw * ch);
  for (int y = 0; y < ch; ++y) {
    const uint8_t* row = src + (size_t)std::min(y, sh - 1) * sw;
    for (int x = 0; x < cw; ++x) out[(size_t)y * cw + x] = row[std::min(x, sw - 1)];
  }
  return true;
}

bool surface_ok(const hmSurface& s, int channels, bool need_vec_align) {
  if (!s.d_ptr || s.width == 0 || s.height == 0) return false;
  if (s.pitch < s.width * (uint32_t)channels) return false;
  if (need_vec_align && channels == 4 && ((reinterpret_cast<uintptr_t>(s.d_ptr) & 3u) || (s.pitch & 3u)))
    return false;
  return true;
}

}  // namespace hm

hipError_t hmStitchPano::init_pyramid(int kind, int bw, int bh, const std::vector<float>& mask_host) {
  HM_RETURN_IF_ERROR(pyr.init(kind, n_images, channels, bw, bh, num_levels, batch, debug));
  HM_RETURN_IF_ERROR(mask0.alloc(mask_host.size() * sizeof(float)));
  HM_RETURN_IF_ERROR(hipMemcpy(mask0.p, mask_host.data(), mask_host.size() * sizeof(float), hipMemcpyHostToDevice));
  // The mask pyramid depends only on the seam: built once and kept resident
  // (the reference rebuilds it every frame with the same kernel arithmetic).
  HM_RETURN_IF_ERROR(pyr.build_mask_pyramid(mask0.as<float>(), nullptr));
  HM_RETURN_IF_ERROR(hipStreamSynchronize(nullptr));
  initialized_pyramid = true;
  return hipSuccess;
}

// ===========================================================================
// CudaStitchPano
// ===========================================================================
namespace {

struct Pano2 : hmStitchPano {
  int x1 = 0, y1 = 0, x2 = 0, y2 = 0;
  int mw[2] = {0, 0}, mh[2] = {0, 0};
  bool minimize = false;
  int bw = 0, bh = 0;
  hm::DevBuf map[2];
  hm::DevBuf seam;  // hard seam, canvas-sized
  std::vector<float> mask_host;
  std::vector<uint8_t> mask_u8_host;  // the same level-0 mask as uint8 (the seam values)
  hm::DevBuf mask_u8;
  Pano2L0 l0{};

  hipError_t setup(const hmControlMasks2& m, int minimize_blend);
  hipError_t ensure_pyramid();
  hipError_t process(const hmSurface& in1, const hmSurface& in2, const hmSurface& canvas, hipStream_t s);
};

hipError_t Pano2::setup(const hmControlMasks2& m, int minimize_blend) {
  if (!m.img1_col || !m.img1_row || !m.img2_col || !m.img2_row || !m.whole_seam_mask_image || m.img1_w < 1 ||
      m.img1_h < 1 || m.img2_w < 1 || m.img2_h < 1)
    return hipErrorFileNotFound;  // ControlMasks::is_valid() failed (cudaPano.inl:28-31)
  hard_seam = num_levels == 0;
  // ControlMasks::canvas_width/height in float, then size_t (controlMasks.cpp:265-271).
  canvas_w = (int)(size_t)std::max(m.xpos[0] + (float)m.img1_w, m.xpos[1] + (float)m.img2_w);
  canvas_h = (int)(size_t)std::max(m.ypos[0] + (float)m.img1_h, m.ypos[1] + (float)m.img2_h);
  // CanvasInfo positions are cv::Point (int) (cudaPano.inl:47-52).
  x1 = (int)m.xpos[0];
  y1 = (int)m.ypos[0];
  x2 = (int)m.xpos[1];
  y2 = (int)m.ypos[1];
  mw[0] = m.img1_w;
  mh[0] = m.img1_h;
  mw[1] = m.img2_w;
  mh[1] = m.img2_h;
  minimize = minimize_blend && !hard_seam;

  std::vector<uint8_t> padded;
  if (!hm::pad_replicate(m.whole_seam_mask_image, m.seam_w, m.seam_h, canvas_w, canvas_h, padded))
    return hipErrorInvalidValue;
  HM_RETURN_IF_ERROR(hm::upload_packed_map(m.img1_col, m.img1_row, mw[0], mh[0], map[0]));
  HM_RETURN_IF_ERROR(hm::upload_packed_map(m.img2_col, m.img2_row, mw[1], mh[1], map[1]));

  if (hard_seam) {
    HM_RETURN_IF_ERROR(seam.alloc(padded.size()));
    return hipMemcpy(seam.p, padded.data(), padded.size(), hipMemcpyHostToDevice);
  }

  // CanvasManager::updateMinimizeBlend (canvasManager.cpp:16-61).
  const int overlap = mw[0] - x2;
  const int blend_width = overlap + 2 * kOverlapPad;
  int bx0 = 0;
  l0 = Pano2L0{};
  if (minimize) {
    // The reference asserts these (canvasManager.cpp:29, 51-52); the first
    // ROI must also start inside the canvas.
    if (overlap <= 0 || x2 - kOverlapPad < 0 || mw[0] + kOverlapPad > canvas_w || x2 - kOverlapPad - x1 < 0)
      return hipErrorInvalidValue;
    bx0 = x2 - kOverlapPad;
    bw = blend_width;
    bh = canvas_h;
    const int roi_w = blend_width - kOverlapPad;
    // full1[x1 + x] <- canvas[(x2 - 128 - x1) + x]; full2[x1 + 128 + x] <- canvas[x2 + x]
    // (make_full calls, cudaPano.inl:225-236 and :306-317).
    l0.d[0] = (x2 - kOverlapPad - x1) - x1;
    l0.ulo[0] = std::max(0, x1);
    l0.uhi[0] = std::min(bw, x1 + roi_w);
    l0.d[1] = x2 - (x1 + kOverlapPad);
    l0.ulo[1] = std::max(0, x1 + kOverlapPad);
    l0.uhi[1] = std::min(bw, x1 + kOverlapPad + roi_w);
    l0.g2_canvas = 1;
    l0.write_canvas = 1;
    l0.skip_x0 = bx0;
    l0.skip_x1 = bx0 + bw;
  } else {
    bw = canvas_w;
    bh = canvas_h;
    l0.d[0] = l0.d[1] = 0;
    l0.ulo[0] = l0.ulo[1] = 0;
    l0.uhi[0] = l0.uhi[1] = canvas_w;
    l0.g2_canvas = 0;
    l0.write_canvas = 0;
  }
  // convertMaskMat crop + convertTo(float) (cudaPano.inl:62-77).
  mask_host.resize((size_t)bw * bh);
  mask_u8_host.resize((size_t)bw * bh);
  for (int y = 0; y < bh; ++y)
    for (int x = 0; x < bw; ++x) {
      const uint8_t v = padded[(size_t)y * canvas_w + bx0 + x];
      mask_u8_host[(size_t)y * bw + x] = v;
      mask_host[(size_t)y * bw + x] = (float)v;
    }
  for (int i = 0; i < 2; ++i) {
    HM_RETURN_IF_ERROR(g0[i].alloc((size_t)bw * bh * channels * batch));
    // Columns the per-frame kernel never writes are the zero of the
    // reference's per-frame memset; clear them once.
    HM_RETURN_IF_ERROR(hipMemset(g0[i].p, 0, g0[i].bytes));
  }
  return hipSuccess;
}

hipError_t Pano2::ensure_pyramid() {
  if (hard_seam || initialized_pyramid) return hipSuccess;
  HM_RETURN_IF_ERROR(init_pyramid(kBlend2, bw, bh, mask_host));
  HM_RETURN_IF_ERROR(mask_u8.alloc(mask_u8_host.size()));
  HM_RETURN_IF_ERROR(hipMemcpy(mask_u8.p, mask_u8_host.data(), mask_u8_host.size(), hipMemcpyHostToDevice));
  pyr.mask0_fast = mask_u8.p;  // float(seam byte) == the float mask, 1 byte per pixel
  pyr.mask0_mode = kMask0U8;
  return hipSuccess;
}

hipError_t Pano2::process(const hmSurface& in1, const hmSurface& in2, const hmSurface& canvas, hipStream_t s) {
  if (!hm::surface_ok(in1, channels, true) || !hm::surface_ok(in2, channels, true) ||
      !hm::surface_ok(canvas, channels, true))
    return hipErrorInvalidValue;
  if ((int)canvas.width != canvas_w || (int)canvas.height != canvas_h) return hipErrorInvalidValue;
  const Surf s1 = to_surf(in1), s2 = to_surf(in2);
  if (hard_seam) {
    const hm::TileGrid tg = hm::tile_grid(canvas_w, canvas_h, batch, dim3(kBlock.x * kPx, kBlock.y, 1));
    const dim3 grid = tg.grid;
    Pano2Hard a{};
    a.src[0] = s1;
    a.src[1] = s2;
    for (int i = 0; i < 2; ++i) {
      a.map[i] = map[i].as<uint32_t>();
      a.mw[i] = mw[i];
      a.mh[i] = mh[i];
    }
    a.px[0] = x1;
    a.py[0] = y1;
    a.px[1] = x2;
    a.py[1] = y2;
    a.canvas = static_cast<uint8_t*>(canvas.d_ptr);
    a.cpitch = canvas.pitch;
    a.cw = canvas_w;
    a.ch = canvas_h;
    a.seam = seam.as<uint8_t>();
    a.tiles_x = tg.tiles_x;
    a.tiles_y = tg.tiles_y;
    timer.begin(s);
    if (channels == 4)
      hipLaunchKernelGGL(pano2_hard_kernel<4>, grid, kBlock, 0, s, a);
    else
      hipLaunchKernelGGL(pano2_hard_kernel<3>, grid, kBlock, 0, s, a);
    timer.end(s, "hard_seam");
    return hipGetLastError();
  }
  HM_RETURN_IF_ERROR(ensure_pyramid());
  Pano2Fused f{};
  Pano2L0& a = f.l0;
  a = l0;
  a.src[0] = s1;
  a.src[1] = s2;
  for (int i = 0; i < 2; ++i) {
    a.map[i] = map[i].as<uint32_t>();
    a.mw[i] = mw[i];
    a.mh[i] = mh[i];
    a.g0[i] = g0[i].as<uint8_t>();
  }
  a.px[0] = x1;
  a.py[0] = y1;
  a.px[1] = x2;
  a.py[1] = y2;
  a.canvas = static_cast<uint8_t*>(canvas.d_ptr);
  a.cpitch = canvas.pitch;
  a.cw = canvas_w;
  a.ch = canvas_h;
  a.bw = bw;
  a.bh = bh;
  f.lv = pyr.levels_table();
  f.n_down = std::min(5, pyr.levels - 1);
  f.roi_tiles_x = (bw + 31) / 32;
  f.roi_tiles_y = (bh + 31) / 32;
  const int strip = 64 * kPx;
  f.strips_x0 = a.write_canvas ? (a.skip_x0 + strip - 1) / strip : 0;
  f.strips_x1 = a.write_canvas ? (canvas_w - a.skip_x1 + strip - 1) / strip : 0;
  f.strips_y = (canvas_h + 3) / 4;
  const int n_strips = (f.strips_x0 + f.strips_x1) * f.strips_y;
  const dim3 fgrid((unsigned)(f.roi_tiles_x * f.roi_tiles_y + n_strips), 1u, (unsigned)batch);
  timer.begin(s);
  if (channels == 4)
    hipLaunchKernelGGL(pano2_fused_kernel<4>, fgrid, dim3(256), 0, s, f);
  else
    hipLaunchKernelGGL(pano2_fused_kernel<3>, fgrid, dim3(256), 0, s, f);
  timer.end(s, "remap_down");
  HM_RETURN_IF_ERROR(hipGetLastError());

  const void* g[2] = {g0[0].p, g0[1].p};
  hm::Level0Out out;
  out.u8 = static_cast<uint8_t*>(canvas.d_ptr);
  out.u8_pitch = canvas.pitch;
  out.u8_w = canvas_w;
  out.u8_h = canvas_h;
  out.ox = minimize ? x2 - kOverlapPad : 0;  // copy-back offset (cudaPano.inl:408)
  out.oy = 0;
  if (pyr.debug) return pyr.run(g, true, out, s, timer_or_null());  // per-level path keeping every level
  // Levels beyond the first five continue from the last level produced above.
  HM_RETURN_IF_ERROR(pyr.run_down(g, true, f.n_down, s, timer_or_null()));
  return pyr.run_up(g, true, out, s, timer_or_null());
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" int hmStitchPanoCreate(int batch_size, int num_levels, const hmControlMasks2* masks, int quiet,
                                  int minimize_blend, int max_output_width, int channels, hmStitchPano** out) {
  (void)quiet;
  (void)max_output_width;  // unused by the reference too (cudaPano.inl:27)
  if (!out) return hipErrorInvalidValue;
  auto* p = new Pano2();
  *out = p;
  p->n_images = 2;
  p->batch = batch_size;
  p->num_levels = num_levels;
  p->channels = channels;
  if (!masks) {
    p->status = hipErrorFileNotFound;
  } else if (batch_size < 1 || num_levels < 0 || num_levels > hmdev::kMaxLevels || (channels != 3 && channels != 4)) {
    p->status = hipErrorInvalidValue;
  } else {
    p->status = p->setup(*masks, minimize_blend);
  }
  return p->status;
}

extern "C" int hmStitchPanoProcess(hmStitchPano* pano, hmSurface input1, hmSurface input2, hmSurface canvas,
                                   hipStream_t stream) {
  if (!pano) return hipErrorInvalidValue;
  if (pano->status != hipSuccess) return pano->status;
  auto* p = dynamic_cast<Pano2*>(pano);
  if (!p) return hipErrorInvalidValue;
  const hipError_t e = p->process(input1, input2, canvas, stream);
  if (e != hipSuccess) pano->status = e;  // CudaStitchPano::process makes errors sticky (cudaPano.inl:468-470)
  return e;
}
