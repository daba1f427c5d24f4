// hm_pano2.hip -- CudaStitchPano (2-image) on gfx950.
//
// Reference: src/pano/cudaPano.inl (ctor :19-94, process_impl :167-418),
// src/pano/canvasManager.cpp (:16-101).  The reference runs, per frame,
// memset(canvas, full1, full2) -> remap1 -> make_full1 -> remap2 -> make_full2
// -> 34-launch Laplacian blend -> copy-back (about 44 stream operations).
// Here, per frame:
//   1. one launch remaps both images over the canvas columns that feed the
//      blend and stores the two level-0 blend inputs as exact uint8;
//   2. the rest of the canvas (which does not depend on the blend) is
//      remapped on a second stream, overlapping the pyramid;
//   3. the Gaussian levels come from one tile-fused down launch, and the
//      reconstruction from one launch per level in 2x2 quads, the level-0
//      launch writing the truncated result straight into the canvas.
// The hard seam (levels == 0) is one launch.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "hm_pano.hpp"
#include "hm_pano2_device.hpp"
#include "hm_pano_device.hpp"

using namespace hmdev;

namespace {

constexpr int kOverlapPad = 128;  // CanvasManager overlap_pad default (canvasManager.h:65)
// Calls f(C, PX) with std::integral_constant arguments for the channel count
// and the remap's pixels per thread (a tuning knob: 2, 4 or 8).
template <typename F>
void dispatch_c_px(int channels, int px, int ry, F&& f) {
  using std::integral_constant;
  auto per_ry = [&](auto CV, auto PV) {
    if (ry == 2)
      f(CV, PV, integral_constant<int, 2>{});
    else if (ry == 4)
      f(CV, PV, integral_constant<int, 4>{});
    else
      f(CV, PV, integral_constant<int, 1>{});
  };
  auto per_c = [&](auto CV) {
    if (px == 2)
      per_ry(CV, integral_constant<int, 2>{});
    else if (px == 8)
      per_ry(CV, integral_constant<int, 8>{});
    else
      per_ry(CV, integral_constant<int, 4>{});
  };
  if (channels == 4)
    per_c(integral_constant<int, 4>{});
  else
    per_c(integral_constant<int, 3>{});
}

// The blend-input remap fused into the first down launch.  A workgroup owns
// a 32x32 tile of the level-0 blend inputs: each thread remaps its 2x2 block
// for both inputs (G1 = the canvas after remap 1, G2 = after remap 2 or
// image 2 alone, as level0_tile computes them), stores the two uint8 blocks
// (the level-0 pass reads them), and takes the 2x2 mean of each straight from
// registers into the tile of level 1; levels 2.. follow in LDS as in
// down_fused_kernel.  Used when the remap of the blend columns writes nothing
// else (no canvas pixels), so this replaces that launch and the first down
// launch's read of the level-0 inputs.
struct RemapDownArgs {
  Pano2L0 r;                     // remap geometry and the G0 outputs (bw x bh)
  DownFusedArgs<uint8_t, 2> dn;  // the down pass over G0 (w0 = bw, h0 = bh)
};

template <int C>
__device__ __forceinline__ uint32_t remap_at(const Pano2L0& r, int i, int b, int X, int Y) {
  const int rx = X - r.px[i], ry = Y - r.py[i];
  const bool ok = rx >= 0 && ry >= 0 && rx < r.mw[i] && ry < r.mh[i];
  const uint32_t m = r.map[i][ok ? (size_t)ry * r.mw[i] + rx : 0];
  return remap_px_sel<C>(r.src[i], b, m, ok);
}

template <int C>
__device__ __forceinline__ bool covers(const Pano2L0& r, int i, int X, int Y) {
  const int rx = X - r.px[i], ry = Y - r.py[i];
  return rx >= 0 && ry >= 0 && rx < r.mw[i] && ry < r.mh[i];
}

template <int C>
__global__ void __launch_bounds__(256) remap_down_kernel(RemapDownArgs A) {
  __shared__ float4 s1[2][16 * 16];
  const Pano2L0& r = A.r;
  const DownFusedArgs<uint8_t, 2>& a = A.dn;
  const int t = xcd_tile((int)blockIdx.x, a.tiles_x * a.tiles_y);
  const int tx = t % a.tiles_x, ty = t / a.tiles_x;
  const int b = blockIdx.z;
  const int tid = threadIdx.x;
  const DownLevel<2>& d = a.lv[0];
  const int lx = tid & 15, ly = tid >> 4;
  const int x = tx * 16 + lx, y = ty * 16 + ly;  // level-1 pixel
  if (x < d.w && y < d.h) {
    const int u0 = 2 * x, v0 = 2 * y;  // its 2x2 block of level 0
    const bool has_x1 = u0 + 1 < a.w0, has_y1 = v0 + 1 < a.h0;
    // Remaps first, every load of the block in flight at once: image 1 at
    // the G1 column, image 2 (and image 1, for the canvas after both) at the
    // G2 column.  Pixels past the edge repeat an in-range one (not stored,
    // and ignored by down_px).
    uint32_t g1[4], g2[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int u = (e & 1) && has_x1 ? u0 + 1 : u0, v = (e >> 1) && has_y1 ? v0 + 1 : v0;
      const int X1 = u + r.d[0], X2 = u + r.d[1];
      const uint32_t c1 = remap_at<C>(r, 0, b, X1, v);
      const uint32_t i2 = remap_at<C>(r, 1, b, X2, v);
      const bool cov1 = covers<C>(r, 0, X1, v), cov2 = covers<C>(r, 1, X2, v);
      uint32_t c1b = c1;
      bool cov1b = cov1;
      if (r.d[1] != r.d[0]) {  // uniform: image 1 under G2's column too
        c1b = remap_at<C>(r, 0, b, X2, v);
        cov1b = covers<C>(r, 0, X2, v);
      }
      const uint32_t g1v = cov1 ? c1 : 0u;
      const uint32_t c12 = cov2 ? i2 : (cov1b ? c1b : 0u);
      const uint32_t g2v = r.g2_canvas ? c12 : (cov2 ? i2 : 0u);
      g1[e] = (u >= r.ulo[0] && u < r.uhi[0]) ? g1v : 0u;
      g2[e] = (u >= r.ulo[1] && u < r.uhi[1]) ? g2v : 0u;
    }
    // the uint8 blend inputs
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (((e & 1) && !has_x1) || ((e >> 1) && !has_y1)) continue;
      const int u = u0 + (e & 1), v = v0 + (e >> 1);
      const size_t o = (((size_t)b * a.h0 + v) * a.w0 + u) * C;
      store_px<C>(r.g0[0] + o, g1[e]);
      store_px<C>(r.g0[1] + o, g2[e]);
    }
    // level 1: the 2x2 means (ld_px's uint8 -> float conversion)
    const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
    float4 p[2][4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t q[2] = {g1[e], g2[e]};
#pragma unroll
      for (int i = 0; i < 2; ++i)
        p[i][e] = make_float4((float)(q[i] & 0xffu), (float)((q[i] >> 8) & 0xffu), (float)((q[i] >> 16) & 0xffu),
                              C == 4 ? (float)(q[i] >> 24) : 0.0f);
    }
    const size_t o1 = ((size_t)b * d.h + y) * d.w + x;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float4 v = down_px<C>(p[i], in_range);
      s1[i][ly * 16 + lx] = v;
      if (in_rect(d.z[i], x, y)) st_px<C>(d.g_out[i], o1, v);
    }
  }
  down_rest<C, 2>(a, tx, ty, b, tid, s1);
}

}  // namespace

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
  int roi_x0 = 0, roi_x1 = 0;  // canvas columns whose remap feeds the blend inputs
  bool fuse_ok = false;        // that remap writes no canvas pixel: it can run inside the first down launch

  hipError_t setup(const hmControlMasks2& m, int minimize_blend);
  hipError_t ensure_pyramid();
  hipError_t prepare() override { return ensure_pyramid(); }
  bool tunes_remap_px() const override { return true; }
  hipError_t process_any(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, int,
                         hipStream_t s) override {
    if (status != hipSuccess) return (hipError_t)status;
    const hipError_t e = !(n_inputs == 2 && inputs) ? hipErrorInvalidValue
                         : pipeline_f32             ? process_f32(inputs[0], inputs[1], canvas, s)
                                                    : process(inputs[0], inputs[1], canvas, s);
    if (e != hipSuccess) status = e;  // CudaStitchPano::process makes errors sticky (cudaPano.inl:468-470)
    return e;
  }
  size_t device_bytes() const override { return base_device_bytes() + map[0].bytes + map[1].bytes + seam.bytes + mask_u8.bytes; }
  hipError_t process(const hmSurface& in1, const hmSurface& in2, const hmSurface& canvas, hipStream_t s);
  hipError_t process_f32(const hmSurface& in1, const hmSurface& in2, const hmSurface& canvas, hipStream_t s);
  hipError_t run_blend(const hmSurface& canvas, hipStream_t s, int down_from = 0);
  hm::Level0Out level0_out(const hmSurface& canvas) const;  // the level-0 reconstruction into the canvas
  hipError_t launch_remap_down(const Pano2L0& a, hipStream_t s, int* n_first);
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
    roi_x0 = std::max(0, std::min(l0.ulo[0] + l0.d[0], l0.ulo[1] + l0.d[1]));
    roi_x1 = std::min(canvas_w, std::max(l0.uhi[0] + l0.d[0], l0.uhi[1] + l0.d[1]));
    fuse_ok = roi_x0 >= l0.skip_x0 && roi_x1 <= l0.skip_x1;
  } else {
    bw = canvas_w;
    bh = canvas_h;
    l0.d[0] = l0.d[1] = 0;
    l0.ulo[0] = l0.ulo[1] = 0;
    l0.uhi[0] = l0.uhi[1] = canvas_w;
    l0.g2_canvas = 0;
    l0.write_canvas = 0;
    roi_x0 = 0;
    roi_x1 = canvas_w;
    fuse_ok = true;  // no canvas writes at all
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
    // uint8 level-0 inputs, or the float pipeline's full buffers
    HM_RETURN_IF_ERROR(g0[i].alloc((size_t)bw * bh * channels * batch * (pipeline_f32 ? sizeof(float) : 1)));
    // Columns the per-frame kernel never writes are the zero of the
    // reference's per-frame memset; clear them once.
    HM_RETURN_IF_ERROR(hipMemset(g0[i].p, 0, g0[i].bytes));
  }
  // The float pipeline's blend output (the reference writes it over full
  // buffer 1: the same values, but the level-0 inputs stay for the dumps).
  if (pipeline_f32) HM_RETURN_IF_ERROR(fout.alloc((size_t)bw * bh * channels * batch * sizeof(float)));
  return hipSuccess;
}

hipError_t Pano2::ensure_pyramid() {
  if (hard_seam || initialized_pyramid) return hipSuccess;
  HM_RETURN_IF_ERROR(init_pyramid(kBlend2, bw, bh, mask_host));
  HM_RETURN_IF_ERROR(mask_u8.alloc(mask_u8_host.size()));
  HM_RETURN_IF_ERROR(hipMemcpy(mask_u8.p, mask_u8_host.data(), mask_u8_host.size(), hipMemcpyHostToDevice));
  pyr.mask0_fast = mask_u8.p;  // float(seam byte) == the float mask, 1 byte per pixel
  pyr.mask0_mode = kMask0U8;
  if (l0.write_canvas && !pipeline_f32) HM_RETURN_IF_ERROR(ensure_side());
  return hipSuccess;
}

hipError_t Pano2::process(const hmSurface& in1, const hmSurface& in2, const hmSurface& canvas, hipStream_t s) {
  if (!hm::surface_ok(in1, channels, true) || !hm::surface_ok(in2, channels, true) ||
      !hm::surface_ok(canvas, channels, true))
    return hipErrorInvalidValue;
  if ((int)canvas.width != canvas_w || (int)canvas.height != canvas_h) return hipErrorInvalidValue;
  const Surf s1 = to_surf(in1), s2 = to_surf(in2);
  if (hard_seam) {
    const int px = tuning.remap_px, ry = tuning.remap_rows;
    const hm::TileGrid tg = hm::tile_grid(canvas_w, canvas_h, batch, dim3(kBlock.x * px, kBlock.y * ry, 1));
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
    hipError_t rec = hipSuccess;
    auto go = [&](auto CV, auto PV, auto RV) {
      constexpr int kC = decltype(CV)::value, kP = decltype(PV)::value, kR = decltype(RV)::value;
      if (hm::g_recorder) {
        rec = hm::g_recorder->add(hm::op_pano2_hard<kC, kP, kR>(), grid, kBlock, s, a, "pano2_hard");
        return;
      }
      hipLaunchKernelGGL((pano2_hard_kernel<kC, kP, kR>), grid, kBlock, 0, s, a);
    };
    dispatch_c_px(channels, px, ry, go);
    if (hm::g_recorder) return rec;
    timer.end(s, "hard_seam");
    return hipGetLastError();
  }
  HM_RETURN_IF_ERROR(ensure_pyramid());
  Pano2L0 a = l0;
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
  const int px = tuning.remap_px;
  const int tile_w = (int)kBlock.x * px;
  const int ry = tuning.remap_rows;
  a.tiles_y = (canvas_h + (int)kBlock.y * ry - 1) / ((int)kBlock.y * ry);
  a.xcd = tuning.remap_xcd;
  auto launch_spans = [&](int x0a, int x1a, int x0b, int x1b, hipStream_t st, const char* name) -> hipError_t {
    Pano2L0 k = a;
    k.span_x0[0] = x0a;
    k.span_x1[0] = x1a;
    k.span_x0[1] = x0b;
    k.span_x1[1] = x1b;
    k.span_tiles[0] = std::max(0, (x1a - x0a + tile_w - 1) / tile_w);
    k.span_tiles[1] = std::max(0, (x1b - x0b + tile_w - 1) / tile_w);
    k.tiles_x = k.span_tiles[0] + k.span_tiles[1];
    if (k.tiles_x == 0) return hipSuccess;
    const int tiles = k.tiles_x * k.tiles_y;
    const int wgs = st == side && tuning.canvas_grid > 0 ? std::min(tiles, tuning.canvas_grid) : tiles;
    const dim3 grid((unsigned)wgs, 1u, (unsigned)batch);
    hipError_t rec = hipSuccess;
    auto go = [&](auto CV, auto PV, auto RV) {
      constexpr int kC = decltype(CV)::value, kP = decltype(PV)::value, kR = decltype(RV)::value;
      if (hm::g_recorder) {
        rec = hm::g_recorder->add(hm::op_pano2_level0<kC, kP, kR>(), grid, kBlock, st, k, name);
        return;
      }
      hipLaunchKernelGGL((pano2_level0_kernel<kC, kP, kR>), grid, kBlock, 0, st, k);
    };
    if (hm::g_recorder) {
      dispatch_c_px(channels, px, ry, go);
      return rec;
    }
    timer.begin(st);
    dispatch_c_px(channels, px, ry, go);
    timer.end(st, name);
    return hipGetLastError();
  };
  // Blend-input columns first, on the caller's stream: the pyramid waits on
  // them.  The canvas remap forks after them (overlap 1), so it does not
  // delay them, after the down pass (overlap 3), or at the start of the
  // frame (overlap 2): it reads nothing the
  // blend columns write, and the fork's event then sits next to the previous
  // frame's join instead of between two launches of the chain.
  // With fuse_down the columns are remapped inside the first down launch.
  const bool fork_first = side && tuning.overlap == 2;
  hipError_t e_first = hipSuccess;
  if (fork_first) {
    HM_RETURN_IF_ERROR(fork(s));
    e_first = launch_spans(0, roi_x0, roi_x1, canvas_w, side, "remap_canvas");
    if (e_first != hipSuccess) {
      (void)join(s);
      return e_first;
    }
  }
  int down_from = 0;
  const hipError_t e_roi = tuning.fuse_down && fuse_ok && !debug && pyr.levels >= 2
                               ? launch_remap_down(a, s, &down_from)
                               : launch_spans(roi_x0, roi_x1, 0, 0, s, "remap_roi");
  if (e_roi != hipSuccess) {
    if (fork_first) (void)join(s);  // the caller's stream still orders the side work
    return e_roi;
  }
  if (!side) return run_blend(canvas, s, down_from);
  if (tuning.overlap == 3 && !pyr.debug) {
    // After the down pass: the canvas remap then shares the GPU with the up
    // pass's latency-bound small levels rather than with the bandwidth-bound
    // blend columns and down pass.
    const void* g[2] = {g0[0].p, g0[1].p};
    HM_RETURN_IF_ERROR(pyr.run_down(g, true, down_from, s, timer_or_null()));
    HM_RETURN_IF_ERROR(fork(s));
    const hipError_t e1 = launch_spans(0, roi_x0, roi_x1, canvas_w, side, "remap_canvas");
    const hipError_t e2 = e1 == hipSuccess ? pyr.run_up(g, true, level0_out(canvas), s, timer_or_null()) : e1;
    const hipError_t j = join(s);
    return e2 != hipSuccess ? e2 : j;
  }
  if (fork_first) {
    const hipError_t e2 = run_blend(canvas, s, down_from);
    const hipError_t j = join(s);
    return e2 != hipSuccess ? e2 : j;
  }
  if (!tuning.overlap) {  // everything on the caller's stream
    HM_RETURN_IF_ERROR(launch_spans(0, roi_x0, roi_x1, canvas_w, s, "remap_canvas"));
    return run_blend(canvas, s, down_from);
  }
  HM_RETURN_IF_ERROR(fork(s));
  const hipError_t e1 = launch_spans(0, roi_x0, roi_x1, canvas_w, side, "remap_canvas");
  const hipError_t e2 = e1 == hipSuccess ? run_blend(canvas, s, down_from) : e1;
  // Join after the pyramid's launches (and after an error too, so the
  // caller's stream still orders the side work).
  const hipError_t j = join(s);
  return e2 != hipSuccess ? e2 : j;
}

hipError_t Pano2::launch_remap_down(const Pano2L0& a, hipStream_t s, int* n_first) {
  RemapDownArgs ra;
  ra.r = a;
  const int n_out = std::min(5, pyr.levels - 1);
  for (int i = 0; i < 2; ++i) ra.dn.g0[i] = g0[i].as<uint8_t>();
  fill_down_args(ra.dn, pyr.level_table, bw, bh, n_out);
  const dim3 grid((unsigned)(ra.dn.tiles_x * ra.dn.tiles_y), 1u, (unsigned)batch);
  if (hm::g_recorder) return hm::g_recorder->fail("remap_down");
  timer.begin(s);
  if (channels == 4)
    hipLaunchKernelGGL((remap_down_kernel<4>), grid, dim3(256), 0, s, ra);
  else
    hipLaunchKernelGGL((remap_down_kernel<3>), grid, dim3(256), 0, s, ra);
  timer.end(s, "remap_down");
  *n_first = n_out;
  return hipGetLastError();
}

// The float pipeline, in the reference's order (cudaPano.inl:167-411): the
// canvas cleared, each remap onto the canvas and its blend columns copied
// into a float full buffer (or, without minimize_blend, each image remapped
// straight into its buffer), the blend into full buffer 1, and its copy back
// into the canvas.
hipError_t Pano2::process_f32(const hmSurface& in1, const hmSurface& in2, const hmSurface& canvas, hipStream_t s) {
  const int c = channels;
  if (!hm::surface_ok_f32(in1, c) || !hm::surface_ok_f32(in2, c) || !hm::surface_ok_f32(canvas, c))
    return hipErrorInvalidValue;
  if ((int)canvas.width != canvas_w || (int)canvas.height != canvas_h) return hipErrorInvalidValue;
  if (hm::g_recorder) return hipErrorNotSupported;  // the batched pass carries the uint8 pipeline only
  const Surf s1 = to_surf(in1), s2 = to_surf(in2), cv = to_surf(canvas);
  const uint32_t* m1 = map[0].as<uint32_t>();
  const uint32_t* m2 = map[1].as<uint32_t>();
  HM_RETURN_IF_ERROR(hm::f32_zero(cv, c, batch, s));  // cudaPano.inl:186-187
  if (hard_seam) {
    // cudaPano.inl:262-276, :345-359: image 1 where the seam is 1, image 2 where it is 0
    HM_RETURN_IF_ERROR(hm::f32_remap(1, c, s1, m1, mw[0], mh[0], x1, y1, 0, 0, mw[0], mh[0], cv, 1,
                                     seam.as<uint8_t>(), batch, s));
    return hm::f32_remap(1, c, s2, m2, mw[1], mh[1], x2, y2, 0, 0, mw[1], mh[1], cv, 0, seam.as<uint8_t>(), batch, s);
  }
  HM_RETURN_IF_ERROR(ensure_pyramid());
  const Surf f1 = hm::f32_tight(g0[0].as<float>(), bw, bh, c), f2 = hm::f32_tight(g0[1].as<float>(), bw, bh, c);
  HM_RETURN_IF_ERROR(hm::f32_zero(f1, c, batch, s));
  HM_RETURN_IF_ERROR(hm::f32_zero(f2, c, batch, s));
  if (minimize) {
    const int roi_w = bw - kOverlapPad;
    // cudaPano.inl:210-237: remap 1 onto the canvas, its blend columns into full 1
    HM_RETURN_IF_ERROR(hm::f32_remap(0, c, s1, m1, mw[0], mh[0], x1, y1, 0, 0, mw[0], mh[0], cv, 0, nullptr, batch, s));
    HM_RETURN_IF_ERROR(hm::f32_copy(c, cv, roi_w, bh, x2 - kOverlapPad - x1, 0, f1, x1, 0, batch, s));
    // :291-318: remap 2 onto the canvas, its blend columns into full 2
    HM_RETURN_IF_ERROR(hm::f32_remap(0, c, s2, m2, mw[1], mh[1], x2, y2, 0, 0, mw[1], mh[1], cv, 0, nullptr, batch, s));
    HM_RETURN_IF_ERROR(hm::f32_copy(c, cv, roi_w, bh, x2, 0, f2, x1 + kOverlapPad, 0, batch, s));
  } else {
    // :242-255, :323-336
    HM_RETURN_IF_ERROR(hm::f32_remap(0, c, s1, m1, mw[0], mh[0], x1, y1, 0, 0, mw[0], mh[0], f1, 0, nullptr, batch, s));
    HM_RETURN_IF_ERROR(hm::f32_remap(0, c, s2, m2, mw[1], mh[1], x2, y2, 0, 0, mw[1], mh[1], f2, 0, nullptr, batch, s));
  }
  // The blend (cudaPano.inl:384-392), then the copy back (:401-411).
  const void* g[2] = {g0[0].p, g0[1].p};
  hm::Level0Out out;
  out.f32 = fout.as<float>();
  HM_RETURN_IF_ERROR(pyr.run(g, false, out, s, timer_or_null()));
  return hm::f32_copy(c, hm::f32_tight(fout.as<float>(), bw, bh, c), bw, bh, 0, 0, cv,
                      minimize ? x2 - kOverlapPad : 0, 0, batch, s);
}

hm::Level0Out Pano2::level0_out(const hmSurface& canvas) const {
  hm::Level0Out out;
  out.u8 = static_cast<uint8_t*>(canvas.d_ptr);
  out.u8_pitch = canvas.pitch;
  out.u8_w = canvas_w;
  out.u8_h = canvas_h;
  out.ox = minimize ? x2 - kOverlapPad : 0;  // copy-back offset (cudaPano.inl:408)
  out.oy = 0;
  return out;
}

hipError_t Pano2::run_blend(const hmSurface& canvas, hipStream_t s, int down_from) {
  const void* g[2] = {g0[0].p, g0[1].p};
  const hm::Level0Out out = level0_out(canvas);
  // Debug: the per-level path keeping every level; else the fused passes
  // (the first down_from levels already built by the fused remap).
  if (pyr.debug) return pyr.run(g, true, out, s, timer_or_null());
  HM_RETURN_IF_ERROR(pyr.run_down(g, true, down_from, s, timer_or_null()));
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
  const bool channels_ok = hm::parse_channels(channels, &p->channels, &p->pipeline_f32);
  if (!masks) {
    p->status = hipErrorFileNotFound;
  } else if (batch_size < 1 || num_levels < 0 || num_levels > hmdev::kMaxLevels || !channels_ok) {
    p->status = hipErrorInvalidValue;
  } else {
    p->status = p->setup(*masks, minimize_blend);
  }
  return p->status;
}

extern "C" int hmStitchPanoProcess(hmStitchPano* pano, hmSurface input1, hmSurface input2, hmSurface canvas,
                                   hipStream_t stream) {
  if (!pano) return hipErrorInvalidValue;
  auto* p = dynamic_cast<Pano2*>(pano);
  if (!p) return hipErrorInvalidValue;
  const hmSurface in[2] = {input1, input2};
  return p->process_any(in, 2, canvas, 1, stream);
}
