// hm_panoN.hip -- CudaStitchPanoN (2..8 images) on gfx950.
//
// Reference: src/pano/cudaPanoN.inl (ctor :108-370, process :489-685),
// src/pano/canvasManagerN.h (:15-79), src/pano/controlMasksN.cpp (:195-232).
// Per frame the reference clears the canvas, runs the N-image hard seam over
// it, remaps every image into a float buffer over the blend ROI, runs the
// N-image Laplacian blend (about 7 launches per level) and copies the write
// ROI back.  Here:
//   1. one launch remaps all N images over the blend ROI into exact uint8
//      level-0 blend inputs (zero outside each image's box);
//   2. one launch on a second stream writes the canvas outside the write ROI
//      (what the clear + hard seam leave there), overlapping the pyramid;
//   3. the pyramid (tile-fused down, per-level quad up, one-hot level-0 mask
//      read as one label byte per pixel) writes the write ROI of the
//      reconstruction straight into the canvas.
// levels == 0: one launch, the hard seam with the cleared default.
#include <algorithm>
#include <vector>

#include "hm_pano.hpp"
#include "hm_pano_device.hpp"

using namespace hmdev;

namespace {

constexpr int kOverlapPadN = 128;  // CanvasManagerN overlap_pad default (canvasManagerN.h:15)

struct Rect {
  int x = 0, y = 0, w = 0, h = 0;
  bool empty() const { return w <= 0 || h <= 0; }
};

// detailN::pyramid_margin / pyramid_alignment (cudaPanoN.inl:20-36).
int pyramid_margin(int levels) { return levels <= 0 ? 0 : 1 << std::min(levels, 30); }
int pyramid_alignment(int levels) { return levels <= 1 ? 1 : 1 << std::min(levels - 1, 30); }

// detailN::seam_boundary_bbox (cudaPanoN.inl:38-77): the box around every
// pixel whose label differs from its right or lower neighbour, together with
// that neighbour.
bool seam_boundary_bbox(const std::vector<uint8_t>& s, int w, int h, Rect& out) {
  if (w <= 1 || h <= 1) return false;
  int x0 = w, y0 = h, x1 = -1, y1 = -1;
  for (int y = 0; y < h; ++y) {
    const uint8_t* r = &s[(size_t)y * w];
    const uint8_t* d = y + 1 < h ? r + w : nullptr;
    for (int x = 0; x < w; ++x) {
      const bool right = x + 1 < w && r[x] != r[x + 1];
      const bool down = d && r[x] != d[x];
      if (!right && !down) continue;
      x0 = std::min(x0, x);
      y0 = std::min(y0, y);
      x1 = std::max(x1, right ? x + 1 : x);
      y1 = std::max(y1, down ? y + 1 : y);
    }
  }
  if (x1 < 0 || y1 < 0) return false;
  out = Rect{x0, y0, x1 - x0 + 1, y1 - y0 + 1};
  return true;
}

// detailN::expand_and_clamp / align_and_clamp (cudaPanoN.inl:79-105).
Rect expand_and_clamp(const Rect& r, int pad, int max_w, int max_h) {
  if (r.empty()) return Rect{};
  const int x0 = std::max(0, r.x - pad), y0 = std::max(0, r.y - pad);
  const int x1 = std::min(max_w, r.x + r.w + pad), y1 = std::min(max_h, r.y + r.h + pad);
  return Rect{x0, y0, std::max(0, x1 - x0), std::max(0, y1 - y0)};
}
Rect align_and_clamp(const Rect& r, int align, int max_w, int max_h) {
  if (r.empty()) return Rect{};
  if (align <= 1) return r;
  const int x0 = std::max(0, (r.x / align) * align), y0 = std::max(0, (r.y / align) * align);
  const int x1 = std::min(max_w, ((r.x + r.w + align - 1) / align) * align);
  const int y1 = std::min(max_h, ((r.y + r.h + align - 1) / align) * align);
  return Rect{x0, y0, std::max(0, x1 - x0), std::max(0, y1 - y0)};
}

struct PanoN : hmStitchPano {
  int pos_x[kMaxImages] = {}, pos_y[kMaxImages] = {}, mw[kMaxImages] = {}, mh[kMaxImages] = {};
  hm::DevBuf map[kMaxImages];
  hm::DevBuf seam;  // padded labels, canvas-sized (hard seam and the canvas outside the write ROI)
  bool use_min = false;
  Rect write_roi, blend_roi;
  int bx0 = 0, by0 = 0, bw = 0, bh = 0;
  std::vector<float> mask_host;      // one-hot, bw x bh x n
  std::vector<uint8_t> label_host;   // the same mask as one label byte per pixel
  hm::DevBuf labels;

  hipError_t setup(const hmControlMasksN& m, int minimize_blend);
  hipError_t ensure_pyramid();
  hipError_t prepare() override { return ensure_pyramid(); }
  hipError_t process_any(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, int,
                         hipStream_t s) override {
    if (status != hipSuccess) return (hipError_t)status;
    // CudaStitchPanoN::process: errors are not sticky
    return pipeline_f32 ? process_f32(inputs, n_inputs, canvas, s) : process(inputs, n_inputs, canvas, s);
  }
  size_t device_bytes() const override {
    size_t b = base_device_bytes() + seam.bytes + labels.bytes;
    for (const auto& m : map) b += m.bytes;
    return b;
  }
  hipError_t process(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, hipStream_t s);
  hipError_t process_f32(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, hipStream_t s);
  template <int N>
  hipError_t launch_region(const MultiRegion& r, const hmSurface* inputs, const hmSurface& canvas, hipStream_t s,
                           const char* name);
};

hipError_t PanoN::setup(const hmControlMasksN& m, int minimize_blend) {
  const int n = m.n_images;
  if (n < 2 || n > kMaxImages) return hipErrorInvalidValue;  // cudaBlendN dispatches N in [2, 8]
  // ControlMasksN::is_valid (controlMasksN.cpp:195-203).
  if (!m.whole_seam_mask_indexed || m.seam_w < 1 || m.seam_h < 1) return hipErrorFileNotFound;
  for (int i = 0; i < n; ++i)
    if (!m.img_col[i] || !m.img_row[i] || m.img_w[i] < 1 || m.img_h[i] < 1) return hipErrorFileNotFound;
  n_images = n;
  hard_seam = num_levels == 0;
  // ControlMasksN::canvas_width/height: float max, then size_t
  // (controlMasksN.cpp:205-219); positions become cv::Point (int).
  float maxw = 0.0f, maxh = 0.0f;
  for (int i = 0; i < n; ++i) {
    maxw = std::max(maxw, m.xpos[i] + (float)m.img_w[i]);
    maxh = std::max(maxh, m.ypos[i] + (float)m.img_h[i]);
    pos_x[i] = (int)m.xpos[i];
    pos_y[i] = (int)m.ypos[i];
    mw[i] = m.img_w[i];
    mh[i] = m.img_h[i];
  }
  canvas_w = (int)(size_t)maxw;
  canvas_h = (int)(size_t)maxh;
  for (int i = 0; i < n; ++i) HM_RETURN_IF_ERROR(hm::upload_packed_map(m.img_col[i], m.img_row[i], mw[i], mh[i], map[i]));
  // CanvasManagerN::convertMaskMat (canvasManagerN.h:33-55).
  std::vector<uint8_t> padded;
  if (!hm::pad_replicate(m.whole_seam_mask_indexed, m.seam_w, m.seam_h, canvas_w, canvas_h, padded))
    return hipErrorInvalidValue;
  HM_RETURN_IF_ERROR(seam.alloc(padded.size()));
  HM_RETURN_IF_ERROR(hipMemcpy(seam.p, padded.data(), padded.size(), hipMemcpyHostToDevice));
  if (hard_seam) return hipSuccess;

  // ROI choice of the ctor (cudaPanoN.inl:151-190).
  const bool minimize = minimize_blend != 0;  // minimize_blend_ = minimize_blend && num_levels > 0
  Rect bb;
  if (minimize && seam_boundary_bbox(padded, canvas_w, canvas_h, bb)) {
    write_roi = expand_and_clamp(bb, kOverlapPadN, canvas_w, canvas_h);
    blend_roi = expand_and_clamp(write_roi, pyramid_margin(num_levels), canvas_w, canvas_h);
    blend_roi = align_and_clamp(blend_roi, pyramid_alignment(num_levels), canvas_w, canvas_h);
  }
  // process: use_minimized_blend (cudaPanoN.inl:541-543).
  use_min = minimize && !blend_roi.empty() && !write_roi.empty();
  bx0 = use_min ? blend_roi.x : 0;
  by0 = use_min ? blend_roi.y : 0;
  bw = use_min ? blend_roi.w : canvas_w;
  bh = use_min ? blend_roi.h : canvas_h;
  // split_to_channels + convertTo(float) of the (cropped) seam
  // (controlMasksN.cpp:221-232, cudaPanoN.inl:192-194).
  mask_host.assign((size_t)bw * bh * n, 0.0f);
  label_host.resize((size_t)bw * bh);
  for (int y = 0; y < bh; ++y)
    for (int x = 0; x < bw; ++x) {
      const uint8_t l = padded[(size_t)(by0 + y) * canvas_w + bx0 + x];
      label_host[(size_t)y * bw + x] = l;
      if (l < n) mask_host[((size_t)y * bw + x) * n + l] = 1.0f;
    }
  // Cleared once: the sparse fast path writes only each image's valid region,
  // so the rest stays zero (what the dense pyramid sees there).
  const size_t es = pipeline_f32 ? sizeof(float) : 1;  // uint8 inputs, or the float pipeline's full buffers
  for (int i = 0; i < n; ++i) {
    HM_RETURN_IF_ERROR(g0[i].alloc((size_t)bw * bh * channels * batch * es));
    HM_RETURN_IF_ERROR(hipMemset(g0[i].p, 0, g0[i].bytes));
  }
  if (pipeline_f32) HM_RETURN_IF_ERROR(fout.alloc((size_t)bw * bh * channels * batch * sizeof(float)));
  return hipSuccess;
}

hipError_t PanoN::ensure_pyramid() {
  if (hard_seam || initialized_pyramid) return hipSuccess;
  HM_RETURN_IF_ERROR(init_pyramid(kBlendN, bw, bh, mask_host));
  HM_RETURN_IF_ERROR(labels.alloc(label_host.size()));
  HM_RETURN_IF_ERROR(hipMemcpy(labels.p, label_host.data(), label_host.size(), hipMemcpyHostToDevice));
  pyr.mask0_fast = labels.p;  // one-hot of the label: the float mask's values
  pyr.mask0_mode = kMask0Label;
  // Each image's input is zero outside its box: its pyramid is sparse.
  ImgRect s0[kMaxImages];
  for (int i = 0; i < n_images; ++i)
    s0[i] = ImgRect{pos_x[i] - bx0, pos_y[i] - by0, pos_x[i] + mw[i] - 1 - bx0, pos_y[i] + mh[i] - 1 - by0};
  HM_RETURN_IF_ERROR(pyr.set_supports(s0));
  if (use_min && !pipeline_f32) HM_RETURN_IF_ERROR(ensure_side());
  return hipSuccess;
}

// The float pipeline, in the reference's order (cudaPanoN.inl:489-685).
hipError_t PanoN::process_f32(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, hipStream_t s) {
  const int c = channels;
  if (!canvas.d_ptr) return hipErrorInvalidDevicePointer;
  if (!inputs || n_inputs != n_images) return hipErrorInvalidValue;
  for (int i = 0; i < n_images; ++i)
    if (!hm::surface_ok_f32(inputs[i], c)) return hipErrorInvalidValue;
  if (!hm::surface_ok_f32(canvas, c) || (int)canvas.width != canvas_w || (int)canvas.height != canvas_h)
    return hipErrorInvalidValue;
  if (hm::g_recorder) return hipErrorNotSupported;  // the batched pass carries the uint8 pipeline only
  const Surf cv = to_surf(canvas);
  if (hard_seam || use_min) {  // cleared canvas, hard seam (:508-541, :545-588)
    HM_RETURN_IF_ERROR(hm::f32_zero(cv, c, batch, s));
    HM_RETURN_IF_ERROR(hm::f32_hard_n(n_images, c, inputs, map, pos_x, pos_y, mw, mh, seam.as<uint8_t>(), canvas_w,
                                      cv, batch, s));
    if (hard_seam) return hipSuccess;
  }
  HM_RETURN_IF_ERROR(ensure_pyramid());
  for (int i = 0; i < n_images; ++i) {
    const Surf full = hm::f32_tight(g0[i].as<float>(), bw, bh, c);
    HM_RETURN_IF_ERROR(hm::f32_zero(full, c, batch, s));
    const Surf src = to_surf(inputs[i]);
    const uint32_t* mp = map[i].as<uint32_t>();
    if (use_min) {  // only the image's intersection with the blend ROI (:590-618)
      const int ix0 = std::max(pos_x[i], blend_roi.x), iy0 = std::max(pos_y[i], blend_roi.y);
      const int ix1 = std::min(pos_x[i] + mw[i], blend_roi.x + blend_roi.w);
      const int iy1 = std::min(pos_y[i] + mh[i], blend_roi.y + blend_roi.h);
      if (ix1 > ix0 && iy1 > iy0)
        HM_RETURN_IF_ERROR(hm::f32_remap(0, c, src, mp, mw[i], mh[i], pos_x[i] - blend_roi.x, pos_y[i] - blend_roi.y,
                                         ix0 - pos_x[i], iy0 - pos_y[i], ix1 - ix0, iy1 - iy0, full, 0, nullptr, batch,
                                         s));
    } else {  // remap_soft (:644-660)
      HM_RETURN_IF_ERROR(hm::f32_remap(0, c, src, mp, mw[i], mh[i], pos_x[i], pos_y[i], 0, 0, mw[i], mh[i], full, 0,
                                       nullptr, batch, s));
    }
  }
  const void* g[kMaxImages];
  for (int i = 0; i < n_images; ++i) g[i] = g0[i].p;
  hm::Level0Out out;
  out.f32 = fout.as<float>();
  HM_RETURN_IF_ERROR(pyr.run(g, false, out, s, timer_or_null()));
  const Surf fo = hm::f32_tight(fout.as<float>(), bw, bh, c);
  if (use_min)  // the write ROI back into the canvas (:628-640)
    return hm::f32_copy(c, fo, write_roi.w, write_roi.h, write_roi.x - blend_roi.x, write_roi.y - blend_roi.y, cv,
                        write_roi.x, write_roi.y, batch, s);
  return hm::f32_copy(c, fo, bw, bh, 0, 0, cv, 0, 0, batch, s);  // (:668-681)
}

template <int N>
hipError_t PanoN::launch_region(const MultiRegion& r, const hmSurface* inputs, const hmSurface& canvas, hipStream_t s,
                                const char* name) {
  MultiGeom g;
  for (int i = 0; i < N; ++i) {
    g.map[i] = map[i].as<uint32_t>();
    g.mw[i] = mw[i];
    g.mh[i] = mh[i];
    g.px[i] = pos_x[i];
    g.py[i] = pos_y[i];
    g.g0[i] = g0[i].as<uint8_t>();
    g.z0[i] = pyr.Z[0][i];
  }
  g.bw = bw;
  g.bh = bh;
  g.seam = seam.as<uint8_t>();
  g.cw = canvas_w;
  g.ch = canvas_h;
  timer.begin(s);
  const hipError_t e = launch_multi<N>(g, r, channels, inputs, canvas, batch, s);
  timer.end(s, name);
  return e;
}

hipError_t PanoN::process(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, hipStream_t s) {
  // Validation errors of CudaStitchPanoN::process (cudaPanoN.inl:494-504).
  if (!canvas.d_ptr) return hipErrorInvalidDevicePointer;
  if (!inputs || n_inputs != n_images) return hipErrorInvalidValue;
  for (int i = 0; i < n_images; ++i)
    if (!hm::surface_ok(inputs[i], channels, true)) return hipErrorInvalidValue;
  if (!hm::surface_ok(canvas, channels, true) || (int)canvas.width != canvas_w || (int)canvas.height != canvas_h)
    return hipErrorInvalidValue;
  auto per_n = [&](auto NV) -> hipError_t {
    constexpr int N = decltype(NV)::value;
    // The canvas: clear + hard seam (cudaPanoN.inl:508-541, :545-560).
    MultiRegion cr;
    cr.kernel = kMultiCanvas;
    cr.x1 = canvas_w;
    cr.y1 = canvas_h;
    if (hard_seam) return launch_region<N>(cr, inputs, canvas, s, "hard_seam");
    HM_RETURN_IF_ERROR(ensure_pyramid());
    MultiRegion ir;
    ir.kernel = kMultiInputs;
    ir.x0 = bx0;
    ir.x1 = bx0 + bw;
    ir.y0 = by0;
    ir.y1 = by0 + bh;
    HM_RETURN_IF_ERROR(launch_region<N>(ir, inputs, canvas, s, "remap_inputs"));
    hipError_t e_side = hipSuccess;
    const bool forked = use_min && tuning.overlap;
    if (use_min) {  // outside the write ROI, which the blend writes back
      // The bands above and below the ROI and the strips left and right of
      // it, one launch each: no workgroup is spent on the ROI itself.
      const int rx0 = write_roi.x, ry0 = write_roi.y, rx1 = write_roi.x + write_roi.w, ry1 = write_roi.y + write_roi.h;
      const int bands[4][4] = {{0, canvas_w, 0, ry0}, {0, canvas_w, ry1, canvas_h}, {0, rx0, ry0, ry1},
                               {rx1, canvas_w, ry0, ry1}};  // x0, x1, y0, y1
      hipStream_t cs = s;
      if (forked) {
        HM_RETURN_IF_ERROR(fork(s));
        cs = side;
      }
      for (const auto& bd : bands) {
        MultiRegion br = cr;
        br.x0 = bd[0];
        br.x1 = bd[1];
        br.y0 = bd[2];
        br.y1 = bd[3];
        if (br.x1 <= br.x0 || br.y1 <= br.y0) continue;
        const hipError_t e = launch_region<N>(br, inputs, canvas, cs, "remap_canvas");
        if (e != hipSuccess) {
          if (!forked) return e;
          e_side = e;
          break;
        }
      }
    }
    const void* g[kMaxImages];
    for (int i = 0; i < n_images; ++i) g[i] = g0[i].p;
    hm::Level0Out out;
    out.u8 = static_cast<uint8_t*>(canvas.d_ptr);
    out.u8_pitch = canvas.pitch;
    out.u8_w = canvas_w;
    out.u8_h = canvas_h;
    out.ox = bx0;
    out.oy = by0;
    if (use_min) {  // copy-back of the write ROI (cudaPanoN.inl:628-640)
      out.win_x = write_roi.x - bx0;
      out.win_y = write_roi.y - by0;
      out.win_w = write_roi.w;
      out.win_h = write_roi.h;
    }
    const hipError_t e = e_side == hipSuccess ? pyr.run(g, true, out, s, timer_or_null()) : e_side;
    if (!forked) return e;
    // Join after the pyramid's launches (and after an error too).
    const hipError_t j = join(s);
    return e != hipSuccess ? e : j;
  };
  using std::integral_constant;
  switch (n_images) {
    case 2: return per_n(integral_constant<int, 2>{});
    case 3: return per_n(integral_constant<int, 3>{});
    case 4: return per_n(integral_constant<int, 4>{});
    case 5: return per_n(integral_constant<int, 5>{});
    case 6: return per_n(integral_constant<int, 6>{});
    case 7: return per_n(integral_constant<int, 7>{});
    case 8: return per_n(integral_constant<int, 8>{});
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" int hmStitchPanoNCreate(int batch_size, int num_levels, const hmControlMasksN* masks, int minimize_blend,
                                   int quiet, int channels, hmStitchPano** out) {
  (void)quiet;
  if (!out) return hipErrorInvalidValue;
  auto* p = new PanoN();
  *out = p;
  p->batch = batch_size;
  p->num_levels = num_levels;
  const bool channels_ok = hm::parse_channels(channels, &p->channels, &p->pipeline_f32);
  if (!masks) {
    p->status = hipErrorFileNotFound;
  } else if (batch_size < 1 || num_levels < 0 || num_levels > kMaxLevels || !channels_ok) {
    p->status = hipErrorInvalidValue;
  } else {
    p->status = p->setup(*masks, minimize_blend && num_levels > 0);
    // Levels 1 and 0 in one launch by default for three images only: it
    // holds more LDS per workgroup (R1 and every slot's G1 and G2 taps), which
    // pays where tiles meet few images (3 x 3840x2160 -> 6200x2300: 51 us
    // against 57 us for the two launches) and not with eight (14000x2600:
    // 252 us against 213 us).  The small levels from level 3 on in one launch
    // (three or four images): 0.152 ms per 6200x2300 frame against 0.156 ms
    // with a launch per level, 0.167 ms from level 2.  The autotuner tries
    // every setting.
    p->tuning.fuse_up = p->n_images <= 3 ? 1 : 0;
    p->tuning.tail_base = 3;
    p->apply_tuning();
  }
  return p->status;
}

extern "C" int hmStitchPanoNProcess(hmStitchPano* pano, const hmSurface* inputs, int num_inputs, hmSurface canvas,
                                    hipStream_t stream) {
  if (!pano) return hipErrorInvalidValue;
  auto* p = dynamic_cast<PanoN*>(pano);
  if (!p) return hipErrorInvalidValue;
  return p->process_any(inputs, num_inputs, canvas, 1, stream);
}
