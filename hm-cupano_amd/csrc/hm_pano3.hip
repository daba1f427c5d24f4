// hm_pano3.hip -- CudaStitchPano3 (3 images) on gfx950.
//
// Reference: src/pano/cudaPano3.inl (ctor :26-110, process_impl :183-402,
// process :404-422), src/pano/cudaPano3_optimized.inl (:19-163, the default
// fused path), src/cuda/cudaFusedKernels3.cu (:25-217),
// src/pano/canvasManager3.cpp, src/pano/controlMasks3.cpp (:297-317).
// The 3-image class blends over the whole seam extent (seam width x canvas
// height).  Per frame the reference clears the canvas (RGBA only), fills three
// float "full" buffers (one fused kernel, or remap-to-canvas + box copies),
// runs the 3-image Laplacian blend and copies the result back.  Here:
//   1. one launch remaps all three images into exact uint8 level-0 blend
//      inputs (with the fused kernel's or the remap's default pixel);
//   2. canvas columns beyond the seam width (if any) get what the reference
//      leaves there, on a second stream;
//   3. the pyramid writes the reconstruction straight into the canvas.
// levels == 0: one launch of the hard seam, with the fused kernel's or the
// dest-map remaps' handling of unmapped pixels.
#include <algorithm>
#include <vector>

#include "hm_pano.hpp"
#include "hm_pano_device.hpp"

using namespace hmdev;

namespace {

struct Pano3 : hmStitchPano {
  int pos_x[3] = {}, pos_y[3] = {}, mw[3] = {}, mh[3] = {};
  hm::DevBuf map[3];
  hm::DevBuf seam;  // hard seam: canvas-sized labels
  int bw = 0, bh = 0;
  std::vector<float> mask_host;     // one-hot, bw x bh x 3
  std::vector<uint8_t> label_host;  // the same mask as one label byte per pixel
  hm::DevBuf labels;

  hipError_t setup(const hmControlMasks3& m);
  hipError_t ensure_pyramid();
  hipError_t prepare() override { return ensure_pyramid(); }
  hipError_t process_any(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, int fused,
                         hipStream_t s) override {
    if (status != hipSuccess) return (hipError_t)status;
    const hipError_t e = !(n_inputs == 3 && inputs) ? hipErrorInvalidValue
                         : pipeline_f32             ? process_f32(inputs, canvas, fused != 0, s)
                                                    : process(inputs, canvas, fused != 0, s);
    if (e != hipSuccess) status = e;  // both process paths make errors sticky (cudaPano3.inl:415-421)
    return e;
  }
  size_t device_bytes() const override { return base_device_bytes() + map[0].bytes + map[1].bytes + map[2].bytes + seam.bytes + labels.bytes; }
  hipError_t process(const hmSurface* inputs, const hmSurface& canvas, bool fused, hipStream_t s);
  hipError_t process_f32(const hmSurface* inputs, const hmSurface& canvas, bool fused, hipStream_t s);
  MultiGeom geom() const;
};

hipError_t Pano3::setup(const hmControlMasks3& m) {
  // ControlMasks3::is_valid (controlMasks3.cpp:297-301).
  if (!m.whole_seam_mask_image || m.seam_w < 1 || m.seam_h < 1) return hipErrorFileNotFound;
  for (int i = 0; i < 3; ++i)
    if (!m.img_col[i] || !m.img_row[i] || m.img_w[i] < 1 || m.img_h[i] < 1) return hipErrorFileNotFound;
  n_images = 3;
  hard_seam = num_levels == 0;
  // ControlMasks3::canvas_width/height (controlMasks3.cpp:303-317); the
  // positions become cv::Point (int) (cudaPano3.inl:49-56).
  float maxw = 0.0f, maxh = 0.0f;
  for (int i = 0; i < 3; ++i) {
    maxw = std::max(maxw, m.xpos[i] + (float)m.img_w[i]);
    maxh = std::max(maxh, m.ypos[i] + (float)m.img_h[i]);
    pos_x[i] = (int)m.xpos[i];
    pos_y[i] = (int)m.ypos[i];
    mw[i] = m.img_w[i];
    mh[i] = m.img_h[i];
  }
  canvas_w = (int)(size_t)maxw;
  canvas_h = (int)(size_t)maxh;
  for (int i = 0; i < 3; ++i) HM_RETURN_IF_ERROR(hm::upload_packed_map(m.img_col[i], m.img_row[i], mw[i], mh[i], map[i]));
  if (hard_seam) {
    // The hard-seam kernels index the seam with the canvas width
    // (cudaFusedKernels3.cu:151, cudaRemap.cu:233): it must be canvas-sized.
    if (m.seam_w != canvas_w || m.seam_h < canvas_h) return hipErrorInvalidValue;
    const size_t bytes = (size_t)canvas_w * canvas_h;
    HM_RETURN_IF_ERROR(seam.alloc(bytes));
    return hipMemcpy(seam.p, m.whole_seam_mask_image, bytes, hipMemcpyHostToDevice);
  }
  // The full buffers and the blend take the seam's width and the canvas
  // height (cudaPano3.inl:74-89; CanvasManager3::convertMaskMat crops rows).
  if (m.seam_h < canvas_h) return hipErrorInvalidValue;
  bw = m.seam_w;
  bh = canvas_h;
  mask_host.assign((size_t)bw * bh * 3, 0.0f);
  label_host.resize((size_t)bw * bh);
  for (int y = 0; y < bh; ++y)
    for (int x = 0; x < bw; ++x) {
      const uint8_t l = m.whole_seam_mask_image[(size_t)y * m.seam_w + x];
      label_host[(size_t)y * bw + x] = l;
      if (l < 3) mask_host[((size_t)y * bw + x) * 3 + l] = 1.0f;  // split_to_channels (controlMasks3.cpp:187-224)
    }
  // Cleared once: the sparse fast path writes only each image's valid region,
  // so the rest stays zero (what the dense pyramid sees there).
  const size_t es = pipeline_f32 ? sizeof(float) : 1;  // uint8 inputs, or the float pipeline's full buffers
  for (int i = 0; i < 3; ++i) {
    HM_RETURN_IF_ERROR(g0[i].alloc((size_t)bw * bh * channels * batch * es));
    HM_RETURN_IF_ERROR(hipMemset(g0[i].p, 0, g0[i].bytes));
  }
  if (pipeline_f32) HM_RETURN_IF_ERROR(fout.alloc((size_t)bw * bh * channels * batch * sizeof(float)));
  return hipSuccess;
}

hipError_t Pano3::ensure_pyramid() {
  if (hard_seam || initialized_pyramid) return hipSuccess;
  HM_RETURN_IF_ERROR(init_pyramid(kBlend3, bw, bh, mask_host));
  HM_RETURN_IF_ERROR(labels.alloc(label_host.size()));
  HM_RETURN_IF_ERROR(hipMemcpy(labels.p, label_host.data(), label_host.size(), hipMemcpyHostToDevice));
  pyr.mask0_fast = labels.p;  // one-hot of the label: the float mask's values
  pyr.mask0_mode = kMask0Label;
  // Each image's input is zero outside its box: its pyramid is sparse.
  ImgRect s0[3];
  for (int i = 0; i < 3; ++i) s0[i] = ImgRect{pos_x[i], pos_y[i], pos_x[i] + mw[i] - 1, pos_y[i] + mh[i] - 1};
  HM_RETURN_IF_ERROR(pyr.set_supports(s0));
  if (bw < canvas_w && !pipeline_f32) HM_RETURN_IF_ERROR(ensure_side());
  return hipSuccess;
}

// The float pipeline, in the reference's order: process_optimized
// (cudaPano3_optimized.inl:19-163, fused) or the remap + copy path
// (cudaPano3.inl:186-402).
hipError_t Pano3::process_f32(const hmSurface* inputs, const hmSurface& canvas, bool fused, hipStream_t s) {
  const int c = channels;
  for (int i = 0; i < 3; ++i)
    if (!hm::surface_ok_f32(inputs[i], c)) return hipErrorInvalidValue;
  if (!hm::surface_ok_f32(canvas, c) || (int)canvas.width != canvas_w || (int)canvas.height != canvas_h)
    return hipErrorInvalidValue;
  if (hm::g_recorder) return hipErrorNotSupported;  // the batched pass carries the uint8 pipeline only
  const Surf cv = to_surf(canvas);
  if (c == 4) HM_RETURN_IF_ERROR(hm::f32_zero(cv, c, batch, s));  // RGBA canvases are cleared first
  if (hard_seam) {
    if (fused)  // FusedRemapHardSeam3: unmapped pixels are left alone
      return hm::f32_hard_n(3, c, inputs, map, pos_x, pos_y, mw, mh, seam.as<uint8_t>(), canvas_w, cv, batch, s);
    for (int i = 0; i < 3; ++i)  // dest-map remaps in image order
      HM_RETURN_IF_ERROR(hm::f32_remap(1, c, to_surf(inputs[i]), map[i].as<uint32_t>(), mw[i], mh[i], pos_x[i],
                                       pos_y[i], 0, 0, mw[i], mh[i], cv, i, seam.as<uint8_t>(), batch, s));
    return hipSuccess;
  }
  HM_RETURN_IF_ERROR(ensure_pyramid());
  for (int i = 0; i < 3; ++i) {
    const Surf full = hm::f32_tight(g0[i].as<float>(), bw, bh, c);
    HM_RETURN_IF_ERROR(hm::f32_zero(full, c, batch, s));
    const Surf src = to_surf(inputs[i]);
    const uint32_t* mp = map[i].as<uint32_t>();
    if (fused) {  // FusedRemapToFullKernel3: mapped, in-range pixels of each box
      HM_RETURN_IF_ERROR(hm::f32_remap(2, c, src, mp, mw[i], mh[i], pos_x[i], pos_y[i], 0, 0, mw[i], mh[i], full, 0,
                                       nullptr, batch, s));
    } else {  // remap onto the canvas, then the image's box into its full buffer
      HM_RETURN_IF_ERROR(hm::f32_remap(0, c, src, mp, mw[i], mh[i], pos_x[i], pos_y[i], 0, 0, mw[i], mh[i], cv, 0,
                                       nullptr, batch, s));
      HM_RETURN_IF_ERROR(hm::f32_copy(c, cv, mw[i], mh[i], pos_x[i], pos_y[i], full, pos_x[i], pos_y[i], batch, s));
    }
  }
  const void* g[3] = {g0[0].p, g0[1].p, g0[2].p};
  hm::Level0Out out;
  out.f32 = fout.as<float>();
  HM_RETURN_IF_ERROR(pyr.run(g, false, out, s, timer_or_null()));
  return hm::f32_copy(c, hm::f32_tight(fout.as<float>(), bw, bh, c), bw, bh, 0, 0, cv, 0, 0, batch, s);
}

MultiGeom Pano3::geom() const {
  MultiGeom g;
  for (int i = 0; i < 3; ++i) {
    g.map[i] = map[i].as<uint32_t>();
    g.mw[i] = mw[i];
    g.mh[i] = mh[i];
    g.px[i] = pos_x[i];
    g.py[i] = pos_y[i];
    g.g0[i] = const_cast<uint8_t*>(g0[i].as<uint8_t>());
    g.z0[i] = pyr.Z[0][i];
  }
  g.bw = bw;
  g.bh = bh;
  g.seam = seam.as<uint8_t>();
  g.cw = canvas_w;
  g.ch = canvas_h;
  return g;
}

hipError_t Pano3::process(const hmSurface* inputs, const hmSurface& canvas, bool fused, hipStream_t s) {
  for (int i = 0; i < 3; ++i)
    if (!hm::surface_ok(inputs[i], channels, true)) return hipErrorInvalidValue;
  if (!hm::surface_ok(canvas, channels, true) || (int)canvas.width != canvas_w || (int)canvas.height != canvas_h)
    return hipErrorInvalidValue;
  const int clear = channels == 4;  // RGBA canvases are cleared first (cudaPano3_optimized.inl:35-39)
  if (hard_seam) {
    const MultiGeom g = geom();
    MultiRegion r;
    r.kernel = kMultiCanvas;
    r.x1 = canvas_w;
    r.y1 = canvas_h;
    r.mode = kCanvasLabel;
    // Fused (FusedRemapHardSeam3) leaves unmapped pixels alone; the dest-map
    // remaps (cudaPano3.inl:330-380) write their zero default pixel.
    r.invalid_zero = fused ? clear : 1;
    r.uncovered_zero = clear;
    timer.begin(s);
    const hipError_t e = launch_multi<3>(g, r, channels, inputs, canvas, batch, s);
    timer.end(s, "hard_seam");
    return e;
  }
  HM_RETURN_IF_ERROR(ensure_pyramid());
  const MultiGeom g = geom();  // after ensure_pyramid: it carries the blend inputs' valid regions
  MultiRegion ir;
  ir.kernel = fused ? kMultiInputsDM : kMultiInputs;
  ir.x1 = bw;
  ir.y1 = bh;
  timer.begin(s);
  HM_RETURN_IF_ERROR(launch_multi<3>(g, ir, channels, inputs, canvas, batch, s));
  timer.end(s, "remap_inputs");
  const bool extra = bw < canvas_w;
  const bool forked = extra && tuning.overlap;
  hipError_t e_side = hipSuccess;
  if (extra) {
    // Columns the blend does not write back: cleared (RGBA), and on the
    // unfused path overwritten by the remaps of every image in order.
    MultiRegion cr;
    cr.kernel = kMultiCanvas;
    cr.x0 = bw;
    cr.x1 = canvas_w;
    cr.y1 = canvas_h;
    cr.mode = fused ? kCanvasClear : kCanvasLast;
    cr.uncovered_zero = clear;
    hipStream_t st = s;
    if (forked) {
      HM_RETURN_IF_ERROR(fork(s));
      st = side;
    }
    timer.begin(st);
    e_side = launch_multi<3>(g, cr, channels, inputs, canvas, batch, st);
    timer.end(st, "remap_canvas");
  }
  const void* g0p[3] = {g0[0].p, g0[1].p, g0[2].p};
  hm::Level0Out out;
  out.u8 = static_cast<uint8_t*>(canvas.d_ptr);
  out.u8_pitch = canvas.pitch;
  out.u8_w = canvas_w;
  out.u8_h = canvas_h;
  const hipError_t e = e_side == hipSuccess ? pyr.run(g0p, true, out, s, timer_or_null()) : e_side;
  if (!forked) return e;
  // Join after the pyramid's launches (and after an error too).
  const hipError_t j = join(s);
  return e != hipSuccess ? e : j;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" int hmStitchPano3Create(int batch_size, int num_levels, const hmControlMasks3* masks, int quiet,
                                   int channels, hmStitchPano** out) {
  (void)quiet;
  if (!out) return hipErrorInvalidValue;
  auto* p = new Pano3();
  *out = p;
  p->batch = batch_size;
  p->num_levels = num_levels;
  const bool channels_ok = hm::parse_channels(channels, &p->channels, &p->pipeline_f32);
  if (!masks) {
    p->status = hipErrorFileNotFound;
  } else if (batch_size < 1 || num_levels < 0 || num_levels > kMaxLevels || !channels_ok) {
    p->status = hipErrorInvalidValue;
  } else {
    p->status = p->setup(*masks);
    // The up pass's small levels in one launch from level 3 on, as the
    // N-image class (levels 2.. in one launch is slower: 0.202 against
    // 0.189 ms per 6000x2000 frame at levels 4).
    p->tuning.tail_base = 3;
    p->apply_tuning();
  }
  return p->status;
}

extern "C" int hmStitchPano3Process(hmStitchPano* pano, hmSurface input0, hmSurface input1, hmSurface input2,
                                    hmSurface canvas, int fused, hipStream_t stream) {
  if (!pano) return hipErrorInvalidValue;
  auto* p = dynamic_cast<Pano3*>(pano);
  if (!p) return hipErrorInvalidValue;
  const hmSurface in[3] = {input0, input1, input2};
  return p->process_any(in, 3, canvas, fused, stream);
}
