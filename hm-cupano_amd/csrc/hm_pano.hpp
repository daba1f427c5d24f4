// hm_pano.hpp -- state shared by the 2-, 3- and N-image pano objects.
#pragma once

#include <atomic>
#include <string>
#include <vector>

#include "hm_internal.hpp"
#include "hm_stitch.h"

struct hmStitchPano {
  virtual ~hmStitchPano();

  int status = hipSuccess;  // sticky, like CudaStitchPano::status_
  int n_images = 0;
  int batch = 1;
  int num_levels = 0;
  int channels = 4;
  // Float pipeline (CudaStitchPano*<float4|float3, ...>): float inputs and
  // canvas; g0 then holds the float full buffers of the reference.
  bool pipeline_f32 = false;
  hm::DevBuf fout;  // float pipeline: the blend's output (N- and 3-image classes)
  int canvas_w = 0, canvas_h = 0;
  bool hard_seam = false;
  bool debug = false;
  bool initialized_pyramid = false;

  hm::Pyramid pyr;                       // soft-seam pyramid (levels >= 1 owned)
  hm::DevBuf g0[hmdev::kMaxImages];      // level-0 images, uint8 tight, resident
  hm::DevBuf mask0;                      // level-0 mask, float, resident
  hm::LaunchTimer timer;

  hm::LaunchTimer* timer_or_null() { return timer.enabled() ? &timer : nullptr; }

  // Launch-shape knobs (hmTuning); every setting computes the same output.
  hmTuning tuning{4, 2, 1, 0, 1, 0, 0, 1};
  // Set while a server holds the object (hmServerAddRig .. RemoveRig/Destroy):
  // its scheduler thread then owns the launches and the captured graphs bake
  // in the launch shapes, so tuning and batch recording are refused.
  std::atomic<bool> served{false};
  // Which knobs matter for this object (the autotuner sweeps only those).
  virtual bool tunes_remap_px() const { return false; }
  bool tunes_tail() const {
    return !hard_seam && num_levels >= 4 && n_images <= hm::kMaxTailImages;
  }
  bool tunes_overlap() const { return side != nullptr; }
  bool tunes_fuse_up() const {
    return !hard_seam && ((n_images == 2 && num_levels >= 2) ||
                          (pyr.kind == hmdev::kBlendN && n_images > 2 && num_levels >= 3 && !pipeline_f32));
  }
  void apply_tuning() {
    pyr.tail_base = tuning.tail_base;
    pyr.fuse_up = tuning.fuse_up != 0;
  }

  // Work that does not feed the blend (the canvas outside the written-back
  // region) runs on a second stream, forked from and joined back into the
  // caller's stream, so it overlaps the pyramid's latency-bound launches.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipError_t ensure_side();
  hipError_t fork(hipStream_t s);  // side waits for the caller's stream
  hipError_t join(hipStream_t s);  // the caller's stream waits for side

  // Allocates the pyramid and the resident mask pyramid for a bw x bh blend
  // region; mask_host is bw x bh x M float.
  hipError_t init_pyramid(int kind, int bw, int bh, const std::vector<float>& mask_host);

  // Allocates everything a soft-seam process needs (done by the first process
  // otherwise), so that later calls allocate nothing.
  virtual hipError_t prepare() = 0;
  // Device memory held by the object.
  virtual size_t device_bytes() const = 0;
  // One frame with the class's own error rules (sticky or not).  `fused`
  // selects the 3-image class's path and is ignored by the others.
  virtual hipError_t process_any(const hmSurface* inputs, int n_inputs, const hmSurface& canvas, int fused,
                                 hipStream_t s) = 0;
  size_t base_device_bytes() const {
    size_t b = mask0.bytes + pyr.allocation_size + fout.bytes;
    for (const auto& g : g0) b += g.bytes;
    return b;
  }
};

namespace hm {

// Packs a CV_16U map pair into one 32-bit word per pixel (x | y << 16) so the
// remap reads one dword per lane; the values are unchanged.
hipError_t upload_packed_map(const uint16_t* map_x, const uint16_t* map_y, int w, int h, DevBuf& dst);

// cv::copyMakeBorder(BORDER_REPLICATE) on the right/bottom edges, as
// CanvasManager::convertMaskMat (src/pano/canvasManager.cpp:63-101).
bool pad_replicate(const uint8_t* src, int sw, int sh, int cw, int ch, std::vector<uint8_t>& out);

bool surface_ok(const hmSurface& s, int channels, bool need_vec_align);

// ---- float pipelines (hm_pano_f32.hip) ------------------------------------
// channels | kPipelineF32 selects the float pipeline at construction.
constexpr int kPipelineF32 = HM_PIPELINE_F32;
// Splits a constructor's channels argument: false when it is not 3 or 4
// (optionally with kPipelineF32).
inline bool parse_channels(int arg, int* channels, bool* f32) {
  *f32 = (arg & kPipelineF32) != 0;
  *channels = arg & ~kPipelineF32;
  return *channels == 3 || *channels == 4;
}
// A float surface is valid for c channels: pointer, pitch and 4-byte alignment.
bool surface_ok_f32(const hmSurface& s, int channels);
hmdev::Surf f32_tight(float* p, int w, int h, int channels);
hipError_t f32_zero(const hmdev::Surf& s, int channels, int batch, hipStream_t st);
// mode 0: BatchedRemapKernelExOffset(ROI); 1: ...WithDestMap (labels are the
// canvas-wide seam, image_index the label written); 2: mapped pixels only
// (FusedRemapToFullKernel3).  roi is in map coordinates.
hipError_t f32_remap(int mode, int channels, const hmdev::Surf& src, const uint32_t* map, int mw, int mh, int off_x,
                     int off_y, int roi_x, int roi_y, int roi_w, int roi_h, const hmdev::Surf& dst, int image_index,
                     const uint8_t* labels, int batch, hipStream_t st);
hipError_t f32_hard_n(int n, int channels, const hmSurface* inputs, const DevBuf* maps, const int* px, const int* py,
                      const int* mw, const int* mh, const uint8_t* labels, int labels_w, const hmdev::Surf& dst,
                      int batch, hipStream_t st);
hipError_t f32_copy(int channels, const hmdev::Surf& src, int region_w, int region_h, int sx, int sy,
                    const hmdev::Surf& dst, int ox, int oy, int batch, hipStream_t st);

}  // namespace hm
