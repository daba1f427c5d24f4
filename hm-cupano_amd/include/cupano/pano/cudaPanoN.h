// cudaPanoN.h -- hm::pano::cuda::CudaStitchPanoN<T_pipeline, T_compute>.
//
// The constructor (note the argument order: minimize_blend before quiet),
// process(), status(), canvas_width/height() and batch_size() of hm-cupano's
// src/pano/cudaPanoN.h:79-146, over the gfx950 library's C ABI
// (include/hm_stitch.h).  N is 2..8.  Constructor errors are sticky;
// process() validation errors are returned without being stored, as in
// cudaPanoN.inl:494-504.
#pragma once

#include <memory>
#include <type_traits>
#include <vector>

#include "cupano/cuda/cudaStatus.h"
#include "cupano/cuda/cudaTypes.h"
#include "cupano/pano/controlMasks.h"
#include "cupano/pano/cudaMat.h"
#include "hm_stitch.h"

namespace hm {
namespace pano {
namespace cuda {

template <typename T_pipeline, typename T_compute>
class CudaStitchPanoN {
  static_assert(pano_types_ok<T_pipeline, T_compute>(),
                "CudaStitchPanoN supports <uchar4, float4>, <uchar3, float3>, <float4, float4> and <float3, float3>");

 public:
  CudaStitchPanoN(int batch_size, int num_levels, const ControlMasksN& control_masks, bool minimize_blend,
                  bool quiet) {
    if (!control_masks.is_valid()) {
      status_ = CudaStatus(hipErrorFileNotFound, "Stitching masks (N-image) could not be loaded");
      return;
    }
    const int n = (int)control_masks.img_col.size();
    if (n > HM_MAX_IMAGES) {
      status_ = CudaStatus(hipErrorInvalidValue, "CudaStitchPanoN supports up to 8 images");
      return;
    }
    hmControlMasksN m{};
    m.n_images = n;
    for (int i = 0; i < n; ++i) {
      m.img_col[i] = control_masks.img_col[i].data.data();
      m.img_row[i] = control_masks.img_row[i].data.data();
      m.img_w[i] = control_masks.img_col[i].cols;
      m.img_h[i] = control_masks.img_col[i].rows;
      m.xpos[i] = control_masks.positions[i].xpos;
      m.ypos[i] = control_masks.positions[i].ypos;
    }
    m.whole_seam_mask_indexed = control_masks.whole_seam_mask_indexed.data.data();
    m.seam_w = control_masks.whole_seam_mask_indexed.cols;
    m.seam_h = control_masks.whole_seam_mask_indexed.rows;
    hmStitchPano* h = nullptr;
    const int rc = hmStitchPanoNCreate(batch_size, num_levels, &m, minimize_blend, quiet,
                                       pano_channels_arg<T_pipeline>(), &h);
    handle_.reset(h);
    n_images_ = n;
    if (rc != hipSuccess) status_ = CudaStatus(static_cast<hipError_t>(rc), "CudaStitchPanoN construction failed");
  }

  int canvas_width() const {
    int w = 0, h = 0;
    if (handle_) hmStitchPanoCanvasSize(handle_.get(), &w, &h);
    return w;
  }
  int canvas_height() const {
    int w = 0, h = 0;
    if (handle_) hmStitchPanoCanvasSize(handle_.get(), &w, &h);
    return h;
  }
  int batch_size() const { return handle_ ? hmStitchPanoBatchSize(handle_.get()) : 0; }
  const CudaStatus status() const { return status_; }

  CudaStatusOr<std::unique_ptr<CudaMat<T_pipeline>>> process(const std::vector<const CudaMat<T_pipeline>*>& inputs,
                                                              hipStream_t stream,
                                                              std::unique_ptr<CudaMat<T_pipeline>>&& canvas) {
    CUDA_RETURN_IF_ERROR(status_);
    if (!canvas) return CudaStatus(hipErrorInvalidDevicePointer, "Canvas must be provided");
    if ((int)inputs.size() != n_images_) return CudaStatus(hipErrorInvalidValue, "inputs size != N");
    std::vector<hmSurface> surfs;
    for (const auto* in : inputs) {
      if (!in || in->batch_size() != batch_size()) return CudaStatus(hipErrorInvalidValue, "Mismatched batch sizes");
      surfs.push_back(to_hm_surface(in->surface()));
    }
    if (canvas->batch_size() != batch_size()) return CudaStatus(hipErrorInvalidValue, "Canvas batch mismatch");
    const int rc = hmStitchPanoNProcess(handle_.get(), surfs.data(), (int)surfs.size(),
                                        to_hm_surface(canvas->surface()), stream);
    if (rc != hipSuccess) return CudaStatus(static_cast<hipError_t>(rc));
    return std::move(canvas);
  }

  hmStitchPano* native_handle() const { return handle_.get(); }

 private:
  struct Deleter {
    void operator()(hmStitchPano* p) const { hmStitchPanoDestroy(p); }
  };
  std::unique_ptr<hmStitchPano, Deleter> handle_;
  CudaStatus status_;
  int n_images_ = 0;
};

}  // namespace cuda
}  // namespace pano
}  // namespace hm
