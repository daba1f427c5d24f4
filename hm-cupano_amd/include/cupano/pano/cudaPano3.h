// cudaPano3.h -- hm::pano::cuda::CudaStitchPano3<T_pipeline, T_compute>.
//
// The constructor, process() (fused = true is the reference's default
// process_optimized path), status(), canvas_width/height() and batch_size()
// of hm-cupano's src/pano/cudaPano3.h:64-120, over the gfx950 library's C ABI
// (include/hm_stitch.h).  Constructor and process errors are sticky
// (cudaPano3.inl:415-421).
#pragma once

#include <memory>
#include <type_traits>

#include "cupano/cuda/cudaStatus.h"
#include "cupano/cuda/cudaTypes.h"
#include "cupano/pano/controlMasks.h"
#include "cupano/pano/cudaMat.h"
#include "hm_stitch.h"

namespace hm {
namespace pano {
namespace cuda {

template <typename T_pipeline, typename T_compute>
class CudaStitchPano3 {
  static_assert(pano_types_ok<T_pipeline, T_compute>(),
                "CudaStitchPano3 supports <uchar4, float4>, <uchar3, float3>, <float4, float4> and <float3, float3>");

 public:
  using pipeline_type = T_pipeline;
  using compute_type = T_compute;

  CudaStitchPano3(int batch_size, int num_levels, const ControlMasks3& control_masks, bool quiet = false) {
    if (!control_masks.is_valid()) {
      status_ = CudaStatus(hipErrorFileNotFound, "Stitching masks (3-image) were not able to be loaded");
      return;
    }
    hmControlMasks3 m{};
    const HostMat<uint16_t>* cols[3] = {&control_masks.img0_col, &control_masks.img1_col, &control_masks.img2_col};
    const HostMat<uint16_t>* rows[3] = {&control_masks.img0_row, &control_masks.img1_row, &control_masks.img2_row};
    for (int i = 0; i < 3; ++i) {
      m.img_col[i] = cols[i]->data.data();
      m.img_row[i] = rows[i]->data.data();
      m.img_w[i] = cols[i]->cols;
      m.img_h[i] = cols[i]->rows;
      m.xpos[i] = control_masks.positions[i].xpos;
      m.ypos[i] = control_masks.positions[i].ypos;
    }
    m.whole_seam_mask_image = control_masks.whole_seam_mask_image.data.data();
    m.seam_w = control_masks.whole_seam_mask_image.cols;
    m.seam_h = control_masks.whole_seam_mask_image.rows;
    hmStitchPano* h = nullptr;
    const int rc = hmStitchPano3Create(batch_size, num_levels, &m, quiet, pano_channels_arg<T_pipeline>(), &h);
    handle_.reset(h);
    if (rc != hipSuccess) status_ = CudaStatus(static_cast<hipError_t>(rc), "CudaStitchPano3 construction failed");
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

  CudaStatusOr<std::unique_ptr<CudaMat<T_pipeline>>> process(const CudaMat<T_pipeline>& inputImage0,
                                                              const CudaMat<T_pipeline>& inputImage1,
                                                              const CudaMat<T_pipeline>& inputImage2,
                                                              hipStream_t stream,
                                                              std::unique_ptr<CudaMat<T_pipeline>>&& canvas,
                                                              bool fused = true) {
    CUDA_RETURN_IF_ERROR(status_);
    if (!canvas) return CudaStatus(hipErrorInvalidValue, "Canvas must be provided");
    const int rc = hmStitchPano3Process(handle_.get(), to_hm_surface(inputImage0.surface()),
                                        to_hm_surface(inputImage1.surface()), to_hm_surface(inputImage2.surface()),
                                        to_hm_surface(canvas->surface()), fused ? 1 : 0, stream);
    if (rc != hipSuccess) {
      const CudaStatus s(static_cast<hipError_t>(rc));
      status_.Update(s);
      return s;
    }
    return std::move(canvas);
  }

  hmStitchPano* native_handle() const { return handle_.get(); }

 private:
  struct Deleter {
    void operator()(hmStitchPano* p) const { hmStitchPanoDestroy(p); }
  };
  std::unique_ptr<hmStitchPano, Deleter> handle_;
  CudaStatus status_;
};

}  // namespace cuda
}  // namespace pano
}  // namespace hm
