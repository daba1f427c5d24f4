// cudaPano.h -- hm::pano::cuda::CudaStitchPano<T_pipeline, T_compute>.
//
// The constructor, process(), status(), canvas_width/height(), batch_size()
// and dump_soft_blend_pyramid() of hm-cupano's src/pano/cudaPano.h:68-115, over the gfx950
// library's C ABI (include/hm_stitch.h).  T_pipeline is uchar4 / uchar3 with
// T_compute float4 / float3, or float4 / float3 for both (the float pipeline).  Like the reference, constructor errors
// are sticky in status() and process() moves the canvas in and back out.
#pragma once

#include <memory>
#include <string>
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
class CudaStitchPano {
  static_assert(pano_types_ok<T_pipeline, T_compute>(),
                "CudaStitchPano supports <uchar4, float4>, <uchar3, float3>, <float4, float4> and <float3, float3>");

 public:
  CudaStitchPano(int batch_size, int num_levels, const ControlMasks& control_masks, bool quiet = false,
                 bool minimize_blend = true, int max_output_width = 0) {
    if (!control_masks.is_valid()) {
      status_ = CudaStatus(hipErrorFileNotFound, "Stitching masks were not able to be loaded");
      return;
    }
    hmControlMasks2 m{};
    m.img1_col = control_masks.img1_col.data.data();
    m.img1_row = control_masks.img1_row.data.data();
    m.img1_w = control_masks.img1_col.cols;
    m.img1_h = control_masks.img1_col.rows;
    m.img2_col = control_masks.img2_col.data.data();
    m.img2_row = control_masks.img2_row.data.data();
    m.img2_w = control_masks.img2_col.cols;
    m.img2_h = control_masks.img2_col.rows;
    m.whole_seam_mask_image = control_masks.whole_seam_mask_image.data.data();
    m.seam_w = control_masks.whole_seam_mask_image.cols;
    m.seam_h = control_masks.whole_seam_mask_image.rows;
    for (int i = 0; i < 2; ++i) {
      m.xpos[i] = control_masks.positions[i].xpos;
      m.ypos[i] = control_masks.positions[i].ypos;
    }
    hmStitchPano* h = nullptr;
    const int rc = hmStitchPanoCreate(batch_size, num_levels, &m, quiet, minimize_blend, max_output_width,
                                      pano_channels_arg<T_pipeline>(), &h);
    handle_.reset(h);
    if (rc != hipSuccess) status_ = CudaStatus(static_cast<hipError_t>(rc), "CudaStitchPano construction failed");
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

  CudaStatusOr<std::unique_ptr<CudaMat<T_pipeline>>> process(const CudaMat<T_pipeline>& inputImage1,
                                                              const CudaMat<T_pipeline>& inputImage2,
                                                              hipStream_t stream,
                                                              std::unique_ptr<CudaMat<T_pipeline>>&& canvas) {
    CUDA_RETURN_IF_ERROR(status_);
    if (!canvas) return CudaStatus(hipErrorInvalidValue, "Canvas must be provided");
    const int rc = hmStitchPanoProcess(handle_.get(), to_hm_surface(inputImage1.surface()),
                                       to_hm_surface(inputImage2.surface()), to_hm_surface(canvas->surface()),
                                       stream);
    if (rc != hipSuccess) {
      const CudaStatus s(static_cast<hipError_t>(rc));
      status_.Update(s);
      return s;
    }
    return std::move(canvas);
  }

  // src/pano/cudaPano.h:101 / cudaPano.inl:420-458: metadata.txt and one
  // float TIFF per level of the last processed frame.
  CudaStatus dump_soft_blend_pyramid(const std::string& directory, hipStream_t stream) const {
    CUDA_RETURN_IF_ERROR(status_);
    const int rc = hmStitchPanoDumpPyramid(handle_.get(), directory.c_str(), stream);
    if (rc != hipSuccess)
      return CudaStatus(static_cast<hipError_t>(rc), "Soft-blend pyramid is only available when num_levels > 0 "
                                                     "and after the first process");
    return CudaStatus::OkStatus();
  }

  // The library handle, for pyramid inspection (hmStitchPanoReadLevel ...).
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
