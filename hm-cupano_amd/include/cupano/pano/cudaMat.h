// cudaMat.h -- hm::CudaMat<T>, a batched device image (B x H x W of T).
//
// The accessors of hm-cupano's src/pano/cudaMat.h:354-498 (width, height,
// batch_size, channels, pitch, size, surface, data, data_raw); OpenCV is not
// used, so upload/download take raw host buffers (tight rows of W pixels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <utility>

#include "cupano/cuda/cudaStatus.h"
#include "cupano/cuda/cudaTypes.h"

namespace hm {

template <typename T>
class CudaMat {
 public:
  using element_type = T;
  using base_element_type = BaseScalar_t<T>;

  CudaMat(const CudaMat&) = delete;
  CudaMat& operator=(const CudaMat&) = delete;

  // Allocates B images of W x H, tightly pitched (cudaMat.h:449-451).
  CudaMat(int B, int W, int H) : rows_(H), cols_(W), batch_(B) {
    if (hipMalloc(reinterpret_cast<void**>(&d_data_), size()) != hipSuccess) d_data_ = nullptr;
  }
  // Borrows existing device memory (not freed by the destructor).
  CudaMat(T* dataptr, int B, int W, int H, int pitch = 0)
      : d_data_(dataptr), rows_(H), cols_(W), pitch_(pitch), batch_(B), owns_(false) {}
  // Allocates and uploads one host image (tight rows).
  CudaMat(const T* host, int W, int H) : CudaMat(1, W, H) {
    if (d_data_ && upload(host) != hipSuccess) {
      (void)hipFree(d_data_);
      d_data_ = nullptr;
    }
  }
  ~CudaMat() {
    if (owns_ && d_data_) (void)hipFree(d_data_);
  }

  bool is_valid() const { return d_data_ != nullptr; }

  hipError_t upload(const T* host, int batch_item = 0, hipStream_t stream = nullptr) {
    const hipError_t e = hipMemcpy2DAsync(item_ptr(batch_item), pitch(), host, sizeof(T) * cols_,
                                          sizeof(T) * cols_, rows_, hipMemcpyHostToDevice, stream);
    return e != hipSuccess ? e : hipStreamSynchronize(stream);
  }
  hipError_t download(T* host, int batch_item = 0, hipStream_t stream = nullptr) const {
    const hipError_t e = hipMemcpy2DAsync(host, sizeof(T) * cols_, item_ptr(batch_item), pitch(), sizeof(T) * cols_,
                                          rows_, hipMemcpyDeviceToHost, stream);
    return e != hipSuccess ? e : hipStreamSynchronize(stream);
  }

  T* data() { return d_data_; }
  const T* data() const { return d_data_; }
  base_element_type* data_raw() { return reinterpret_cast<base_element_type*>(d_data_); }
  const base_element_type* data_raw() const { return reinterpret_cast<const base_element_type*>(d_data_); }

  int width() const { return cols_; }
  int height() const { return rows_; }
  int batch_size() const { return batch_; }
  constexpr int channels() const { return num_channels_v<T>; }
  int pitch() const { return pitch_ ? pitch_ : (int)sizeof(T) * cols_; }
  size_t size() const { return (size_t)pitch() * rows_ * batch_; }

  CudaSurface<T> surface() const {
    return CudaSurface<T>{d_data_, (uint32_t)cols_, (uint32_t)rows_, (uint32_t)pitch()};
  }
  operator CudaSurface<T>() const { return surface(); }

 private:
  void* item_ptr(int b) const { return reinterpret_cast<char*>(d_data_) + (size_t)b * pitch() * rows_; }
  T* d_data_ = nullptr;
  int rows_ = 0, cols_ = 0, pitch_ = 0, batch_ = 0;
  bool owns_ = true;
};

}  // namespace hm
