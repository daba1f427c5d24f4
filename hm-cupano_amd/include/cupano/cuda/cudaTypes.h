// cudaTypes.h -- pitched device image descriptor and pixel traits.
//
// CudaSurface<T> keeps the field names of hm-cupano's src/cuda/cudaTypes.h:195-201
// (d_ptr, width, height, pitch in bytes); BaseScalar_t<T> is the channel type.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "hm_stitch.h"

template <typename T>
struct CudaSurface final {
  T* d_ptr;
  uint32_t width;
  uint32_t height;
  uint32_t pitch;
};

template <typename T>
struct BaseScalar {
  using type = T;
};
template <> struct BaseScalar<uchar3> { using type = unsigned char; };
template <> struct BaseScalar<uchar4> { using type = unsigned char; };
template <> struct BaseScalar<float3> { using type = float; };
template <> struct BaseScalar<float4> { using type = float; };
template <typename T>
using BaseScalar_t = typename BaseScalar<T>::type;

template <typename T>
constexpr int num_channels_v = (int)(sizeof(T) / sizeof(BaseScalar_t<T>));

// The pano constructors' channels argument for a pipeline pixel type: the
// channel count, with HM_PIPELINE_F32 for a float pipeline.
template <typename T>
constexpr int pano_channels_arg() {
  return num_channels_v<T> | (std::is_same_v<BaseScalar_t<T>, float> ? HM_PIPELINE_F32 : 0);
}

// The pipeline / compute pairs the pano classes instantiate: uchar with the
// float of the same channel count, or float throughout.
template <typename T_pipeline, typename T_compute>
constexpr bool pano_types_ok() {
  return (std::is_same_v<T_pipeline, uchar4> && std::is_same_v<T_compute, float4>) ||
         (std::is_same_v<T_pipeline, uchar3> && std::is_same_v<T_compute, float3>) ||
         (std::is_same_v<T_pipeline, float4> && std::is_same_v<T_compute, float4>) ||
         (std::is_same_v<T_pipeline, float3> && std::is_same_v<T_compute, float3>);
}

// The C-ABI pixel code of a pixel type.
template <typename T>
constexpr hmPixelType hm_pixel_type() {
  if constexpr (std::is_same_v<T, uchar3>) return HM_PIX_U8C3;
  if constexpr (std::is_same_v<T, uchar4>) return HM_PIX_U8C4;
  if constexpr (std::is_same_v<T, float3>) return HM_PIX_F32C3;
  return HM_PIX_F32C4;
}

template <typename T>
inline hmSurface to_hm_surface(const CudaSurface<T>& s) {
  return hmSurface{const_cast<void*>(static_cast<const void*>(s.d_ptr)), s.width, s.height, s.pitch};
}
