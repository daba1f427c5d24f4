// hm_batch.hpp -- recording a pano call's launches for the batched pass.
//
// A batched pass (hm_batch.hip) runs the pipelines of several pano calls --
// rigs of different classes, image counts, canvas sizes and level counts --
// as one persistent launch per stage.  To build it, each call's process()
// runs once with a LaunchRecorder installed: the launch sites below append
// (kernel body, grid, arguments) instead of launching.  Only the bodies the
// batched kernel carries can be recorded; any other launch makes the
// recording fail (hipErrorNotSupported), so a recorded call never loses work.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <type_traits>
#include <vector>

#include "hm_pyramid_kernels.hpp"

namespace hm {

// The kernel bodies of the batched pass: 4-channel pipelines of the 2-image
// class (any levels, remap tuning 4 pixels x 1 or 2 rows), the 3-image class
// and the N-image class with 3 or 4 images.
enum BatchOp : int {
  kOpNone = -1,
  kOpP2L0_4_1 = 0,  // pano2_level0_body<4, 4, 1>
  kOpP2L0_4_2,      // pano2_level0_body<4, 4, 2>
  kOpP2Hard_4_1,    // pano2_hard_body<4, 4, 1>
  kOpP2Hard_4_2,    // pano2_hard_body<4, 4, 2>
  kOpDown2U8,       // down_fused_body<uint8_t, 4, 2>
  kOpDown2F,        // down_fused_body<float, 4, 2>
  kOpDown3U8,       // down_fused_slots_body<uint8_t, 4, 3>
  kOpDown3F,
  kOpDown4U8,       // down_fused_slots_body<uint8_t, 4, 4>
  kOpDown4F,
  kOpTail2,         // up_tail_body<4, 2, kBlend2>
  kOpUq2F,          // up_quad_body<float, 4, 2, kBlend2, kMask0F32, NEXT>
  kOpUq2FLast,
  kOpUq2U8,         // up_quad_body<uint8_t, 4, 2, kBlend2, kMask0U8, NEXT>
  kOpUq2U8Last,
  kOpUq3F,          // up_quad_body<float, 4, 3, kBlend3, kMask0F32, NEXT>
  kOpUq3FLast,
  kOpUq3L,          // up_quad_body<uint8_t, 4, 3, kBlend3, kMask0Label, NEXT>
  kOpUq3LLast,
  kOpUqN3F,         // up_quad_body<float, 4, 3, kBlendN, kMask0F32, NEXT>
  kOpUqN3FLast,
  kOpUqN3L,         // up_quad_body<uint8_t, 4, 3, kBlendN, kMask0Label, NEXT>
  kOpUqN3LLast,
  kOpUqN4F,         // the same with 4 images
  kOpUqN4FLast,
  kOpUqN4L,
  kOpUqN4LLast,
  kOpMi3,           // multi_inputs_body<4, 3, false>
  kOpMi3DM,         // multi_inputs_body<4, 3, true>
  kOpMi4,           // multi_inputs_body<4, 4, false>
  kOpMc3,           // multi_canvas_body<4, 3>
  kOpMc4,           // multi_canvas_body<4, 4>
  kNumBatchOps
};

// ---- template parameters -> op (kOpNone: not carried) ----------------------
template <int C, int PX, int RY>
constexpr int op_pano2_level0() {
  return (C == 4 && PX == 4) ? (RY == 1 ? kOpP2L0_4_1 : (RY == 2 ? kOpP2L0_4_2 : kOpNone)) : kOpNone;
}
template <int C, int PX, int RY>
constexpr int op_pano2_hard() {
  return (C == 4 && PX == 4) ? (RY == 1 ? kOpP2Hard_4_1 : (RY == 2 ? kOpP2Hard_4_2 : kOpNone)) : kOpNone;
}
template <typename T0, int C, int N>
constexpr int op_down() {
  constexpr bool u8 = std::is_same<T0, uint8_t>::value;
  if (C != 4) return kOpNone;
  if (N == 2) return u8 ? kOpDown2U8 : kOpDown2F;
  if (N == 3) return u8 ? kOpDown3U8 : kOpDown3F;
  if (N == 4) return u8 ? kOpDown4U8 : kOpDown4F;
  return kOpNone;
}
template <int C, int N, int KIND>
constexpr int op_up_tail() {
  return (C == 4 && N == 2 && KIND == hmdev::kBlend2) ? kOpTail2 : kOpNone;
}
template <typename T0, int C, int N, int KIND, int MODE, bool NEXT>
constexpr int op_up_quad() {
  constexpr bool u8 = std::is_same<T0, uint8_t>::value;
  if (C != 4) return kOpNone;
  // MODE values: 0 float mask, 1 uint8 mask, 2 labels (hmdev::Mask0Mode)
  if (N == 2 && KIND == hmdev::kBlend2) {
    if (!u8 && MODE == 0) return NEXT ? kOpUq2F : kOpUq2FLast;
    if (u8 && MODE == 1) return NEXT ? kOpUq2U8 : kOpUq2U8Last;
    return kOpNone;
  }
  if (N == 3 && KIND == hmdev::kBlend3) {
    if (!u8 && MODE == 0) return NEXT ? kOpUq3F : kOpUq3FLast;
    if (u8 && MODE == 2) return NEXT ? kOpUq3L : kOpUq3LLast;
    return kOpNone;
  }
  if ((N == 3 || N == 4) && KIND == hmdev::kBlendN) {
    if (!u8 && MODE == 0) return N == 3 ? (NEXT ? kOpUqN3F : kOpUqN3FLast) : (NEXT ? kOpUqN4F : kOpUqN4FLast);
    if (u8 && MODE == 2) return N == 3 ? (NEXT ? kOpUqN3L : kOpUqN3LLast) : (NEXT ? kOpUqN4L : kOpUqN4LLast);
  }
  return kOpNone;
}
template <int C, int N, bool DM>
constexpr int op_multi_inputs() {
  if (C != 4) return kOpNone;
  if (N == 3) return DM ? kOpMi3DM : kOpMi3;
  if (N == 4 && !DM) return kOpMi4;
  return kOpNone;
}
template <int C, int N>
constexpr int op_multi_canvas() {
  if (C != 4) return kOpNone;
  return N == 3 ? kOpMc3 : (N == 4 ? kOpMc4 : kOpNone);
}

// ---- the recorder -----------------------------------------------------------
struct RecordedLaunch {
  int op = kOpNone;
  uint32_t wgs = 0;    // workgroups along x (tiles)
  uint32_t batch = 1;  // along z
  hipStream_t stream = nullptr;  // the stream process() put it on
  int main_before = 0;           // launches on other streams recorded before it
  std::vector<uint8_t> args;
};

class LaunchRecorder {
 public:
  std::vector<RecordedLaunch> launches;
  hipError_t error = hipSuccess;
  std::string what;  // the first launch that could not be recorded

  template <typename A>
  hipError_t add(int op, dim3 grid, dim3 block, hipStream_t stream, const A& a, const char* name) {
    static_assert(std::is_trivially_copyable<A>::value, "kernel arguments are copied as bytes");
    if (op == kOpNone || grid.y != 1 || block.x * block.y * block.z != 256) return fail(name);
    if (grid.x == 0 || grid.z == 0) return hipSuccess;
    RecordedLaunch r;
    r.op = op;
    r.wgs = grid.x;
    r.batch = grid.z;
    r.stream = stream;
    for (const auto& q : launches) r.main_before += q.stream != stream;
    r.args.assign(reinterpret_cast<const uint8_t*>(&a), reinterpret_cast<const uint8_t*>(&a) + sizeof(A));
    launches.push_back(std::move(r));
    return hipSuccess;
  }
  hipError_t fail(const char* name) {
    if (error == hipSuccess) {
      error = hipErrorNotSupported;
      what = name ? name : "";
    }
    return error;
  }
};

// Installed by the batch builder around one process() call (this thread only).
extern thread_local LaunchRecorder* g_recorder;

}  // namespace hm
