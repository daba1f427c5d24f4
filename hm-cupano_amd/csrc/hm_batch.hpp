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

// The kernel bodies of the batched pass: the uint8 pipelines of the 2-image
// class (remap tuning 4 pixels x 1 or 2 rows), the 3-image class and the
// N-image class with 2 to 8 images, with 3 or 4 channels.  An op names one
// body with its template parameters: family (bits 8+), NEXT (bit 6), a
// family-specific flag A (bit 5), 4 channels (bit 4) and an image or row
// count n (bits 0-3).
enum BatchFamily : int {
  kFamP2L0 = 1,  // pano2_level0_body<C, 4, RY>                 n = RY
  kFamP2Hard,    // pano2_hard_body<C, 4, RY>                   n = RY
  kFamDown,      // down_fused_body (N = 2) or down_fused_slots_body<T0, C, N>; A: T0 = uint8_t, else float
  kFamTail2,     // up_tail_body<C, 2, kBlend2>
  kFamUq2,       // up_quad_body<T0, C, 2, kBlend2, M, NEXT>; A: uint8 level and mask (M = kMask0U8), else float (kMask0F32)
  kFamUq3,       // up_quad_body<T0, C, 3, kBlend3, M, NEXT>; A: uint8 level and labels (kMask0Label), else float
  kFamUqN,       // up_quad_body<T0, C, N, kBlendN, M, NEXT>; A as for kFamUq3
  kFamMi,        // multi_inputs_body<C, N, DM>; A: DM
  kFamMc,        // multi_canvas_body<C, N>
  kFamTail2L,    // up_tail_lds_body<C, 2, kBlend2, kTailLdsTile> (the tail in one load burst)
  kFamL10,       // up_l10_body<T0, C, M, NEXT>: levels 1 and 0 of the 2-image up pass; A as for kFamUq2
};
constexpr int kOpNone = -1;
constexpr int batch_op(int fam, int c, int a, int next, int n) {
  return fam << 8 | next << 6 | a << 5 | (c == 4 ? 16 : 0) | n;
}
constexpr int op_fam(int op) { return op >> 8; }
constexpr int op_c(int op) { return (op & 16) ? 4 : 3; }
constexpr int op_a(int op) { return (op >> 5) & 1; }
constexpr int op_next(int op) { return (op >> 6) & 1; }
constexpr int op_n(int op) { return op & 15; }

// Every op the batched pass carries, in a fixed order (the order of a
// stage's launches).
template <typename F>
constexpr void for_each_batch_op(F&& f) {
  for (int c = 3; c <= 4; ++c) {
    for (int ry = 1; ry <= 2; ++ry) {
      f(batch_op(kFamP2L0, c, 0, 0, ry));
      f(batch_op(kFamP2Hard, c, 0, 0, ry));
    }
    for (int n = 2; n <= 8; ++n) {
      f(batch_op(kFamMi, c, 0, 0, n));
      f(batch_op(kFamMc, c, 0, 0, n));
    }
    f(batch_op(kFamMi, c, 1, 0, 3));
    for (int a = 0; a <= 1; ++a)
      for (int n = 2; n <= 8; ++n) f(batch_op(kFamDown, c, a, 0, n));
    f(batch_op(kFamTail2, c, 0, 0, 2));
    f(batch_op(kFamTail2L, c, 0, 0, 2));
    for (int a = 0; a <= 1; ++a)
      for (int next = 0; next <= 1; ++next) {
        f(batch_op(kFamL10, c, a, next, 2));
        f(batch_op(kFamUq2, c, a, next, 2));
        f(batch_op(kFamUq3, c, a, next, 3));
        for (int n = 2; n <= 8; ++n) f(batch_op(kFamUqN, c, a, next, n));
      }
  }
}
constexpr int count_batch_ops() {
  int k = 0;
  for_each_batch_op([&](int) { ++k; });
  return k;
}
constexpr int kNumBatchOps = count_batch_ops();

// ---- template parameters -> op (kOpNone: not carried) ----------------------
constexpr bool batch_channels(int c) { return c == 3 || c == 4; }
template <int C, int PX, int RY>
constexpr int op_pano2_level0() {
  return (batch_channels(C) && PX == 4 && (RY == 1 || RY == 2)) ? batch_op(kFamP2L0, C, 0, 0, RY) : kOpNone;
}
template <int C, int PX, int RY>
constexpr int op_pano2_hard() {
  return (batch_channels(C) && PX == 4 && (RY == 1 || RY == 2)) ? batch_op(kFamP2Hard, C, 0, 0, RY) : kOpNone;
}
template <typename T0, int C, int N>
constexpr int op_down() {
  constexpr bool u8 = std::is_same<T0, uint8_t>::value;
  return (batch_channels(C) && N >= 2 && N <= 8) ? batch_op(kFamDown, C, u8, 0, N) : kOpNone;
}
template <int C, int N, int KIND>
constexpr int op_up_tail() {
  return (batch_channels(C) && N == 2 && KIND == hmdev::kBlend2) ? batch_op(kFamTail2, C, 0, 0, 2) : kOpNone;
}
template <int C, int N, int KIND>
constexpr int op_up_tail_lds() {
  return (batch_channels(C) && N == 2 && KIND == hmdev::kBlend2) ? batch_op(kFamTail2L, C, 0, 0, 2) : kOpNone;
}
template <typename T0, int C, int MODE, bool NEXT>
constexpr int op_up_l10() {
  constexpr bool u8 = std::is_same<T0, uint8_t>::value;
  // MODE values: 0 float mask, 1 uint8 mask (hmdev::Mask0Mode)
  return (batch_channels(C) && (u8 ? MODE == 1 : MODE == 0)) ? batch_op(kFamL10, C, u8, NEXT, 2) : kOpNone;
}
template <typename T0, int C, int N, int KIND, int MODE, bool NEXT>
constexpr int op_up_quad() {
  constexpr bool u8 = std::is_same<T0, uint8_t>::value;
  if (!batch_channels(C)) return kOpNone;
  // MODE values: 0 float mask, 1 uint8 mask, 2 labels (hmdev::Mask0Mode)
  if (KIND == hmdev::kBlend2)
    return (N == 2 && (u8 ? MODE == 1 : MODE == 0)) ? batch_op(kFamUq2, C, u8, NEXT, 2) : kOpNone;
  const bool mode_ok = u8 ? MODE == 2 : MODE == 0;
  if (KIND == hmdev::kBlend3) return (N == 3 && mode_ok) ? batch_op(kFamUq3, C, u8, NEXT, 3) : kOpNone;
  if (KIND == hmdev::kBlendN && N >= 2 && N <= 8 && mode_ok) return batch_op(kFamUqN, C, u8, NEXT, N);
  return kOpNone;
}
template <int C, int N, bool DM>
constexpr int op_multi_inputs() {
  return (batch_channels(C) && N >= 2 && N <= 8 && (!DM || N == 3)) ? batch_op(kFamMi, C, DM, 0, N) : kOpNone;
}
template <int C, int N>
constexpr int op_multi_canvas() {
  return (batch_channels(C) && N >= 2 && N <= 8) ? batch_op(kFamMc, C, 0, 0, N) : kOpNone;
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
