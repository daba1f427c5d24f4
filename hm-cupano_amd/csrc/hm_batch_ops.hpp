// hm_batch_ops.hpp -- the persistent work-list kernel of the batched pass.
//
// batch_op_kernel<OP, PERSIST> runs one kernel body (the op, see hm_batch.hpp)
// over a stage's work-list: every tile (x) of every batch entry (z) of every
// recorded launch of that body, walked by the workgroups with a stride of the
// grid size.  The items and their argument blocks travel in the kernel
// arguments.  The kernels are compiled in several translation units, one per
// group of families (hm_batch_k_*.hip), which each export a lookup from op to
// kernel.
#pragma once

#include <array>
#include <utility>

#include "hm_batch.hpp"
#include "hm_pyramid_fused.hpp"  // ImgRect, the pyramid bodies
#include "hm_pano2_device.hpp"
#include "hm_pano_device.hpp"

namespace hm {

constexpr int kStageItems = 16;
constexpr int kStageBlob = 3584;

struct StageItem {
  int op;
  uint32_t off;    // argument block in the blob
  uint32_t first;  // first work unit
  uint32_t wgs;    // workgroups along x of the recorded launch
};

// Kernel arguments of one stage launch (under the 4 KiB argument limit).
struct StageArgs {
  int n_items;
  uint32_t total;  // work units: sum of wgs * batch
  int pad[2];
  StageItem items[kStageItems];
  alignas(16) uint8_t blob[kStageBlob];
};
static_assert(sizeof(StageArgs) <= 4000, "stage arguments must fit the kernel-argument segment");

typedef void (*StageKernel)(StageArgs);

// The kernel of `op` from the translation unit that compiles its family, or
// nullptr (not carried); persist: the form whose workgroups walk the list
// (a grid smaller than the work units), else one workgroup per unit.
StageKernel batch_kernel_remap(int op, bool persist);  // kFamP2L0, kFamP2Hard, kFamMi, kFamMc
StageKernel batch_kernel_down(int op, bool persist);   // kFamDown, kFamTail2, kFamTail2L
StageKernel batch_kernel_up3(int op, bool persist);    // kFamUq2, kFamUq3, kFamUqN, kFamL10 with 3 channels
StageKernel batch_kernel_up4(int op, bool persist);    // the same with 4 channels
inline StageKernel batch_kernel(int op, bool persist) {
  switch (op_fam(op)) {
    case kFamP2L0:
    case kFamP2Hard:
    case kFamMi:
    case kFamMc:
      return batch_kernel_remap(op, persist);
    case kFamDown:
    case kFamTail2:
    case kFamTail2L:
      return batch_kernel_down(op, persist);
    case kFamUq2:
    case kFamUq3:
    case kFamUqN:
    case kFamL10:
      return op_c(op) == 4 ? batch_kernel_up4(op, persist) : batch_kernel_up3(op, persist);
    default:
      return nullptr;
  }
}

}  // namespace hm

namespace hmdev {

// The stage's arguments are read through the kernel-argument segment pointer
// (constant memory: scalar loads).  Taking typed views of different bodies'
// argument blocks in the by-value parameter instead makes the compiler copy
// the whole parameter into scratch memory.
template <typename T>
using kptr = const __attribute__((address_space(4))) T*;

template <typename A>
__device__ __forceinline__ const A& arg_at(kptr<hm::StageArgs> s, uint32_t off) {
  return *(const A*)(s->blob + off);  // constant memory seen through a generic pointer
}

// The argument type of an op's body.
template <int OP, int FAM = hm::op_fam(OP)>
struct OpArgsOf;
template <int OP> struct OpArgsOf<OP, hm::kFamP2L0> { using type = Pano2L0; };
template <int OP> struct OpArgsOf<OP, hm::kFamP2Hard> { using type = Pano2Hard; };
template <int OP> struct OpArgsOf<OP, hm::kFamMi> { using type = MultiArgs<hm::op_n(OP)>; };
template <int OP> struct OpArgsOf<OP, hm::kFamMc> { using type = MultiArgs<hm::op_n(OP)>; };
template <int OP>
struct OpArgsOf<OP, hm::kFamDown> {
  using type = DownFusedArgs<std::conditional_t<hm::op_a(OP) == 1, uint8_t, float>, hm::op_n(OP)>;
};
template <int OP> struct OpArgsOf<OP, hm::kFamTail2> { using type = UpTailArgs<2>; };
template <int OP> struct OpArgsOf<OP, hm::kFamTail2L> { using type = UpTailArgs<2>; };
template <int OP> struct OpArgsOf<OP, hm::kFamL10> {
  using type = UpL10Args<std::conditional_t<hm::op_a(OP) == 1, uint8_t, float>>;
};
template <int OP> struct OpArgsOf<OP, hm::kFamUq2> {
  using type = UpQuadArgs<std::conditional_t<hm::op_a(OP) == 1, uint8_t, float>, 2>;
};
template <int OP> struct OpArgsOf<OP, hm::kFamUq3> {
  using type = UpQuadArgs<std::conditional_t<hm::op_a(OP) == 1, uint8_t, float>, 3>;
};
template <int OP> struct OpArgsOf<OP, hm::kFamUqN> {
  using type = UpQuadArgs<std::conditional_t<hm::op_a(OP) == 1, uint8_t, float>, hm::op_n(OP)>;
};

// LDS of an op's body, in float4.
template <int OP>
constexpr int op_lds_float4() {
  constexpr int fam = hm::op_fam(OP), n = hm::op_n(OP);
  if constexpr (fam == hm::kFamDown) return down_lds_float4<down_slots_count<n>()>();
  if constexpr (fam == hm::kFamTail2) return kTailLdsFloat4;
  if constexpr (fam == hm::kFamTail2L) return up_tail_lds_float4<2, kBlend2, kTailLdsTile>();
  if constexpr (fam == hm::kFamL10) return up_l10_lds_float4();
  if constexpr (fam == hm::kFamMc) return kCanvasGeoFloat4;
  if constexpr (fam == hm::kFamUq2 || fam == hm::kFamUq3 || fam == hm::kFamUqN) return quad_lds_float4<n>();
  return 1;
}

// One work unit of an op's body.
template <int OP>
__device__ __forceinline__ void op_body(kptr<hm::StageArgs> s, uint32_t off, Blk blk, uint32_t wgs, float4* lds) {
  constexpr int fam = hm::op_fam(OP), C = hm::op_c(OP), N = hm::op_n(OP);
  constexpr bool A = hm::op_a(OP) == 1, NEXT = hm::op_next(OP) == 1;
  using Args = typename OpArgsOf<OP>::type;
  const Args& a = arg_at<Args>(s, off);
  if constexpr (fam == hm::kFamP2L0) {
    pano2_level0_body<C, 4, N>(a, blk, (int)wgs);
  } else if constexpr (fam == hm::kFamP2Hard) {
    pano2_hard_body<C, 4, N>(a, blk);
  } else if constexpr (fam == hm::kFamMi) {
    multi_inputs_body<C, N, A>(a, blk);
  } else if constexpr (fam == hm::kFamMc) {
    multi_canvas_body<C, N>(a, blk, lds);
  } else if constexpr (fam == hm::kFamDown) {
    using T0 = std::conditional_t<A, uint8_t, float>;
    down_any_body<T0, C, N>(a, blk, lds);
  } else if constexpr (fam == hm::kFamTail2) {
    up_tail_body<C, 2, kBlend2>(a, blk, lds);
  } else if constexpr (fam == hm::kFamTail2L) {
    up_tail_lds_body<C, 2, kBlend2, kTailLdsTile>(a, blk, lds);
  } else if constexpr (fam == hm::kFamL10) {
    using T0 = std::conditional_t<A, uint8_t, float>;
    up_l10_body<T0, C, A ? kMask0U8 : kMask0F32, NEXT>(a, blk, lds);
  } else if constexpr (fam == hm::kFamUq2) {
    using T0 = std::conditional_t<A, uint8_t, float>;
    up_quad_body<T0, C, 2, kBlend2, A ? kMask0U8 : kMask0F32, NEXT>(a, blk, lds);
  } else if constexpr (fam == hm::kFamUq3) {
    using T0 = std::conditional_t<A, uint8_t, float>;
    up_quad_body<T0, C, 3, kBlend3, A ? kMask0Label : kMask0F32, NEXT>(a, blk, lds);
  } else if constexpr (fam == hm::kFamUqN) {
    using T0 = std::conditional_t<A, uint8_t, float>;
    up_quad_body<T0, C, N, kBlendN, A ? kMask0Label : kMask0F32, NEXT>(a, blk, lds);
  }
}

// Work unit w of the stage: its item, and the body on the item's tile.
template <int OP>
__device__ __forceinline__ void run_unit(kptr<hm::StageArgs> s, uint32_t w, float4* lds) {
  int j = 0;
  for (int q = 1; q < s->n_items; ++q)
    if (w >= s->items[q].first) j = q;
  const uint32_t off = s->items[j].off, wgs = s->items[j].wgs;
  const uint32_t local = w - s->items[j].first;
  op_body<OP>(s, off, Blk{(int)(local % wgs), (int)(local / wgs)}, wgs, lds);
}

// Runs on 64 x 4 threads; every body uses all 256 (see Blk).  PERSIST: the
// workgroups walk the launch's work-list with a stride of the grid size;
// otherwise the grid has one workgroup per work unit.  (The loop costs the
// bodies registers -- the 8-image canvas writer 98 VGPRs against 34 without
// it -- so the one-per-unit grid, the default, runs the loop-free form.)
template <int OP, bool PERSIST>
__global__ void __launch_bounds__(256) batch_op_kernel(hm::StageArgs stage) {
  (void)stage;  // read through the segment pointer below
  const kptr<hm::StageArgs> s = (kptr<hm::StageArgs>)__builtin_amdgcn_kernarg_segment_ptr();
  __shared__ float4 lds[op_lds_float4<OP>()];
  (void)lds;
  if constexpr (PERSIST) {
    for (uint32_t w = blockIdx.x; w < s->total; w += gridDim.x) {
      run_unit<OP>(s, w, lds);
      __syncthreads();  // the next item may reuse the LDS
    }
  } else {
    run_unit<OP>(s, blockIdx.x, lds);
  }
}

// The ops of the families a translation unit compiles (PRED(op) true), in
// the order of hm::for_each_batch_op, and their kernels.
template <typename Pred>
constexpr int count_ops_where(Pred pred) {
  int k = 0;
  hm::for_each_batch_op([&](int op) { k += pred(op) ? 1 : 0; });
  return k;
}
template <int K, typename Pred>
constexpr std::array<int, K> ops_where(Pred pred) {
  std::array<int, K> a{};
  int k = 0;
  hm::for_each_batch_op([&](int op) {
    if (pred(op)) a[k++] = op;
  });
  return a;
}
template <const auto& OPS, bool PERSIST, size_t... I>
std::array<hm::StageKernel, sizeof...(I)> op_kernels(std::index_sequence<I...>) {
  static_assert(((sizeof(typename OpArgsOf<OPS[I]>::type) <= hm::kStageBlob) && ...),
                "every body's arguments fit a stage's blob");
  return {{batch_op_kernel<OPS[I], PERSIST>...}};
}
template <const auto& OPS>
hm::StageKernel find_op_kernel(int op, bool persist) {
  static const auto once = op_kernels<OPS, false>(std::make_index_sequence<OPS.size()>{});
  static const auto walk = op_kernels<OPS, true>(std::make_index_sequence<OPS.size()>{});
  for (size_t i = 0; i < OPS.size(); ++i)
    if (OPS[i] == op) return persist ? walk[i] : once[i];
  return nullptr;
}

}  // namespace hmdev
