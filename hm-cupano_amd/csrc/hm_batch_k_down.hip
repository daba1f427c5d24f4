// hm_batch_k_down.hip -- batched-pass kernels of the families
// kFamDown, kFamTail2 and kFamTail2L: the fused down pass and the 2-image
// up-pass tail (both forms).
#include "hm_batch_ops.hpp"

namespace {

constexpr bool in_part(int op) {
  const int f = hm::op_fam(op);
  return f == hm::kFamDown || f == hm::kFamTail2 || f == hm::kFamTail2L;
}
constexpr int kCount = hmdev::count_ops_where(in_part);
constexpr std::array<int, kCount> kOps = hmdev::ops_where<kCount>(in_part);

}  // namespace

hm::StageKernel hm::batch_kernel_down(int op, bool persist) { return hmdev::find_op_kernel<kOps>(op, persist); }
