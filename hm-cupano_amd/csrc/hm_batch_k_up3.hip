// hm_batch_k_up3.hip -- batched-pass kernels of the families
// kFamUq2, kFamUq3, kFamUqN and kFamL10 with 3 channels: the up-pass levels.
#include "hm_batch_ops.hpp"

namespace {

constexpr bool in_part(int op) {
  const int f = hm::op_fam(op);
  return (f == hm::kFamUq2 || f == hm::kFamUq3 || f == hm::kFamUqN || f == hm::kFamL10) && hm::op_c(op) == 3;
}
constexpr int kCount = hmdev::count_ops_where(in_part);
constexpr std::array<int, kCount> kOps = hmdev::ops_where<kCount>(in_part);

}  // namespace

hm::StageKernel hm::batch_kernel_up3(int op, bool persist) { return hmdev::find_op_kernel<kOps>(op, persist); }
