// hm_batch_k_remap.hip -- batched-pass kernels of the families
// kFamP2L0, kFamP2Hard, kFamMi and kFamMc: the 2-image remaps and the N/3-image input and canvas writers.
#include "hm_batch_ops.hpp"

namespace {

constexpr bool in_part(int op) {
  const int f = hm::op_fam(op);
  return f == hm::kFamP2L0 || f == hm::kFamP2Hard || f == hm::kFamMi || f == hm::kFamMc;
}
constexpr int kCount = hmdev::count_ops_where(in_part);
constexpr std::array<int, kCount> kOps = hmdev::ops_where<kCount>(in_part);

}  // namespace

hm::StageKernel hm::batch_kernel_remap(int op, bool persist) { return hmdev::find_op_kernel<kOps>(op, persist); }
