// blas1_inst.hip -- the kernel instances of one BLAS1 term for one element
// type (blas1_impl.hpp). The build compiles this unit once per term and type
// (RB_INST_TERM = 0 sum, 1 asum, 2 dot, 3 ssq; RB_INST_TYPE = double or
// float), so the instances compile in parallel.
#include "blas1_impl.hpp"

#if !defined(RB_INST_TERM) || !defined(RB_INST_TYPE)
#error "blas1_inst.hip needs RB_INST_TERM and RB_INST_TYPE"
#endif

namespace rb {

#if RB_INST_TERM == 0
bool blas1_launch_sum(int mode, const Blas1Args<RB_INST_TYPE>& a, int blocks, hipStream_t s) {
  return launch_fold_term<kSum>(mode, a, blocks, s);
}
#elif RB_INST_TERM == 1
bool blas1_launch_asum(int mode, const Blas1Args<RB_INST_TYPE>& a, int blocks, hipStream_t s) {
  return launch_fold_term<kASum>(mode, a, blocks, s);
}
#elif RB_INST_TERM == 2
bool blas1_launch_dot(int mode, const Blas1Args<RB_INST_TYPE>& a, int blocks, hipStream_t s) {
  return launch_fold_term<kDot>(mode, a, blocks, s);
}
#else
bool blas1_launch_ssq(int mode, const Blas1Args<RB_INST_TYPE>& a, int blocks, hipStream_t s) {
  return launch_fold_term<kSSq>(mode, a, blocks, s);
}
#endif

}  // namespace rb
