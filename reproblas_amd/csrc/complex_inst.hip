// complex_inst.hip -- the kernel instances of one complex BLAS1 term for one
// element type (complex_impl.hpp). The build compiles this unit once per term
// and type (RB_INST_TERM = 0 sum, 1 dotu, 2 dotc, 3 asum, 4 ssq; RB_INST_TYPE
// = double or float), so the instances compile in parallel.
#include "complex_impl.hpp"

#if !defined(RB_INST_TERM) || !defined(RB_INST_TYPE)
#error "complex_inst.hip needs RB_INST_TERM and RB_INST_TYPE"
#endif

namespace rb {

#if RB_INST_TERM == 0
#define RB_INST_NAME zblas1_launch_sum
#define RB_INST_T kZSum
#elif RB_INST_TERM == 1
#define RB_INST_NAME zblas1_launch_dotu
#define RB_INST_T kZDotU
#elif RB_INST_TERM == 2
#define RB_INST_NAME zblas1_launch_dotc
#define RB_INST_T kZDotC
#elif RB_INST_TERM == 3
#define RB_INST_NAME zblas1_launch_asum
#define RB_INST_T kZASum
#else
#define RB_INST_NAME zblas1_launch_ssq
#define RB_INST_T kZSSq
#endif

bool RB_INST_NAME(int fold, int mode, const ZArgs<RB_INST_TYPE>& a, int blocks, hipStream_t s) {
  return zlaunch_fold_term<RB_INST_T>(fold, mode, a, blocks, s);
}

}  // namespace rb
