// blas1_f.hip -- the float BLAS1 entry of blas1_impl.hpp (argument set-up and
// grid; the kernel instances are in the blas1_inst.hip units).
#include "blas1_impl.hpp"

namespace rb {

bool blas1_s(int fold, Term term, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy,
             const float* state_in, float scale_in, const float* scale_in_dev, float* state_out, float* scale_out,
             float* result, hipStream_t s) {
  return blas1_impl<float>(fold, term, n, x, incx, y, incy, state_in, scale_in, scale_in_dev, state_out, scale_out,
                           result, s);
}

}  // namespace rb
