// complex_f.hip -- the float complex entry of complex_impl.hpp (argument set-up
// and grid of zblas1; the kernel instances are in the complex_inst.hip
// units) and the float complex row-dot gemv.
#include "complex_impl.hpp"

namespace rb {

bool zblas1(int fold, ZTerm term, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy,
            const float* state_in, float scale_in, const float* scale_in_dev, float* state_out, float* scale_out,
            float* result, hipStream_t s) {
  return zblas1_impl<float>(fold, term, n, x, incx, y, incy, state_in, scale_in, scale_in_dev, state_out,
                            scale_out, result, s);
}

bool zgemv_rows(int fold, int64_t Mo, int64_t No, const float* A, int64_t lda, int conj, const float* xs, float* R,
                hipStream_t s) {
  return zrows_impl<float>(fold, Mo, No, A, lda, conj, xs, R, s);
}

}  // namespace rb
