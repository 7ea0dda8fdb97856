// complex_d.hip -- the double complex entry of complex_impl.hpp (argument set-up
// and grid of zblas1; the kernel instances are in the complex_inst.hip
// units) and the double complex row-dot gemv.
#include "complex_impl.hpp"

namespace rb {

bool zblas1(int fold, ZTerm term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
            const double* state_in, double scale_in, const double* scale_in_dev, double* state_out,
            double* scale_out, double* result, hipStream_t s) {
  return zblas1_impl<double>(fold, term, n, x, incx, y, incy, state_in, scale_in, scale_in_dev, state_out,
                             scale_out, result, s);
}

bool zgemv_rows(int fold, int64_t Mo, int64_t No, const double* A, int64_t lda, int conj, const double* xs,
                double* R, hipStream_t s) {
  return zrows_impl<double>(fold, Mo, No, A, lda, conj, xs, R, s);
}

}  // namespace rb
