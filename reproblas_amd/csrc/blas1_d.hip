// blas1_d.hip -- the double BLAS1 entry of blas1_impl.hpp (argument set-up
// and grid; the kernel instances are in the blas1_inst.hip units, the ragged
// segment kernels in blas1_seg.hip).
#include "blas1_impl.hpp"

namespace rb {

bool fold_supported(int fold) { return fold >= 2 && fold <= kMaxFold; }

bool blas1(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
           const double* state_in, double scale_in, const double* scale_in_dev, double* state_out, double* scale_out,
           double* result, hipStream_t s) {
  return blas1_impl<double>(fold, term, n, x, incx, y, incy, state_in, scale_in, scale_in_dev, state_out, scale_out,
                            result, s);
}

bool blas1_scratch(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
                   double* result, void* scratch, size_t scratch_bytes, unsigned* ticket, hipStream_t s) {
  return blas1_impl<double>(fold, term, n, x, incx, y, incy, nullptr, 0.0, nullptr, nullptr, nullptr, result, s,
                            scratch, scratch_bytes, ticket);
}

}  // namespace rb
