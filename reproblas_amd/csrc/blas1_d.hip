// blas1_d.hip -- the double instances of blas1_impl.hpp (blas1, the ragged
// segment kernels); blas1_f.hip holds the float ones, so the two compile in
// parallel.
#include "blas1_impl.hpp"

namespace rb {

bool fold_supported(int fold) { return fold >= 2 && fold <= kMaxFold; }

bool blas1(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
           const double* state_in, double scale_in, const double* scale_in_dev, double* state_out, double* scale_out,
           double* result, hipStream_t s) {
  return blas1_impl<double>(fold, term, n, x, incx, y, incy, state_in, scale_in, scale_in_dev, state_out, scale_out,
                            result, s);
}

bool segments(int fold, Term term, const double* x, const double* y, const SegChunk* chunks, int64_t nchunks,
              double* partials, const int64_t* first_chunk, int64_t nseg, double* results, int grid, hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  SegArgs a;
  a.x = x;
  a.y = y;
  a.chunks = chunks;
  a.nchunks = nchunks;
  a.partials = partials;
  a.first_chunk = first_chunk;
  a.nseg = nseg;
  a.results = results;
  a.fold = fold;
  if (grid < 1) grid = 1;
  switch (fold) {
    case 2: return launch_seg_term<2, true>(term, a, grid, s);
    case 3: return launch_seg_term<3, true>(term, a, grid, s);
    case 4: return launch_seg_term<4, true>(term, a, grid, s);
    default:
      if (fold <= 8) return launch_seg_term<8, false>(term, a, grid, s);
      return launch_seg_term<kMaxFold, false>(term, a, grid, s);
  }
}

}  // namespace rb
