// blas1_seg.hip -- the ragged segment kernels of blas1_impl.hpp (double),
// behind the batch plans of batch.hip.
#include "blas1_impl.hpp"

namespace rb {

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
