// kernels.hpp -- launchers of the HIP kernels (host-callable, device operands).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rb {

// Folds with a dedicated (fully register-resident) kernel instance; every
// other fold in [2, kMaxFold] runs the general instance.
bool fold_supported(int fold);

enum Term : int { kSum = 0, kASum = 1, kDot = 2, kSSq = 3 };

// One reduction of the binned BLAS1 family over device vectors:
//   kSum: t_i = x_i, kASum: |x_i|, kDot: fl(x_i*y_i), kSSq: fl(fl(x_i/scl)^2)
// state_in  -- [2*fold] binned value to accumulate into, or null for zero
// scale_in  -- kSSq only: the scale of state_in (0 when none)
// scale_in_dev -- kSSq only: device copy of that scale; used instead when set
// state_out -- [2*fold] final binned value, or null
// scale_out -- kSSq only: the final scale, or null
// result    -- converted double (kSSq: scl*sqrt(conv)), or null
// All device pointers; runs on stream s without synchronising.
bool blas1(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
           const double* state_in, double scale_in, const double* scale_in_dev, double* state_out,
           double* scale_out, double* result, hipStream_t s);
// The converted double result of a BLAS1 reduction with the caller's scratch
// instead of the library workspace: scratch holds scratch_bytes of device
// memory for the workgroup records of nrm2 (the launch fails with
// kBadArgument when the grid needs more; the other terms need none), ticket
// a zeroed device block of kBlas1SyncBytes (completion ticket and grid
// accumulators) that the launch leaves zeroed. Allocates nothing.
constexpr size_t kBlas1SyncBytes = 32768;
bool blas1_scratch(int fold, Term term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
                   double* result, void* scratch, size_t scratch_bytes, unsigned* ticket, hipStream_t s);
// The same for float operands and float_binned values (folds 2..21).
bool blas1_s(int fold, Term term, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy,
             const float* state_in, float scale_in, const float* scale_in_dev, float* state_out, float* scale_out,
             float* result, hipStream_t s);

// Complex reductions over interleaved (real, imaginary) device vectors of n
// complex elements (increments in complex elements):
//   kZSum, kZDotU, kZDotC -- complex binned result: state_in/state_out hold
//     4*fold values (interleaved primaries, then interleaved carries),
//     result two values;
//   kZASum, kZSSq -- real binned result (2*fold), result one value (kZSSq:
//     scl*sqrt(conv)); scale_in/scale_in_dev/scale_out as for kSSq.
enum ZTerm : int { kZSum = 0, kZDotU = 1, kZDotC = 2, kZASum = 3, kZSSq = 4 };
bool zblas1(int fold, ZTerm term, int64_t n, const double* x, int64_t incx, const double* y, int64_t incy,
            const double* state_in, double scale_in, const double* scale_in_dev, double* state_out,
            double* scale_out, double* result, hipStream_t s);
bool zblas1(int fold, ZTerm term, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy,
            const float* state_in, float scale_in, const float* scale_in_dev, float* state_out, float* scale_out,
            float* result, hipStream_t s);

// Ragged batch of BLAS1 reductions over CSR segments of x (and y for dot),
// unit stride. chunks[c] = [begin, end) element range (at most a few
// thousand terms); segment s owns chunks [first_chunk[s], first_chunk[s+1]).
// partials: nchunks*(2*fold+1) doubles of scratch; results: nseg doubles.
struct SegChunk {
  int64_t begin;
  int64_t end;
};
bool segments(int fold, Term term, const double* x, const double* y, const SegChunk* chunks, int64_t nchunks,
              double* partials, const int64_t* first_chunk, int64_t nseg, double* results, int grid, hipStream_t s);

// Elementwise binned helpers on device arrays of binned values.
// dbdbaddv / sbsbaddv: Y[i] += X[i] for i < n (stride incX/incY in binned elements).
bool dbdbaddv(int fold, int64_t n, const double* X, int64_t incX, double* Y, int64_t incY, hipStream_t s);
bool sbsbaddv(int fold, int64_t n, const float* X, int64_t incX, float* Y, int64_t incY, hipStream_t s);
// zaddv: the same for complex binned values (4*fold each, both components).
bool zaddv(int fold, int64_t n, const double* X, int64_t incX, double* Y, int64_t incY, hipStream_t s);
bool zaddv(int fold, int64_t n, const float* X, int64_t incX, float* Y, int64_t incY, hipStream_t s);
// conv_to_binned: Y[i] = dbdconv/sbsconv(beta == 1 ? y[i] : y[i]*beta) or zero when beta == 0.
bool conv_to_binned(int fold, int64_t n, const double* y, int64_t incy, double beta, double* Yb, hipStream_t s);
bool conv_to_binned(int fold, int64_t n, const float* y, int64_t incy, float beta, float* Yb, hipStream_t s);
// conv_from_binned: y[i] = ddbconv/ssbconv(Yb[i]); conv_to_double is the double one.
bool conv_from_binned(int fold, int64_t n, const double* Yb, double* y, int64_t incy, hipStream_t s);
bool conv_from_binned(int fold, int64_t n, const float* Yb, float* y, int64_t incy, hipStream_t s);
bool conv_to_double(int fold, int64_t n, const double* Yb, double* y, int64_t incy, hipStream_t s);
// beta_only: y(i, o) = conv(binned(beta*y(i, o))) in place for i < inner,
// o < outer, element (i, o) at y[i*inc_i + o*inc_o] (+0 when beta == 0): the
// rdgemv / rdgemm result when no product is added. One launch, no scratch.
bool beta_only(int fold, int64_t inner, int64_t outer, double* y, int64_t inc_i, int64_t inc_o, double beta,
               hipStream_t s);
bool beta_only(int fold, int64_t inner, int64_t outer, float* y, int64_t inc_i, int64_t inc_o, float beta,
               hipStream_t s);

// Binned gemv on device operands (reference binnedBLAS_dbdgemv semantics,
// column- or row-major): Yb[i*incY] (+)= sum_j fl(op(A)_ij * fl(alpha*x_j)).
bool dbdgemv(int fold, char order, char trans, int64_t M, int64_t N, double alpha, const double* A, int64_t lda,
             const double* x, int64_t incx, double* Yb, int64_t incYb, hipStream_t s);

// Binned gemm on device operands (reference binnedBLAS_dbdgemm semantics):
// C(i,j) (+)= sum_k fl(fl(alpha*op(A)_ik) * op(B)_kj), C binned with
// leading dimension ldc (in binned elements).
bool dbdgemm(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double* Cb, int64_t ldc, hipStream_t s);

// Float counterparts (binnedBLAS_sbsgemv / sbsgemm, reproBLAS_rsgemv / rsgemm).
bool sbsgemv(int fold, char order, char trans, int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
             const float* x, int64_t incx, float* Yb, int64_t incYb, hipStream_t s);
bool sbsgemm(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, float alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float* Cb, int64_t ldc, hipStream_t s);
bool rsgemv_device(int fold, char order, char trans, int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
                   const float* x, int64_t incx, float beta, float* y, int64_t incy, hipStream_t s);
bool rsgemm_device(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, float alpha,
                   const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc,
                   hipStream_t s);

// In-place y = conv(binned(beta*y) (+) sum_j fl(op(A)_ij * fl(alpha*x_j)))
// (reproBLAS_rdgemv) on device operands.
bool rdgemv_device(int fold, char order, char trans, int64_t M, int64_t N, double alpha, const double* A,
                   int64_t lda, const double* x, int64_t incx, double beta, double* y, int64_t incy, hipStream_t s);

// In-place C = conv(binned(beta*C) (+) sum_k fl(fl(alpha*op(A)_ik) * op(B)_kj))
// (reproBLAS_rdgemm) on device operands.
bool rdgemm_device(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, double alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                   hipStream_t s);

// Complex gemv outputs whose terms are lda apart (column-major N, row-major
// T/C) in one pass: R[2i], R[2i+1] (real binned, accumulated into) get the
// real and imaginary parts of sum_k op(A)_ik * xs[k]; A interleaved complex
// with op(A)_ik = A[k*lda + i] (conjugated when conj), xs = alpha*x contiguous.
bool zgemv_cols(int fold, int64_t Mo, int64_t No, const double* A, int64_t lda, int conj, const double* xs,
                double* R, hipStream_t s);
bool zgemv_cols(int fold, int64_t Mo, int64_t No, const float* A, int64_t lda, int conj, const float* xs, float* R,
                hipStream_t s);

// The same for outputs whose terms are contiguous (row-major N, column-major
// T/C): op(A)_ik = A[i*lda + k], one complex dot per row, one pass over A.
bool zgemv_rows(int fold, int64_t Mo, int64_t No, const double* A, int64_t lda, int conj, const double* xs,
                double* R, hipStream_t s);
bool zgemv_rows(int fold, int64_t Mo, int64_t No, const float* A, int64_t lda, int conj, const float* xs, float* R,
                hipStream_t s);

// Complex gemm / gemv on device operands (interleaved pairs; alpha and beta
// are host pairs): binnedBLAS_zbzgemm / zbzgemv semantics into complex
// binned arrays (Zb with leading dimension ldz, Yb with increment incy, in
// complex binned elements), reproBLAS_rzgemm / rzgemv semantics in place.
// The float overloads are the c forms.
bool zbzgemm(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const double* alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double* Zb, int64_t ldz, hipStream_t s);
bool zbzgemm(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const float* alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float* Zb, int64_t ldz, hipStream_t s);
bool rzgemm_device(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const double* alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, const double* beta, double* C,
                   int64_t ldc, hipStream_t s);
bool rzgemm_device(int fold, char order, char ta, char tb, int64_t M, int64_t N, int64_t K, const float* alpha,
                   const float* A, int64_t lda, const float* B, int64_t ldb, const float* beta, float* C, int64_t ldc,
                   hipStream_t s);
bool zbzgemv(int fold, char order, char trans, int64_t M, int64_t N, const double* alpha, const double* A,
             int64_t lda, const double* x, int64_t incx, double* Yb, int64_t incy, hipStream_t s);
bool zbzgemv(int fold, char order, char trans, int64_t M, int64_t N, const float* alpha, const float* A, int64_t lda,
             const float* x, int64_t incx, float* Yb, int64_t incy, hipStream_t s);
bool rzgemv_device(int fold, char order, char trans, int64_t M, int64_t N, const double* alpha, const double* A,
                   int64_t lda, const double* x, int64_t incx, const double* beta, double* y, int64_t incy,
                   hipStream_t s);
bool rzgemv_device(int fold, char order, char trans, int64_t M, int64_t N, const float* alpha, const float* A,
                   int64_t lda, const float* x, int64_t incx, const float* beta, float* y, int64_t incy,
                   hipStream_t s);

// One gemm on device operands in the kernels' terms:
//   op(A)(i,k) = A[i*sai + k*sak], op(B)(k,j) = B[k*sbk + j*sbj],
//   C(i,j) = C[i*sci + j*scj] (double output, beta applied) or, when Cb is
//   set, the binned output Cb[(i*cbi + j*cbj)*2*fold..] accumulated into.
// rowmax/colmax: scratch of ceil(K/32)*M and ceil(K/32)*N doubles.
template <class F>
struct GemmDescT {
  int64_t M, N, K;
  const F* A;
  int64_t sai, sak;
  const F* B;
  int64_t sbk, sbj;
  F alpha;
  F* C;
  int64_t sci, scj;
  F beta;
  F* Cb;
  int64_t cbi, cbj;
  const F* rowmax;
  const F* colmax;
  int fold;
};
using GemmDesc = GemmDescT<double>;
struct GemmTile {
  int32_t gemm, ti, tj, pad;
};
// Fills a GemmDesc from BLAS arguments (Order/Trans characters).
GemmDesc make_gemm_desc(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                        double* Cb, int64_t ldcb);
// Tile edge of the batch kernel for a tiling: rows 16*TM, columns 16*TN.
void gemm_tile_shape(int tiling, int fold, int* bm, int* bn);
// Ragged gemm batch: all descs share one fold and one layout (A/B contiguity
// flags taken from descs[0]); tiles list (gemm, tile row, tile column).
bool gemm_batch(int fold, int tiling, const GemmDesc* d_descs, const GemmDesc& desc0, int count, int64_t max_work,
                const GemmTile* d_tiles, int64_t ntiles, int grid, hipStream_t s);

// gemm tilings (outputs per lane 4x4, 4x2, 2x2); all give identical bits
int gemm_tiling_count();
bool gemm_set_tiling(int t);
int gemm_get_tiling();

}  // namespace rb
