// gemm.hip -- binned gemm on device operands.
//
// Reference: binnedBLAS_dbdgemm (src/binnedBLAS/dbdgemm.ccog:61-420) under
// reproBLAS_rdgemm (src/reproBLAS/rdgemm.c:41-105):
//   C_ij = ddbconv( dbdconv(beta*C_ij) (+) sum_k fl(fl(alpha*A_ik) * B_kj) )
// alpha is applied to A (every branch of dbdgemm.ccog), beta == 0 ignores C,
// beta == 1 converts C unscaled. Every output is one binned dot of K terms,
// so the tile shape and the chunking of K are free.
//
// Work layout (FP64 VALU bound; MFMA is not used because its fused
// accumulation would round products and partial sums differently):
//   * a 256-thread workgroup owns a (16*TM) x (16*TN) tile of C; each lane
//     owns TM x TN outputs, each a fold-F binned value: primaries in VGPRs,
//     carries as int32 in VGPRs (a carry changes only at a renorm, by -2..1);
//   * K is walked in chunks of KC = 32: the chunk of A (times alpha) and B is
//     staged in LDS, then every lane runs the deposit of all TM*TN products
//     for each k -- exactly 1 mul + (4F-2) add/or per product;
//   * an output's index must equal dindex(max_k |fl(a*b)|). Per chunk a lane
//     compares a bound fl(fl(|alpha| max|a|) max|b|) (row and column maxima
//     from a small pre-pass) with the admissible limit of its outputs; only
//     when the bound reaches it does the lane take the exact per-product max
//     of the chunk and raise indices. Both give the reference's index.
//   * renormalisation every 2048 deposits; the epilogue folds in beta*C (or
//     the caller's binned C), converts and writes.
// Non-finite products poison an output's top primary; the epilogue then
// recomputes that output's class from its K terms (NaN, +inf or -inf).
#include <limits.h>

#include <algorithm>

#include "binned_core.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace rb {

namespace {

constexpr int kThreads = 256;
constexpr int KC = 32;



// ---- int-carry variants of renorm / raise (carries are small integers) ----
template <int CAP, class F>
__device__ __forceinline__ void renorm_ic(int fold, F (&p)[CAP], int (&c)[CAP]) {
  if (p[0] == F(0) || nan_or_inf(p[0])) return;
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      if constexpr (sizeof(F) == 8) {
        // bits 51:50 of the mantissa live in bits 19:18 of the high word
        uint32_t hi = (uint32_t)(bits(p[j]) >> 32);
        c[j] += (int)((hi >> 18) & 3u) - 2;
        hi = (hi & ~(1u << 18)) | (1u << 19);
        p[j] = from_bits(((uint64_t)hi << 32) | (bits(p[j]) & 0xffffffffull));
      } else {
        uint32_t u = bits(p[j]);  // float: bits 22:21
        c[j] += (int)((u >> 21) & 3u) - 2;
        u = (u & ~(1u << 21)) | (1u << 22);
        p[j] = fp_from_bits<float>(u);
      }
    }
  }
}

template <int CAP, class F>
__device__ __forceinline__ void raise_ic(int fold, int shift, int I, F (&p)[CAP], int (&c)[CAP]) {
  if (shift <= 0) return;
  F np[CAP];
  int nc[CAP];
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j) {
    if (j < fold) {
      if (j >= shift) {
        F v = F(0);
        int w = 0;
RB_UNROLL_CAP
        for (int t = 0; t < CAP; ++t)
          if (t == j - shift) { v = p[t]; w = c[t]; }
        np[j] = v;
        nc[j] = w;
      } else {
        np[j] = bin<F>(I + j);
        nc[j] = 0;
      }
    }
  }
RB_UNROLL_CAP
  for (int j = 0; j < CAP; ++j)
    if (j < fold) { p[j] = np[j]; c[j] = nc[j]; }
}

// ---------------------------------------------------------------- prepass --
// rowmax[ch][i] = max_{k in chunk ch} |A(i,k)|, colmax[ch][j] likewise for B.
template <class F>
__global__ __launch_bounds__(256) void chunk_max_kernel(int64_t rows, int64_t K, const F* X, int64_t s_row,
                                                        int64_t s_k, F* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t chunks = (K + KC - 1) / KC;
  if (t >= rows * chunks) return;
  const int64_t r = t % rows, ch = t / rows;
  const int64_t k0 = ch * KC;
  const int64_t k1 = (k0 + KC < K) ? k0 + KC : K;
  F m = F(0);
  for (int64_t k = k0; k < k1; ++k) m = fmax(m, fabs(X[r * s_row + k * s_k]));
  out[ch * rows + r] = m;
}

// The same maxima for an operand whose k index is contiguous (s_k == 1):
// each half-wave reads one chunk of a row (32 consecutive elements, one per
// lane; KC == 32) and reduces it across its lanes, so the reads coalesce.
// (Thread-per-(row, chunk) reads would be a whole leading dimension apart.)
static_assert(KC == 32, "one chunk per half-wave");
template <class F>
__global__ __launch_bounds__(256) void chunk_max_kcontig_kernel(int64_t rows, int64_t K, const F* X, int64_t s_row,
                                                                F* out) {
  const int lane = threadIdx.x & 63;
  const int64_t chunks = (K + KC - 1) / KC;
  const int64_t segs = (chunks + 1) / 2;  // 64-element segments per row
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; w < rows * segs; w += nw) {
    const int64_t r = w / segs, sg = w - r * segs;
    const int64_t k = sg * 64 + lane;
    F m = k < K ? fabs(X[r * s_row + k]) : F(0);
#pragma unroll
    for (int off = 1; off < 32; off *= 2) m = fmax(m, __shfl_xor(m, off, 64));
    m = fmax(F(0), m);  // as the running max from 0 of chunk_max_kernel
    const int64_t ch = 2 * sg + (lane >> 5);
    if ((lane & 31) == 0 && ch < chunks) out[ch * rows + r] = m;
  }
}

// Chunk maxima of one operand (rows x K, elements s_row / s_k apart).
template <class F>
void launch_chunk_max(int64_t rows, int64_t K, const F* X, int64_t s_row, int64_t s_k, F* out, hipStream_t s) {
  if (s_k == 1) {
    const int64_t segs = rows * ((K + 2 * KC - 1) / (2 * KC));
    const int64_t blocks = std::min<int64_t>((segs + 3) / 4, 65536);
    chunk_max_kcontig_kernel<<<(unsigned)blocks, 256, 0, s>>>(rows, K, X, s_row, out);
  } else {
    const int64_t chunks = (K + KC - 1) / KC;
    chunk_max_kernel<<<(unsigned)((rows * chunks + 255) / 256), 256, 0, s>>>(rows, K, X, s_row, s_k, out);
  }
}

template <int CAP, class F>
__device__ __forceinline__ void beta_init(int fold, F beta, F yv, F (&p)[CAP], F (&c)[CAP]) {
  set_zero(fold, p, c);
  if (beta == F(0)) return;
  const F v = (beta == F(1)) ? yv : yv * beta;
  update(fold, v, p, c);
  if (nan_or_inf(v) || nan_or_inf(p[0])) {
    p[0] += v;
  } else {
    deposit(fold, dmindex0(p[0]), v, p);
  }
  renorm(fold, p, c);
}

// ------------------------------------------------------------------ kernel --
// Row padding of the LDS tiles: keeps rows 16-byte aligned and spreads the
// staging writes over banks.
template <class F>
constexpr int pad() { return 16 / (int)sizeof(F); }

// One (16*TM) x (16*TN) tile of C at (i0, j0), all of K.
template <int CAP, bool FIXED, int TM, int TN, bool AI, bool BK, class F>
__device__ __forceinline__ void gemm_tile(const GemmDescT<F>& g, int64_t i0, int64_t j0,
                                          F (&As)[KC][16 * TM + pad<F>()], F (&Bs)[KC][16 * TN + pad<F>()]) {
  constexpr int BM = 16 * TM, BN = 16 * TN;
  const int fold = FIXED ? CAP : g.fold;
  const int tid = threadIdx.x;
  const int tr = tid / 16, tc = tid % 16;

  F p[TM][TN][CAP];
  int c[TM][TN][CAP];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
RB_UNROLL_CAP
      for (int f = 0; f < CAP; ++f) {
        p[i][j][f] = bin<F>(Fmt<F>::max_index + f);
        c[i][j][f] = 0;
      }
  F lmin = admit_limit<F>(Fmt<F>::max_index);  // smallest admissible limit over the lane's outputs
  bool anytop = false;                   // some output of the lane at index 0
  int dep = 0;
  const F aalpha = fabs(g.alpha);
  const int64_t chunks = (g.K + KC - 1) / KC;

  // Staging: this thread's element r of a chunk sits at (li0 + r*dli, lka0 +
  // r*dlka) of the A tile and (lj0 + r*dlj, lkb0 + r*dlkb) of the B tile, the
  // same in every chunk: its offset from the chunk's corner is off + r*step,
  // with the step the same for every thread. Interior tiles and chunks load
  // without bounds tests.
  constexpr int RA = (BM * KC) / kThreads, RB = (BN * KC) / kThreads;
  const int li0 = AI ? tid % BM : tid / KC, lka0 = AI ? tid / BM : tid % KC;
  constexpr int dli = AI ? 0 : kThreads / KC, dlka = AI ? kThreads / BM : 0;
  const int lj0 = BK ? tid / KC : tid % BN, lkb0 = BK ? tid % KC : tid / BN;
  constexpr int dlj = BK ? kThreads / KC : 0, dlkb = BK ? 0 : kThreads / BN;
  const int64_t offA = li0 * g.sai + lka0 * g.sak, stepA = dli * g.sai + dlka * g.sak;
  const int64_t offB = lkb0 * g.sbk + lj0 * g.sbj, stepB = dlkb * g.sbk + dlj * g.sbj;
  const bool rows_in = i0 + BM <= g.M, cols_in = j0 + BN <= g.N;  // wave-uniform

  for (int64_t ch = 0; ch < chunks; ++ch) {
    const int64_t k0 = ch * KC;
    // ---- stage A(i0.., k0..) * alpha and B(k0.., j0..) into LDS
    const bool k_in = k0 + KC <= g.K;
    const F* Ac = g.A + i0 * g.sai + k0 * g.sak + offA;
    const F* Bc = g.B + k0 * g.sbk + j0 * g.sbj + offB;
    if (rows_in && k_in) {
#pragma unroll
      for (int r = 0; r < RA; ++r) {
        F v = Ac[r * stepA];
        if (g.alpha != F(1)) v = v * g.alpha;
        As[lka0 + r * dlka][li0 + r * dli] = v;
      }
    } else {
#pragma unroll
      for (int r = 0; r < RA; ++r) {
        const int li = li0 + r * dli, lk = lka0 + r * dlka;
        F v = (i0 + li < g.M && k0 + lk < g.K) ? Ac[r * stepA] : F(0);
        if (g.alpha != F(1)) v = v * g.alpha;
        As[lk][li] = v;
      }
    }
    if (cols_in && k_in) {
#pragma unroll
      for (int r = 0; r < RB; ++r) Bs[lkb0 + r * dlkb][lj0 + r * dlj] = Bc[r * stepB];
    } else {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int lj = lj0 + r * dlj, lk = lkb0 + r * dlkb;
        Bs[lk][lj] = (j0 + lj < g.N && k0 + lk < g.K) ? Bc[r * stepB] : F(0);
      }
    }
    __syncthreads();

    // ---- index guard for this chunk
    F ra = F(0), cb = F(0);
    if (rows_in) {  // all loads in flight together
#pragma unroll
      for (int i = 0; i < TM; ++i) ra = fmax(ra, g.rowmax[ch * g.M + i0 + tr * TM + i]);
    } else {
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int64_t gi = i0 + tr * TM + i;
        if (gi < g.M) ra = fmax(ra, g.rowmax[ch * g.M + gi]);
      }
    }
    if (cols_in) {
#pragma unroll
      for (int j = 0; j < TN; ++j) cb = fmax(cb, g.colmax[ch * g.N + j0 + tc * TN + j]);
    } else {
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int64_t gj = j0 + tc * TN + j;
        if (gj < g.N) cb = fmax(cb, g.colmax[ch * g.N + gj]);
      }
    }
    const F bound = (aalpha * ra) * cb;
    if (!(bound < lmin)) {
      // exact max |fl(a*b)| of each output over the chunk, then raise; one
      // output at a time keeps this rare path's registers few
      lmin = (F)__builtin_inf();
      anytop = false;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          F mx = F(0);
#pragma unroll 1
          for (int k = 0; k < KC; ++k) mx = fmax(mx, fabs(As[k][tr * TM + i] * Bs[k][tc * TN + j]));
          const int cur = dmindex(p[i][j][0]);
          const int nI = dindex(mx);
          if (!nan_or_inf(p[i][j][0]) && nI < cur) raise_ic(fold, cur - nI, nI, p[i][j], c[i][j]);
          const bool live = (i0 + tr * TM + i < g.M) && (j0 + tc * TN + j < g.N);
          if (live && !nan_or_inf(p[i][j][0])) lmin = fmin(lmin, admit_limit<F>(dmindex(p[i][j][0])));
          anytop |= dmindex0(p[i][j][0]);
        }
    }
    if (dep + KC > Fmt<F>::endurance) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) renorm_ic(fold, p[i][j], c[i][j]);
      dep = 0;
    }

    // ---- deposits: 1 mul + (4F-2) add/or per product
    if (!anytop) {
      // two k steps per trip: the bins then alternate between two register
      // sets instead of being copied back at the end of every trip
#pragma unroll 2
      for (int k = 0; k < KC; ++k) {
        F a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = As[k][tr * TM + i];
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = Bs[k][tc * TN + j];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) deposit(fold, false, a[i] * b[j], p[i][j]);
      }
    } else {
#pragma unroll 1
      for (int k = 0; k < KC; ++k) {
        F a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) a[i] = As[k][tr * TM + i];
#pragma unroll
        for (int j = 0; j < TN; ++j) b[j] = Bs[k][tc * TN + j];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) deposit(fold, dmindex0(p[i][j][0]), a[i] * b[j], p[i][j]);
      }
    }
    dep += KC;
    __syncthreads();
  }

  // ---- epilogue
#pragma unroll
  for (int i = 0; i < TM; ++i) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int64_t gi = i0 + tr * TM + i, gj = j0 + tc * TN + j;
      if (gi >= g.M || gj >= g.N) continue;
      renorm_ic(fold, p[i][j], c[i][j]);
      F ps[CAP], cs[CAP];
RB_UNROLL_CAP
      for (int f = 0; f < CAP; ++f) {
        ps[f] = p[i][j][f];
        cs[f] = (F)c[i][j][f];
      }
      if (nan_or_inf(ps[0])) {
        bool nan = false, pinf = false, ninf = false;
        for (int64_t k = 0; k < g.K; ++k) {
          F av = g.A[gi * g.sai + k * g.sak];
          if (g.alpha != F(1)) av = av * g.alpha;
          const F t = av * g.B[k * g.sbk + gj * g.sbj];
          nan |= (t != t);
          pinf |= (t == (F)__builtin_inf());
          ninf |= (t == -(F)__builtin_inf());
        }
        ps[0] = (nan || (pinf && ninf) || !(pinf || ninf)) ? (F)__builtin_nan("")
                                                            : (pinf ? (F)__builtin_inf() : -(F)__builtin_inf());
      }
      if (g.Cb) {
        F* cbp = g.Cb + (gi * g.cbi + gj * g.cbj) * 2 * fold;
        F py[CAP], cy[CAP];
RB_UNROLL_CAP
        for (int f = 0; f < CAP; ++f)
          if (f < fold) { py[f] = cbp[f]; cy[f] = cbp[fold + f]; }
        renorm(fold, py, cy);
        add(fold, py, cy, ps, cs);
        renorm(fold, py, cy);
RB_UNROLL_CAP
        for (int f = 0; f < CAP; ++f)
          if (f < fold) { cbp[f] = py[f]; cbp[fold + f] = cy[f]; }
      } else {
        F* cp = g.C + gi * g.sci + gj * g.scj;
        F py[CAP], cy[CAP];
        beta_init(fold, g.beta, g.beta == F(0) ? F(0) : *cp, py, cy);
        add(fold, py, cy, ps, cs);
        renorm(fold, py, cy);
        *cp = convert(fold, py, cy);
      }
    }
  }
}

template <int CAP, bool FIXED, int TM, int TN, bool AI, bool BK, class F>
__global__ __launch_bounds__(kThreads) void gemm_kernel(GemmDescT<F> g) {
  __shared__ __attribute__((aligned(16))) F As[KC][16 * TM + pad<F>()];
  __shared__ __attribute__((aligned(16))) F Bs[KC][16 * TN + pad<F>()];
  // one linear tile index (row tiles on grid.y would cap M at 65535 tiles);
  // column tiles vary fastest, the dispatch order of a 2-D grid
  const unsigned tiles_n = (unsigned)((g.N + 16 * TN - 1) / (16 * TN));
  const unsigned ti = blockIdx.x / tiles_n, tj = blockIdx.x - ti * tiles_n;
  gemm_tile<CAP, FIXED, TM, TN, AI, BK>(g, (int64_t)ti * 16 * TM, (int64_t)tj * 16 * TN, As, Bs);
}

// Ragged gemm batch: one workgroup per entry of a tile list (gemm, tile row,
// tile column) sorted heaviest K first, so the longest tiles are dispatched
// first. (One tile per workgroup keeps the register count of gemm_kernel; a
// persistent loop over tiles costs the compiler ~80 more VGPRs.)
template <int CAP, bool FIXED, int TM, int TN, bool AI, bool BK>
__global__ __launch_bounds__(kThreads) void gemm_batch_kernel(const GemmDesc* descs, const GemmTile* tiles,
                                                              int64_t ntiles) {
  __shared__ __attribute__((aligned(16))) double As[KC][16 * TM + pad<double>()];
  __shared__ __attribute__((aligned(16))) double Bs[KC][16 * TN + pad<double>()];
  if ((int64_t)blockIdx.x >= ntiles) return;
  const GemmTile tl = tiles[blockIdx.x];
  const GemmDesc g = descs[tl.gemm];
  gemm_tile<CAP, FIXED, TM, TN, AI, BK>(g, (int64_t)tl.ti * 16 * TM, (int64_t)tl.tj * 16 * TN, As, Bs);
}

// Chunk maxima for every gemm of a batch (blockIdx.y = gemm).
__global__ __launch_bounds__(256) void chunk_max_batch_kernel(const GemmDesc* descs) {
  const GemmDesc g = descs[blockIdx.y];
  const int64_t chunks = (g.K + KC - 1) / KC;
  const int64_t total = (g.M + g.N) * chunks;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const bool isA = t < g.M * chunks;
    const int64_t u = isA ? t : t - g.M * chunks;
    const int64_t rows = isA ? g.M : g.N;
    const int64_t r = u % rows, ch = u / rows;
    const int64_t k0 = ch * KC, k1 = (k0 + KC < g.K) ? k0 + KC : g.K;
    const double* X = isA ? g.A : g.B;
    const int64_t s_row = isA ? g.sai : g.sbj, s_k = isA ? g.sak : g.sbk;
    double m = 0.0;
    for (int64_t k = k0; k < k1; ++k) m = fmax(m, fabs(X[r * s_row + k * s_k]));
    (isA ? const_cast<double*>(g.rowmax) : const_cast<double*>(g.colmax))[ch * rows + r] = m;
  }
}

template <int CAP, bool FIXED, int TM, int TN, class F>
bool launch_tm(bool ai, bool bk, const GemmDescT<F>& g, hipStream_t s) {
  const int64_t tiles = ((g.N + 16 * TN - 1) / (16 * TN)) * ((g.M + 16 * TM - 1) / (16 * TM));
  if (tiles > INT_MAX) {
    set_error(kBadArgument, "gemm of %lld output tiles exceeds one launch", (long long)tiles);
    return false;
  }
  const dim3 grid((unsigned)tiles);
  if (ai && bk) gemm_kernel<CAP, FIXED, TM, TN, true, true, F><<<grid, kThreads, 0, s>>>(g);
  else if (ai) gemm_kernel<CAP, FIXED, TM, TN, true, false, F><<<grid, kThreads, 0, s>>>(g);
  else if (bk) gemm_kernel<CAP, FIXED, TM, TN, false, true, F><<<grid, kThreads, 0, s>>>(g);
  else gemm_kernel<CAP, FIXED, TM, TN, false, false, F><<<grid, kThreads, 0, s>>>(g);
  return RB_HIP(hipGetLastError());
}

template <int CAP, bool FIXED, int TM, int TN>
bool launch_batch_tm(bool ai, bool bk, const GemmDesc* d, const GemmTile* t, int64_t n, int grid, hipStream_t s) {
  if (ai && bk) gemm_batch_kernel<CAP, FIXED, TM, TN, true, true><<<grid, kThreads, 0, s>>>(d, t, n);
  else if (ai) gemm_batch_kernel<CAP, FIXED, TM, TN, true, false><<<grid, kThreads, 0, s>>>(d, t, n);
  else if (bk) gemm_batch_kernel<CAP, FIXED, TM, TN, false, true><<<grid, kThreads, 0, s>>>(d, t, n);
  else gemm_batch_kernel<CAP, FIXED, TM, TN, false, false><<<grid, kThreads, 0, s>>>(d, t, n);
  return RB_HIP(hipGetLastError());
}

// Tiling of the calling host thread (0: 4x4 outputs per lane, 1: 4x2, 2: 2x2);
// every tiling gives the same bits, so threads may differ freely.
thread_local int g_tiling = 1;

// Tiling actually used for a fold (folds without a dedicated instance use 2x2).
int effective_tiling(int tiling, int fold) { return (fold >= 2 && fold <= 4) ? tiling : 2; }

template <int CAP, bool FIXED, class F>
bool launch_fold(int tiling, bool ai, bool bk, const GemmDescT<F>& g, hipStream_t s) {
  if constexpr (!FIXED) {
    return launch_tm<CAP, FIXED, 2, 2>(ai, bk, g, s);  // folds 5..8: fewest registers
  } else {
    switch (tiling) {
      case 0: return launch_tm<CAP, FIXED, 4, 4>(ai, bk, g, s);
      case 1: return launch_tm<CAP, FIXED, 4, 2>(ai, bk, g, s);
      default: return launch_tm<CAP, FIXED, 2, 2>(ai, bk, g, s);
    }
  }
}

template <int CAP, bool FIXED>
bool launch_batch_fold(int tiling, bool ai, bool bk, const GemmDesc* d, const GemmTile* t, int64_t n, int grid,
                       hipStream_t s) {
  if constexpr (!FIXED) {
    return launch_batch_tm<CAP, FIXED, 2, 2>(ai, bk, d, t, n, grid, s);
  } else {
    switch (tiling) {
      case 0: return launch_batch_tm<CAP, FIXED, 4, 4>(ai, bk, d, t, n, grid, s);
      case 1: return launch_batch_tm<CAP, FIXED, 4, 2>(ai, bk, d, t, n, grid, s);
      default: return launch_batch_tm<CAP, FIXED, 2, 2>(ai, bk, d, t, n, grid, s);
    }
  }
}

template <class F>
bool gemm_common(const GemmDescT<F>& g0, hipStream_t s) {
  const int fold = g0.fold;
  if (fold < 2 || fold > Fmt<F>::max_fold) {
    set_error(kBadFold, "fold %d is outside the supported range [2, %d]", fold, Fmt<F>::max_fold);
    return false;
  }
  GemmDescT<F> g = g0;
  const int64_t chunks = (g.K + KC - 1) / KC;
  F* mx = (F*)workspace(s, kSlotGemmMax, sizeof(F) * (size_t)chunks * (g.M + g.N));
  if (!mx) return false;
  g.rowmax = mx;
  g.colmax = mx + chunks * g.M;
  launch_chunk_max(g.M, g.K, g.A, g.sai, g.sak, mx, s);
  launch_chunk_max(g.N, g.K, g.B, g.sbj, g.sbk, mx + chunks * g.M, s);
  if (!RB_HIP(hipGetLastError())) return false;
  const bool ai = (g.sai == 1), bk = (g.sbk == 1);
  switch (fold) {
    case 2: return launch_fold<2, true>(g_tiling, ai, bk, g, s);
    case 3: return launch_fold<3, true>(g_tiling, ai, bk, g, s);
    case 4: return launch_fold<4, true>(g_tiling, ai, bk, g, s);
    default:
      if (fold <= 8) return launch_fold<8, false>(g_tiling, ai, bk, g, s);
      return launch_fold<Fmt<F>::max_fold, false>(g_tiling, ai, bk, g, s);
  }
}

}  // namespace

namespace {
template <class F>
GemmDescT<F> make_desc(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, F alpha,
                       const F* A, int64_t lda, const F* B, int64_t ldb, F beta, F* C, int64_t ldc, F* Cb,
                       int64_t ldcb) {
  const bool rm = (order == 'r' || order == 'R');
  const bool na = (transa == 'n' || transa == 'N');
  const bool nb = (transb == 'n' || transb == 'N');
  GemmDescT<F> g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.B = B;
  // op(A)(i,k): column-major N -> A[i + k*lda]; T -> A[k + i*lda]; row-major swaps
  if (na != rm) { g.sai = 1; g.sak = lda; } else { g.sai = lda; g.sak = 1; }
  if (nb != rm) { g.sbk = 1; g.sbj = ldb; } else { g.sbk = ldb; g.sbj = 1; }
  if (rm) { g.sci = ldc; g.scj = 1; g.cbi = ldcb; g.cbj = 1; } else { g.sci = 1; g.scj = ldc; g.cbi = 1; g.cbj = ldcb; }
  g.alpha = alpha;
  g.C = C;
  g.beta = beta;
  g.Cb = Cb;
  g.rowmax = g.colmax = nullptr;
  g.fold = fold;
  return g;
}

// rdgemm.c:41-105 / rsgemm.c: C = conv(binned(beta*C) (+) sum) in place
template <class F>
bool rgemm_impl(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, F alpha, const F* A,
                int64_t lda, const F* B, int64_t ldb, F beta, F* C, int64_t ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return true;  // rdgemm.c:50-52
  if (K <= 0 || alpha == F(0)) {
    // no products (dbdgemm.ccog:77-80 adds nothing): C still goes through
    // binned(beta*C) and back, one in-place launch over the whole matrix
    const bool rm = (order == 'r' || order == 'R');
    const int64_t outer = rm ? M : N, inner = rm ? N : M;
    return beta_only(fold, inner, outer, C, 1, ldc, beta, s);
  }
  return gemm_common(make_desc(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                               (F*)nullptr, 0), s);
}
}  // namespace

GemmDesc make_gemm_desc(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                        double* Cb, int64_t ldcb) {
  return make_desc<double>(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, Cb, ldcb);
}

void gemm_tile_shape(int tiling, int fold, int* bm, int* bn) {
  const int t = effective_tiling(tiling, fold);
  *bm = 16 * (t == 2 ? 2 : 4);
  *bn = 16 * (t == 0 ? 4 : 2);
}

bool gemm_batch(int fold, int tiling, const GemmDesc* d_descs, const GemmDesc& desc0, int count, int64_t max_work,
                const GemmTile* d_tiles, int64_t ntiles, int grid, hipStream_t s) {
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return false;
  }
  if (count <= 0 || ntiles <= 0) return true;
  (void)grid;  // one workgroup per tile
  if (ntiles > INT_MAX) {
    set_error(kBadArgument, "gemm batch of %lld tiles exceeds one launch", (long long)ntiles);
    return false;
  }
  int64_t mgrid = (max_work + 255) / 256;
  if (mgrid > 1024) mgrid = 1024;
  if (mgrid < 1) mgrid = 1;
  chunk_max_batch_kernel<<<dim3((unsigned)mgrid, (unsigned)count), 256, 0, s>>>(d_descs);
  if (!RB_HIP(hipGetLastError())) return false;
  const bool ai = (desc0.sai == 1), bk = (desc0.sbk == 1);
  switch (fold) {
    case 2: return launch_batch_fold<2, true>(tiling, ai, bk, d_descs, d_tiles, ntiles, (int)ntiles, s);
    case 3: return launch_batch_fold<3, true>(tiling, ai, bk, d_descs, d_tiles, ntiles, (int)ntiles, s);
    case 4: return launch_batch_fold<4, true>(tiling, ai, bk, d_descs, d_tiles, ntiles, (int)ntiles, s);
    default:
      if (fold <= 8) return launch_batch_fold<8, false>(tiling, ai, bk, d_descs, d_tiles, ntiles, (int)ntiles, s);
      return launch_batch_fold<kMaxFold, false>(tiling, ai, bk, d_descs, d_tiles, ntiles, (int)ntiles, s);
  }
}

int gemm_tiling_count() { return 3; }
bool gemm_set_tiling(int t) {
  if (t < 0 || t >= 3) {
    set_error(kBadArgument, "gemm tiling %d does not exist (0..2)", t);
    return false;
  }
  g_tiling = t;
  return true;
}
int gemm_get_tiling() { return g_tiling; }

bool dbdgemm(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double* Cb, int64_t ldc, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0 || alpha == 0.0) return true;  // dbdgemm.ccog:77-80
  return gemm_common(make_desc<double>(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, 0.0, nullptr, 0,
                                       Cb, ldc), s);
}

bool sbsgemm(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, float alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float* Cb, int64_t ldc, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0 || alpha == 0.0f) return true;  // sbsgemm.ccog:78-80
  return gemm_common(make_desc<float>(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, 0.0f, nullptr, 0,
                                      Cb, ldc), s);
}

bool rdgemm_device(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, double alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                   hipStream_t s) {
  return rgemm_impl<double>(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

bool rsgemm_device(int fold, char order, char transa, char transb, int64_t M, int64_t N, int64_t K, float alpha,
                   const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc,
                   hipStream_t s) {
  return rgemm_impl<float>(fold, order, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace rb
