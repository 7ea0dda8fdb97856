// batch.hip -- ragged batches behind one launch (optionally one hipGraph).
//
// A batch plan holds any number of segment groups (CSR segments of one vector
// or vector pair: binned sums, asums, dots or nrm2s) and at most one gemm
// group (many small gemms of one fold and layout). Planning happens once on
// the host: segments are cut into chunks, gemms into output tiles sorted by K
// (heaviest first); the tables live in device memory owned by the plan.
// Launching runs, per segment group, a persistent grid that walks its chunk
// table plus a small finishing kernel, and, for the gemm group, one workgroup
// per output tile on a forked side stream (overlapping the HBM-bound segments
// with the VALU-bound gemms); with graph launch the whole plan is captured
// once into a hipGraph and replayed. No kernel of a plan writes through memset or
// atomics, so replays need no resets.
//
// None of this has a counterpart in the reference, which offers one call per
// reduction; the bits of every result equal that call's (one binned value per
// segment or output, merged exactly).
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/reproblas_hip.h"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace rb;

namespace {

struct SegGroup {
  Term term;
  const double* x;
  const double* y;
  int64_t nseg;
  SegChunk* d_chunks = nullptr;
  int64_t nchunks = 0;
  int64_t* d_first = nullptr;
  double* d_partials = nullptr;
  double* results;
  int grid;
};

struct GemmGroup {
  int count = 0;
  GemmDesc desc0;
  GemmDesc* d_descs = nullptr;
  GemmTile* d_tiles = nullptr;
  int64_t ntiles = 0;
  double* d_maxima = nullptr;
  int64_t max_work = 0;
  int tiling = 1;
  int grid = 1;
};

}  // namespace

struct reproblas_hip_batch {
  int fold;
  int device;
  std::vector<SegGroup> segs;
  std::vector<GemmGroup> gemms;
  hipStream_t capture = nullptr;
  hipStream_t side = nullptr;  // gemm branch of a plan that also has segments
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// Persistent workgroups per CU walking the chunk table of a segment group
// (compile-time overridable for tuning builds). Two leave VGPR room on every
// SIMD for gemm tiles running beside them.
#ifndef RB_SEG_BLOCKS_PER_CU
#define RB_SEG_BLOCKS_PER_CU 2
#endif
constexpr int kSegBlocksPerCU = RB_SEG_BLOCKS_PER_CU;

// Segment groups run on `s`. When the plan also has gemms, they run on a side
// stream forked from `s` and joined back into it: the segments are HBM-bound
// and the gemm tiles VALU-bound, so the two overlap on the same CUs. Under
// graph capture the fork and join become parallel branches of the graph.
bool launch_all(reproblas_hip_batch* b, hipStream_t s) {
  const bool fork = !b->segs.empty() && !b->gemms.empty();
  hipStream_t gs = s;
  if (fork) {
    if (!b->side && !RB_HIP(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking))) return false;
    if (!b->ev_fork && !RB_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming))) return false;
    if (!b->ev_join && !RB_HIP(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming))) return false;
    if (!RB_HIP(hipEventRecord(b->ev_fork, s))) return false;
    if (!RB_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0))) return false;
    gs = b->side;
  }
  bool ok = true;
  for (const SegGroup& g : b->segs)
    if (ok && !segments(b->fold, g.term, g.x, g.y, g.d_chunks, g.nchunks, g.d_partials, g.d_first, g.nseg,
                        g.results, g.grid, s))
      ok = false;
  for (const GemmGroup& g : b->gemms)
    if (ok && !gemm_batch(b->fold, g.tiling, g.d_descs, g.desc0, g.count, g.max_work, g.d_tiles, g.ntiles, g.grid,
                          gs))
      ok = false;
  if (fork) {
    // join even after a failed launch, so a capture in progress stays whole
    if (!RB_HIP(hipEventRecord(b->ev_join, gs))) return false;
    if (!RB_HIP(hipStreamWaitEvent(s, b->ev_join, 0))) return false;
  }
  return ok;
}

void free_group(SegGroup& g) {
  (void)hipFree(g.d_chunks);
  (void)hipFree(g.d_first);
  (void)hipFree(g.d_partials);
  g.d_chunks = nullptr;
  g.d_first = nullptr;
  g.d_partials = nullptr;
}

void free_group(GemmGroup& g) {
  (void)hipFree(g.d_descs);
  (void)hipFree(g.d_tiles);
  (void)hipFree(g.d_maxima);
  g.d_descs = nullptr;
  g.d_tiles = nullptr;
  g.d_maxima = nullptr;
}

// Frees a group that failed to build and returns the error already set.
template <class G>
int fail_group(G& g) {
  free_group(g);
  return last_error();
}

void drop_graph(reproblas_hip_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

template <class T>
T* upload(const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* d = nullptr;
  if (!RB_HIP(hipMalloc((void**)&d, sizeof(T) * v.size()))) return nullptr;
  if (!RB_HIP(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice))) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

}  // namespace

extern "C" {

reproblas_hip_batch* reproblas_hip_batch_create(int fold) {
  clear_error();
  if (!fold_supported(fold)) {
    set_error(kBadFold, "fold %d is outside the supported range [2, 52]", fold);
    return nullptr;
  }
  reproblas_hip_batch* b = new reproblas_hip_batch();
  b->fold = fold;
  if (!RB_HIP(hipGetDevice(&b->device))) {
    delete b;
    return nullptr;
  }
  return b;
}

int reproblas_hip_batch_add_segments(reproblas_hip_batch* b, int op, int64_t nseg, const int64_t* offsets,
                                     const double* X, const double* Y, double* results, int64_t chunk) {
  clear_error();
  if (!b || nseg < 0 || !offsets || (nseg > 0 && (!X || !results)) || op < 0 || op > 3 || (op == kDot && !Y)) {
    set_error(kBadArgument, "reproblas_hip_batch_add_segments: bad argument");
    return last_error();
  }
  if (nseg > 0 && (!require_device_accessible(X, "X") || (op == kDot && !require_device_accessible(Y, "Y")) ||
                   !require_device_accessible(results, "results")))
    return last_error();
  if (chunk <= 0) chunk = 16384;
  chunk = (chunk + 127) / 128 * 128;
  drop_graph(b);
  SegGroup g;
  g.term = (Term)op;
  g.x = X;
  g.y = Y;
  g.nseg = nseg;
  g.results = results;
  std::vector<SegChunk> chunks;
  std::vector<int64_t> first((size_t)nseg + 1);
  for (int64_t sgi = 0; sgi < nseg; ++sgi) {
    first[sgi] = (int64_t)chunks.size();
    const int64_t lo = offsets[sgi], hi = offsets[sgi + 1];
    if (hi < lo) {
      set_error(kBadArgument, "reproblas_hip_batch_add_segments: offsets must not decrease (segment %lld)",
                (long long)sgi);
      return last_error();
    }
    for (int64_t a = lo; a < hi; a += chunk) chunks.push_back(SegChunk{a, std::min(hi, a + chunk)});
  }
  first[nseg] = (int64_t)chunks.size();
  g.nchunks = (int64_t)chunks.size();
  g.d_chunks = upload(chunks);
  g.d_first = upload(first);
  if ((g.nchunks && !g.d_chunks) || !g.d_first) return fail_group(g);
  if (g.nchunks && !RB_HIP(hipMalloc((void**)&g.d_partials, sizeof(double) * g.nchunks * (2 * b->fold + 1))))
    return fail_group(g);
  const DeviceInfo* d = device_info();
  if (!d) return fail_group(g);
  int64_t grid = (g.nchunks + 3) / 4;
  grid = std::min<int64_t>(grid, (int64_t)d->cus * kSegBlocksPerCU);
  g.grid = (int)std::max<int64_t>(grid, 1);
  b->segs.push_back(g);
  return kOk;
}

int reproblas_hip_batch_add_gemms(reproblas_hip_batch* b, char Order, char TransA, char TransB, int count,
                                  const int64_t* M, const int64_t* N, const int64_t* K, const double* alpha,
                                  const double* const* A, const int64_t* lda, const double* const* B,
                                  const int64_t* ldb, const double* beta, double* const* C, const int64_t* ldc,
                                  int tiling) {
  clear_error();
  if (!b || count < 0 || (count > 0 && (!M || !N || !K || !alpha || !A || !lda || !B || !ldb || !beta || !C || !ldc))) {
    set_error(kBadArgument, "reproblas_hip_batch_add_gemms: bad argument");
    return last_error();
  }
  if (tiling < 0) tiling = gemm_get_tiling();
  if (tiling > 2) {
    set_error(kBadArgument, "gemm tiling %d does not exist (0..2)", tiling);
    return last_error();
  }
  drop_graph(b);
  GemmGroup g;
  g.count = count;
  g.tiling = tiling;
  std::vector<GemmDesc> descs(count);
  int64_t maxima = 0;
  for (int i = 0; i < count; ++i) {
    if (M[i] <= 0 || N[i] <= 0 || K[i] <= 0 || alpha[i] == 0.0) {
      set_error(kUnsupported, "reproblas_hip_batch_add_gemms: gemm %d has an empty product (M, N, K > 0 and alpha != 0 "
                "are required in a batch; call reproBLAS_rdgemm for it)", i);
      return last_error();
    }
    if (!require_device_accessible(A[i], "A") || !require_device_accessible(B[i], "B") ||
        !require_device_accessible(C[i], "C"))
      return last_error();
    descs[i] = make_gemm_desc(b->fold, Order, TransA, TransB, M[i], N[i], K[i], alpha[i], A[i], lda[i], B[i], ldb[i],
                              beta[i], C[i], ldc[i], nullptr, 0);
    maxima += ((K[i] + 31) / 32) * (M[i] + N[i]);
    g.max_work = std::max<int64_t>(g.max_work, ((K[i] + 31) / 32) * (M[i] + N[i]));
  }
  if (count == 0) return kOk;
  if (!RB_HIP(hipMalloc((void**)&g.d_maxima, sizeof(double) * maxima))) return last_error();
  int64_t off = 0;
  for (int i = 0; i < count; ++i) {
    const int64_t ch = (K[i] + 31) / 32;
    descs[i].rowmax = g.d_maxima + off;
    descs[i].colmax = g.d_maxima + off + ch * M[i];
    off += ch * (M[i] + N[i]);
  }
  // tiles, heaviest (largest K) first so the longest tiles are dispatched first
  int bm = 64, bn = 32;
  gemm_tile_shape(tiling, b->fold, &bm, &bn);
  std::vector<GemmTile> tiles;
  for (int i = 0; i < count; ++i)
    for (int32_t ti = 0; ti < (int32_t)((M[i] + bm - 1) / bm); ++ti)
      for (int32_t tj = 0; tj < (int32_t)((N[i] + bn - 1) / bn); ++tj) tiles.push_back(GemmTile{i, ti, tj, 0});
  std::stable_sort(tiles.begin(), tiles.end(),
                   [&](const GemmTile& x, const GemmTile& y) { return K[x.gemm] > K[y.gemm]; });
  g.ntiles = (int64_t)tiles.size();
  g.desc0 = descs[0];
  g.d_descs = upload(descs);
  g.d_tiles = upload(tiles);
  if (!g.d_descs || !g.d_tiles) return fail_group(g);
  const DeviceInfo* d = device_info();
  if (!d) return fail_group(g);
  g.grid = (int)std::min<int64_t>(g.ntiles, INT_MAX);  // one workgroup per tile
  b->gemms.push_back(g);
  return kOk;
}

int reproblas_hip_batch_launch(reproblas_hip_batch* b, void* stream) {
  clear_error();
  if (!b) {
    set_error(kBadArgument, "reproblas_hip_batch_launch: null plan");
    return last_error();
  }
  return launch_all(b, (hipStream_t)stream) ? kOk : last_error();
}

int reproblas_hip_batch_graph_launch(reproblas_hip_batch* b, void* stream) {
  clear_error();
  if (!b) {
    set_error(kBadArgument, "reproblas_hip_batch_graph_launch: null plan");
    return last_error();
  }
  if (!b->exec) {
    if (!b->capture && !RB_HIP(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking))) return last_error();
    if (!RB_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal))) return last_error();
    const bool ok = launch_all(b, b->capture);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->capture, &graph);
    if (!ok) {
      if (graph) (void)hipGraphDestroy(graph);
      return last_error();
    }
    if (!hip_ok(e, "hipStreamEndCapture")) return last_error();
    b->graph = graph;
    if (!RB_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0))) {
      drop_graph(b);
      return last_error();
    }
  }
  return RB_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream)) ? kOk : last_error();
}

void reproblas_hip_batch_destroy(reproblas_hip_batch* b) {
  if (!b) return;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();
  drop_graph(b);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  for (SegGroup& g : b->segs) free_group(g);
  for (GemmGroup& g : b->gemms) free_group(g);
  if (prev >= 0) (void)hipSetDevice(prev);
  delete b;
}

}  // extern "C"
