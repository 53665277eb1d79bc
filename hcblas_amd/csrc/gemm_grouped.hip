// Grouped GEMM (no counterpart in the reference): thousands of independent
// problems of uneven size -- fp32, or fp16 in/out with fp32 accumulation --
// each with its own order, transposes and leading dimensions, executed by ONE
// persistent kernel launch that is captured once in a hipGraph and replayed.
//
// Plan (host, once per problem table):
//   - every problem is put in canonical column-major form (row-major is the
//     operand swap C^T = op(B)^T op(A)^T), validated, and copied to a device
//     table;
//   - every problem is cut into macro tiles; the work list of
//     (problem, tile) entries is ordered longest-first (LPT: tile cost ~ K,
//     fp32 tiles weighted by the FP32:FP16 MFMA rate ratio) or in table order;
//   - the persistent launch is captured into a hipGraph.
// Run: the graph launch. Workgroups (a few per CU) take work entries from a
// device counter or by a static stride; the last workgroup to finish resets
// the counters in the kernel, so the graph holds a single kernel node and no
// memset nodes.
// Autotune: times every (tile shape, order, workgroups per CU, queue or
// static schedule) candidate on the plan's own problems and keeps the
// fastest.
//
// Each tile runs the plain GEMM's main loop with the same K step and an
// epilogue with the same arithmetic (staged through LDS so that C is written
// with 16-byte column runs), so a grouped result is bit-identical to
// per-problem hcblas_sgemm / hcblas_hgemm launches.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gemm_kernel.hpp"

struct hcblasGroupedPlan;

namespace hcb {
namespace {

struct ProbDev {
  int dtype;  // 0: fp32, 1: fp16
  int flags;  // bit 0: A' k-contiguous, bit 1: B' k-contiguous, bit 2: C
              // allows 16-byte vector accesses
  int M, N, K;
  int tiles_n;
  float alpha, beta;
  const void *A;
  int64_t lda;
  const void *B;
  int64_t ldb;
  void *C;
  int64_t ldc;
};

constexpr int kF32BK = 32;  // same K steps as the plain kernels
constexpr int kF16BK = 64;

template <int TM>
struct GroupCfg {
  static constexpr int BM = TM, BN = TM;
  static constexpr int LDS =
      tile_lds_bytes<Elem::F32, TM, TM, kF32BK>() >
              tile_lds_bytes<Elem::F16, TM, TM, kF16BK>()
          ? tile_lds_bytes<Elem::F32, TM, TM, kF32BK>()
          : tile_lds_bytes<Elem::F16, TM, TM, kF16BK>();
};

// One tile: the plain kernels' main loop (same K step, so the sums are
// bit-identical to per-problem launches), then the LDS-staged epilogue.
template <Elem E, typename OutT, int TM, bool VL, bool A_K, bool B_K>
__device__ __forceinline__ void tile_body(const ProbDev &p, int m0, int n0,
                                          char *lds) {
  using T = typename ElemT<E>::T;
  constexpr int BK = E == Elem::F32 ? kF32BK : kF16BK;
  using WT = WaveTile<E, TM, TM, 2, 2, 16>;
  typename WT::AccV acc[WT::FM][WT::FN];
  gemm_mainloop<E, A_K, B_K, TM, TM, BK, 2, 2, VL, VL, 16>(
      static_cast<const T *>(p.A), p.lda, static_cast<const T *>(p.B), p.ldb,
      p.M, p.N, p.K, m0, n0, lds, acc);
  store_tile_lds<E, OutT, TM, TM, 2, 2, 16, GroupCfg<TM>::LDS>(
      acc, static_cast<OutT *>(p.C), p.ldc, p.M, p.N, m0, n0, p.alpha,
      p.beta, lds, (p.flags & 4) != 0);
}

template <Elem E, typename OutT, int TM, bool VL>
__device__ __forceinline__ void run_tile(const ProbDev &p, int m0, int n0,
                                         char *lds) {
  switch (p.flags & 3) {
    case 0:
      tile_body<E, OutT, TM, VL, false, false>(p, m0, n0, lds);
      break;
    case 1:
      tile_body<E, OutT, TM, VL, true, false>(p, m0, n0, lds);
      break;
    case 2:
      tile_body<E, OutT, TM, VL, false, true>(p, m0, n0, lds);
      break;
    default:
      tile_body<E, OutT, TM, VL, true, true>(p, m0, n0, lds);
      break;
  }
}

// Dynamic schedule: counters[0] is the next work entry, taken with one
// atomic per tile. Static schedule: workgroup g takes entries g, g + G,
// g + 2G, ... (G = grid size); over a longest-first list that balances like
// the queue without its atomic per tile. counters[1] counts finished
// workgroups; the last one resets both counters.
// OCC: workgroups per CU the launch is sized for (the register budget
// follows from it: 4 waves per workgroup, one per SIMD).
template <int TM, bool VL, int OCC>
__global__ __launch_bounds__(256, OCC) void grouped_gemm_kernel(
    const ProbDev *__restrict__ probs, const int2 *__restrict__ work,
    int total, unsigned int *counters, int dynamic) {
  __shared__ __attribute__((aligned(16))) char lds[GroupCfg<TM>::LDS];
  __shared__ int next;
  for (int s = blockIdx.x;; s += gridDim.x) {
    int w = s;
    if (dynamic) {
      if (threadIdx.x == 0)
        next = (int)__hip_atomic_fetch_add(&counters[0], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      w = next;
      __syncthreads();
    }
    // uniform across the workgroup: the descriptor comes through scalar loads
    w = __builtin_amdgcn_readfirstlane(w);
    if (w >= total) break;
    const int2 it = work[w];
    const ProbDev p = probs[__builtin_amdgcn_readfirstlane(it.x)];
    const int t = __builtin_amdgcn_readfirstlane(it.y);
    const int tm = t / p.tiles_n, tn = t - tm * p.tiles_n;
    if (p.dtype == 0)
      run_tile<Elem::F32, float, TM, VL>(p, tm * TM, tn * TM, lds);
    else
      run_tile<Elem::F16, uint16_t, TM, VL>(p, tm * TM, tn * TM, lds);
  }
  if (threadIdx.x == 0) {
    const unsigned int t = __hip_atomic_fetch_add(
        &counters[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (t == gridDim.x - 1) {
      __hip_atomic_store(&counters[0], 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&counters[1], 0u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

constexpr int kTileShapes[] = {64, 32};
constexpr int kNumShapes = 2;
constexpr int kOccupancies[] = {2, 3, 4};  // workgroups per CU
constexpr int kNumOcc = 3;

}  // namespace
}  // namespace hcb

using namespace hcb;

struct hcblasGroupedPlan {
  int device = 0;
  int count = 0;
  bool vload = true;
  std::vector<ProbDev> probs;  // host copy (canonical)
  ProbDev *d_probs = nullptr;
  int2 *d_work = nullptr;
  size_t work_cap = 0;
  unsigned int *d_counters = nullptr;
  int total = 0;
  int grid = 0;
  int shape = 0;   // index into kTileShapes
  int order = 0;   // 0: longest first, 1: table order
  int occ = 0;     // index into kOccupancies
  int sched = 0;   // 0: atomic queue, 1: static stride
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void free_graph(hcblasGroupedPlan *p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

template <int TM, bool VL>
void launch_occ(hcblasGroupedPlan *p, hipStream_t s) {
  dim3 g(p->grid);
  switch (kOccupancies[p->occ]) {
    case 2:
      grouped_gemm_kernel<TM, VL, 2><<<g, 256, 0, s>>>(
          p->d_probs, p->d_work, p->total, p->d_counters, p->sched == 0);
      break;
    case 3:
      grouped_gemm_kernel<TM, VL, 3><<<g, 256, 0, s>>>(
          p->d_probs, p->d_work, p->total, p->d_counters, p->sched == 0);
      break;
    default:
      grouped_gemm_kernel<TM, VL, 4><<<g, 256, 0, s>>>(
          p->d_probs, p->d_work, p->total, p->d_counters, p->sched == 0);
      break;
  }
}

void launch_grouped(hcblasGroupedPlan *p, hipStream_t s) {
  const int tm = kTileShapes[p->shape];
  if (tm == 64) {
    if (p->vload)
      launch_occ<64, true>(p, s);
    else
      launch_occ<64, false>(p, s);
  } else {
    if (p->vload)
      launch_occ<32, true>(p, s);
    else
      launch_occ<32, false>(p, s);
  }
}

// Builds the work list for (shape, order, occupancy), uploads it and
// captures the graph.
hcblasStatus_t build(hcblasHandle_t h, hcblasGroupedPlan *p, int shape,
                     int order, int occ, int sched) {
  const int tm = kTileShapes[shape];
  std::vector<int2> work;
  std::vector<int64_t> cost;
  for (int i = 0; i < p->count; ++i) {
    ProbDev &q = p->probs[i];
    if (q.M == 0 || q.N == 0) continue;
    const int tmi = (q.M + tm - 1) / tm;
    q.tiles_n = (q.N + tm - 1) / tm;
    const int64_t c = (int64_t)(q.K + 64) * (q.dtype == 0 ? 8 : 1);
    for (int t = 0; t < tmi * q.tiles_n; ++t) {
      work.push_back(make_int2(i, t));
      cost.push_back(c);
    }
  }
  if (order == 0) {
    std::vector<int> idx(work.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int2> w2(work.size());
    for (size_t i = 0; i < idx.size(); ++i) w2[i] = work[idx[i]];
    work.swap(w2);
  }
  p->total = (int)work.size();
  if (work.size() > p->work_cap) {
    if (p->d_work) (void)hipFree(p->d_work);
    p->d_work = nullptr;
    if (hipMalloc(&p->d_work, sizeof(int2) * std::max<size_t>(1, work.size())) !=
        hipSuccess)
      return HCBLAS_STATUS_ALLOC_FAILED;
    p->work_cap = work.size();
  }
  if (!work.empty() &&
      hipMemcpy(p->d_work, work.data(), sizeof(int2) * work.size(),
                hipMemcpyHostToDevice) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  if (hipMemcpy(p->d_probs, p->probs.data(), sizeof(ProbDev) * p->count,
                hipMemcpyHostToDevice) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  p->shape = shape;
  p->order = order;
  p->occ = occ;
  p->sched = sched;
  p->grid =
      std::max(1, std::min(p->total, h->num_cus * kOccupancies[occ]));
  free_graph(p);
  if (p->total == 0) return HCBLAS_STATUS_SUCCESS;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  hcblasStatus_t st = HCBLAS_STATUS_SUCCESS;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) !=
      hipSuccess) {
    st = HCBLAS_STATUS_EXECUTION_FAILED;
  } else {
    launch_grouped(p, cs);
    if (hipStreamEndCapture(cs, &p->graph) != hipSuccess ||
        hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) !=
            hipSuccess)
      st = HCBLAS_STATUS_EXECUTION_FAILED;
  }
  (void)hipStreamDestroy(cs);
  return st;
}

// N or T; the reference's 'n' / 't' codes are accepted as aliases.
bool valid_op(int t) {
  return canon_op(t) == HCBLAS_OP_N || canon_op(t) == HCBLAS_OP_T;
}

}  // namespace

extern "C" {

hcblasStatus_t hcblasGroupedPlanCreate(hcblasHandle_t handle,
                                       const hcblasGemmProblem *problems,
                                       int count, hcblasGroupedPlan_t *plan) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (plan == nullptr || count < 0 || (count > 0 && problems == nullptr))
    return HCBLAS_STATUS_INVALID_VALUE;
  *plan = nullptr;
  hcblasGroupedPlan *p = new (std::nothrow) hcblasGroupedPlan();
  if (p == nullptr) return HCBLAS_STATUS_ALLOC_FAILED;
  p->device = handle->device;
  p->count = count;
  p->probs.resize(count);
  bool vload = true;
  for (int i = 0; i < count; ++i) {
    const hcblasGemmProblem &u = problems[i];
    ProbDev &q = p->probs[i];
    if ((u.dtype != HCBLAS_R_32F && u.dtype != HCBLAS_R_16F) ||
        (u.order != HCBLAS_ROW_MAJOR && u.order != HCBLAS_COL_MAJOR) ||
        !valid_op(u.transA) || !valid_op(u.transB) || u.M < 0 || u.N < 0 ||
        u.K < 0) {
      delete p;
      return HCBLAS_STATUS_INVALID_VALUE;
    }
    const bool col = u.order == HCBLAS_COL_MAJOR;
    const bool an = canon_op(u.transA) == HCBLAS_OP_N;
    const bool bn = canon_op(u.transB) == HCBLAS_OP_N;
    const int64_t a_rows = col ? (an ? u.M : u.K) : (an ? u.K : u.M);
    const int64_t b_rows = col ? (bn ? u.K : u.N) : (bn ? u.N : u.K);
    const int64_t c_rows = col ? u.M : u.N;
    if (u.lda < std::max<int64_t>(1, a_rows) ||
        u.ldb < std::max<int64_t>(1, b_rows) ||
        u.ldc < std::max<int64_t>(1, c_rows) ||
        (u.M > 0 && u.N > 0 && (u.C == nullptr ||
                                (u.K > 0 && (u.A == nullptr ||
                                             u.B == nullptr))))) {
      delete p;
      return HCBLAS_STATUS_INVALID_VALUE;
    }
    q.dtype = u.dtype == HCBLAS_R_32F ? 0 : 1;
    q.alpha = u.alpha;
    q.beta = u.beta;
    q.C = u.C;
    q.ldc = u.ldc;
    if (col) {
      q.M = u.M, q.N = u.N, q.K = u.K;
      q.flags = (an ? 0 : 1) | (bn ? 2 : 0);
      q.A = u.A, q.lda = u.lda, q.B = u.B, q.ldb = u.ldb;
    } else {
      q.M = u.N, q.N = u.M, q.K = u.K;
      q.flags = (bn ? 0 : 1) | (an ? 2 : 0);
      q.A = u.B, q.lda = u.ldb, q.B = u.A, q.ldb = u.lda;
    }
    // The K == 0 / alpha == 0 forms still run the tile (an empty K loop
    // leaves acc = 0, so C = beta*C).
    if (q.alpha == 0.f) q.K = 0;
    const int vec = q.dtype == 0 ? 4 : 8;
    const size_t es = q.dtype == 0 ? 4 : 2;
    auto al = [&](const void *ptr) {
      return ptr == nullptr ||
             (reinterpret_cast<uintptr_t>(ptr) % 16 == 0);
    };
    // bit 2: 16-byte vector accesses of C in the epilogue
    if (q.ldc % vec == 0 && al(q.C)) q.flags |= 4;
    (void)es;
    if (q.lda % vec || q.ldb % vec || !al(q.A) || !al(q.B)) vload = false;
    // r-contiguous operands: whole vectors of rows (see TileLoader)
    if ((!(q.flags & 1) && q.M % vec) || (!(q.flags & 2) && q.N % vec))
      vload = false;
  }
  p->vload = vload;
  if (hipMalloc(&p->d_probs, sizeof(ProbDev) * std::max(1, count)) !=
          hipSuccess ||
      hipMalloc(&p->d_counters, 2 * sizeof(unsigned int)) != hipSuccess ||
      hipMemset(p->d_counters, 0, 2 * sizeof(unsigned int)) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    hcblasGroupedPlanDestroy(p);
    return HCBLAS_STATUS_ALLOC_FAILED;
  }
  hcblasStatus_t st = build(handle, p, 0, 0, 0, 0);
  if (st != HCBLAS_STATUS_SUCCESS) {
    hcblasGroupedPlanDestroy(p);
    return st;
  }
  *plan = p;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasGroupedGemm(hcblasHandle_t handle,
                                 hcblasGroupedPlan_t plan) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (plan == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  if (plan->total == 0) return HCBLAS_STATUS_SUCCESS;
  return hipGraphLaunch(plan->exec, handle->stream) == hipSuccess
             ? HCBLAS_STATUS_SUCCESS
             : HCBLAS_STATUS_EXECUTION_FAILED;
}

hcblasStatus_t hcblasGroupedPlanAutotune(hcblasHandle_t handle,
                                         hcblasGroupedPlan_t plan, int reps) {
  if (!handle_ok(handle)) return HCBLAS_STATUS_NOT_INITIALIZED;
  if (plan == nullptr || reps <= 0) return HCBLAS_STATUS_INVALID_VALUE;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return HCBLAS_STATUS_EXECUTION_FAILED;
  float best = 1e30f;
  int best_shape = 0, best_order = 0, best_occ = 0, best_sched = 0;
  hcblasStatus_t st = HCBLAS_STATUS_SUCCESS;
  for (int shape = 0; shape < kNumShapes && st == HCBLAS_STATUS_SUCCESS;
       ++shape) {
    for (int order = 0; order < 2 && st == HCBLAS_STATUS_SUCCESS; ++order) {
      for (int occ = 0; occ < kNumOcc && st == HCBLAS_STATUS_SUCCESS; ++occ) {
        for (int sched = 0; sched < 2; ++sched) {
          st = build(handle, plan, shape, order, occ, sched);
          if (st != HCBLAS_STATUS_SUCCESS) break;
          if (plan->total == 0) break;
          (void)hipGraphLaunch(plan->exec, handle->stream);  // warm-up
          (void)hipEventRecord(e0, handle->stream);
          for (int r = 0; r < reps; ++r)
            (void)hipGraphLaunch(plan->exec, handle->stream);
          (void)hipEventRecord(e1, handle->stream);
          if (hipEventSynchronize(e1) != hipSuccess) {
            st = HCBLAS_STATUS_EXECUTION_FAILED;
            break;
          }
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (ms < best) {
            best = ms;
            best_shape = shape;
            best_order = order;
            best_occ = occ;
            best_sched = sched;
          }
        }
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != HCBLAS_STATUS_SUCCESS) return st;
  return build(handle, plan, best_shape, best_order, best_occ, best_sched);
}

hcblasStatus_t hcblasGroupedPlanInfo(hcblasGroupedPlan_t plan, int *tile,
                                     int *order, int *tiles, int *workgroups) {
  if (plan == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  if (tile) *tile = kTileShapes[plan->shape];
  if (order) *order = plan->order;
  if (tiles) *tiles = plan->total;
  if (workgroups) *workgroups = plan->grid;
  return HCBLAS_STATUS_SUCCESS;
}

hcblasStatus_t hcblasGroupedPlanDestroy(hcblasGroupedPlan_t plan) {
  if (plan == nullptr) return HCBLAS_STATUS_INVALID_VALUE;
  (void)hipDeviceSynchronize();
  free_graph(plan);
  if (plan->d_probs) (void)hipFree(plan->d_probs);
  if (plan->d_work) (void)hipFree(plan->d_work);
  if (plan->d_counters) (void)hipFree(plan->d_counters);
  delete plan;
  return HCBLAS_STATUS_SUCCESS;
}

}  // extern "C"
