// Batched uneven evaluations in one persistent launch.
//
// A plan holds P independent problems (own sources, own targets, own outputs).  The host
// cuts every problem into work items (target tile x source chunk) of about equal cost and
// sorts them by cost, largest first.  One persistent grid (a few workgroups per CU) pulls
// items from a global counter until the list is empty, so uneven problems balance like a
// longest-processing-time schedule.  Problems whose sources were chunked write chunk
// partial sums; k_batch_reduce sums them per target in chunk order (deterministic) and
// applies G, the stores and the fused FP64 kick.  The sequence reset -> pack -> eval ->
// reduce uses no memset, so it can be captured in a hipGraph and replayed.
#pragma once
#include <type_traits>

#include "kernels.hpp"
#include "grid.hpp"

namespace tg {

struct BSrc {              // one problem's caller-owned source arrays
  const double *x, *y, *z, *m, *eps;
};

struct BProb {
  Tgt T;                   // targets (explicit)
  int64_t src_off;         // first packed source of this problem
  int64_t ns;
  int64_t chunk;           // sources per chunk
  int nchunk;              // 1: results written by the eval kernel; >1 or 0: by k_batch_reduce
  int pad;
  int64_t part_off;        // this problem's chunk partials in the scratch array (doubles)
  Out O;                   // final outputs (+ kick)
};

struct BItem {             // work item: problem, target tile (or 8^3 target block), source chunk
  int prob, tile, chunk, pad;
};

// counters = 0 (generic and block-target work lists), per-problem stats = {+inf bits, 0}
__global__ void k_batch_reset(int nprob, unsigned int* counter, unsigned long long* stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    counter[0] = 0u;
    counter[1] = 0u;
  }
  if (i < nprob) {
    stats[2 * i] = 0x7ff0000000000000ULL;
    stats[2 * i + 1] = 0ULL;
  }
}

__device__ __forceinline__ int find_segment(const int64_t* __restrict__ prefix, int n, int64_t i) {
  int lo = 0, hi = n;   // prefix[lo] <= i < prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Packs every problem's sources into one {x,y,z,m} + eps^2 array and folds each problem's
// smallest eps^2 into stats[2p].  A workgroup inside one problem reduces first and issues
// one atomic; a workgroup straddling problems falls back to one atomic per lane.
__global__ void k_batch_pack(int nprob, const BSrc* __restrict__ srcs,
                             const int64_t* __restrict__ prefix, double4* __restrict__ s4,
                             double* __restrict__ se2, unsigned long long* stats) {
  __shared__ unsigned long long smin[256];
  __shared__ int sp[2];
  const int64_t total = prefix[nprob];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int p = -1;
  unsigned long long mine = 0x7ff0000000000000ULL;
  if (i < total) {
    p = find_segment(prefix, nprob, i);
    const int64_t j = i - prefix[p];
    const BSrc S = srcs[p];
    s4[i] = make_double4(S.x[j], S.y[j], S.z[j], S.m[j]);
    const double e = S.eps ? S.eps[j] : 0.0;
    se2[i] = e * e;
    mine = (unsigned long long)__double_as_longlong(e * e);
  }
  if (threadIdx.x == 0) {
    sp[0] = p;
    const int64_t last = min(total, (int64_t)(blockIdx.x + 1) * blockDim.x) - 1;
    sp[1] = last >= 0 ? find_segment(prefix, nprob, last) : -1;
  }
  smin[threadIdx.x] = mine;
  __syncthreads();
  if (sp[0] == sp[1]) {
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0 && sp[0] >= 0) atomicMin(&stats[2 * sp[0]], smin[0]);
  } else if (p >= 0) {
    atomicMin(&stats[2 * p], mine);
  }
}

template <int K, int NT, bool MIXED>
__global__ __launch_bounds__(NT) void k_batch_eval(const BProb* __restrict__ probs,
                                                   const BItem* __restrict__ items, int n_items,
                                                   unsigned int* counter,
                                                   const double4* __restrict__ s4,
                                                   const double* __restrict__ se2,
                                                   const unsigned long long* __restrict__ stats,
                                                   double* __restrict__ part) {
  typedef typename std::conditional<MIXED, LdsF32<NT>, LdsF64<NT>>::type Lds;
  __shared__ Lds L;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();                    // all lanes hold `it` before lane 0 overwrites it
    if (it >= n_items) break;           // every lane of the workgroup leaves together
    const BItem w = items[it];
    const BProb& P = probs[w.prob];
    const int64_t sb = P.src_off + (int64_t)w.chunk * P.chunk;
    const int64_t se = min(sb + P.chunk, P.src_off + P.ns);
    const double e2s_min = __longlong_as_double((long long)stats[2 * w.prob]);
    Out o = P.O;
    if (P.nchunk != 1) {
      o.part = part + P.part_off + (int64_t)w.chunk * 4 * P.T.n;
      o.pstride = P.T.n;
    } else {
      o.part = nullptr;
    }
    const int64_t tile0 = (int64_t)w.tile * (K * NT);
    if constexpr (MIXED)
      process_mixed<K, NT>(L, P.T, tile0, s4, se2, sb, se, e2s_min, o);
    else
      process_fp64<K, NT>(L, P.T, tile0, s4, se2, sb, se, e2s_min, o);
  }
}

// Problems whose targets are 8^3 leaf blocks, or explicit targets with their column structure
// (found at plan creation): the same persistent scheme with 64-lane workgroups and one target
// block (group of 512 targets) per work item, evaluated by the column code of k_eval_grid.
template <bool MIXED>
__global__ __launch_bounds__(64) void k_batch_eval_grid(const BProb* __restrict__ probs,
                                                        const BItem* __restrict__ items,
                                                        int n_items, unsigned int* counter,
                                                        const double4* __restrict__ s4,
                                                        const double* __restrict__ se2,
                                                        const unsigned long long* __restrict__ stats,
                                                        double* __restrict__ part) {
  typedef typename std::conditional<MIXED, LdsGridF<64>, LdsGrid<64>>::type Lds;
  __shared__ Lds L;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int it = s_item;
    __syncthreads();                    // all lanes hold `it` before lane 0 overwrites it
    if (it >= n_items) break;           // every lane of the workgroup leaves together
    const BItem w = items[it];
    const BProb& P = probs[w.prob];
    const int64_t sb = P.src_off + (int64_t)w.chunk * P.chunk;
    const int64_t se = min(sb + P.chunk, P.src_off + P.ns);
    const double e2s_min = __longlong_as_double((long long)stats[2 * w.prob]);
    Out o = P.O;
    if (P.nchunk != 1) {
      o.part = part + P.part_off + (int64_t)w.chunk * 4 * P.T.n;
      o.pstride = P.T.n;
    } else {
      o.part = nullptr;
    }
    if (P.T.blocks) grid_block<MIXED>(L, P.T, w.tile, s4, se2, sb, se, e2s_min, o);
    else cols_group<MIXED>(L, P.T, w.tile, s4, se2, sb, se, e2s_min, o);   // explicit columns
  }
}

// Chunk reduction of the problems listed in `ids` (nchunk != 1), one lane per target.
__global__ void k_batch_reduce(int nred, const int* __restrict__ ids,
                               const int64_t* __restrict__ prefix, const BProb* __restrict__ probs,
                               const double* __restrict__ part) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= prefix[nred]) return;
  const int q = find_segment(prefix, nred, i);
  const BProb& P = probs[ids[q]];
  const int64_t t = i - prefix[q];
  const int64_t n = P.T.n;
  const double* base = part + P.part_off;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int c = 0; c < P.nchunk; ++c) {
    const double* pc = base + (int64_t)c * 4 * n;
    s0 += pc[t];
    s1 += pc[n + t];
    s2 += pc[2 * n + t];
    s3 += pc[3 * n + t];
  }
  Out o = P.O;
  o.part = nullptr;
  store_result(o, t, s0, s1, s2, s3);
}

}  // namespace tg
