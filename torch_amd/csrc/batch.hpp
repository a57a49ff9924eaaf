// Batched uneven evaluations in one persistent launch.
//
// A plan holds P independent problems (own sources, own targets, own outputs).  The host
// cuts every problem into work items (target tile x source chunk) of about equal cost and
// sorts them by cost, largest first.  One persistent grid (a few workgroups per CU) pulls
// items from a global counter until the list is empty, so uneven problems balance like a
// longest-processing-time schedule.  Problems whose sources were chunked write chunk
// partial sums; k_batch_reduce sums them per target in chunk order (deterministic) and
// applies G, the stores and the fused FP64 kick.  The sequence reset -> pack -> column
// check -> evals -> reduce uses no memset, so it can be captured in a hipGraph and replayed.
//
// Three work lists, one persistent kernel each: generic (explicit targets, explicit sources),
// columns (targets that are 8^3 leaf blocks, or explicit targets in their C order) and block
// sources (sources that are 8^3 leaf blocks: the column walk of k_eval_gridsrc).  Explicit
// targets whose count is a multiple of 512 may or may not be in the blocks' C order; the plan
// tests that on the device at every run (k_batch_check_cols) and holds items of both forms for
// them, with the same source chunks (so the partial sums line up): the items of the form the
// test did not pick return at once.
#pragma once
#include <type_traits>

#include "kernels.hpp"
#include "grid.hpp"

namespace tg {

struct BSrc {              // one problem's caller-owned source arrays
  const double *x, *y, *z, *m, *eps;
  const double *bx, *by, *bz, *bd;   // bd != null: 8^3 source blocks (corners, cell widths; eps and
  int64_t blk_off;                   //   m per block / per cell), the plan's blocks blk_off ...
};

// How a problem's work items are evaluated.
enum { FORM_GENERIC = 0, FORM_COLS = 1, FORM_CAND = 2, FORM_GRIDSRC = 3 };

struct BProb {
  Tgt T;                   // targets (explicit or 8^3 blocks)
  int64_t src_off;         // first packed source of this problem
  int64_t ns;
  int64_t chunk;           // sources per chunk (FORM_GRIDSRC: source blocks per chunk)
  int nchunk;              // 1: results written by the eval kernel; >1 or 0: by k_batch_reduce
  int form;                // FORM_*; FORM_CAND: columns if colflag[p] == 1 at this run, else generic
  int64_t part_off;        // this problem's chunk partials in the scratch array (doubles)
  int64_t sblk_off, nsblk; // FORM_GRIDSRC: the plan's source blocks [sblk_off, sblk_off + nsblk)
  Out O;                   // final outputs (+ kick)
};

struct BItem {             // work item: problem, target tile (or 8^3 target block), source chunk
  int prob, tile, chunk, pad;
};

// counters = 0 (counter[0] the generic and counter[1] the column work list, counter[4 ... 11] the
// block-source groups; counter[2] is unused), counter[3] = skip_generic (1 when
// every generic item belongs to a column candidate: then the generic list has nothing to do unless
// a candidate fails its test, which clears the word), per-problem source statistics (ST_WORDS
// each, see kernels.hpp) reset, column flags = 1 (k_batch_check_cols clears them)
__global__ void k_batch_reset(int nprob, unsigned int* counter, unsigned long long* stats,
                              int* colflag, unsigned int skip_generic) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    counter[0] = 0u;
    counter[1] = 0u;
    counter[2] = 0u;
    counter[3] = skip_generic;
#pragma unroll
    for (int q = 4; q < 12; ++q) counter[q] = 0u;
  }
  if (i < nprob) {
    colflag[i] = 1;
    unsigned long long* st = stats + (int64_t)ST_WORDS * i;
    st[ST_E2MIN] = 0x7ff0000000000000ULL;
    st[ST_BAD] = 0ULL;
    st[ST_RMAX] = 0ULL;
    st[ST_E2MAX] = 0ULL;
    st[ST_MMAX] = 0ULL;
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
// source statistics (ST_*) into stats[ST_WORDS p ...].  Source blocks are expanded to their cell
// centres (rounded as k_pack_source_blocks rounds them) and also give their cell-centre axes
// (sax, 24 per block), eps^2 (be2) and contiguous cell masses (bms, 512 per block) for the
// block-source kernel.  Each workgroup reduces the statistics of every problem it touches and issues
// at most one atomic per word and problem.  Workgroups of 256 lanes.
__global__ __launch_bounds__(256) void k_batch_pack(int nprob, const BSrc* __restrict__ srcs,
                                                    const int64_t* __restrict__ prefix,
                                                    double4* __restrict__ s4, double* __restrict__ se2,
                                                    double* __restrict__ sax, double* __restrict__ be2,
                                                    double* __restrict__ bms, unsigned long long* stats) {
  __shared__ int sp[2];
  const int64_t total = prefix[nprob];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int p = -1;
  unsigned long long e2min = 0x7ff0000000000000ULL, rmax = 0ULL, e2max = 0ULL, mmax = 0ULL;
  if (i < total) {
    p = find_segment(prefix, nprob, i);
    const int64_t j = i - prefix[p];
    const BSrc S = srcs[p];
    double4 v;
    double e;
    if (S.bd) {
      const int64_t b = j >> 9;
      const int c = (int)(j & 511), ci = c >> 6, cj = (c >> 3) & 7, ck = c & 7;
      const double d = S.bd[b];
      {
#pragma clang fp contract(off)
        v.x = S.bx[b] + ((double)ci + 0.5) * d;
        v.y = S.by[b] + ((double)cj + 0.5) * d;
        v.z = S.bz[b] + ((double)ck + 0.5) * d;
      }
      v.w = S.m[j];
      e = S.eps ? S.eps[b] : 0.0;
      const int64_t gb = S.blk_off + b;
      if (cj == 0 && ck == 0) sax[gb * 24 + ci] = v.x;
      if (ci == 0 && ck == 0) sax[gb * 24 + 8 + cj] = v.y;
      if (ci == 0 && cj == 0) sax[gb * 24 + 16 + ck] = v.z;
      if (c == 0) be2[gb] = e * e;
      bms[gb * 512 + c] = v.w;
    } else {
      v = make_double4(S.x[j], S.y[j], S.z[j], S.m[j]);
      e = S.eps ? S.eps[j] : 0.0;
    }
    s4[i] = v;
    se2[i] = e * e;
    e2min = e2max = dbits(e * e);
    rmax = dbits(max_abs3(v.x, v.y, v.z));
    mmax = dbits(fabs(v.w));
  }
  if (threadIdx.x == 0) {
    sp[0] = p;
    const int64_t last = min(total, (int64_t)(blockIdx.x + 1) * blockDim.x) - 1;
    sp[1] = last >= 0 ? find_segment(prefix, nprob, last) : -1;
  }
  __syncthreads();
  // The workgroup's sources belong to problems sp[0] .. sp[1] (contiguous in the packed order): one
  // workgroup reduction per problem, each lane contributing to its own problem only (per-lane
  // atomics on the few words of a problem would queue up behind each other).
  const int p0 = sp[0], p1 = sp[1];   // uniform over the workgroup
  if (p0 >= 0) {
    for (int q = p0; q <= p1; ++q) {
      const bool mine = p == q;
      fold_source_stats(mine ? e2min : 0x7ff0000000000000ULL, 0, mine ? rmax : 0ULL, mine ? e2max : 0ULL,
                        mine ? mmax : 0ULL, stats + (int64_t)ST_WORDS * q);
    }
  }
}

// The column-structure test of k_check_cols over the plan's candidate problems (ids[q], explicit
// targets, a multiple of 512): lane i checks target i - prefix[q] of candidate q = the segment of
// i in prefix (targets per candidate) and clears colflag[ids[q]] and the skip word counter[3] on a
// violation.  Runs after k_batch_reset set the flags.
__global__ void k_batch_check_cols(int ncand, const int* __restrict__ ids,
                                   const int64_t* __restrict__ prefix, const BProb* __restrict__ probs,
                                   int* colflag, unsigned int* counter) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= prefix[ncand]) return;
  const int q = find_segment(prefix, ncand, i);
  const int p = ids[q];
  if (!column_ok(probs[p].T, i - prefix[q])) {
    colflag[p] = 0;
    counter[3] = 0u;
  }
}

template <int K, int NT, bool MIXED>
__global__ __launch_bounds__(NT) void k_batch_eval(const BProb* __restrict__ probs,
                                                   const BItem* __restrict__ items, int n_items,
                                                   unsigned int* counter,
                                                   const double4* __restrict__ s4,
                                                   const double* __restrict__ se2,
                                                   const unsigned long long* __restrict__ stats,
                                                   const int* __restrict__ colflag,
                                                   double* __restrict__ part) {
  typedef typename std::conditional<MIXED, LdsF32<NT>, LdsF64<NT>>::type Lds;
  __shared__ Lds L;
  __shared__ int s_item;
  if (counter[3] != 0u) return;   // every generic item belongs to a candidate that took the columns
  for (;;) {
    if (threadIdx.x == 0) {   // items of problems whose targets took the column kernel are skipped here
      int i;
      do {
        i = (int)atomicAdd(counter, 1u);
      } while (i < n_items && probs[items[i].prob].form == FORM_CAND && colflag[items[i].prob] != 0);
      s_item = i;
    }
    __syncthreads();
    const int it = s_item;
    __syncthreads();                    // all lanes hold `it` before lane 0 overwrites it
    if (it >= n_items) break;           // every lane of the workgroup leaves together
    const BItem w = items[it];
    const BProb& P = probs[w.prob];
    const int64_t sb = P.src_off + (int64_t)w.chunk * P.chunk;
    const int64_t se = min(sb + P.chunk, P.src_off + P.ns);
    const double* sst = (const double*)(stats + (int64_t)ST_WORDS * w.prob);
    Out o = P.O;
    if (P.nchunk != 1) {
      o.part = part + P.part_off + (int64_t)w.chunk * 4 * P.T.n;
      o.pstride = P.T.n;
    } else {
      o.part = nullptr;
    }
    const int64_t tile0 = (int64_t)w.tile * (K * NT);
    if constexpr (MIXED)
      process_mixed<K, NT>(L, P.T, tile0, s4, se2, sb, se, sst, o);
    else
      process_fp64<K, NT>(L, P.T, tile0, s4, se2, sb, se, sst, o);
  }
}

// Problems whose targets are 8^3 leaf blocks, or explicit targets with their column structure
// (found at plan creation): the same persistent scheme with workgroups of W waves and one
// target block (group of 512 targets) per work item, evaluated by the column code of k_eval_grid
// (the W waves split the item's sources and sum their partial results in LDS).
template <bool MIXED, int W>
__global__ __launch_bounds__(64 * W, 3) void k_batch_eval_grid(const BProb* __restrict__ probs,
                                                        const BItem* __restrict__ items,
                                                        int n_items, unsigned int* counter,
                                                        const double4* __restrict__ s4,
                                                        const double* __restrict__ se2,
                                                        const unsigned long long* __restrict__ stats,
                                                        const int* __restrict__ colflag,
                                                        double* __restrict__ part) {
  __shared__ LdsCol<MIXED, W> L;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) {   // items of problems whose targets are not in C order at this run are skipped
      int i;
      do {
        i = (int)atomicAdd(counter, 1u);
      } while (i < n_items && probs[items[i].prob].form == FORM_CAND && colflag[items[i].prob] != 1);
      s_item = i;
    }
    __syncthreads();
    const int it = s_item;
    __syncthreads();                    // all lanes hold `it` before lane 0 overwrites it
    if (it >= n_items) break;           // every lane of the workgroup leaves together
    const BItem w = items[it];
    const BProb& P = probs[w.prob];
    const int64_t sb = P.src_off + (int64_t)w.chunk * P.chunk;
    const int64_t se = min(sb + P.chunk, P.src_off + P.ns);
    const double* sst = (const double*)(stats + (int64_t)ST_WORDS * w.prob);
    Out o = P.O;
    if (P.nchunk != 1) {
      o.part = part + P.part_off + (int64_t)w.chunk * 4 * P.T.n;
      o.pstride = P.T.n;
    } else {
      o.part = nullptr;
    }
    if (P.T.blocks) grid_block<MIXED, W>(L, P.T, w.tile, s4, se2, sb, se, sst, o);
    else cols_group<MIXED, W>(L, P.T, w.tile, s4, se2, sb, se, sst, o);   // explicit columns
  }
}

// Problems whose sources are 8^3 leaf blocks: the same persistent scheme over (target tile of
// K x 64, chunk of source blocks) items, evaluated by the column walk of k_eval_gridsrc (the W
// waves split the chunk's blocks; the per-block source loop of Gravity_getAccelAtPoint.F90:168-204).
// The items come in 8 groups, group q = the items whose chunk is q mod 8 (items goff[q] ...
// goff[q + 1] - 1, each group cost-sorted and chunk-major), pulled through counters gcount[q].
// Workgroup b starts on group b mod 8 -- the XCD the hardware dispatches it to -- so each chunk's
// masses are read into one XCD's L2, and moves on to the next group when its own is empty, so the
// load still balances.  Which workgroup runs an item changes no result.
template <int K, int W, bool MIXED>
__global__ __launch_bounds__(64 * W) void k_batch_eval_gridsrc(const BProb* __restrict__ probs,
                                                           const BItem* __restrict__ items,
                                                           const int* __restrict__ goff,
                                                           unsigned int* gcount,
                                                           const double* __restrict__ bms,
                                                           const double* __restrict__ sax,
                                                           const double* __restrict__ be2,
                                                           const unsigned long long* __restrict__ stats,
                                                           double* __restrict__ part) {
  __shared__ LdsGSW<MIXED, W> L;
  __shared__ int s_item;
  int cur = blockIdx.x & 7;             // lane 0's group
  for (;;) {
    if (threadIdx.x == 0) {
      int it = -1;
      for (int k = 0; k < 8 && it < 0; ++k) {   // its group, then the others in turn
        const int q = (cur + k) & 7;
        if (goff[q] == goff[q + 1]) continue;
        const int i = goff[q] + (int)atomicAdd(&gcount[q], 1u);
        if (i < goff[q + 1]) {
          it = i;
          cur = q;
        }
      }
      s_item = it;                      // -1: every group is empty
    }
    __syncthreads();
    const int it = s_item;
    __syncthreads();                    // all lanes hold `it` before lane 0 overwrites it
    if (it < 0) break;                  // every lane of the workgroup leaves together
    const BItem w = items[it];
    const BProb& P = probs[w.prob];
    const int64_t b0 = (int64_t)w.chunk * P.chunk;
    const int64_t b1 = min(b0 + P.chunk, P.nsblk);
    const double* sst = (const double*)(stats + (int64_t)ST_WORDS * w.prob);
    Out o = P.O;
    if (P.nchunk != 1) {
      o.part = part + P.part_off + (int64_t)w.chunk * 4 * P.T.n;
      o.pstride = P.T.n;
    } else {
      o.part = nullptr;
    }
    gridsrc_tile<K, W, MIXED>(L, P.T, (int64_t)w.tile * (K * 64), sax + P.sblk_off * 24,
                               be2 + P.sblk_off, bms + P.sblk_off * 512, b0, b1, sst, o);
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
