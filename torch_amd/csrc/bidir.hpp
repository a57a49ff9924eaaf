// Fused bidirectional gas <-> particle direct sum, the GPU form of
// Particles_sinkAccelGasOnSinksAndSinksOnGas.F90:204-316: every (cell, particle) pair is
// evaluated once and feeds both the cell's field (from the particles) and the particle's
// field (from the cells), with the symmetric softening eps2 = eps_c^2 + eps_p^2.
//
// Mapping: a workgroup holds K*NT cells (K per lane, like the directed kernels) and streams
// the particles through LDS in tiles of NT.  The particle-side sums cannot stay in
// registers (each lane sees every particle), so each wave owns a private FP64 accumulator
// copy of the tile in LDS and walks the tile in a rotated order: at step s of a 64-particle
// block, lane l takes particle (l + s) mod 64.  In every step the 64 lanes touch 64
// different accumulators, so the read-add-write needs no atomics and the order in which a
// particle's contributions are added is fixed.  After the tile the wave copies are summed in
// wave order and written as this cell tile's partial sums; a second kernel adds the cell
// tiles' partials per particle in tile order.  Everything is deterministic.
#pragma once
#include "kernels.hpp"

namespace tg {

struct Cells {
  const double *x, *y, *z, *m, *eps;
  int64_t n;
};

template <int NT, bool MIXED>
struct LdsBi {
  typename std::conditional<MIXED, float4, double4>::type p4[NT];   // particle x, y, z, m
  typename std::conditional<MIXED, float, double>::type pe2[NT];    // particle eps^2 (+ uniform cell eps^2)
  double acc[NT / 64][NT][4];                                        // per-wave particle sums
  double red[2 * 4 * (NT / 64)];
};

// One pair, both directions.  d = x_p - x_c.  Cell side: (cx, cy, cz, cp) += m_p y^3 d, m_p y;
// particle side: (ux, uy, uz, up) += m_c y^3 d, m_c y (the particle's acceleration is -G u).
template <class T, bool GUARD>
__device__ __forceinline__ void bi_pair(T px, T py, T pz, T pm, T r2e, T qx, T qy, T qz, T qm,
                                        T& cx, T& cy, T& cz, T& cp, T& ux, T& uy, T& uz, T& up) {
  const T dx = px - qx, dy = py - qy, dz = pz - qz;
  T r2;
  T y;
  if constexpr (sizeof(T) == 4) {
    r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, r2e)));
    y = __builtin_amdgcn_rsqf(r2);
    if (GUARD) y = r2 > 0.0f ? y : 0.0f;
  } else {
    r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, r2e)));
    y = rsqrt_f64<GUARD ? F_GUARD : 0>(r2);
  }
  const T y3 = (y * y) * y;
  const T wp = pm * y3;
  cp = fma(pm, y, cp);
  cx = fma(wp, dx, cx);
  cy = fma(wp, dy, cy);
  cz = fma(wp, dz, cz);
  const T wc = qm * y3;
  up = fma(qm, y, up);
  ux = fma(wc, dx, ux);
  uy = fma(wc, dy, uy);
  uz = fma(wc, dz, uz);
}

template <int K, int NT, bool MIXED, bool GUARD, bool UNIF>
__device__ __forceinline__ void bi_particles(
    LdsBi<NT, MIXED>& L, const double4* __restrict__ p4, const double* __restrict__ pe2,
    int64_t p_begin, int64_t p_end, double e2u, double ax_, double ay_, double az_,
    const typename std::conditional<MIXED, float, double>::type (&qx)[K],
    const typename std::conditional<MIXED, float, double>::type (&qy)[K],
    const typename std::conditional<MIXED, float, double>::type (&qz)[K],
    const typename std::conditional<MIXED, float, double>::type (&qm)[K],
    const typename std::conditional<MIXED, float, double>::type (&qe2)[K],
    double (&hx)[K], double (&hy)[K], double (&hz)[K], double (&hp)[K],
    double* __restrict__ ppart, int64_t pstride) {
  typedef typename std::conditional<MIXED, float, double>::type T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int64_t base = p_begin; base < p_end; base += NT) {
    __syncthreads();
    const int64_t j = base + tid;
    if (j < p_end) {
      const double4 s = p4[j];
      if constexpr (MIXED)
        L.p4[tid] = make_float4((float)(s.x - ax_), (float)(s.y - ay_), (float)(s.z - az_), (float)s.w);
      else
        L.p4[tid] = s;
      L.pe2[tid] = (T)(pe2[j] + e2u);
    } else {                      // padding particle: zero mass, eps^2 = 1
      if constexpr (MIXED) L.p4[tid] = make_float4(0.f, 0.f, 0.f, 0.f);
      else L.p4[tid] = make_double4(0.0, 0.0, 0.0, 0.0);
      L.pe2[tid] = (T)1.0;
    }
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
      L.acc[w][tid][0] = 0.0;
      L.acc[w][tid][1] = 0.0;
      L.acc[w][tid][2] = 0.0;
      L.acc[w][tid][3] = 0.0;
    }
    __syncthreads();
    for (int blk = 0; blk < NT; blk += 64) {
      T cx[K], cy[K], cz[K], cp[K];
#pragma unroll
      for (int k = 0; k < K; ++k) cx[k] = cy[k] = cz[k] = cp[k] = (T)0;
      for (int s = 0; s < 64; ++s) {
        const int jj = blk + ((lane + s) & 63);
        const auto P = L.p4[jj];
        const T e2 = L.pe2[jj];
        T ux = 0, uy = 0, uz = 0, up = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
          bi_pair<T, GUARD>(P.x, P.y, P.z, P.w, UNIF ? e2 : e2 + qe2[k], qx[k], qy[k], qz[k], qm[k],
                            cx[k], cy[k], cz[k], cp[k], ux, uy, uz, up);
        double* a = L.acc[wave][jj];
        a[0] += (double)ux;
        a[1] += (double)uy;
        a[2] += (double)uz;
        a[3] += (double)up;
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        hx[k] += (double)cx[k];
        hy[k] += (double)cy[k];
        hz[k] += (double)cz[k];
        hp[k] += (double)cp[k];
      }
    }
    __syncthreads();
    if (j < p_end) {              // this cell tile's partial sums for particle j, wave order
      double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) {
        s0 += L.acc[w][tid][0];
        s1 += L.acc[w][tid][1];
        s2 += L.acc[w][tid][2];
        s3 += L.acc[w][tid][3];
      }
      ppart[j] = s0;
      ppart[pstride + j] = s1;
      ppart[2 * pstride + j] = s2;
      ppart[3 * pstride + j] = s3;
    }
  }
}

// blockIdx.x = cell tile, blockIdx.y = particle chunk.  Cell results go to CO (final, or
// chunk partials when CO.part is set); particle partials go to
// ppart[(tile * 4 + q) * np + j].
template <int K, int NT, bool MIXED>
__global__ __launch_bounds__(NT) void k_bidir(Cells C, const double4* __restrict__ p4,
                                              const double* __restrict__ pe2, int64_t np,
                                              int64_t chunk, const double* e2p_min_p, Out CO,
                                              double* __restrict__ ppart) {
  typedef typename std::conditional<MIXED, float, double>::type T;
  __shared__ LdsBi<NT, MIXED> L;
  if (gated_off(CO)) return;    // the cells turned out to be 8^3 block columns
  const int tid = threadIdx.x;
  const int64_t tile0 = (int64_t)blockIdx.x * (K * NT);
  const int64_t pb = (int64_t)blockIdx.y * chunk;
  const int64_t pe = pb + chunk < np ? pb + chunk : np;
  const double e2p_min = *e2p_min_p;
  double px[K], py[K], pz[K], pm[K], pq[K];
  double mn[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  double mx[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    const bool ok = t < C.n;
    const int64_t tc = ok ? t : C.n - 1;
    px[k] = C.x[tc];
    py[k] = C.y[tc];
    pz[k] = C.z[tc];
    pm[k] = ok ? C.m[tc] : 0.0;          // a padding cell pulls on nothing
    const double e = C.eps ? C.eps[tc] : 0.0;
    pq[k] = e * e;
    if (ok) {
      mn[0] = fmin(mn[0], px[k]); mx[0] = fmax(mx[0], px[k]);
      mn[1] = fmin(mn[1], py[k]); mx[1] = fmax(mx[1], py[k]);
      mn[2] = fmin(mn[2], pz[k]); mx[2] = fmax(mx[2], pz[k]);
      mn[3] = fmin(mn[3], pq[k]); mx[3] = fmax(mx[3], pq[k]);
    }
  }
  block_minmax<NT, 4>(mn, mx, L.red);
  double ax_ = 0.0, ay_ = 0.0, az_ = 0.0;
  if (MIXED) {
    ax_ = 0.5 * (mn[0] + mx[0]);
    ay_ = 0.5 * (mn[1] + mx[1]);
    az_ = 0.5 * (mn[2] + mx[2]);
  }
  const bool unif = mn[3] == mx[3];
  const bool guard = MIXED ? !(e2p_min + mn[3] > TG_F32_TINY_E2) : !(e2p_min + mn[3] > 0.0);
  const double e2u = unif ? mn[3] : 0.0;
  T qx[K], qy[K], qz[K], qm[K], qe2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    qx[k] = (T)(px[k] - ax_);
    qy[k] = (T)(py[k] - ay_);
    qz[k] = (T)(pz[k] - az_);
    qm[k] = (T)pm[k];
    qe2[k] = (T)pq[k];
  }
  double hx[K], hy[K], hz[K], hp[K];
#pragma unroll
  for (int k = 0; k < K; ++k) hx[k] = hy[k] = hz[k] = hp[k] = 0.0;
  double* pp = ppart + (int64_t)blockIdx.x * 4 * np;
  if (unif) {
    if (guard) bi_particles<K, NT, MIXED, true, true>(L, p4, pe2, pb, pe, e2u, ax_, ay_, az_, qx, qy, qz, qm, qe2, hx, hy, hz, hp, pp, np);
    else bi_particles<K, NT, MIXED, false, true>(L, p4, pe2, pb, pe, e2u, ax_, ay_, az_, qx, qy, qz, qm, qe2, hx, hy, hz, hp, pp, np);
  } else {
    if (guard) bi_particles<K, NT, MIXED, true, false>(L, p4, pe2, pb, pe, 0.0, ax_, ay_, az_, qx, qy, qz, qm, qe2, hx, hy, hz, hp, pp, np);
    else bi_particles<K, NT, MIXED, false, false>(L, p4, pe2, pb, pe, 0.0, ax_, ay_, az_, qx, qy, qz, qm, qe2, hx, hy, hz, hp, pp, np);
  }
  Out o = CO;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    if (t < C.n) store_result(o, t, hx[k], hy[k], hz[k], hp[k]);
  }
}

// Particle side, first level (many cell tiles): blockIdx.y = group g of tiles
// [g*per, min((g+1)*per, ntiles)); writes the group's sums (tile order) to out[(g*4 + q)*np + j].
__global__ void k_bidir_group(const double* __restrict__ ppart, int ntiles, int per, int64_t np,
                              double* __restrict__ out, Out O) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np || gated_off(O)) return;
  const int t0 = blockIdx.y * per;
  const int t1 = t0 + per < ntiles ? t0 + per : ntiles;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int t = t0; t < t1; ++t) {
    const double* p = ppart + (int64_t)t * 4 * np;
    s0 += p[j];
    s1 += p[np + j];
    s2 += p[2 * np + j];
    s3 += p[3 * np + j];
  }
  double* o = out + (int64_t)blockIdx.y * 4 * np;
  o[j] = s0;
  o[np + j] = s1;
  o[2 * np + j] = s2;
  o[3 * np + j] = s3;
}

// Particle side: sum the ntiles cell-tile partials of each particle (tile order), then the
// particle's field is (-G u, -G up).
__global__ void k_bidir_reduce(const double* __restrict__ ppart, int ntiles, int64_t np, Out O) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np || gated_off(O)) return;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int t = 0; t < ntiles; ++t) {
    const double* p = ppart + (int64_t)t * 4 * np;
    s0 += p[j];
    s1 += p[np + j];
    s2 += p[2 * np + j];
    s3 += p[3 * np + j];
  }
  Out o = O;
  o.part = nullptr;
  store_result(o, j, -s0, -s1, -s2, s3);
}

// ------------------------------------------------------------- 8^3 cell blocks ----
// The fused pass with the cells handed over as 8^3 leaf blocks: one 64-lane workgroup per
// block, lane l owns the cell column (i, j) = (l / 8, l % 8) and its 8 cells along z (as
// k_eval_grid).  Particles stream through LDS 64 at a time and are walked in the rotated
// order of k_bidir (lane l takes particle (l + s) mod 64 at step s), so each lane meets one
// particle per step with its whole column: dx, dy and dx^2 + dy^2 + eps^2 are formed once
// per (column, particle), the cell side accumulates in registers, and the particle side sums
// the column's 8 cells in registers before one conflict-free read-add-write of the
// particle's LDS accumulator (FP32 in mixed mode: one tile's 512 cells, folded to FP64 in the
// partial sums, like the cell side's per-tile fold).  Cell eps^2 is uniform per block and
// folded into the staged particle eps^2.
struct CellBlocks {
  const double *ox, *oy, *oz, *dx, *m, *eps;   // per block; m per cell (C order)
  int64_t nblk;
  // Explicit cells in the blocks' C order instead (ox .. dx unused): centres and softening per
  // cell; each aligned 8 share x and y, each aligned 512 share eps (tested by k_check_cols).
  const double *cx, *cy, *cz, *ceps;
};

#ifndef TG_BIGRID_UNROLL
#define TG_BIGRID_UNROLL 1
#endif
// Particle side of the fused block pass: the column's 8 cells share dx and dy, so the x, y
// sums can take dx, dy times the column's sum of weights (TG_BIGRID_SUMW: 0 never, 1 FP64
// only, 2 both modes, 3 mixed only).  Measured on c2: mixed 4.55 -> 4.35 ms, FP64 7.93 ->
// 8.05 ms, hence mixed only.
#ifndef TG_BIGRID_SUMW
#define TG_BIGRID_SUMW 3
#endif

template <bool MIXED>
struct LdsBiGrid {
  typedef typename std::conditional<MIXED, float, double>::type T;
  typename std::conditional<MIXED, float4, double4>::type p4[64];
  T pe2[64];
  typename std::conditional<MIXED, float4, double4>::type acc[64];   // particle sums of the tile
  // ZROW (every lane holds the same 8 cell z): z_p - z_k of each staged particle, cells 0-3 in
  // dz[0], 4-7 in dz[1], formed once by the lane that stages the particle
  typename std::conditional<MIXED, float4, double4>::type dz[2][64];
};

template <bool MIXED, bool GUARD, bool ZROW>
__device__ __forceinline__ void bi_grid_tiles(
    LdsBiGrid<MIXED>& L, const double4* __restrict__ p4, const double* __restrict__ pe2,
    int64_t pb, int64_t pe, double e2c, double ax_, double ay_, double az_,
    typename std::conditional<MIXED, float, double>::type qx,
    typename std::conditional<MIXED, float, double>::type qy,
    const typename std::conditional<MIXED, float, double>::type (&qz)[8],
    const typename std::conditional<MIXED, float, double>::type (&qm)[8],
    double (&hx)[8], double (&hy)[8], double (&hz)[8], double (&hp)[8],
    double* __restrict__ ppart, int64_t np) {
  typedef typename std::conditional<MIXED, float, double>::type T;
  const int lane = threadIdx.x;
  for (int64_t base = pb; base < pe; base += 64) {
    __syncthreads();
    const int64_t j = base + lane;
    T pz = (T)0;
    if (j < pe) {
      const double4 s = p4[j];
      if constexpr (MIXED)
        L.p4[lane] = make_float4((float)(s.x - ax_), (float)(s.y - ay_), (float)(s.z - az_), (float)s.w);
      else
        L.p4[lane] = s;
      pz = MIXED ? (T)(s.z - az_) : (T)s.z;
      L.pe2[lane] = (T)(pe2[j] + e2c);
    } else {                      // padding particle: zero mass, eps^2 = 1
      if constexpr (MIXED) L.p4[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
      else L.p4[lane] = make_double4(0.0, 0.0, 0.0, 0.0);
      L.pe2[lane] = (T)1.0;
    }
    if constexpr (ZROW) {         // this lane's qz are every lane's: the dz the loop would form
      L.dz[0][lane] = {pz - qz[0], pz - qz[1], pz - qz[2], pz - qz[3]};
      L.dz[1][lane] = {pz - qz[4], pz - qz[5], pz - qz[6], pz - qz[7]};
    }
    if constexpr (MIXED) L.acc[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    else L.acc[lane] = make_double4(0.0, 0.0, 0.0, 0.0);
    __syncthreads();
    T cx[8], cy[8], cz[8], cp[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) cx[k] = cy[k] = cz[k] = cp[k] = (T)0;
#pragma unroll TG_BIGRID_UNROLL
    for (int s = 0; s < 64; ++s) {
      const int jj = (lane + s) & 63;
      const auto P = L.p4[jj];
      const T dx = P.x - qx, dy = P.y - qy;
      const T base2 = fma(dx, dx, fma(dy, dy, L.pe2[jj]));
      constexpr bool sumw = TG_BIGRID_SUMW == 2 || (TG_BIGRID_SUMW == 1 && !MIXED) ||
                            (TG_BIGRID_SUMW == 3 && MIXED);
      T ux = 0, uy = 0, uz = 0, up = 0, uw = 0;
      T zr[8];
      if constexpr (ZROW) {
        const auto d0 = L.dz[0][jj], d1 = L.dz[1][jj];
        zr[0] = d0.x; zr[1] = d0.y; zr[2] = d0.z; zr[3] = d0.w;
        zr[4] = d1.x; zr[5] = d1.y; zr[6] = d1.z; zr[7] = d1.w;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const T dz = ZROW ? zr[k] : P.z - qz[k];
        const T r2 = fma(dz, dz, base2);
        T y;
        if constexpr (MIXED) {
          y = __builtin_amdgcn_rsqf(r2);
          if (GUARD) y = r2 > 0.0f ? y : 0.0f;
        } else {
          y = rsqrt_f64<GUARD ? F_GUARD : 0>(r2);
        }
        const T y3 = (y * y) * y;
        const T wp = P.w * y3;
        cp[k] = fma(P.w, y, cp[k]);
        cx[k] = fma(wp, dx, cx[k]);
        cy[k] = fma(wp, dy, cy[k]);
        cz[k] = fma(wp, dz, cz[k]);
        const T wc = qm[k] * y3;
        up = fma(qm[k], y, up);
        if constexpr (sumw) {
          uw += wc;
        } else {
          ux = fma(wc, dx, ux);
          uy = fma(wc, dy, uy);
        }
        uz = fma(wc, dz, uz);
      }
      if constexpr (sumw) {
        ux = uw * dx;
        uy = uw * dy;
      }
      auto a = L.acc[jj];         // the tile's sum for this particle (FP32 in mixed mode)
      a.x += ux;
      a.y += uy;
      a.z += uz;
      a.w += up;
      L.acc[jj] = a;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      hx[k] += (double)cx[k];
      hy[k] += (double)cy[k];
      hz[k] += (double)cz[k];
      hp[k] += (double)cp[k];
    }
    __syncthreads();
    if (j < pe) {                 // this block's partial sums for particle j
      const auto a = L.acc[lane];
      ppart[j] = (double)a.x;
      ppart[np + j] = (double)a.y;
      ppart[2 * np + j] = (double)a.z;
      ppart[3 * np + j] = (double)a.w;
    }
  }
}

// blockIdx.x = leaf block, blockIdx.y = particle chunk.  Cell results (C order inside the
// block) go to CO (final, or chunk partials when CO.part is set); particle partials go to
// ppart[(block * 4 + q) * np + j], summed by k_bidir_reduce in block order.
template <bool MIXED>
__global__ __launch_bounds__(64) void k_bidir_grid(CellBlocks C, const double4* __restrict__ p4,
                                                   const double* __restrict__ pe2, int64_t np,
                                                   int64_t chunk, const double* e2p_min_p, Out CO,
                                                   double* __restrict__ ppart) {
  typedef typename std::conditional<MIXED, float, double>::type T;
  __shared__ LdsBiGrid<MIXED> L;
  if (gated_off(CO)) return;    // explicit cells that are not 8^3 block columns
  const int lane = threadIdx.x;
  const int64_t b = blockIdx.x;
  const int i = lane >> 3, jj = lane & 7;
  const int64_t pb = (int64_t)blockIdx.y * chunk;
  const int64_t pe = pb + chunk < np ? pb + chunk : np;
  const int64_t t0 = b * 512 + (int64_t)lane * 8;
  double x, y, z[8], e;
  double ax_ = 0.0, ay_ = 0.0, az_ = 0.0;
  if (C.cx) {                     // explicit cells (uniform branch)
    x = C.cx[t0];
    y = C.cy[t0];
#pragma unroll
    for (int k = 0; k < 8; ++k) z[k] = C.cz[t0 + k];
    e = C.ceps ? C.ceps[b * 512] : 0.0;
    if (MIXED) {                  // anchor: the centre of the group's bounding box
      double mn[3] = {x, y, z[0]}, mx[3] = {x, y, z[0]};
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        mn[2] = fmin(mn[2], z[k]);
        mx[2] = fmax(mx[2], z[k]);
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          mn[v] = fmin(mn[v], __shfl_xor(mn[v], off));
          mx[v] = fmax(mx[v], __shfl_xor(mx[v], off));
        }
      }
      ax_ = 0.5 * (mn[0] + mx[0]);
      ay_ = 0.5 * (mn[1] + mx[1]);
      az_ = 0.5 * (mn[2] + mx[2]);
    }
  } else {
    const double d = C.dx[b];
    e = C.eps ? C.eps[b] : 0.0;
    {
#pragma clang fp contract(off)
      x = C.ox[b] + ((double)i + 0.5) * d;
      y = C.oy[b] + ((double)jj + 0.5) * d;
#pragma unroll
      for (int k = 0; k < 8; ++k) z[k] = C.oz[b] + ((double)k + 0.5) * d;
    }
    if (MIXED) {                  // anchor: the centre of the cell centres' bounding box, as
#pragma clang fp contract(off)      // the explicit form computes it from the same centres
      ax_ = 0.5 * ((C.ox[b] + 0.5 * d) + (C.ox[b] + 7.5 * d));
      ay_ = 0.5 * ((C.oy[b] + 0.5 * d) + (C.oy[b] + 7.5 * d));
      az_ = 0.5 * ((C.oz[b] + 0.5 * d) + (C.oz[b] + 7.5 * d));
    }
  }
  const double e2c = e * e;
  T qx = (T)(x - ax_), qy = (T)(y - ay_), qz[8], qm[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    qz[k] = (T)(z[k] - az_);
    qm[k] = (T)C.m[t0 + k];
  }
  const double lo = *e2p_min_p + e2c;
  const bool guard = MIXED ? !(lo > TG_F32_TINY_E2) : !(lo > 0.0);
  double hx[8], hy[8], hz[8], hp[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) hx[k] = hy[k] = hz[k] = hp[k] = 0.0;
  double* pp = ppart + b * 4 * np;
  // the same 8 z in every lane: leaf blocks always, explicit cells in their C order (wave vote)
  bool same = true;
#pragma unroll
  for (int k = 0; k < 8; ++k) same = same && z[k] == __shfl(z[k], 0);
  const bool zrow = !C.cx || __all(same) != 0;
  if (zrow) {
    if (guard) bi_grid_tiles<MIXED, true, true>(L, p4, pe2, pb, pe, e2c, ax_, ay_, az_, qx, qy, qz, qm, hx, hy, hz, hp, pp, np);
    else bi_grid_tiles<MIXED, false, true>(L, p4, pe2, pb, pe, e2c, ax_, ay_, az_, qx, qy, qz, qm, hx, hy, hz, hp, pp, np);
  } else {
    if (guard) bi_grid_tiles<MIXED, true, false>(L, p4, pe2, pb, pe, e2c, ax_, ay_, az_, qx, qy, qz, qm, hx, hy, hz, hp, pp, np);
    else bi_grid_tiles<MIXED, false, false>(L, p4, pe2, pb, pe, e2c, ax_, ay_, az_, qx, qy, qz, qm, hx, hy, hz, hp, pp, np);
  }
  Out o = CO;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
#pragma unroll
  for (int k = 0; k < 8; ++k) store_result(o, t0 + k, hx[k], hy[k], hz[k], hp[k]);
}

}  // namespace tg
