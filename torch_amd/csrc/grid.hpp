// Leaf-block kernels for 8^3 blocks (FLASH's NXB = NYB = NZB = 8).
//
// Block targets (k_eval_grid): one wave = one leaf block.  Lane l owns the cell column
// (i, j) = (l / 8, l % 8) and its 8 cells along z, so for every source the lane forms
// dx, dy and dx^2 + dy^2 + eps^2 once and only dz and r^2 per cell:
//   per interaction  dz (1) + r^2 (1) + rsqrt + m y, y^2, w (3) + phi (1) + a (3)
//   per source/lane  dx, dy (2) + dx^2 + dy^2 + eps^2 (2), shared by 8 cells
// i.e. ~9.5 VALU instructions instead of 13 (FP32) per interaction.  The 64 columns of a leaf
// block have the same 8 cell z, so a source's 8 dz are the same in every lane: the workgroup
// forms them once per source when it stages the source in LDS (F_ZROW), and each lane reads them
// (two broadcast LDS reads per source) instead of forming them, 8.5 instructions per interaction.
// Output index b*512 + (i*8 + j)*8 + k is the C order of the block format.
//
// Block sources (k_eval_gridsrc): the same sharing with the roles swapped -- the cells of a
// source block are walked column by column, and every lane forms dx, dy and the partial r^2
// of its K targets once per column; dz depends only on (cell row, target), so it is formed
// once per block, and the x, y accelerations take dx, dy times the column's sum of weights:
//   per interaction  r^2 (1) + rsqrt + m y, y^2, w (3) + phi (1) + sum w (1) + a_z (1)
//   per column/target  dx, dy (2) + base (2) + a_x, a_y (2), shared by 8 cells
// (the sum of weights in FP64; the mixed kernel keeps a_x, a_y per cell, see gs_block).
#pragma once
#include "kernels.hpp"

// Mixed-precision column kernel: 64-source LDS tiles whose FP32 partial sums are added to FP64
// accumulators every TG_GRID_FOLD_TILES tiles; TG_GRID_RCP = 1 forms m / r^3 as (m y) rcp(r^2)
// (a transcendental op in place of a multiply).
#ifndef TG_GRID_FOLD_TILES
#define TG_GRID_FOLD_TILES 1
#endif
#ifndef TG_GRID_RCP
#define TG_GRID_RCP 0
#endif

namespace tg {

// Block-source kernels (mixed): the next source block's masses loaded into registers while the
// current block is walked (c5 reverse 91.9 -> 90.4 ms).
#ifndef TG_GS_PREFETCH
#define TG_GS_PREFETCH 1
#endif
#ifndef TG_WAVE_LOCAL_STAGE
#define TG_WAVE_LOCAL_STAGE 1
#endif

// The column kernels' source rounds: the lanes of wave v stage exactly the sources that wave v
// then reads, so a round needs ordering only within the wave, not a workgroup barrier: a wave's LDS
// operations are carried out in order, and the fences keep the compiler from moving its LDS reads
// and writes across the point.  The waves of a workgroup then run their rounds independently, and
// one wave's wait for its next sources no longer holds the others (c4 mixed kernel 11.91 -> 11.78
// ms).  The LDS the rounds use is shared with the waves' final sum (LdsCol), which begins with a
// workgroup barrier.
template <int W>
__device__ __forceinline__ void stage_sync() {
#if TG_WAVE_LOCAL_STAGE
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#else
  __syncthreads();
#endif
}

// Column flag (set by grid_columns only, never by tile_flags): every lane of the workgroup holds
// the same 8 target z, and the staged sources carry their 8 dz.
enum { F_ZROW = 16 };

template <int NT>
struct LdsGrid {
  double4 s4[NT];
  double se2[NT];
  double dz[NT][8];   // F_ZROW: z_s - z_k of the 8 target cells
};
template <int NT>
struct LdsGridF {
  float4 s4[NT];
  float se2[NT];
  float4 dz[NT][2];   // F_ZROW: fp32(z_s - c) - fp32(z_k - c) of the 8 target cells
};

// LDS of a column workgroup of W waves that share one target block and split its sources: the
// source tile (64 sources per wave, staged by the whole workgroup) and, after the source loop, in
// the same space, the waves' partial sums on their way to one total (32 FP64 values per lane).
template <bool MIXED, int W>
struct LdsCol {
  typedef typename std::conditional<MIXED, LdsGridF<64 * W>, LdsGrid<64 * W>>::type Tile;
  union {
    Tile t;
    double red[W > 1 ? 32 * 64 : 1];
  };
};

template <int FL>
__device__ __forceinline__ void grid_tile_f64(const double4* __restrict__ s4, const double* __restrict__ se2,
                                              const double (*__restrict__ sdz)[8], double tx, double ty,
                                              const double (&tz)[8], double (&ax)[8], double (&ay)[8],
                                              double (&az)[8], double (&ap)[8]) {
  // The 8 cells' dependency chains are written step by step across the cells (each step for all 8
  // before the next), so that independent FP64 operations sit next to each other: written cell by
  // cell, the compiler keeps each chain of ~16 dependent FP64 operations in one piece (measured 2-4%
  // slower).
#pragma unroll 1
  for (int j = 0; j < 64; ++j) {
    const double4 s = s4[j];
    const double dx = s.x - tx, dy = s.y - ty;
    const double base = fma(dx, dx, fma(dy, dy, se2[j]));
    double dz[8], r2[8], y[8], e[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      dz[k] = (FL & F_ZROW) ? sdz[j][k] : s.z - tz[k];
      r2[k] = fma(dz[k], dz[k], base);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (FL & F_SEED32) y[k] = (double)__builtin_amdgcn_rsqf((float)r2[k]);
      else y[k] = __builtin_amdgcn_rsq(r2[k]);
      if (FL & F_GUARD) y[k] = r2[k] > 0.0 ? y[k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = fma(-(r2[k] * y[k]), y[k], 1.0);
#pragma unroll
    for (int k = 0; k < 8; ++k) y[k] = fma(y[k] * e[k], fma(e[k], 0.375, 0.5), y[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double mr = s.w * y[k];
      if (!(FL & F_NOPOT)) ap[k] += mr;
      const double w = mr * (y[k] * y[k]);
      ax[k] = fma(w, dx, ax[k]);
      ay[k] = fma(w, dy, ay[k]);
      az[k] = fma(w, dz[k], az[k]);
    }
  }
}

template <int FL>
__device__ __forceinline__ void grid_tile_f32(const float4* __restrict__ s4, const float* __restrict__ se2,
                                              const float4 (*__restrict__ sdz)[2], float tx, float ty,
                                              const float (&tz)[8], float (&ax)[8], float (&ay)[8],
                                              float (&az)[8], float (&ap)[8]) {
#pragma unroll 2
  for (int j = 0; j < 64; ++j) {
    const float4 s = s4[j];
    const float dx = s.x - tx, dy = s.y - ty;
    const float base = fmaf(dx, dx, fmaf(dy, dy, se2[j]));
    float zr[8];
    if (FL & F_ZROW) {
      const float4 d0 = sdz[j][0], d1 = sdz[j][1];
      zr[0] = d0.x; zr[1] = d0.y; zr[2] = d0.z; zr[3] = d0.w;
      zr[4] = d1.x; zr[5] = d1.y; zr[6] = d1.z; zr[7] = d1.w;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float dz = (FL & F_ZROW) ? zr[k] : s.z - tz[k];
      const float r2 = fmaf(dz, dz, base);
      float y = __builtin_amdgcn_rsqf(r2);
      if (FL & F_GUARD) y = r2 > 0.0f ? y : 0.0f;
      const float mr = s.w * y;
      if (!(FL & F_NOPOT)) ap[k] += mr;
#if TG_GRID_RCP
      float ri = __builtin_amdgcn_rcpf(r2);
      if (FL & F_GUARD) ri = r2 > 0.0f ? ri : 0.0f;
      const float w = mr * ri;
#else
      const float w = (mr * y) * y;
#endif
      ax[k] = fmaf(w, dx, ax[k]);
      ay[k] = fmaf(w, dy, ay[k]);
      az[k] = fmaf(w, dz, az[k]);
    }
  }
}

// Sources [sb, se) in rounds of 64 W, staged by all 64 W lanes of the workgroup; wave v takes
// sources base + 64 v ... + 63 of each round (a wave whose slice starts past se skips the round;
// a partial slice is padded with massless sources, which add exactly zero).
template <int FL, int W>
__device__ __forceinline__ void grid_sources_f64(LdsGrid<64 * W>& L, const double4* __restrict__ src4,
                                                 const double* __restrict__ srce2, int64_t sb,
                                                 int64_t se, double e2t, double tx, double ty,
                                                 const double (&tz)[8], double (&ax)[8],
                                                 double (&ay)[8], double (&az)[8], double (&ap)[8]) {
  const int tid = threadIdx.x, v = tid >> 6;
  for (int64_t base = sb; base < se; base += 64 * W) {
    stage_sync<W>();
    const int64_t j = base + tid;
    const double4 s = j < se ? src4[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    L.s4[tid] = s;
    L.se2[tid] = j < se ? srce2[j] + e2t : 1.0;
    if (FL & F_ZROW) {   // this lane's tz are every lane's
#pragma unroll
      for (int k = 0; k < 8; ++k) L.dz[tid][k] = s.z - tz[k];
    }
    stage_sync<W>();
    if (W == 1 || base + 64 * v < se)   // uniform over the wave
      grid_tile_f64<FL>(L.s4 + 64 * v, L.se2 + 64 * v, L.dz + 64 * v, tx, ty, tz, ax, ay, az, ap);
  }
}

template <int FL, int W>
__device__ __forceinline__ void grid_sources_f32(LdsGridF<64 * W>& L, const double4* __restrict__ src4,
                                                 const double* __restrict__ srce2, int64_t sb,
                                                 int64_t se, double e2t, double cx, double cy,
                                                 double cz, float tx, float ty, const float (&tz)[8],
                                                 double (&hx)[8], double (&hy)[8], double (&hz)[8],
                                                 double (&hp)[8]) {
  const int tid = threadIdx.x, v = tid >> 6;
  float ax[8], ay[8], az[8], ap[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) ax[k] = ay[k] = az[k] = ap[k] = 0.f;
  int tiles = 0;
  for (int64_t base = sb; base < se; base += 64 * W) {
    stage_sync<W>();
    const int64_t j = base + tid;
    float4 s4 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j < se) {
      const double4 s = src4[j];
      s4 = make_float4((float)(s.x - cx), (float)(s.y - cy), (float)(s.z - cz), (float)s.w);
    }
    L.s4[tid] = s4;
    L.se2[tid] = j < se ? (float)(srce2[j] + e2t) : 1.0f;
    if (FL & F_ZROW) {   // this lane's tz are every lane's: the dz the tile loop would form
      L.dz[tid][0] = make_float4(s4.z - tz[0], s4.z - tz[1], s4.z - tz[2], s4.z - tz[3]);
      L.dz[tid][1] = make_float4(s4.z - tz[4], s4.z - tz[5], s4.z - tz[6], s4.z - tz[7]);
    }
    stage_sync<W>();
    if (W > 1 && base + 64 * v >= se) continue;   // uniform over the wave: nothing in its slice
    grid_tile_f32<FL>(L.s4 + 64 * v, L.se2 + 64 * v, L.dz + 64 * v, tx, ty, tz, ax, ay, az, ap);
    if (++tiles == TG_GRID_FOLD_TILES) {   // uniform: every lane of the wave counts the same tiles
      tiles = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        hx[k] += (double)ax[k];
        hy[k] += (double)ay[k];
        hz[k] += (double)az[k];
        if (!(FL & F_NOPOT)) hp[k] += (double)ap[k];
        ax[k] = ay[k] = az[k] = ap[k] = 0.f;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    hx[k] += (double)ax[k];
    hy[k] += (double)ay[k];
    hz[k] += (double)az[k];
    if (!(FL & F_NOPOT)) hp[k] += (double)ap[k];
  }
}

template <int W>
struct GridF64Run {
  template <int FL>
  static __device__ __forceinline__ void run(LdsGrid<64 * W>& L, const double4* src4, const double* srce2,
                                             int64_t sb, int64_t se, double e2t, double tx,
                                             double ty, const double (&tz)[8], double (&ax)[8],
                                             double (&ay)[8], double (&az)[8], double (&ap)[8]) {
    grid_sources_f64<FL | F_UNIF, W>(L, src4, srce2, sb, se, e2t, tx, ty, tz, ax, ay, az, ap);
  }
};

template <int W>
struct GridF32Run {
  template <int FL>
  static __device__ __forceinline__ void run(LdsGridF<64 * W>& L, const double4* src4, const double* srce2,
                                             int64_t sb, int64_t se, double e2t, double cx,
                                             double cy, double cz, float tx, float ty,
                                             const float (&tz)[8], double (&hx)[8], double (&hy)[8],
                                             double (&hz)[8], double (&hp)[8]) {
    grid_sources_f32<FL | F_UNIF, W>(L, src4, srce2, sb, se, e2t, cx, cy, cz, tx, ty, tz, hx, hy, hz,
                                     hp);
  }
};

#define TG_DISPATCH_GRID_(FN, Z, fl, ...)                                     \
  switch (fl & (F_GUARD | F_NOPOT | F_SEED32)) {                              \
    case 0: FN<Z>(__VA_ARGS__); break;                                        \
    case F_GUARD: FN<Z | F_GUARD>(__VA_ARGS__); break;                        \
    case F_NOPOT: FN<Z | F_NOPOT>(__VA_ARGS__); break;                        \
    case F_NOPOT | F_GUARD: FN<Z | F_NOPOT | F_GUARD>(__VA_ARGS__); break;    \
    case F_SEED32: FN<Z | F_SEED32>(__VA_ARGS__); break;                      \
    default: FN<Z | F_SEED32 | F_NOPOT>(__VA_ARGS__); break;                  \
  }
#define TG_DISPATCH_GRID(FN, fl, ...)                                         \
  if (fl & F_ZROW) {                                                          \
    TG_DISPATCH_GRID_(FN, F_ZROW, fl, __VA_ARGS__)                            \
  } else {                                                                    \
    TG_DISPATCH_GRID_(FN, 0, fl, __VA_ARGS__)                                 \
  }

// Sums the W waves' partial results (a, b, c, d: N values per lane each) of a workgroup whose
// waves split one sum, in wave order, ((s_0 + s_1) + s_2) + ..., through red (4 N 64 doubles of
// LDS); returns true on wave W - 1, which then holds the total.  Every wave of the workgroup must
// call it, after its own loop.
template <int W, int N>
__device__ __forceinline__ bool wave_sum(double* red, double (&a)[N], double (&b)[N], double (&c)[N],
                                         double (&d)[N]) {
  if constexpr (W == 1) {
    return true;
  } else {
    const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
    __syncthreads();                       // the loop's LDS (possibly the same space) is no longer read
    for (int w = 0; w < W; ++w) {
      if (v == w) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          double* r = red + k * 256 + lane;
          if (w > 0) {
            a[k] = r[0] + a[k];
            b[k] = r[64] + b[k];
            c[k] = r[128] + c[k];
            d[k] = r[192] + d[k];
          }
          if (w < W - 1) {
            r[0] = a[k];
            r[64] = b[k];
            r[128] = c[k];
            r[192] = d[k];
          }
        }
      }
      if (w < W - 1) __syncthreads();
    }
    return v == W - 1;
  }
}

// One target column group against sources [sb, se), by a workgroup of W waves that all hold the
// same targets and split the sources (grid_sources_*): lane l (of every wave) owns 8 targets
// t0 + k that share x, y and eps^2 (e2t, uniform over the wave) and differ in z[k]; (cx, cy, cz) is
// the FP64 anchor of the mixed form (the centre of the group's bounding box); targets t0 + k with
// k >= nk are not stored.  `o` already points at this source chunk's partials if the sources are
// split over workgroups.  sst: the source statistics; rt: the largest |coordinate| of the targets
// (FP64 seed choice); zrow: every lane of the workgroup holds the same z[8] (F_ZROW).
template <bool MIXED, int W>
__device__ __forceinline__ void grid_columns(
    LdsCol<MIXED, W>& L, double x, double y,
    const double (&z)[8], double e2t, double cx, double cy, double cz, int64_t t0, int nk,
    const double4* __restrict__ src4, const double* __restrict__ srce2, int64_t sb, int64_t se,
    const double* sst, double rt, const Out& o, bool zrow) {
  const int fl = tile_flags(sst[ST_E2MIN], e2t, e2t, o, MIXED, r2_bound(sst, rt, e2t)) |
                 (zrow ? F_ZROW : 0);
  if constexpr (MIXED) {
    const float tx = (float)(x - cx), ty = (float)(y - cy);
    float tz[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) tz[k] = (float)(z[k] - cz);
    double hx[8], hy[8], hz[8], hp[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) hx[k] = hy[k] = hz[k] = hp[k] = 0.0;
    TG_DISPATCH_GRID(GridF32Run<W>::template run, fl & ~F_SEED32, L.t, src4, srce2, sb, se, e2t, cx,
                     cy, cz, tx, ty, tz, hx, hy, hz, hp)
    if (wave_sum<W, 8>(L.red, hx, hy, hz, hp)) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < nk) store_result(o, t0 + k, hx[k], hy[k], hz[k], hp[k]);
    }
  } else {
    double ax[8], ay[8], az[8], ap[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) ax[k] = ay[k] = az[k] = ap[k] = 0.0;
    TG_DISPATCH_GRID(GridF64Run<W>::template run, fl, L.t, src4, srce2, sb, se, e2t, x, y, z, ax, ay,
                     az, ap)
    if (wave_sum<W, 8>(L.red, ax, ay, az, ap)) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < nk) store_result(o, t0 + k, ax[k], ay[k], az[k], ap[k]);
    }
  }
}

// One leaf block b of block-structured targets against sources [sb, se), by a workgroup of W
// waves (lane l of each = cell column (i, j) = (l / 8, l % 8)).  Shared by k_eval_grid and the
// persistent batch kernel.  The mixed anchor is the centre of the bounding box of the block's
// cell centres, (c_0 + c_7) / 2 per axis, the same value the explicit-target form below
// computes from the same centres.
template <bool MIXED, int W>
__device__ __forceinline__ void grid_block(LdsCol<MIXED, W>& L, const Tgt& T, int64_t b,
                                           const double4* __restrict__ src4,
                                           const double* __restrict__ srce2, int64_t sb, int64_t se,
                                           const double* sst, const Out& o) {
  const int lane = threadIdx.x & 63;
  const int i = lane >> 3, jj = lane & 7;
  const double d = T.bdx[b];
  const double e = T.beps ? T.beps[b] : 0.0;
  double x, y, z[8], cx, cy, cz;
  {
#pragma clang fp contract(off)
    const double ox = T.ox[b], oy = T.oy[b], oz = T.oz[b];
    x = ox + ((double)i + 0.5) * d;
    y = oy + ((double)jj + 0.5) * d;
#pragma unroll
    for (int k = 0; k < 8; ++k) z[k] = oz + ((double)k + 0.5) * d;
    cx = 0.5 * ((ox + 0.5 * d) + (ox + 7.5 * d));
    cy = 0.5 * ((oy + 0.5 * d) + (oy + 7.5 * d));
    cz = 0.5 * ((oz + 0.5 * d) + (oz + 7.5 * d));
  }
  // every cell centre lies within [o, o + 8 d] on each axis
  const double rt = fmax(max_abs3(T.ox[b], T.oy[b], T.oz[b]),
                         max_abs3(T.ox[b] + 8.0 * d, T.oy[b] + 8.0 * d, T.oz[b] + 8.0 * d));
  grid_columns<MIXED, W>(L, x, y, z, e * e, cx, cy, cz, b * 512 + (int64_t)lane * 8, 8, src4, srce2,
                         sb, se, sst, rt, o, true);
}

// blockIdx.x = leaf block, blockIdx.y = source chunk.  Requires T.blocks && T.nb == 8 and a
// workgroup of 64 W lanes.
template <bool MIXED, int W>
__global__ __launch_bounds__(64 * W) void k_eval_grid(Tgt T, const double4* __restrict__ src4,
                                                      const double* __restrict__ srce2, int64_t ns,
                                                      int64_t chunk, const double* sst, Out O) {
  __shared__ LdsCol<MIXED, W> L;
  const int64_t sb = (int64_t)blockIdx.y * chunk;
  const int64_t se = sb + chunk < ns ? sb + chunk : ns;
  Out o = O;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
  grid_block<MIXED, W>(L, T, blockIdx.x, src4, srce2, sb, se, sst, o);
}

// Explicit targets with column structure (every aligned group of 8 consecutive targets shares
// x and y, every aligned group of 512 shares eps: the C order of 8^3 leaf-block meshgrids, as
// checked by k_check_cols): the column code of k_eval_grid on explicit coordinates.  Group g =
// targets 512 g ... 512 g + 511, a workgroup of W waves, lane l of each owns targets
// 512 g + 8 l ... + 7.  Shared by k_eval_cols and the persistent batch kernel.
template <bool MIXED, int W>
__device__ __forceinline__ void cols_group(LdsCol<MIXED, W>& L, const Tgt& T, int64_t g,
                                           const double4* __restrict__ src4,
                                           const double* __restrict__ srce2, int64_t sb, int64_t se,
                                           const double* sst, const Out& o) {
  const int lane = threadIdx.x & 63;
  const int64_t t0 = g * 512 + (int64_t)lane * 8;
  const int64_t last = T.n - 1;
  const int64_t tc = t0 < last ? t0 : last;
  const double x = T.x[tc], y = T.y[tc];
  double z[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) z[k] = T.z[t0 + k < last ? t0 + k : last];
  const int64_t te = g * 512 < last ? g * 512 : last;
  const double e = T.eps ? T.eps[te] : 0.0;
  double cx = 0.0, cy = 0.0, cz = 0.0;
  if constexpr (MIXED) {   // centre of the bounding box of the group's targets
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
    cx = 0.5 * (mn[0] + mx[0]);
    cy = 0.5 * (mn[1] + mx[1]);
    cz = 0.5 * (mn[2] + mx[2]);
  }
  double rt = 0.0;   // the group's largest |coordinate| (the FP64 seed choice)
  if constexpr (!MIXED) {
    rt = fmax(fabs(x), fabs(y));
#pragma unroll
    for (int k = 0; k < 8; ++k) rt = fmax(rt, fabs(z[k]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rt = fmax(rt, __shfl_xor(rt, off));
  }
  // the same 8 z in every lane (every wave of the workgroup holds these targets): the C order of
  // 8^3 leaf blocks has it; a ragged group's clamped lanes and other layouts take the per-lane dz
  bool same = true;
#pragma unroll
  for (int k = 0; k < 8; ++k) same = same && z[k] == __shfl(z[k], 0);
  const bool zrow = __all(same) != 0;
  const int64_t left = T.n - t0;
  grid_columns<MIXED, W>(L, x, y, z, e * e, cx, cy, cz, t0, left < 8 ? (int)left : 8, src4, srce2,
                         sb, se, sst, rt, o, zrow);
}

// blockIdx.x = group of 512 targets, blockIdx.y = source chunk.  Gated (Out::gate) against
// the generic kernel on the check's flag.
template <bool MIXED, int W>
__global__ __launch_bounds__(64 * W) void k_eval_cols(Tgt T, const double4* __restrict__ src4,
                                                      const double* __restrict__ srce2, int64_t ns,
                                                      int64_t chunk, const double* sst, Out O) {
  __shared__ LdsCol<MIXED, W> L;
  if (gated_off(O)) return;
  const int64_t sb = (int64_t)blockIdx.y * chunk;
  const int64_t se = sb + chunk < ns ? sb + chunk : ns;
  Out o = O;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
  cols_group<MIXED, W>(L, T, blockIdx.x, src4, srce2, sb, se, sst, o);
}

// ---------------------------------------------------------------- block sources ----
template <bool MIXED>
struct LdsGS {
  typedef typename std::conditional<MIXED, float, double>::type T;
  T m[512];                    // cell masses of the staged block, C order
  T cx[8], cy[8], cz[8];       // cell-centre coordinates (mixed: relative to the anchor)
  T e2;                        // block eps^2 (+ the tile's uniform target eps^2)
  double red[2 * 4 * 4];
};

template <int K, bool MIXED, int FL>
__device__ __forceinline__ void gs_block(const LdsGS<MIXED>& L,
                                         const typename LdsGS<MIXED>::T (&tx)[K],
                                         const typename LdsGS<MIXED>::T (&ty)[K],
                                         const typename LdsGS<MIXED>::T (&tz)[K],
                                         const typename LdsGS<MIXED>::T (&te2)[K],
                                         typename LdsGS<MIXED>::T (&ax)[K],
                                         typename LdsGS<MIXED>::T (&ay)[K],
                                         typename LdsGS<MIXED>::T (&az)[K],
                                         typename LdsGS<MIXED>::T (&ap)[K]) {
  typedef typename LdsGS<MIXED>::T T;
  const T e2 = L.e2;
  for (int col = 0; col < 64; ++col) {
    const T xs = L.cx[col >> 3], ys = L.cy[col & 7];
    T dx[K], dy[K], base[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      dx[k] = xs - tx[k];
      dy[k] = ys - ty[k];
      base[k] = fma(dx[k], dx[k], fma(dy[k], dy[k], (FL & F_UNIF) ? e2 : e2 + te2[k]));
    }
    // The 8 cells of the column share dx and dy, so sum_cells w dx = dx sum_cells w: one add
    // per cell instead of two FMAs, then two FMAs per column.  FP64 only: in the mixed kernel the
    // shorter loop measured slower (7% more cycles at the same occupancy, also with the FP64
    // accumulators moved to LDS for a fourth wave per SIMD), so it keeps the per-cell FMAs.
    T sw[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sw[k] = (T)0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const T zs = L.cz[kk];
      const T m = L.m[col * 8 + kk];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const T dz = zs - tz[k];
        const T r2 = fma(dz, dz, base[k]);
        T y;
        if constexpr (MIXED) {
          y = __builtin_amdgcn_rsqf(r2);
          if (FL & F_GUARD) y = r2 > 0.0f ? y : 0.0f;
        } else {
          y = rsqrt_f64<FL>(r2);
        }
        const T mr = m * y;
        if (!(FL & F_NOPOT)) ap[k] += mr;
        const T w = (mr * y) * y;
        if constexpr (MIXED) {
          ax[k] = fma(w, dx[k], ax[k]);
          ay[k] = fma(w, dy[k], ay[k]);
        } else {
          sw[k] += w;
        }
        az[k] = fma(w, dz, az[k]);
      }
    }
    if constexpr (!MIXED) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        ax[k] = fma(sw[k], dx[k], ax[k]);
        ay[k] = fma(sw[k], dy[k], ay[k]);
      }
    }
  }
}

// Stages one source block (masses and cell-centre coordinates) into LDS.
// sax[24 blk + 8 axis + idx] is the block's cell-centre coordinate idx along axis (x, y, z).
template <bool MIXED>
__device__ __forceinline__ void gs_stage(LdsGS<MIXED>& L, const double* __restrict__ sax,
                                         const double* __restrict__ be2,
                                         const double* __restrict__ bm, int64_t blk, double e2u,
                                         double ax_, double ay_, double az_) {
  typedef typename LdsGS<MIXED>::T T;
  const int tid = threadIdx.x;
  for (int q = tid; q < 512; q += blockDim.x) L.m[q] = (T)bm[blk * 512 + q];
  if (tid < 24) {
    const int axis = tid >> 3, idx = tid & 7;
    const double c = sax[blk * 24 + tid];
    const double a = axis == 0 ? ax_ : axis == 1 ? ay_ : az_;
    T v = MIXED ? (T)(c - a) : (T)c;
    if (axis == 0) L.cx[idx] = v;
    else if (axis == 1) L.cy[idx] = v;
    else L.cz[idx] = v;
  }
  if (tid == 0) L.e2 = (T)(be2[blk] + e2u);
}

// The block-source kernels' workgroup: W waves that hold the same K x 64 targets and split the
// source blocks (wave v takes blocks b0 + v, b0 + v + W, ...; each stages its own block), then
// add their sums in wave order.
template <bool MIXED, int W>
struct LdsGSW {
  LdsGS<MIXED> s[W];           // the block each wave is on
  double red[W > 1 ? 16 * 64 : 1];   // cross-wave sums: 4 K doubles per lane (K <= 4)
};

// Stages source block blk into one wave's slot, by that wave's 64 lanes.
template <bool MIXED>
__device__ __forceinline__ void gs_stage_wave(LdsGS<MIXED>& L, const double* __restrict__ sax,
                                              const double* __restrict__ be2,
                                              const double* __restrict__ bm, int64_t blk, double e2u,
                                              double ax_, double ay_, double az_) {
  typedef typename LdsGS<MIXED>::T T;
  const int lane = threadIdx.x & 63;
  for (int q = lane; q < 512; q += 64) L.m[q] = (T)bm[blk * 512 + q];
  if (lane < 24) {
    const int axis = lane >> 3, idx = lane & 7;
    const double c = sax[blk * 24 + lane];
    const double a = axis == 0 ? ax_ : axis == 1 ? ay_ : az_;
    T v = MIXED ? (T)(c - a) : (T)c;
    if (axis == 0) L.cx[idx] = v;
    else if (axis == 1) L.cy[idx] = v;
    else L.cz[idx] = v;
  }
  if (lane == 0) L.e2 = (T)(be2[blk] + e2u);
}

template <int K, int W, bool MIXED>
struct GsRun {
  typedef typename LdsGS<MIXED>::T T;
  template <int FL>
  static __device__ __forceinline__ void run(LdsGSW<MIXED, W>& L, const double* sax, const double* be2,
                                             const double* bm, int64_t b0, int64_t b1, double e2u,
                                             double ax_, double ay_, double az_, const T (&tx)[K],
                                             const T (&ty)[K], const T (&tz)[K], const T (&te2)[K],
                                             double (&hx)[K], double (&hy)[K], double (&hz)[K],
                                             double (&hp)[K]) {
    const int v = threadIdx.x >> 6;
    // PF: the wave's next block (masses, axis value, eps^2) in registers while the current one is
    // walked (mixed only: the FP64 kernel has no registers to spare)
    constexpr bool PF = TG_GS_PREFETCH && MIXED;
    const int lane = threadIdx.x & 63;
    double pm[8], pa = 0.0, pe = 0.0;
    auto fetch = [&](int64_t b) {
#pragma unroll
      for (int i = 0; i < 8; ++i) pm[i] = bm[b * 512 + lane + 64 * i];
      if (lane < 24) pa = sax[b * 24 + lane];
      if (lane == 0) pe = be2[b];
    };
    if (PF && b0 + v < b1) fetch(b0 + v);
    for (int64_t base = b0; base < b1; base += W) {   // uniform over the workgroup
      const int64_t blk = base + v;
      // Each wave stages and reads only its own block here too, but with the waves' rounds left
      // unsynchronised this loop measured 0.3% slower (c5 reverse); it keeps the workgroup barriers.
      __syncthreads();
      if (PF && blk < b1) {             // the values gs_stage_wave would load, from the registers
        typedef typename LdsGS<MIXED>::T T;
        LdsGS<MIXED>& S = L.s[v];
#pragma unroll
        for (int i = 0; i < 8; ++i) S.m[lane + 64 * i] = (T)pm[i];
        if (lane < 24) {
          const int axis = lane >> 3, idx = lane & 7;
          const double a = axis == 0 ? ax_ : axis == 1 ? ay_ : az_;
          const T val = MIXED ? (T)(pa - a) : (T)pa;
          if (axis == 0) S.cx[idx] = val;
          else if (axis == 1) S.cy[idx] = val;
          else S.cz[idx] = val;
        }
        if (lane == 0) S.e2 = (T)(pe + ((FL & F_UNIF) ? e2u : 0.0));
      } else if (!PF && blk < b1) {
        gs_stage_wave<MIXED>(L.s[v], sax, be2, bm, blk, (FL & F_UNIF) ? e2u : 0.0, ax_, ay_, az_);
      }
      __syncthreads();
      if (blk >= b1) continue;          // uniform over the wave: no block left for it this round
      if (PF && blk + W < b1) fetch(blk + W);
      T ax[K], ay[K], az[K], ap[K];
#pragma unroll
      for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = ap[k] = (T)0;
      gs_block<K, MIXED, FL>(L.s[v], tx, ty, tz, te2, ax, ay, az, ap);
      // per-block partials folded into FP64 accumulators (mixed: one FP32 sum per 512 cells)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        hx[k] += (double)ax[k];
        hy[k] += (double)ay[k];
        hz[k] += (double)az[k];
        hp[k] += (double)ap[k];
      }
    }
  }
};

// One target tile (K per lane, targets tile0 + 64 k + lane, the same in every wave) against source
// blocks [b0, b1), split over the W waves: sax / be2 / bm are the set's cell-centre axes, block
// eps^2 and contiguous cell masses; o already points at this block chunk's partials if the blocks
// are split over workgroups.  Shared by k_eval_gridsrc and the persistent batch kernel.
template <int K, int W, bool MIXED>
__device__ __forceinline__ void gridsrc_tile(LdsGSW<MIXED, W>& L, const Tgt& T, int64_t tile0,
                                             const double* __restrict__ sax,
                                             const double* __restrict__ be2,
                                             const double* __restrict__ bm, int64_t b0, int64_t b1,
                                             const double* sst, const Out& o) {
  static_assert(K <= 4, "the cross-wave buffer holds 4 K <= 16 values per lane");
  typedef typename LdsGS<MIXED>::T TT;
  const int lane = threadIdx.x & 63;
  double px[K], py[K], pz[K], pe2[K];
  double mn[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  double mx[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * 64 + lane;
    load_target(T, t < T.n ? t : T.n - 1, px[k], py[k], pz[k], pe2[k]);
    if (t < T.n) {
      mn[0] = fmin(mn[0], px[k]); mx[0] = fmax(mx[0], px[k]);
      mn[1] = fmin(mn[1], py[k]); mx[1] = fmax(mx[1], py[k]);
      mn[2] = fmin(mn[2], pz[k]); mx[2] = fmax(mx[2], pz[k]);
      mn[3] = fmin(mn[3], pe2[k]); mx[3] = fmax(mx[3], pe2[k]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {   // over the wave: every wave holds the same tile
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      mn[q] = fmin(mn[q], __shfl_xor(mn[q], off));
      mx[q] = fmax(mx[q], __shfl_xor(mx[q], off));
    }
  }
  double ax_ = 0.0, ay_ = 0.0, az_ = 0.0;
  if (MIXED) {
    ax_ = 0.5 * (mn[0] + mx[0]);
    ay_ = 0.5 * (mn[1] + mx[1]);
    az_ = 0.5 * (mn[2] + mx[2]);
  }
  const double rt = fmax(fmax(max_abs3(mn[0], mn[1], mn[2]), max_abs3(mx[0], mx[1], mx[2])), 0.0);
  const int fl = tile_flags(sst[ST_E2MIN], mn[3], mx[3], o, MIXED, r2_bound(sst, rt, mx[3]));
  TT tx[K], ty[K], tz[K], te2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    tx[k] = (TT)(px[k] - ax_);
    ty[k] = (TT)(py[k] - ay_);
    tz[k] = (TT)(pz[k] - az_);
    te2[k] = (TT)pe2[k];
  }
  double hx[K], hy[K], hz[K], hp[K];
#pragma unroll
  for (int k = 0; k < K; ++k) hx[k] = hy[k] = hz[k] = hp[k] = 0.0;
  typedef GsRun<K, W, MIXED> R;
  if constexpr (MIXED) {
    TG_DISPATCH8(R::template run, fl, L, sax, be2, bm, b0, b1, mn[3], ax_, ay_, az_, tx, ty, tz, te2,
                 hx, hy, hz, hp)
  } else {
    TG_DISPATCH(R::template run, fl, L, sax, be2, bm, b0, b1, mn[3], ax_, ay_, az_, tx, ty, tz, te2,
                hx, hy, hz, hp)
  }
  if (!wave_sum<W, K>(L.red, hx, hy, hz, hp)) return;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * 64 + lane;
    if (t < T.n) store_result(o, t, hx[k], hy[k], hz[k], hp[k]);
  }
}

// Explicit targets (K per lane, K x 64 per workgroup of W waves) against 8^3 source blocks: a grid
// of (target tiles, block chunks [q chunk, min((q + 1) chunk, nblk))).  Workgroup L = x + y gridDim.x
// is dispatched to XCD L mod 8; with gridDim.y a multiple of 8 the kernel maps its workgroups so that
// all tiles of block chunk q run on XCD q mod 8 -- each chunk's masses are then fetched into one
// L2, not eight -- and workgroups whose chunk starts past the last block (the padding of
// gridDim.y to a multiple of 8) return.  Otherwise tile = x, chunk = y.
template <int K, int W, bool MIXED>
__global__ __launch_bounds__(64 * W) void k_eval_gridsrc(Tgt T, const double* __restrict__ sax,
                                                         const double* __restrict__ be2,
                                                         const double* __restrict__ bm, int64_t nblk,
                                                         int64_t chunk, const double* sst, Out O) {
  __shared__ LdsGSW<MIXED, W> L;
  if (gated_off(O)) return;   // explicit sources that turned out not to be 8^3 blocks
  int64_t tile = blockIdx.x, q = blockIdx.y;
  if (gridDim.y % 8 == 0) {
    const int64_t l = (int64_t)blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    const int64_t s = l >> 3;
    tile = s % gridDim.x;
    q = (s / gridDim.x) * 8 + (l & 7);
  }
  const int64_t b0 = q * chunk;
  if (b0 >= nblk) return;     // uniform over the workgroup
  const int64_t b1 = b0 + chunk < nblk ? b0 + chunk : nblk;
  Out o = O;
  if (o.part) o.part += q * 4 * o.pstride;
  gridsrc_tile<K, W, MIXED>(L, T, tile * (K * 64), sax, be2, bm, b0, b1, sst, o);
}

// Expands 8^3 source blocks into the packed per-cell form (so every kernel can use them),
// writes each block's eps^2 (be2), cell-centre axes (sax) and the cell masses as one contiguous
// array (sm; the three staged by the block-source kernels, which then read 8 B per cell instead
// of a 32 B record) and folds the source statistics (ST_*) into stat; one lane per cell,
// workgroups of 256 lanes.
__global__ __launch_bounds__(256) void k_pack_source_blocks(int64_t nblk, const double* __restrict__ ox,
                                     const double* __restrict__ oy, const double* __restrict__ oz,
                                     const double* __restrict__ dx, const double* __restrict__ m,
                                     const double* __restrict__ eps, double4* __restrict__ s4,
                                     double* __restrict__ se2, double* __restrict__ be2,
                                     double* __restrict__ sax, double* __restrict__ sm,
                                     unsigned long long* stat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long e2min = 0x7ff0000000000000ULL, rmax = 0ULL, e2max = 0ULL, mmax = 0ULL;
  if (i < nblk * 512) {
    const int64_t b = i >> 9;
    const int c = (int)(i & 511);
    const double d = dx[b];
    double x, y, z;
    {
#pragma clang fp contract(off)
      x = ox[b] + ((double)(c >> 6) + 0.5) * d;
      y = oy[b] + ((double)((c >> 3) & 7) + 0.5) * d;
      z = oz[b] + ((double)(c & 7) + 0.5) * d;
    }
    const double e = eps ? eps[b] : 0.0;
    const double mi = m[i];
    s4[i] = make_double4(x, y, z, mi);
    se2[i] = e * e;
    sm[i] = mi;
    const int ci = c >> 6, cj = (c >> 3) & 7, ck = c & 7;
    if (cj == 0 && ck == 0) sax[b * 24 + ci] = x;
    if (ci == 0 && ck == 0) sax[b * 24 + 8 + cj] = y;
    if (ci == 0 && cj == 0) sax[b * 24 + 16 + ck] = z;
    if (c == 0) be2[b] = e * e;
    e2min = e2max = dbits(e * e);
    rmax = dbits(max_abs3(x, y, z));
    mmax = dbits(fabs(mi));
  }
  fold_source_stats(e2min, 0, rmax, e2max, mmax, stat);
}

// Explicit sources that are 8^3 leaf blocks in C order (cell c = 64 i + 8 j + k of block b at
// b * 512 + c): x depends on i only, y on j only, z on k only, and eps^2 is the block's.  Reads
// the packed sources; writes the axis arrays sax, the block eps^2 and the contiguous cell masses
// sm (as k_pack_source_blocks does), and clears flag[0] (set to 1 before) on any violation.  One
// thread per source.
__global__ void k_detect_source_blocks(int64_t nblk, const double4* __restrict__ s4,
                                       const double* __restrict__ se2, double* __restrict__ sax,
                                       double* __restrict__ be2, double* __restrict__ sm, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblk * 512) return;
  const int64_t b = i >> 9, base = b * 512;
  const int c = (int)(i & 511);
  const int ci = c >> 6, cj = (c >> 3) & 7, ck = c & 7;
  const double4 s = s4[i];
  const bool ok = s.x == s4[base + ci * 64].x && s.y == s4[base + cj * 8].y &&
                  s.z == s4[base + ck].z && se2[i] == se2[base];
  if (!ok) flag[0] = 0;
  sm[i] = s.w;
  if (cj == 0 && ck == 0) sax[b * 24 + ci] = s.x;
  if (ci == 0 && ck == 0) sax[b * 24 + 8 + cj] = s.y;
  if (ci == 0 && cj == 0) sax[b * 24 + 16 + ck] = s.z;
  if (c == 0) be2[b] = se2[i];
}

}  // namespace tg
