// Device code of the direct-sum field: target loaders, the per-tile interaction loops
// (FP64 and mixed precision), the per-workgroup driver shared by the grid kernels and
// the persistent batch kernel, and the small elementwise kernels around them.
//
// Execution model (gfx950, wave64):
//   * a workgroup of NT lanes owns a target tile of K*NT targets (K per lane, held in
//     VGPRs; target t = tile0 + k*NT + lane, so every load/store is coalesced);
//   * sources stream through LDS in tiles of NT (one source staged per lane per tile);
//     every lane then reads the same LDS word (broadcast, no bank conflicts) and updates
//     its K targets;
//   * the source range may be a chunk of the full set (split over blockIdx.y or over
//     work items); chunk partial sums then go to a scratch array and a second kernel
//     sums them in fixed order, so results are deterministic (no float atomics).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

// ----------------------------------------------------------------- descriptors ----
struct Tgt {
  const double *x, *y, *z, *eps;           // explicit targets (eps may be null)
  const double *ox, *oy, *oz, *bdx, *beps; // block-structured targets
  int64_t n;
  int nb;
  int blocks;
};

struct Out {
  double *ax, *ay, *az, *phi;   // final outputs (any may be null)
  double *vx, *vy, *vz;         // fused kick v += dt * a (null = no kick)
  double dt;
  double *part;                 // non-null: write unscaled chunk sums to part[q*pstride + t]
  int64_t pstride;
  double G;
  int nopot = 0;                // 1: the potential is not wanted (its accumulator is skipped)
  const int* gate = nullptr;    // non-null: the kernel runs only if *gate == gate_on, else it
  int gate_on = 0;              //   returns at once (one of several launches picked on the device)
};

// Uniform across the grid: every workgroup of a gated-off launch returns before any barrier.
__device__ __forceinline__ bool gated_off(const Out& O) { return O.gate && *O.gate != O.gate_on; }

// Cell centre of block-structured target t: o + (idx + 1/2) dx with the multiply and the
// add rounded separately (no FMA contraction) so the centres are bit-identical to the
// host-side construction (numpy / C: o + (i + 0.5) * dx).
__device__ __forceinline__ void load_target(const Tgt& T, int64_t t, double& x, double& y,
                                            double& z, double& e2) {
  if (!T.blocks) {
    x = T.x[t];
    y = T.y[t];
    z = T.z[t];
    const double e = T.eps ? T.eps[t] : 0.0;
    e2 = e * e;
  } else {
    const int nb = T.nb;
    const int64_t nb3 = (int64_t)nb * nb * nb;
    const int64_t b = t / nb3;
    const int c = (int)(t - b * nb3);
    const int i = c / (nb * nb), j = (c / nb) % nb, k = c % nb;
    const double d = T.bdx[b];
    {
#pragma clang fp contract(off)
      x = T.ox[b] + ((double)i + 0.5) * d;
      y = T.oy[b] + ((double)j + 0.5) * d;
      z = T.oz[b] + ((double)k + 0.5) * d;
    }
    const double e = T.beps ? T.beps[b] : 0.0;
    e2 = e * e;
  }
}

// Min/max over the workgroup of NV values per lane: xor-shuffle within the wave, then
// one LDS slot per wave.  `red` holds 2*NV*(NT/64) doubles and is private to this call.
template <int NT, int NV>
__device__ __forceinline__ void block_minmax(double (&mn)[NV], double (&mx)[NV], double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      mn[v] = fmin(mn[v], __shfl_xor(mn[v], off));
      mx[v] = fmax(mx[v], __shfl_xor(mx[v], off));
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      red[(w * NV + v) * 2] = mn[v];
      red[(w * NV + v) * 2 + 1] = mx[v];
    }
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    mn[v] = red[v * 2];
    mx[v] = red[v * 2 + 1];
  }
#pragma unroll
  for (int w2 = 1; w2 < NT / 64; ++w2) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      mn[v] = fmin(mn[v], red[(w2 * NV + v) * 2]);
      mx[v] = fmax(mx[v], red[(w2 * NV + v) * 2 + 1]);
    }
  }
  __syncthreads();
}

// --------------------------------------------------------------- inner-loop flags ----
// Per-tile code path, chosen at run time once per workgroup (a uniform branch), compiled
// as separate loops:
//   F_GUARD  the tile's smallest possible eps^2 is 0: zero r2 -> the pair contributes 0
//   F_UNIF   all targets of the tile share eps_t^2: folded into the staged source eps^2
//   F_NOPOT  the caller did not ask for the potential
//   F_SEED32 (FP64 only, built with TG_FP64_SEED32 = 1) r2 >= 1e-30 is guaranteed: seed 1/sqrt
//            from v_rsq_f32 instead of the quarter-rate v_rsq_f64 (both seeds are ~2^-24 accurate;
//            the refinement is the same)
enum { F_GUARD = 1, F_UNIF = 2, F_NOPOT = 4, F_SEED32 = 8 };

// --------------------------------------------------------------- FP64 tile loop ----
// 1/sqrt(r2) in FP64: a hardware estimate refined by one third-order Newton/Householder
// step  y <- y + y e (1/2 + 3/8 e),  e = 1 - r2 y^2  (the Taylor series of (1-e)^(-1/2);
// an estimate with relative error d leaves ~2.5 d^3, below FP64 rounding for d ~ 2^-24).
template <int FL>
__device__ __forceinline__ double rsqrt_f64(double r2) {
  double y;
  if (FL & F_SEED32) y = (double)__builtin_amdgcn_rsqf((float)r2);
  else y = __builtin_amdgcn_rsq(r2);
  if (FL & F_GUARD) y = r2 > 0.0 ? y : 0.0;   // |d|^2 + eps^2 == 0: the pair contributes 0
  const double t = r2 * y;
  const double e = fma(-t, y, 1.0);
  const double p = fma(e, 0.375, 0.5);
  const double q = y * e;
  return fma(q, p, y);
}

template <int K, int NT, int FL>
__device__ __forceinline__ void fp64_tile(const double4* __restrict__ s4,
                                          const double* __restrict__ se2,
                                          const double (&tx)[K], const double (&ty)[K],
                                          const double (&tz)[K], const double (&te2)[K],
                                          double (&ax)[K], double (&ay)[K], double (&az)[K],
                                          double (&ap)[K]) {
#pragma unroll 2
  for (int j = 0; j < NT; ++j) {
    const double4 s = s4[j];
    const double e2 = se2[j];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double dx = s.x - tx[k], dy = s.y - ty[k], dz = s.z - tz[k];
      const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, (FL & F_UNIF) ? e2 : e2 + te2[k])));
      const double y = rsqrt_f64<FL>(r2);
      const double mr = s.w * y;
      if (!(FL & F_NOPOT)) ap[k] += mr;
      const double w = mr * (y * y);
      ax[k] = fma(w, dx, ax[k]);
      ay[k] = fma(w, dy, ay[k]);
      az[k] = fma(w, dz, az[k]);
    }
  }
}

// ------------------------------------------------------ mixed-precision tile loop ----
template <int K, int NT, int FL>
__device__ __forceinline__ void f32_tile(const float4* __restrict__ s4,
                                         const float* __restrict__ se2,
                                         const float (&tx)[K], const float (&ty)[K],
                                         const float (&tz)[K], const float (&te2)[K],
                                         float (&ax)[K], float (&ay)[K], float (&az)[K],
                                         float (&ap)[K]) {
#pragma unroll 4
  for (int j = 0; j < NT; ++j) {
    const float4 s = s4[j];
    const float e2 = se2[j];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float dx = s.x - tx[k], dy = s.y - ty[k], dz = s.z - tz[k];
      const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, (FL & F_UNIF) ? e2 : e2 + te2[k])));
      float y = __builtin_amdgcn_rsqf(r2);
      if (FL & F_GUARD) y = r2 > 0.0f ? y : 0.0f;
      const float mr = s.w * y;
      if (!(FL & F_NOPOT)) ap[k] += mr;
      const float w = (mr * y) * y;
      ax[k] = fmaf(w, dx, ax[k]);
      ay[k] = fmaf(w, dy, ay[k]);
      az[k] = fmaf(w, dz, az[k]);
    }
  }
}

// Knuth two-sum: (hi, lo) += v with the rounding error of hi + v kept in lo.
__device__ __forceinline__ void two_sum(float& hi, float& lo, float v) {
  const float s = hi + v;
  const float bp = s - hi;
  const float err = (hi - (s - bp)) + (v - bp);
  hi = s;
  lo += err;
}

// ------------------------------------------------------------------ LDS layout ----
template <int NT>
struct LdsF64 {
  double4 s4[NT];
  double se2[NT];
  double red[2 * 4 * (NT / 64)];
};
template <int NT>
struct LdsF32 {
  float4 s4[NT];
  float se2[NT];
  double red[2 * 4 * (NT / 64)];
};

// Threshold on the smallest possible eps^2 of a tile below which the FP32 loop guards
// r2 == 0 (and keeps rsq far from overflow) and the FP64 loop keeps the v_rsq_f64 seed.
#define TG_F32_TINY_E2 1e-30
// Bound on a tile's largest r^2 above which the FP64 loop keeps the v_rsq_f64 seed: the FP32
// seed converts r^2 to float, and FLT_MAX is 3.4e38 (a separation of 1.8e19 in the caller's
// units; cgs coordinates of a few kpc reach it).
#define TG_F32_HUGE_R2 1e36
// 1: the FP64 kernels seed 1/sqrt from v_rsq_f32 of the converted r2 where the tile's range allows
// (F_SEED32, P.s32); 0: from v_rsq_f64 everywhere.  Measured on MI355X with the block kernels'
// instruction mixes, the v_rsq_f64 seed (one quarter-rate transcendental, which the FP64 arithmetic
// around it hides) is as fast or faster than the two conversions plus v_rsq_f32: c4 FP64 column
// kernel 25.25 against 27.13 ms on one box and 26.55 against 26.69 on another, c5 forward 200.0
// against 202.2 ms, c2 5.20 / 5.03 against 5.29 / 5.08 ms; the FLASH kernels within 1% either way.
#ifndef TG_FP64_SEED32
#define TG_FP64_SEED32 0
#endif

// Statistics of a packed source set, device words written by the pack kernels: doubles stored as
// their bits (non-negative, so integer atomicMin / atomicMax order them like the values).
enum {
  ST_E2MIN = 0,   // smallest eps^2 (read by every kernel)
  ST_BAD = 1,     // non-zero: a NaN/Inf input (validation)
  ST_RMAX = 2,    // largest |x|, |y|, |z|
  ST_E2MAX = 3,   // largest eps^2
  ST_MMAX = 4,    // largest |m|
  ST_WORDS = 5
};

// Upper bound on r^2 = |x_s - x_t|^2 + eps_s^2 + eps_t^2 over the tile's pairs, from the source
// statistics `sst` and the tile's largest |coordinate| `rt` and eps_t^2; `reach` widens the
// displacement (periodic images).
__device__ __forceinline__ double r2_bound(const double* sst, double rt, double e2t_max,
                                           double reach = 0.0) {
  const double d = sst[ST_RMAX] + rt + reach;
  return 3.0 * d * d + sst[ST_E2MAX] + e2t_max;
}

__device__ __forceinline__ double max_abs3(double x, double y, double z) {
  return fmax(fabs(x), fmax(fabs(y), fabs(z)));
}

// v + dt*a with the product and the sum rounded separately (no FMA contraction), as the
// reference's FP64 kick loops compute it (interface.F90:1793-1795).
__device__ __forceinline__ void kick_one(const Out& O, int64_t t, double gx, double gy, double gz) {
#pragma clang fp contract(off)
  O.vx[t] = O.vx[t] + O.dt * gx;
  O.vy[t] = O.vy[t] + O.dt * gy;
  O.vz[t] = O.vz[t] + O.dt * gz;
}

__device__ __forceinline__ void store_result(const Out& O, int64_t t, double sx, double sy,
                                             double sz, double sp) {
  if (O.part) {
    O.part[t] = sx;
    O.part[O.pstride + t] = sy;
    O.part[2 * O.pstride + t] = sz;
    O.part[3 * O.pstride + t] = sp;
    return;
  }
  const double gx = O.G * sx, gy = O.G * sy, gz = O.G * sz;
  if (O.ax) O.ax[t] = gx;
  if (O.ay) O.ay[t] = gy;
  if (O.az) O.az[t] = gz;
  if (O.phi) O.phi[t] = -O.G * sp;
  if (O.vx) kick_one(O, t, gx, gy, gz);
}

// Code path of a tile from its smallest/largest target eps^2, the sources' smallest and (FP64)
// a bound on the tile's largest r^2: the FP32 seed needs r^2 inside [1e-30, 1e36].
__device__ __forceinline__ int tile_flags(double e2s_min, double e2t_min, double e2t_max,
                                          const Out& O, bool f32, double r2max = 0.0) {
  int fl = 0;
  if (e2t_min == e2t_max) fl |= F_UNIF;
  const double lo = e2s_min + e2t_min;
  if (f32) {
    if (!(lo > TG_F32_TINY_E2)) fl |= F_GUARD;
  } else {
    if (!(lo > 0.0)) fl |= F_GUARD;
    else if (TG_FP64_SEED32 && lo > TG_F32_TINY_E2 && r2max < TG_F32_HUGE_R2) fl |= F_SEED32;
  }
  if (O.nopot) fl |= F_NOPOT;
  return fl;
}

// ------------------------------------------------------- per-workgroup drivers ----
template <int K, int NT, int FL>
__device__ __forceinline__ void fp64_sources(LdsF64<NT>& L, const double4* __restrict__ src4,
                                          const double* __restrict__ srce2, int64_t s_begin,
                                          int64_t s_end, double e2u, const double (&tx)[K],
                                          const double (&ty)[K], const double (&tz)[K],
                                          const double (&te2)[K], double (&ax)[K],
                                          double (&ay)[K], double (&az)[K], double (&ap)[K]) {
  const int tid = threadIdx.x;
  for (int64_t base = s_begin; base < s_end; base += NT) {
    __syncthreads();
    const int64_t j = base + tid;
    if (j < s_end) {
      L.s4[tid] = src4[j];
      L.se2[tid] = srce2[j] + e2u;
    } else {   // padding source: zero mass, eps^2 = 1 -> contributes exactly 0
      L.s4[tid] = make_double4(0.0, 0.0, 0.0, 0.0);
      L.se2[tid] = 1.0;
    }
    __syncthreads();
    fp64_tile<K, NT, FL>(L.s4, L.se2, tx, ty, tz, te2, ax, ay, az, ap);
  }
}

#define TG_DISPATCH8(FN, fl, ...)                                             \
  switch (fl & ~F_SEED32) {                                                   \
    case 0: FN<0>(__VA_ARGS__); break;                                        \
    case F_GUARD: FN<F_GUARD>(__VA_ARGS__); break;                            \
    case F_UNIF: FN<F_UNIF>(__VA_ARGS__); break;                              \
    case F_UNIF | F_GUARD: FN<F_UNIF | F_GUARD>(__VA_ARGS__); break;          \
    case F_NOPOT: FN<F_NOPOT>(__VA_ARGS__); break;                            \
    case F_NOPOT | F_GUARD: FN<F_NOPOT | F_GUARD>(__VA_ARGS__); break;        \
    case F_NOPOT | F_UNIF: FN<F_NOPOT | F_UNIF>(__VA_ARGS__); break;          \
    default: FN<F_NOPOT | F_UNIF | F_GUARD>(__VA_ARGS__); break;              \
  }

#define TG_DISPATCH(FN, fl, ...)                                              \
  switch (fl) {                                                               \
    case 0: FN<0>(__VA_ARGS__); break;                                        \
    case F_GUARD: FN<F_GUARD>(__VA_ARGS__); break;                            \
    case F_UNIF: FN<F_UNIF>(__VA_ARGS__); break;                              \
    case F_UNIF | F_GUARD: FN<F_UNIF | F_GUARD>(__VA_ARGS__); break;          \
    case F_NOPOT: FN<F_NOPOT>(__VA_ARGS__); break;                            \
    case F_NOPOT | F_GUARD: FN<F_NOPOT | F_GUARD>(__VA_ARGS__); break;        \
    case F_NOPOT | F_UNIF: FN<F_NOPOT | F_UNIF>(__VA_ARGS__); break;          \
    case F_NOPOT | F_UNIF | F_GUARD: FN<F_NOPOT | F_UNIF | F_GUARD>(__VA_ARGS__); break; \
    case F_SEED32: FN<F_SEED32>(__VA_ARGS__); break;                          \
    case F_SEED32 | F_UNIF: FN<F_SEED32 | F_UNIF>(__VA_ARGS__); break;        \
    case F_SEED32 | F_NOPOT: FN<F_SEED32 | F_NOPOT>(__VA_ARGS__); break;      \
    default: FN<F_SEED32 | F_NOPOT | F_UNIF>(__VA_ARGS__); break;             \
  }

template <int K, int NT>
struct Fp64Run {
  template <int FL>
  static __device__ __forceinline__ void run(LdsF64<NT>& L, const double4* src4, const double* srce2,
                                             int64_t sb, int64_t se, double e2u,
                                             const double (&tx)[K], const double (&ty)[K],
                                             const double (&tz)[K], const double (&te2)[K],
                                             double (&ax)[K], double (&ay)[K], double (&az)[K],
                                             double (&ap)[K]) {
    fp64_sources<K, NT, FL>(L, src4, srce2, sb, se, (FL & F_UNIF) ? e2u : 0.0, tx, ty, tz, te2,
                            ax, ay, az, ap);
  }
};

// sst: the source statistics (ST_*).
template <int K, int NT>
__device__ void process_fp64(LdsF64<NT>& L, const Tgt& T, int64_t tile0,
                             const double4* __restrict__ src4, const double* __restrict__ srce2,
                             int64_t s_begin, int64_t s_end, const double* sst, const Out& O) {
  const int tid = threadIdx.x;
  double tx[K], ty[K], tz[K], te2[K];
  double mn[2] = {INFINITY, INFINITY}, mx[2] = {-INFINITY, 0.0};   // eps^2; |coordinate| (max only)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    load_target(T, t < T.n ? t : T.n - 1, tx[k], ty[k], tz[k], te2[k]);
    if (t < T.n) {
      mn[0] = fmin(mn[0], te2[k]);
      mx[0] = fmax(mx[0], te2[k]);
      mx[1] = fmax(mx[1], max_abs3(tx[k], ty[k], tz[k]));
    }
  }
  block_minmax<NT, 2>(mn, mx, L.red);
  const int fl = tile_flags(sst[ST_E2MIN], mn[0], mx[0], O, false, r2_bound(sst, mx[1], mx[0]));
  double ax[K], ay[K], az[K], ap[K];
#pragma unroll
  for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = ap[k] = 0.0;
  typedef Fp64Run<K, NT> R;
  TG_DISPATCH(R::template run, fl, L, src4, srce2, s_begin, s_end, mn[0], tx, ty, tz,
              te2, ax, ay, az, ap)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    if (t < T.n) store_result(O, t, ax[k], ay[k], az[k], ap[k]);
  }
}

// Mixed precision.  The tile's FP64 anchor c is the centre of its targets' bounding box.
// Sources are staged as fp32(x_s - c) (the FP64 difference rounded once), targets are held
// as fp32(x_t - c); the FP32 difference of the two is exact whenever they are within a
// factor of two of each other (Sterbenz), i.e. for exactly the close pairs where a
// difference of raw FP32 coordinates would cancel catastrophically.  The FP32 partial sums
// of each LDS tile (NT sources) are folded into per-target two-sum (hi, lo) accumulators.
template <int K, int NT, int FL>
__device__ __forceinline__ void f32_sources(LdsF32<NT>& L, const double4* __restrict__ src4,
                                         const double* __restrict__ srce2, int64_t s_begin,
                                         int64_t s_end, double e2u, double cx, double cy,
                                         double cz, const float (&tx)[K], const float (&ty)[K],
                                         const float (&tz)[K], const float (&te2)[K],
                                         float (&hx)[K], float (&hy)[K], float (&hz)[K],
                                         float (&hp)[K], float (&lx)[K], float (&ly)[K],
                                         float (&lz)[K], float (&lp)[K]) {
  const int tid = threadIdx.x;
  for (int64_t base = s_begin; base < s_end; base += NT) {
    __syncthreads();
    const int64_t j = base + tid;
    if (j < s_end) {
      const double4 s = src4[j];
      L.s4[tid] = make_float4((float)(s.x - cx), (float)(s.y - cy), (float)(s.z - cz), (float)s.w);
      L.se2[tid] = (float)(srce2[j] + e2u);
    } else {
      L.s4[tid] = make_float4(0.f, 0.f, 0.f, 0.f);
      L.se2[tid] = 1.0f;
    }
    __syncthreads();
    float ax[K], ay[K], az[K], ap[K];
#pragma unroll
    for (int k = 0; k < K; ++k) ax[k] = ay[k] = az[k] = ap[k] = 0.f;
    f32_tile<K, NT, FL>(L.s4, L.se2, tx, ty, tz, te2, ax, ay, az, ap);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      two_sum(hx[k], lx[k], ax[k]);
      two_sum(hy[k], ly[k], ay[k]);
      two_sum(hz[k], lz[k], az[k]);
      if (!(FL & F_NOPOT)) two_sum(hp[k], lp[k], ap[k]);
    }
  }
}

template <int K, int NT>
struct MixRun {
  template <int FL>
  static __device__ __forceinline__ void run(LdsF32<NT>& L, const double4* src4, const double* srce2,
                                             int64_t sb, int64_t se, double e2u, double cx,
                                             double cy, double cz, const float (&tx)[K],
                                             const float (&ty)[K], const float (&tz)[K],
                                             const float (&te2)[K], float (&hx)[K], float (&hy)[K],
                                             float (&hz)[K], float (&hp)[K], float (&lx)[K],
                                             float (&ly)[K], float (&lz)[K], float (&lp)[K]) {
    f32_sources<K, NT, FL>(L, src4, srce2, sb, se, (FL & F_UNIF) ? e2u : 0.0, cx, cy,
                                       cz, tx, ty, tz, te2, hx, hy, hz, hp, lx, ly, lz, lp);
  }
};

template <int K, int NT>
__device__ void process_mixed(LdsF32<NT>& L, const Tgt& T, int64_t tile0,
                              const double4* __restrict__ src4, const double* __restrict__ srce2,
                              int64_t s_begin, int64_t s_end, const double* sst, const Out& O) {
  const int tid = threadIdx.x;
  double px[K], py[K], pz[K], pe2[K];
  double mn[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  double mx[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    load_target(T, t < T.n ? t : T.n - 1, px[k], py[k], pz[k], pe2[k]);
    if (t < T.n) {
      mn[0] = fmin(mn[0], px[k]); mx[0] = fmax(mx[0], px[k]);
      mn[1] = fmin(mn[1], py[k]); mx[1] = fmax(mx[1], py[k]);
      mn[2] = fmin(mn[2], pz[k]); mx[2] = fmax(mx[2], pz[k]);
      mn[3] = fmin(mn[3], pe2[k]); mx[3] = fmax(mx[3], pe2[k]);
    }
  }
  block_minmax<NT, 4>(mn, mx, L.red);
  const double cx = 0.5 * (mn[0] + mx[0]), cy = 0.5 * (mn[1] + mx[1]), cz = 0.5 * (mn[2] + mx[2]);
  const int fl = tile_flags(sst[ST_E2MIN], mn[3], mx[3], O, true);
  float tx[K], ty[K], tz[K], te2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    tx[k] = (float)(px[k] - cx);
    ty[k] = (float)(py[k] - cy);
    tz[k] = (float)(pz[k] - cz);
    te2[k] = (float)pe2[k];
  }
  float hx[K], hy[K], hz[K], hp[K], lx[K], ly[K], lz[K], lp[K];
#pragma unroll
  for (int k = 0; k < K; ++k) hx[k] = hy[k] = hz[k] = hp[k] = lx[k] = ly[k] = lz[k] = lp[k] = 0.f;
  typedef MixRun<K, NT> R;
  TG_DISPATCH8(R::template run, fl, L, src4, srce2, s_begin, s_end, mn[3], cx, cy, cz, tx,
              ty, tz, te2, hx, hy, hz, hp, lx, ly, lz, lp)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    if (t < T.n)
      store_result(O, t, (double)hx[k] + (double)lx[k], (double)hy[k] + (double)ly[k],
                   (double)hz[k] + (double)lz[k], (double)hp[k] + (double)lp[k]);
  }
}

// ------------------------------------------------------------------ grid kernels ----
// blockIdx.x = target tile, blockIdx.y = source chunk [c*chunk, min((c+1)*chunk, ns)).
template <int K, int NT>
__global__ __launch_bounds__(NT) void k_eval_fp64(Tgt T, const double4* __restrict__ src4,
                                                  const double* __restrict__ srce2, int64_t ns,
                                                  int64_t chunk, const double* sst, Out O) {
  __shared__ LdsF64<NT> L;
  if (gated_off(O)) return;
  const int64_t tile0 = (int64_t)blockIdx.x * (K * NT);
  const int64_t sb = (int64_t)blockIdx.y * chunk;
  const int64_t se = sb + chunk < ns ? sb + chunk : ns;
  Out o = O;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
  process_fp64<K, NT>(L, T, tile0, src4, srce2, sb, se, sst, o);
}

template <int K, int NT>
__global__ __launch_bounds__(NT) void k_eval_mixed(Tgt T, const double4* __restrict__ src4,
                                                   const double* __restrict__ srce2, int64_t ns,
                                                   int64_t chunk, const double* sst, Out O) {
  __shared__ LdsF32<NT> L;
  if (gated_off(O)) return;
  const int64_t tile0 = (int64_t)blockIdx.x * (K * NT);
  const int64_t sb = (int64_t)blockIdx.y * chunk;
  const int64_t se = sb + chunk < ns ? sb + chunk : ns;
  Out o = O;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
  process_mixed<K, NT>(L, T, tile0, src4, srce2, sb, se, sst, o);
}

// Sum of the nsplit chunk partials of each target in chunk order, then the G scaling,
// the stores and the fused kick.
__global__ void k_reduce_chunks(const double* __restrict__ part, int64_t pstride, int nsplit,
                                int64_t n, Out O) {
  if (gated_off(O)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  for (int c = 0; c < nsplit; ++c) {
    const double* p = part + (int64_t)c * 4 * pstride;
    s0 += p[t];
    s1 += p[pstride + t];
    s2 += p[2 * pstride + t];
    s3 += p[3 * pstride + t];
  }
  Out o = O;
  o.part = nullptr;
  store_result(o, t, s0, s1, s2, s3);
}

// The same reduction for few targets and many chunks (k_reduce_chunks would run a handful of
// workgroups, each thread walking every chunk): a workgroup of 16 waves takes 64 targets, wave
// w sums chunks [w*per, (w+1)*per) in order, and wave 0 adds the 16 wave sums in wave order,
// so the result does not depend on scheduling.  Grid: ceil(n / 64) workgroups of 1024 lanes.
__global__ __launch_bounds__(1024) void k_reduce_chunks_wide(const double* __restrict__ part,
                                                             int64_t pstride, int nsplit, int64_t n,
                                                             Out O) {
  __shared__ double red[16][4][64];
  if (gated_off(O)) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t t = (int64_t)blockIdx.x * 64 + lane;
  const int per = (nsplit + 15) / 16;
  const int c0 = w * per, c1 = min(nsplit, c0 + per);
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  if (t < n) {
    for (int c = c0; c < c1; ++c) {
      const double* p = part + (int64_t)c * 4 * pstride;
      s0 += p[t];
      s1 += p[pstride + t];
      s2 += p[2 * pstride + t];
      s3 += p[3 * pstride + t];
    }
  }
  red[w][0][lane] = s0;
  red[w][1][lane] = s1;
  red[w][2][lane] = s2;
  red[w][3][lane] = s3;
  __syncthreads();
  if (w == 0 && t < n) {
    double a0 = red[0][0][lane], a1 = red[0][1][lane], a2 = red[0][2][lane], a3 = red[0][3][lane];
    for (int q = 1; q < 16; ++q) {
      a0 += red[q][0][lane];
      a1 += red[q][1][lane];
      a2 += red[q][2][lane];
      a3 += red[q][3][lane];
    }
    Out o = O;
    o.part = nullptr;
    store_result(o, t, a0, a1, a2, a3);
  }
}

// Bits of a non-negative double: their unsigned-integer order is the numeric order.
__device__ __forceinline__ unsigned long long dbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int off) {
  return (unsigned long long)__shfl_xor((long long)v, off);
}

// Folds one lane's source statistics (ST_* words) into `stat`: reduced within each wave by
// shuffles, across the workgroup's (at most 4) waves through LDS, then at most one atomic per
// word.  Every lane of the workgroup (<= 256 lanes) must call it.
__device__ __forceinline__ void fold_source_stats(unsigned long long e2min, int bad, unsigned long long rmax,
                                                  unsigned long long e2max, unsigned long long mmax,
                                                  unsigned long long* stat) {
  __shared__ unsigned long long sm[4][4];
  __shared__ int sbad[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e2min = min(e2min, shfl_xor_u64(e2min, off));
    rmax = max(rmax, shfl_xor_u64(rmax, off));
    e2max = max(e2max, shfl_xor_u64(e2max, off));
    mmax = max(mmax, shfl_xor_u64(mmax, off));
    bad |= __shfl_xor(bad, off);
  }
  const int t = threadIdx.x, w = t >> 6, nw = (blockDim.x + 63) >> 6;
  if ((t & 63) == 0) {
    sm[0][w] = e2min;
    sm[1][w] = rmax;
    sm[2][w] = e2max;
    sm[3][w] = mmax;
    sbad[w] = bad;
  }
  __syncthreads();
  if (t == 0) {
    for (int q = 1; q < nw; ++q) {
      sm[0][0] = min(sm[0][0], sm[0][q]);
      sm[1][0] = max(sm[1][0], sm[1][q]);
      sm[2][0] = max(sm[2][0], sm[2][q]);
      sm[3][0] = max(sm[3][0], sm[3][q]);
      sbad[0] |= sbad[q];
    }
  }
  __syncthreads();
  // One lane per word; the atomic only where the workgroup's value improves on what the word
  // already holds (a plain read, possibly stale, which only costs an atomic): with thousands of
  // workgroups on the same five words, most then skip their atomics instead of queueing on them.
  if (t < 5) {
    volatile unsigned long long* w = stat + t;
    if (t == ST_E2MIN) {
      if (sm[0][0] < *w) atomicMin(&stat[ST_E2MIN], sm[0][0]);
    } else if (t == ST_RMAX) {
      if (sm[1][0] > *w) atomicMax(&stat[ST_RMAX], sm[1][0]);
    } else if (t == ST_E2MAX) {
      if (sm[2][0] > *w) atomicMax(&stat[ST_E2MAX], sm[2][0]);
    } else if (t == ST_MMAX) {
      if (sm[3][0] > *w) atomicMax(&stat[ST_MMAX], sm[3][0]);
    } else if (sbad[0]) {
      atomicOr(&stat[ST_BAD], 1ULL);
    }
  }
}

// Pack SoA FP64 sources into {x,y,z,m} + eps^2 and fold their statistics into stat (ST_*:
// smallest and largest eps^2, largest |coordinate| and |m|).  With `check`, stat[ST_BAD]
// becomes non-zero if any input is NaN/Inf.  stat must be reset first (k_reset_stats).
// Workgroups of 256 lanes.
__global__ __launch_bounds__(256) void k_pack_sources(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      const double* __restrict__ m,
                                                      const double* __restrict__ eps,
                                                      double4* __restrict__ s4,
                                                      double* __restrict__ se2,
                                                      unsigned long long* stat, int check) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long e2min = 0x7ff0000000000000ULL, rmax = 0ULL, e2max = 0ULL, mmax = 0ULL;
  int bad = 0;
  if (i < n) {
    const double4 v = make_double4(x[i], y[i], z[i], m[i]);
    const double e = eps ? eps[i] : 0.0;
    s4[i] = v;
    const double e2 = e * e;
    se2[i] = e2;
    e2min = e2max = dbits(e2);
    rmax = dbits(max_abs3(v.x, v.y, v.z));
    mmax = dbits(fabs(v.w));
    if (check)
      bad = !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w) && isfinite(e));
  }
  fold_source_stats(e2min, bad, rmax, e2max, mmax, stat);
}

// Range statistics (ST_* words: largest |coordinate|, smallest / largest eps^2, largest |m|) of a
// point set (x, y, z, eps per point; dx == null) or of a set of 8^3 blocks (x, y, z the lower
// corners and dx the cell widths, so the centres lie within [x, x + 8 dx]; eps per block), with nm
// masses m (may be null).  For the mixed mode's range check.  out is reset first (k_reset_stats);
// workgroups of 256 lanes over max(n, nm).
__global__ __launch_bounds__(256) void k_range(int64_t n, const double* __restrict__ x,
                                               const double* __restrict__ y, const double* __restrict__ z,
                                               const double* __restrict__ dx, const double* __restrict__ e,
                                               int64_t nm, const double* __restrict__ m,
                                               unsigned long long* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long e2min = 0x7ff0000000000000ULL, rmax = 0ULL, e2max = 0ULL, mmax = 0ULL;
  if (i < n) {
    double r = max_abs3(x[i], y[i], z[i]);
    if (dx) r = fmax(r, max_abs3(x[i] + 8.0 * dx[i], y[i] + 8.0 * dx[i], z[i] + 8.0 * dx[i]));
    const double ei = e ? e[i] : 0.0;
    rmax = dbits(r);
    e2min = e2max = dbits(ei * ei);
  }
  if (m && i < nm) mmax = dbits(fabs(m[i]));
  fold_source_stats(e2min, 0, rmax, e2max, mmax, out);
}

__global__ void k_reset_stats(unsigned long long* stat) {
  if (threadIdx.x == 0) {
    stat[ST_E2MIN] = 0x7ff0000000000000ULL;
    stat[ST_BAD] = 0ULL;
    stat[ST_RMAX] = 0ULL;
    stat[ST_E2MAX] = 0ULL;
    stat[ST_MMAX] = 0ULL;
  }
}

// NaN/Inf check of up to five target arrays (any may be null) into flag[0].
__global__ void k_check_finite(int64_t n, const double* c0, const double* c1, const double* c2,
                               const double* c3, const double* c4, unsigned long long* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* cs[5] = {c0, c1, c2, c3, c4};
  int bad = 0;
#pragma unroll
  for (int q = 0; q < 5; ++q)
    if (cs[q] && !isfinite(cs[q][i])) bad = 1;
  if (bad) atomicOr(flag, 1ULL);
}

__global__ void k_kick(int64_t n, double dt, const double* __restrict__ ax,
                       const double* __restrict__ ay, const double* __restrict__ az,
                       double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  {
#pragma clang fp contract(off)
    vx[i] = vx[i] + dt * ax[i];
    vy[i] = vy[i] + dt * ay[i];
    vz[i] = vz[i] + dt * az[i];
  }
}

// flag[0] = v (one lane; a kernel rather than a memset, so a captured graph replays it).
__global__ void k_set_flag(int* flag, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = v;
}

// Gate value of an explicit-target evaluation: 2 if the sources are explicit 8^3 blocks
// (*sflag, set by k_detect_source_blocks; sflag null: they are not), else `tcand` (1: the targets
// may have column structure, which k_check_cols then tests; 0: they do not).
__global__ void k_eval_gate(int* gate, const int* sflag, int tcand) {
  if (blockIdx.x == 0 && threadIdx.x == 0) gate[0] = (sflag && sflag[0] == 1) ? 2 : tcand;
}

// Column structure of explicit targets, the C order of 8^3 leaf blocks
// (.deprecated/bridge_multiples.py:566-575 flattens each block's meshgrid this way): every
// aligned group of 8 consecutive targets shares x and y, and every aligned group of 512 shares
// eps.  One lane per target, compared with the first of its group (coalesced reads); any
// violation clears flag[0] (1 before; 2, the block-source path, is left alone); all writers store
// the same value.
__device__ __forceinline__ bool column_ok(const Tgt& T, int64_t t) {
  const int64_t g = t & ~(int64_t)7;
  bool ok = T.x[t] == T.x[g] && T.y[t] == T.y[g];
  if (T.eps) ok = ok && T.eps[t] == T.eps[t & ~(int64_t)511];
  return ok;
}

__global__ void k_check_cols(Tgt T, int* flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T.n || flag[0] != 1) return;   // 2: the block-source kernel runs; nothing to test
  if (!column_ok(T, t)) flag[0] = 0;
}

__global__ void k_fill(int64_t n, double v, double* a, double* b, double* c, double* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (a) a[i] = v;
  if (b) b[i] = v;
  if (c) c[i] = v;
  if (d) d[i] = v;
}

}  // namespace tg
