// FLASH's own gas-gravity kernels, FP64: the spline / linear softening inside a global
// radius (Gravity_getAccelAtPoint.F90:140-145, 206-241; Gravity_getPotentialAtPoint.F90:
// 190-222) and the periodic image sum (Gravity_getAccelAtPoint.F90:250-312,
// Gravity_getPotentialAtPoint.F90:229-283).
//
// Execution model: as the Plummer FP64 kernel (k_eval_fp64) -- K targets per lane in VGPRs,
// sources {x, y, z, m} staged through LDS NT at a time and read as broadcasts, source
// chunks over blockIdx.y with fixed-order chunk reduction.  Two things differ:
//   * branch decisions follow the reference bit for bit: it tests sqrt(r2) < R with
//     r2 = (dx*dx + dy*dy) + dz*dz, every product and sum rounded separately.  The
//     correctly rounded sqrt is monotone, so that test equals r2 <= T(R) for the largest
//     double T(R) whose sqrt rounds below R (computed on the host); the kernels form r2 the
//     reference's way wherever it decides a branch and compare it with T.  The kernels are
//     discontinuous there (the image cutoff; the reference's potential inside rsoft), so a
//     decision that differed from the reference's would change a result by a whole term;
//   * Newtonian pairs (r >= rsoft, nearly all of them) use the refined 1/sqrt of the
//     Plummer kernel: kv = y^3, pk = y; the softened branch recomputes both from r.
//     Isolated boundaries: the inner loop is branch-free -- every pair is summed as
//     Newtonian with r2 clamped to at least rsoft^2 (1 + 1e-12), and pairs below that clamp
//     are flagged (one wave-uniform bit per source of a 64-source slice); for each flagged
//     source the wave then takes that Newtonian term back out of the flagged pairs and adds
//     the exact one (flash_fix).
#pragma once
#include "kernels.hpp"
#include "grid.hpp"

namespace tg {

struct Flash {
  double rsoft, hinv, h2inv, slope;
  double soft_thr;                // r2 <= soft_thr  <=>  fl(sqrt(r2)) < rsoft
  double near2;                   // soft_thr (1 + 1e-12): FMA-formed r2 below it may be softened
  double lx, ly, lz, ilx, ily, ilz;  // periodic box and reciprocals
  double cut_thr;                 // r2 <= cut_thr  <=>  fl(sqrt(r2)) < nrep * min(L)
  int type;                       // 1 spline, 2 linear
  int quirks;                     // potential as the reference writes it
  int nrep;                       // periodic: images n in [-nrep, nrep]^3
  int s32;                        // rsoft^2 > 1e-30: a tile may seed 1/sqrt in FP32 (r2_bound decides)
  double reach;                   // periodic: (nrep + 1) max(L), the farthest image offset; else 0
};

// Largest |coordinate| of the workgroup's targets (NT lanes, K each; t >= n excluded), through
// `red` (2 NT / 64 doubles of LDS).
template <int K, int NT>
__device__ __forceinline__ double tile_rmax(const double (&tx)[K], const double (&ty)[K],
                                            const double (&tz)[K], int64_t tile0, int64_t n,
                                            double* red) {
  double mn[1] = {0.0}, mx[1] = {0.0};
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (tile0 + (int64_t)k * NT + threadIdx.x < n) mx[0] = fmax(mx[0], max_abs3(tx[k], ty[k], tz[k]));
  block_minmax<NT, 1>(mn, mx, red);
  return mx[0];
}

// Kernel values for r < rsoft.  Force: the reference's spline pieces and linear slope;
// q <= 1e-5, where the reference leaves the force kernel unassigned, takes the r -> 0 limit
// h^-3 (4/3 - 6/5 q^2 + 1/2 q^3).  Potential: with quirks the reference's expressions
// (spline pot_kernel with its -q^4/10 term, `pgpot = slope`), otherwise the M4-spline
// potential and that of the linear force, both continuous with 1/r at rsoft.
__device__ __forceinline__ void flash_soft(const Flash& P, double r, double& kv, double& pk) {
  if (P.type == 1) {
    const double q = r * P.hinv, q2 = q * q, q3 = q2 * q, q4 = q2 * q2, q5 = q4 * q;
    if (q < 1.0) {
      kv = q > 1.0e-5 ? P.h2inv * (4.0 / 3.0 * q - 1.2 * q3 + 0.5 * q4) / r
                      : P.h2inv * P.hinv * (4.0 / 3.0 - 1.2 * q2 + 0.5 * q3);
      pk = P.quirks ? P.hinv * (2.0 / 3.0 * q2 - 0.1 * q4 + 0.1 * q5 - 1.4)
                    : -P.hinv * (2.0 / 3.0 * q2 - 0.3 * q4 + 0.1 * q5 - 1.4);
    } else {
      kv = P.h2inv * (8.0 / 3.0 * q - 3.0 * q2 + 1.2 * q3 - q4 / 6.0 - 1.0 / (15.0 * q2)) / r;
      const double p =
          P.hinv * (4.0 / 3.0 * q2 - q3 + 0.3 * q4 - q5 / 30.0 - 1.6 + 1.0 / (15.0 * q));
      pk = P.quirks ? p : -p;
    }
  } else {
    kv = P.slope;
    pk = P.quirks ? P.slope : (3.0 * P.rsoft * P.rsoft - r * r) * (0.5 * P.slope);
  }
}

// Force kernel kv (the pair adds -m kv d, d = x_t - x_s) and potential kernel pk of one pair.
template <bool S32>
__device__ __forceinline__ void flash_pair(const Flash& P, double r2, double& kv, double& pk) {
  const double y = rsqrt_f64<S32 ? F_SEED32 : 0>(r2);
  kv = (y * y) * y;
  pk = y;
  if (r2 <= P.soft_thr) flash_soft(P, sqrt(r2), kv, pk);   // inside rsoft (divergent, rare)
}

// Range [lo, hi] of image indices n in [-N, N] that can satisfy (d + n L)^2 <= rem, i.e. the
// integers in [(-s - d) / L, (s - d) / L] with s = sqrt(rem).  A superset: the exact test is
// applied per image, so the bounds only need to miss nothing (float sqrt: relative error
// ~2^-22, so 1e-4 of slack in units of n covers it for N <= 64).  The integers inside the
// widened interval are ceil of its lower end to floor of its upper end; rounding the ends
// outwards instead would add two images per row that the exact test then rejects (a quarter
// more iterations at N = 6).
__device__ __forceinline__ void image_range(double d, double iL, double rem, int N, int& lo,
                                            int& hi) {
  if (!(rem > 0.0)) {
    lo = 1;
    hi = 0;
    return;
  }
  // raw v_sqrt_f32 (~1 ulp); rem above 1e30 is scaled by 2^-100 first and s by 2^50 after (both
  // exact), so the FP32 range holds rem up to ~4e68 (cgs boxes of kpc: rem ~ 1e44)
  const bool big = rem > 1e30;
  double s = (double)__builtin_amdgcn_sqrtf((float)(big ? rem * 0x1p-100 : rem));
  if (big) s *= 0x1p50;
  double a = (-s - d) * iL - 1e-4, b = (s - d) * iL + 1e-4;
  a = fmax(a, -(double)N - 1.0);   // also keeps NaN / inf out of the int conversion
  b = fmin(b, (double)N + 1.0);
  lo = max(-N, (int)ceil(a));
  hi = min(N, (int)floor(b));
}

// One (target, cell) pair of the periodic branch: the images of the cell inside the cutoff,
// in the reference's loop order (nx, ny, nz ascending, nz fastest).  Returns the image sums
// of kv * d (d = x_t - x_c + n L) and of pk -- or, with quirks, the pk of the last image
// inside the cutoff, as the reference assigns pgpot per image (Gravity_getPotentialAtPoint.F90:
// 262-274).
// Image indices run as doubles (exact for |n| <= 64), so n * L is formed without a conversion.
template <bool S32, bool QUIRKS>
__device__ __forceinline__ void flash_images(const Flash& P, double dxi, double dyi, double dzi,
                                             double& px, double& py, double& pz, double& pp) {
  px = py = pz = pp = 0.0;
  const int N = P.nrep;
  int x0, x1;
  image_range(dxi, P.ilx, P.cut_thr, N, x0, x1);
  for (double fx = (double)x0; fx <= (double)x1; fx += 1.0) {
    double dx, x2;
    {
#pragma clang fp contract(off)
      dx = dxi + fx * P.lx;
      x2 = dx * dx;
    }
    int y0, y1;
    image_range(dyi, P.ily, P.cut_thr - x2, N, y0, y1);
    for (double fy = (double)y0; fy <= (double)y1; fy += 1.0) {
      double dy, xy;
      {
#pragma clang fp contract(off)
        dy = dyi + fy * P.ly;
        xy = x2 + dy * dy;
      }
      int z0, z1;
      image_range(dzi, P.ilz, P.cut_thr - xy, N, z0, z1);
      const double fz1 = (double)z1;
      if (xy > P.soft_thr && P.cut_thr < 1e68) {
        // No image of this row lies inside rsoft (r2 >= xy, also as rounded).  The images strictly
        // between the range's ends lie inside the cutoff: the range is the integers of an interval
        // widened by 1e-4 of an image spacing around the exact one (whose half-width carries a
        // relative error of ~2^-22), so every inner image is more than 0.9999 spacings inside it.
        // (cut_thr < 1e68 keeps that half-width inside image_range's FP32 range, so it is finite.)
        // They are summed without tests, with r2 formed by FMAs; only the two end images decide
        // the cutoff, with the reference's rounding.  The x and y sums take the row's sum of
        // weights (dx, dy are fixed along the row).
        double skv = 0.0;
        if (z0 <= z1) {
          double dz, r2;
          {
#pragma clang fp contract(off)
            dz = dzi + (double)z0 * P.lz;
            r2 = xy + dz * dz;
          }
          if (r2 <= P.cut_thr) {
            const double y = rsqrt_f64<S32 ? F_SEED32 : 0>(r2);
            const double kv = (y * y) * y;
            skv += kv;
            pz = fma(kv, dz, pz);
            if (QUIRKS) pp = y;
            else pp += y;
          }
        }
        for (double fz = (double)z0 + 1.0; fz < fz1; fz += 1.0) {
          const double dz = fma(fz, P.lz, dzi);
          const double y = rsqrt_f64<S32 ? F_SEED32 : 0>(fma(dz, dz, xy));
          const double kv = (y * y) * y;
          skv += kv;
          pz = fma(kv, dz, pz);
          if (QUIRKS) pp = y;
          else pp += y;
        }
        if (z1 > z0) {
          double dz, r2;
          {
#pragma clang fp contract(off)
            dz = dzi + fz1 * P.lz;
            r2 = xy + dz * dz;
          }
          if (r2 <= P.cut_thr) {
            const double y = rsqrt_f64<S32 ? F_SEED32 : 0>(r2);
            const double kv = (y * y) * y;
            skv += kv;
            pz = fma(kv, dz, pz);
            if (QUIRKS) pp = y;
            else pp += y;
          }
        }
        px = fma(skv, dx, px);
        py = fma(skv, dy, py);
        continue;
      }
      for (double fz = (double)z0; fz <= fz1; fz += 1.0) {
        double dz, r2;
        {
#pragma clang fp contract(off)
          dz = dzi + fz * P.lz;
          r2 = xy + dz * dz;
        }
        if (r2 <= P.cut_thr) {
          double kv, pk;
          flash_pair<S32>(P, r2, kv, pk);
          px = fma(kv, dx, px);
          py = fma(kv, dy, py);
          pz = fma(kv, dz, pz);
          if (QUIRKS) pp = pk;
          else pp += pk;
        }
      }
    }
  }
}

// Correction for one source whose pairs the branch-free loop flagged (FMA-formed r2 below
// near2): remove the clamped Newtonian term that loop added and add the exact pair term.
// Any pair the reference puts inside rsoft is flagged: its r2 is at most soft_thr, and the
// FMA-formed r2 differs from the reference's by a few ulps.
template <int K, bool S32>
__device__ __forceinline__ void flash_fix(const Flash& P, const double4 s, const double (&tx)[K],
                                          const double (&ty)[K], const double (&tz)[K],
                                          double (&ax)[K], double (&ay)[K], double (&az)[K],
                                          double (&ap)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double fx = tx[k] - s.x, fy = ty[k] - s.y, fz = tz[k] - s.z;
    const double r2f = fma(fx, fx, fma(fy, fy, fz * fz));
    if (r2f < P.near2) {
      const double y0 = rsqrt_f64<S32 ? F_SEED32 : 0>(P.near2);
      const double mr0 = s.w * y0;
      const double w0 = mr0 * (y0 * y0);
      double dx, dy, dz, r2;
      {
#pragma clang fp contract(off)
        dx = tx[k] - s.x;
        dy = ty[k] - s.y;
        dz = tz[k] - s.z;
        r2 = (dx * dx + dy * dy) + dz * dz;
      }
      double kv, pk;
      flash_pair<S32>(P, r2, kv, pk);
      const double w = s.w * kv;
      ax[k] += w0 * fx - w * dx;
      ay[k] += w0 * fy - w * dy;
      az[k] += w0 * fz - w * dz;
      ap[k] += s.w * pk - mr0;
    }
  }
}

// The source loop of k_eval_flash over [sb, se) for one target tile (tx, ty, tz held by the lanes).
template <int K, int NT, bool PER, bool S32, bool QUIRKS>
__device__ __forceinline__ void flash_sources(const Flash& P, double4* ls, const double4* __restrict__ src4,
                                              int64_t sb, int64_t se, const double (&tx)[K],
                                              const double (&ty)[K], const double (&tz)[K],
                                              double (&ax)[K], double (&ay)[K], double (&az)[K],
                                              double (&ap)[K]) {
  const int tid = threadIdx.x;
  for (int64_t base = sb; base < se; base += NT) {
    __syncthreads();
    if (base + tid < se) ls[tid] = src4[base + tid];
    __syncthreads();
    const int cnt = (int)(se - base < NT ? se - base : NT);
    if (PER) {
      for (int j = 0; j < cnt; ++j) {
        const double4 s = ls[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          double px, py, pz, pp;
          flash_images<S32, QUIRKS>(P, tx[k] - s.x, ty[k] - s.y, tz[k] - s.z, px, py, pz, pp);
          ax[k] = fma(-s.w, px, ax[k]);
          ay[k] = fma(-s.w, py, ay[k]);
          az[k] = fma(-s.w, pz, az[k]);
          ap[k] = fma(s.w, pp, ap[k]);
        }
      }
    } else {
      // 64 sources at a time; fm (wave-uniform) marks the sources that came closer than the
      // clamp to any of the wave's targets, which flash_fix then revisits one by one.
      for (int j0 = 0; j0 < cnt; j0 += 64) {
        const int j1 = j0 + 64 < cnt ? j0 + 64 : cnt;
        unsigned long long fm = 0ULL;
        for (int j = j0; j < j1; ++j) {
          const double4 s = ls[j];
          bool nj = false;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const double dx = tx[k] - s.x, dy = ty[k] - s.y, dz = tz[k] - s.z;
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            nj = nj || r2 < P.near2;
            const double y = rsqrt_f64<S32 ? F_SEED32 : 0>(fmax(r2, P.near2));
            const double mr = s.w * y;
            ap[k] += mr;
            const double w = mr * (y * y);
            ax[k] = fma(-w, dx, ax[k]);
            ay[k] = fma(-w, dy, ay[k]);
            az[k] = fma(-w, dz, az[k]);
          }
          if (__builtin_amdgcn_ballot_w64(nj)) fm |= 1ULL << (j - j0);
        }
        while (fm) {
          const int j = j0 + __builtin_ctzll(fm);
          fm &= fm - 1ULL;
          flash_fix<K, S32>(P, ls[j], tx, ty, tz, ax, ay, az, ap);
        }
      }
    }
  }
}

// blockIdx.x = target tile of K*NT targets, blockIdx.y = source chunk.  Accumulators follow
// store_result's convention (a = G s, phi = -G s_p): s -= m kv d, s_p += m pk.  The FP32 seed of
// 1/sqrt is taken per tile where the tile's r^2 bound (source statistics sst, images included)
// keeps r^2 inside the FP32 range.
template <int K, int NT, bool PER, bool QUIRKS = false>
__global__ __launch_bounds__(NT) void k_eval_flash(Tgt T, const double4* __restrict__ src4,
                                                   int64_t ns, int64_t chunk, Flash P,
                                                   const double* sst, Out O) {
  __shared__ double4 ls[NT];
  __shared__ double red[2 * (NT / 64)];
  if (gated_off(O)) return;
  const int tid = threadIdx.x;
  const int64_t tile0 = (int64_t)blockIdx.x * (K * NT);
  const int64_t sb = (int64_t)blockIdx.y * chunk;
  const int64_t se = sb + chunk < ns ? sb + chunk : ns;
  Out o = O;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
  double tx[K], ty[K], tz[K], ax[K], ay[K], az[K], ap[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    double e2;   // the FLASH kernels use the global radius; per-point softening is ignored
    load_target(T, t < T.n ? t : T.n - 1, tx[k], ty[k], tz[k], e2);
    ax[k] = ay[k] = az[k] = ap[k] = 0.0;
  }
  const double rt = tile_rmax<K, NT>(tx, ty, tz, tile0, T.n, red);
  if (P.s32 && r2_bound(sst, rt, 0.0, P.reach) < TG_F32_HUGE_R2)   // uniform over the workgroup
    flash_sources<K, NT, PER, true, QUIRKS>(P, ls, src4, sb, se, tx, ty, tz, ax, ay, az, ap);
  else
    flash_sources<K, NT, PER, false, QUIRKS>(P, ls, src4, sb, se, tx, ty, tz, ax, ay, az, ap);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    if (t < T.n) store_result(o, t, ax[k], ay[k], az[k], ap[k]);
  }
}

// ------------------------------------------------------- 8^3 block sources ----
// Isolated FLASH kernel against cells handed over as 8^3 leaf blocks (tgrav_set_source_blocks):
// the column walk of k_eval_gridsrc (dx, dy and dx^2 + dy^2 formed once per column and
// target) with the branch-free Newtonian loop; a wave-uniform bit per cell of the column marks
// cells that came closer than the clamp to any of the wave's targets, and those are redone
// exactly (reference-rounded r2, exact threshold) as flash_fix does.
template <int K, bool S32>
__device__ __forceinline__ void flash_gs_block(const Flash& P, const LdsGS<false>& L,
                                               const double (&tx)[K], const double (&ty)[K],
                                               const double (&tz)[K], double (&ax)[K],
                                               double (&ay)[K], double (&az)[K], double (&ap)[K]) {
  for (int col = 0; col < 64; ++col) {
    const double xs = L.cx[col >> 3], ys = L.cy[col & 7];
    double dx[K], dy[K], base[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      dx[k] = tx[k] - xs;
      dy[k] = ty[k] - ys;
      base[k] = fma(dx[k], dx[k], dy[k] * dy[k]);
    }
    unsigned int nm = 0u;   // wave-uniform: bit kk = cell kk of the column was flagged
    double sw[K];           // the column's sum of weights: its cells share dx and dy
#pragma unroll
    for (int k = 0; k < K; ++k) sw[k] = 0.0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const double zs = L.cz[kk], m = L.m[col * 8 + kk];
      bool nj = false;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double dz = tz[k] - zs;
        const double r2 = fma(dz, dz, base[k]);
        nj = nj || r2 < P.near2;
        const double y = rsqrt_f64<S32 ? F_SEED32 : 0>(fmax(r2, P.near2));
        const double mr = m * y;
        ap[k] += mr;
        const double w = mr * (y * y);
        sw[k] += w;
        az[k] = fma(-w, dz, az[k]);
      }
      if (__builtin_amdgcn_ballot_w64(nj)) nm |= 1u << kk;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      ax[k] = fma(-sw[k], dx[k], ax[k]);
      ay[k] = fma(-sw[k], dy[k], ay[k]);
    }
    while (nm) {
      const int kk = __builtin_ctz(nm);
      nm &= nm - 1u;
      const double zs = L.cz[kk], m = L.m[col * 8 + kk];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double dz = tz[k] - zs;
        const double r2f = fma(dz, dz, base[k]);
        if (r2f < P.near2) {
          const double y0 = rsqrt_f64<S32 ? F_SEED32 : 0>(P.near2);
          const double mr0 = m * y0;
          const double w0 = mr0 * (y0 * y0);
          double ex, ey, ez, r2;
          {
#pragma clang fp contract(off)
            ex = tx[k] - xs;
            ey = ty[k] - ys;
            ez = tz[k] - zs;
            r2 = (ex * ex + ey * ey) + ez * ez;
          }
          double kv, pk;
          flash_pair<S32>(P, r2, kv, pk);
          const double wv = m * kv;
          ax[k] += w0 * dx[k] - wv * ex;
          ay[k] += w0 * dy[k] - wv * ey;
          az[k] += w0 * dz - wv * ez;
          ap[k] += m * pk - mr0;
        }
      }
    }
  }
}

// Explicit targets (K per lane) against 8^3 source blocks; blockIdx.x = target tile,
// blockIdx.y = chunk of source blocks [y*chunk, min((y+1)*chunk, nblk)).
template <int K, int NT>
__global__ __launch_bounds__(NT) void k_eval_flash_gridsrc(Tgt T, const double* __restrict__ sax,
                                                           const double* __restrict__ be2,
                                                           const double* __restrict__ bm,
                                                           int64_t nblk, int64_t chunk, Flash P,
                                                           const double* sst, Out O) {
  __shared__ LdsGS<false> L;
  if (gated_off(O)) return;   // explicit sources that turned out not to be 8^3 blocks
  const int tid = threadIdx.x;
  const int64_t tile0 = (int64_t)blockIdx.x * (K * NT);
  const int64_t b0 = (int64_t)blockIdx.y * chunk;
  const int64_t b1 = b0 + chunk < nblk ? b0 + chunk : nblk;
  double tx[K], ty[K], tz[K], ax[K], ay[K], az[K], ap[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    double e2;
    load_target(T, t < T.n ? t : T.n - 1, tx[k], ty[k], tz[k], e2);
    ax[k] = ay[k] = az[k] = ap[k] = 0.0;
  }
  const double rt = tile_rmax<K, NT>(tx, ty, tz, tile0, T.n, L.red);
  const bool s32 = P.s32 && r2_bound(sst, rt, 0.0) < TG_F32_HUGE_R2;   // uniform over the workgroup
  for (int64_t blk = b0; blk < b1; ++blk) {
    __syncthreads();
    gs_stage<false>(L, sax, be2, bm, blk, 0.0, 0.0, 0.0, 0.0);
    __syncthreads();
    if (s32) flash_gs_block<K, true>(P, L, tx, ty, tz, ax, ay, az, ap);
    else flash_gs_block<K, false>(P, L, tx, ty, tz, ax, ay, az, ap);
  }
  Out o = O;
  if (o.part) o.part += (int64_t)blockIdx.y * 4 * o.pstride;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t t = tile0 + (int64_t)k * NT + tid;
    if (t < T.n) store_result(o, t, ax[k], ay[k], az[k], ap[k]);
  }
}

}  // namespace tg
