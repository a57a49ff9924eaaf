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

// ------------------------------------------ periodic pairs grouped by separation ----
// The periodic branch's work is the image walk of each (target, cell) pair, and its rows and
// ranges depend only on the pair's separation d = x_t - x_c.  k_eval_flash gives a wave 64
// targets and one cell: their separations differ, so the lanes' rows and ranges differ (lane
// efficiency ~0.8) and every lane forms its own ranges (~44 instructions per row).  Here every
// pair gets the Morton code of d, quantized to 1/64 of the box and clamped to 8 boxes either way,
// the pairs are sorted by it, and each wave takes 64 consecutive pairs.  Their separations are
// close, so the 64 lanes build one table of rows from the wave's bounding box of d (one candidate
// row per lane at a time), and walk it together:
//   - a row's z images certainly inside the cutoff for every d in the box (and outside rsoft)
//     are summed without tests, r2 formed by FMAs;
//   - the images that some d of the box may have inside are tested per lane with the reference's
//     rounding (Gravity_getAccelAtPoint.F90:262-270);
//   - rows that may come within rsoft take the general per-image loop with the softened kernel.
// The walk keeps the reference's order (nx, ny, nz ascending), so with quirks pp is the pk of the
// pair's last image inside the cutoff.  Each lane writes its pair's image sums (px, py, pz, pp) at
// the pair's index t * ns + c; k_pair_reduce sums them per target with the masses.
#define TG_PAIR_NREP 7                                   // rows of the table: (2 N + 1)^2 <= 225
#ifndef TG_PAIR_BITS
#define TG_PAIR_BITS 10      // separation key: bits per axis (bins of L / 2^(bits - 4), 8 boxes either way)
#endif
#define TG_PAIR_SLACK 1e-4   // image spacings; the square roots (v_rsq_f64, ~2^-24) are off by < 1e-5
#define TG_PAIR_ROWS ((2 * TG_PAIR_NREP + 1) * (2 * TG_PAIR_NREP + 1))
// nrep 8 .. 10: a table of (2 nrep + 1)^2 rows in dynamic LDS (<= 21.2 KB).  On c1's inputs it ran at
// 0.50 / 0.45 of FP64 peak at nrep 8 / 10 against 0.37 / 0.40 for the per-target walk; the gap
// narrows as the table takes more LDS (fewer waves per CU), so larger nrep keep the per-target walk.
#define TG_PAIR_NREP_DYN 10

// The z images of a row: zu0 .. zu0 + nlo - 1 tested, then nc inside the cutoff for every pair of
// the wave, then zh0 .. zh0 + nhi - 1 tested; soft: a row that may come within rsoft, its nlo images
// from zu0 all taking the general test and kernel.  The counts are formed when the table is built,
// one row per lane, and packed into one word (nlo | nc << 8 | nhi << 16 | soft << 24, each < 256)
// that the walk moves to the scalar unit in one instruction.  Soft rows fill the table from its end;
// `before` is the number of the other rows that precede a soft row in (nx, ny) order.
struct __align__(16) PairRow {
  double fxl, fyl;          // fl(nx lx), fl(ny ly): the reference forms d + n L from these
  double zu0, zh0;
  int counts, before, pad[2];
};

__device__ __forceinline__ unsigned int morton_spread10(unsigned int v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ unsigned int sep_bin(double d, double ibin) {
  const double h = (double)(1 << (TG_PAIR_BITS - 1));
  const double q = fmin(fmax(floor(d * ibin) + h, 0.0), 2.0 * h - 1.0);   // NaN -> 0
  return (unsigned int)q;
}

// keys[i], idx[i] for pair i = t * ns + c (target t0 + t): Morton code of the quantized
// separation, and i.
__global__ __launch_bounds__(256) void k_pair_keys(Tgt T, int64_t t0, const double4* __restrict__ src4,
                                                   int64_t ns, int64_t npairs, double ibx, double iby,
                                                   double ibz, unsigned int* __restrict__ keys,
                                                   unsigned int* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const int64_t t = i / ns, c = i - t * ns;
  double tx, ty, tz, e2;
  load_target(T, t0 + t, tx, ty, tz, e2);
  const double4 s = src4[c];
  keys[i] = morton_spread10(sep_bin(tx - s.x, ibx)) | (morton_spread10(sep_bin(ty - s.y, iby)) << 1) |
            (morton_spread10(sep_bin(tz - s.z, ibz)) << 2);
  idx[i] = (unsigned int)i;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Smallest and largest of fl(d + c)^2 over d in [lo, hi] (fl rounds monotonically, so a pair's
// own rounded square lies between them).
__device__ __forceinline__ void sq_range(double lo, double hi, double c, double& mn, double& mx) {
  double a, b;
  {
#pragma clang fp contract(off)
    a = lo + c;
    b = hi + c;
  }
  const double a2 = a * a, b2 = b * b;
  mx = fmax(a2, b2);
  mn = (a <= 0.0 && b >= 0.0) ? 0.0 : fmin(a2, b2);
}

// One wave per workgroup: sorted pairs [64 w, 64 w + 64) of `idx`.  DYN = false: a static table of
// 225 rows (10.8 KB) for nrep <= 7 (sizing it by nrep instead, 8.1 KB at nrep = 6 and more waves per
// CU, was 4% slower); DYN = true: (2 nrep + 1)^2 rows of dynamic LDS, set at launch, for nrep 8 .. 10.
template <bool QUIRKS, bool DYN>
__global__ __launch_bounds__(64) void k_eval_flash_pairs(Tgt T, int64_t t0, const double4* __restrict__ src4,
                                                         int64_t ns, int64_t npairs,
                                                         const unsigned int* __restrict__ idx, Flash P,
                                                         double4* __restrict__ pairbuf) {
  PairRow* rows;
  if constexpr (DYN) {
    extern __shared__ PairRow rows_dyn[];
    rows = rows_dyn;
  } else {
    __shared__ PairRow rows_static[TG_PAIR_ROWS];
    rows = rows_static;
  }
  const int ROWS = DYN ? (2 * P.nrep + 1) * (2 * P.nrep + 1) : TG_PAIR_ROWS;
  const int lane = threadIdx.x;
  const int64_t pi = (int64_t)blockIdx.x * 64 + lane;
  const bool valid = pi < npairs;
  const int64_t pr = valid ? (int64_t)idx[pi] : 0;
  const int64_t t = pr / ns, c = pr - t * ns;
  double tx, ty, tz, e2;
  load_target(T, t0 + t, tx, ty, tz, e2);
  const double4 s = src4[c];
  const double dxi = tx - s.x, dyi = ty - s.y, dzi = tz - s.z;
  const double INF = __builtin_huge_val();
  const double dxl = wave_min(valid ? dxi : INF), dxh = wave_max(valid ? dxi : -INF);
  const double dyl = wave_min(valid ? dyi : INF), dyh = wave_max(valid ? dyi : -INF);
  const double dzl = wave_min(valid ? dzi : INF), dzh = wave_max(valid ? dzi : -INF);
  const int N = P.nrep, W2 = 2 * N + 1;
  const double fN = (double)N, cut = P.cut_thr;
  // the row table, in (nx, ny) order: lane j takes candidate 64 r + j; rows that may come within
  // rsoft (rare) go to the table's end, so the walk over the others has no branch for them
  int nrows = 0, nsoft = 0;
  for (int r0 = 0; r0 < W2 * W2; r0 += 64) {
    const int cand = r0 + lane;
    bool keep = false;
    PairRow R;
    if (cand < W2 * W2) {
      const double fx = (double)(cand / W2 - N), fy = (double)(cand % W2 - N);
      R.fxl = fx * P.lx;
      R.fyl = fy * P.ly;
      double x2l, x2h, y2l, y2h;
      sq_range(dxl, dxh, R.fxl, x2l, x2h);
      sq_range(dyl, dyh, R.fyl, y2l, y2h);
      double xyl, xyh, remh, reml;
      {
#pragma clang fp contract(off)
        xyl = x2l + y2l;
        xyh = x2h + y2h;
        remh = cut - xyl;   // no pair of the box has a larger remainder cut_thr - xy
        reml = cut - xyh;   // nor a smaller one (the same rounded operations, monotone)
      }
      if (remh >= 0.0) {
        const double sh = remh * __builtin_amdgcn_rsq(remh + 0x1p-1022);
        const double zu0 = fmax(ceil((-sh - dzh) * P.ilz - TG_PAIR_SLACK), -fN);
        const double zu1 = fmin(floor((sh - dzl) * P.ilz + TG_PAIR_SLACK), fN);
        keep = zu0 <= zu1;
        double zc0 = zu1 + 1.0, zc1 = zu1;   // no certain images unless shown below
        const bool soft = !(xyl > P.soft_thr);   // some image of the row may lie within rsoft
        if (!soft && reml > 0.0) {
          // every pair of the box: |dz + n L| <= sqrt(reml) - m L keeps the reference's rounded
          // r2 below cut_thr (m: the error of the square root and 2.7u cut / sqrt(reml) for the
          // rounding of dz^2 and r2)
          const double yl = __builtin_amdgcn_rsq(reml);
          const double sl = reml * yl;
          const double m = fma(cut * P.ilz * 6e-16, yl, TG_PAIR_SLACK);
          const double c0 = fmax(ceil((-sl - dzl) * P.ilz + m), zu0);
          const double c1 = fmin(floor((sl - dzh) * P.ilz - m), zu1);
          if (c0 <= c1) {
            zc0 = c0;
            zc1 = c1;
          }
        }
        R.zu0 = zu0;
        R.zh0 = zc1 + 1.0;
        R.counts = soft ? ((int)(zu1 - zu0) + 1) | 1 << 24
                        : (int)(zc0 - zu0) | ((int)(zc1 - zc0) + 1) << 8 | (int)(zu1 - zc1) << 16;
      }
    }
    const bool soft = keep && (R.counts >> 24) != 0;
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    const unsigned long long bh = __builtin_amdgcn_ballot_w64(keep && !soft);
    const unsigned long long bs = __builtin_amdgcn_ballot_w64(soft);
    const int before = nrows + __builtin_popcountll(bh & lt);
    if (keep && !soft) rows[before] = R;
    if (soft) {
      R.before = before;
      rows[ROWS - 1 - (nsoft + __builtin_popcountll(bs & lt))] = R;
    }
    nrows += __builtin_popcountll(bh);
    nsoft += __builtin_popcountll(bs);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double px = 0.0, py = 0.0, pz = 0.0, pp = 0.0;
  // the rows in (nx, ny) order (the quirk's potential is the last image's): the runs of rows without
  // a soft row, each followed by the soft row that ends it
  int r = 0;
  for (int k = 0; k <= nsoft; ++k) {
    int end = nrows;
    if (k < nsoft) end = __builtin_amdgcn_readfirstlane(rows[ROWS - 1 - k].before);
    for (; r < end; ++r) {
      const PairRow R = rows[r];
      double dx, dy, xy;
      {
#pragma clang fp contract(off)
        dx = dxi + R.fxl;
        dy = dyi + R.fyl;
        xy = dx * dx + dy * dy;
      }
      // the row's counts are the same in every lane: kept on the scalar unit
      const int counts = __builtin_amdgcn_readfirstlane(R.counts);
      const int nlo = counts & 0xff, nc = (counts >> 8) & 0xff, nhi = (counts >> 16) & 0xff;
      double skv = 0.0;   // the row's sum of kv: dx and dy are fixed along it
      auto tested = [&](double fz) {
        double dz, r2;
        {
#pragma clang fp contract(off)
          dz = dzi + fz * P.lz;
          r2 = xy + dz * dz;
        }
        if (r2 <= cut) {
          const double y = rsqrt_f64<0>(r2);
          const double kv = (y * y) * y;
          skv += kv;
          pz = fma(kv, dz, pz);
          if (QUIRKS) pp = y;
          else pp += y;
        }
      };
      double fz = R.zu0;
      for (int i = 0; i < nlo; ++i, fz += 1.0) tested(fz);
      // the certain images from fz = zu0 + nlo, dz stepped by lz: one add per image instead of an
      // index step and an FMA (4% of the walk); the steps round dz to within nc/2 ulps of the row's
      // largest |dz|
      double dz = fma(fz, P.lz, dzi);
      for (int i = 0; i < nc; ++i, dz += P.lz) {
        const double y = rsqrt_f64<0>(fma(dz, dz, xy));
        const double kv = (y * y) * y;
        skv += kv;
        pz = fma(kv, dz, pz);
        if (QUIRKS) pp = y;
        else pp += y;
      }
      fz = R.zh0;
      for (int i = 0; i < nhi; ++i, fz += 1.0) tested(fz);
      px = fma(skv, dx, px);
      py = fma(skv, dy, py);
    }
    if (k < nsoft) {   // a soft row: the reference's test and kernel per image
      const PairRow R = rows[ROWS - 1 - k];
      double dx, dy, xy;
      {
#pragma clang fp contract(off)
        dx = dxi + R.fxl;
        dy = dyi + R.fyl;
        xy = dx * dx + dy * dy;
      }
      const int n = __builtin_amdgcn_readfirstlane(R.counts) & 0xff;
      double skv = 0.0;
      double fz = R.zu0;
      for (int i = 0; i < n; ++i, fz += 1.0) {
        double dz, r2;
        {
#pragma clang fp contract(off)
          dz = dzi + fz * P.lz;
          r2 = xy + dz * dz;
        }
        if (r2 <= cut) {
          double kv, pk;
          flash_pair<false>(P, r2, kv, pk);
          skv += kv;
          pz = fma(kv, dz, pz);
          if (QUIRKS) pp = pk;
          else pp += pk;
        }
      }
      px = fma(skv, dx, px);
      py = fma(skv, dy, py);
    }
  }
  if (valid) pairbuf[pr] = make_double4(px, py, pz, pp);
}

// Per target t0 + t (one wave each): s -= m_c (px, py, pz), s_p += m_c pp over its cells c, lane
// j taking c = j, j + 64, ... in order, then the lanes' sums in a fixed butterfly order.
__global__ __launch_bounds__(64) void k_pair_reduce(const double4* __restrict__ pairbuf,
                                                    const double4* __restrict__ src4, int64_t ns, int64_t t0,
                                                    Out O) {
  const int64_t t = blockIdx.x;
  const int lane = threadIdx.x;
  double ax = 0.0, ay = 0.0, az = 0.0, ap = 0.0;
  const double4* pb = pairbuf + t * ns;
  for (int64_t c = lane; c < ns; c += 64) {
    const double4 r = pb[c];
    const double m = src4[c].w;
    ax = fma(-m, r.x, ax);
    ay = fma(-m, r.y, ay);
    az = fma(-m, r.z, az);
    ap = fma(m, r.w, ap);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ax += __shfl_xor(ax, o);
    ay += __shfl_xor(ay, o);
    az += __shfl_xor(az, o);
    ap += __shfl_xor(ap, o);
  }
  if (lane == 0) store_result(O, t0 + t, ax, ay, az, ap);
}

// ------------------------------------------------------- 8^3 block sources ----
// Isolated FLASH kernel against cells handed over as 8^3 leaf blocks (tgrav_set_source_blocks):
// the column walk of k_eval_gridsrc (dx, dy and dx^2 + dy^2 formed once per column and
// target) with the branch-free Newtonian loop; a wave-uniform bit per cell of the column marks
// cells that came closer than the clamp to any of the wave's targets, and those are redone
// exactly (reference-rounded r2, exact threshold) as flash_fix does.  FAR: no target of the wave
// comes within the clamp of the block (block_far), so no pair is clamped or flagged -- the same
// arithmetic without the clamp, the flags and their ballots.
template <int K, bool S32, bool FAR>
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
        if (!FAR) nj = nj || r2 < P.near2;
        const double y = rsqrt_f64<S32 ? F_SEED32 : 0>(FAR ? r2 : fmax(r2, P.near2));
        const double mr = m * y;
        ap[k] += mr;
        const double w = mr * (y * y);
        sw[k] += w;
        az[k] = fma(-w, dz, az[k]);
      }
      if (!FAR && __builtin_amdgcn_ballot_w64(nj)) nm |= 1u << kk;
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

// Whether every target of the wave lies farther than the clamp sqrt(near2) from the box of the
// staged block's cell centres: that distance bounds every pair's exact r^2 from below, which the
// FMA-formed r^2 undercuts by a few ulps at most (the 1e-9 covers it and the rounding here).
template <int K>
__device__ __forceinline__ bool block_far(const Flash& P, const LdsGS<false>& L, const double (&tx)[K],
                                          const double (&ty)[K], const double (&tz)[K]) {
  const double x0 = fmin(L.cx[0], L.cx[7]), x1 = fmax(L.cx[0], L.cx[7]);
  const double y0 = fmin(L.cy[0], L.cy[7]), y1 = fmax(L.cy[0], L.cy[7]);
  const double z0 = fmin(L.cz[0], L.cz[7]), z1 = fmax(L.cz[0], L.cz[7]);
  bool far = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double gx = fmax(fmax(x0 - tx[k], tx[k] - x1), 0.0);
    const double gy = fmax(fmax(y0 - ty[k], ty[k] - y1), 0.0);
    const double gz = fmax(fmax(z0 - tz[k], tz[k] - z1), 0.0);
    far = far && gx * gx + gy * gy + gz * gz > P.near2 * (1.0 + 1e-9);   // NaN: not far
  }
  return __all(far);
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
    const bool far = block_far<K>(P, L, tx, ty, tz);   // uniform over the wave
    if (s32) {
      if (far) flash_gs_block<K, true, true>(P, L, tx, ty, tz, ax, ay, az, ap);
      else flash_gs_block<K, true, false>(P, L, tx, ty, tz, ax, ay, az, ap);
    } else {
      if (far) flash_gs_block<K, false, true>(P, L, tx, ty, tz, ax, ay, az, ap);
      else flash_gs_block<K, false, false>(P, L, tx, ty, tz, ax, ay, az, ap);
    }
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
