// Particle-mesh pieces beside the bridge kick, on guarded FLASH leaf blocks, FP64:
// cloud-in-cell mass deposit, the finite-difference acceleration (optionally fused with the
// grid kick v += g dt) and cloud-in-cell interpolation to particles.
//
// Block layout: one variable of nblk blocks of nbg^3 cells (nbg = nb + 2 ng: nb interior
// cells per axis, ng guard layers), block b at b * nbg^3, cell (i, j, k) at
// (k * nbg + j) * nbg + i -- x fastest, as solnData(VAR, i, j, k) of one variable.  Block b
// has interior lower corner (ox, oy, oz)[b] and cubic physical size bs[b]; the cell width is
// del = bs / nb and guarded cell i has centre o + ((i - ng) + 0.5) del.
//
// All three are HBM / latency-bound elementwise work (no MFMA shape): one lane per cell for
// the gradient (x-fastest lanes, so each neighbour load of a wave is one or two contiguous
// segments), one lane per particle for the deposit and the interpolation.
//
// The deposit is deterministic: every contribution c = (m / del^3) w is split exactly into
// 64-bit integers, c 2^S = H 2^32 + L (0 <= L < 2^32; S per cell, from the largest |c| the
// cell receives), and H and L are summed with integer atomics, which do not depend on the
// order of the adds.  A cell's density (H 2^32 + L) 2^-S is then within an ulp of the exact
// sum of its contributions (each is truncated to ~2^-66 of the cell's largest one).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

struct PmBlocks {
  const double *ox, *oy, *oz, *bs;
  int64_t nblk;
  int nb, ng, nbg;
  int64_t nc;   // nbg^3
};

// Cloud-in-cell weights along one axis (the oracle's cic_axis, same operations in the same
// order): u = (x - o) / del - 0.5 + ng, base cell floor(u), weights (1 - f, f).
__device__ __forceinline__ bool pm_axis(double x, double o, double del, int ng, int nbg, int& i0,
                                        double& w0, double& w1) {
  double u;
  {
#pragma clang fp contract(off)
    u = (x - o) / del - 0.5 + (double)ng;
  }
  const double fl = floor(u);
  if (!(fl >= 0.0 && fl + 1.0 < (double)nbg)) return false;
  i0 = (int)fl;
  w1 = u - fl;
  w0 = 1.0 - w1;
  return true;
}

// Scale exponent of one cell's fixed-point sum from cmax = bits of the largest |c| among the
// cell's contributions: np * cmax < 2^e, S = 94 - e keeps every partial sum of H below 2^62
// and resolves each contribution to 2^-66 of that cell's largest one.
__device__ __forceinline__ int pm_scale(unsigned long long cmax, int64_t np) {
  const double mx = __longlong_as_double((long long)cmax);
  if (!(mx > 0.0 && mx < 1e300)) return 0;
  int e;
  frexp(mx * (double)np, &e);
  return 94 - e;
}

// Particle p's block, densities and weights; false if the block index or the CIC stencil
// is outside the guarded block.
__device__ __forceinline__ bool pm_particle(const PmBlocks& B, const int64_t* blk, int64_t p,
                                            double x, double y, double z, int64_t& b, double& del,
                                            int (&c0)[3], double (&w)[3][2]) {
  b = blk[p];
  if (b < 0 || b >= B.nblk) return false;
  del = B.bs[b] / (double)B.nb;
  return pm_axis(x, B.ox[b], del, B.ng, B.nbg, c0[0], w[0][0], w[0][1]) &&
         pm_axis(y, B.oy[b], del, B.ng, B.nbg, c0[1], w[1][0], w[1][1]) &&
         pm_axis(z, B.oz[b], del, B.ng, B.nbg, c0[2], w[2][0], w[2][1]);
}

// The 8 contributions of particle p: c = (m / del^3) ((wx wy) wz) with the oracle's
// roundings, corners dk slowest, di fastest.  Calls f(cell index, c).
template <class F>
__device__ __forceinline__ bool pm_corners(const PmBlocks& B, const double* x, const double* y,
                                           const double* z, const double* m, const int64_t* blk,
                                           int64_t p, F f) {
  int64_t b;
  double del;
  int c0[3];
  double w[3][2];
  if (!pm_particle(B, blk, p, x[p], y[p], z[p], b, del, c0, w)) return false;
  const double md = m[p] / ((del * del) * del);
  const int64_t base = b * B.nc;
#pragma unroll
  for (int dk = 0; dk < 2; ++dk)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int di = 0; di < 2; ++di) {
#pragma clang fp contract(off)
        const double c = md * ((w[0][di] * w[1][dj]) * w[2][dk]);
        f(base + ((int64_t)(c0[2] + dk) * B.nbg + (c0[1] + dj)) * B.nbg + (c0[0] + di), c);
      }
  return true;
}

// Pass 1: cmax[cell] = max |c| over the cell's contributions (as non-negative double bits,
// whose unsigned order is the numeric order); a bad particle sets *err.
__global__ void k_pm_cell_max(int64_t np, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, const double* __restrict__ m,
                              const int64_t* __restrict__ blk, PmBlocks B, unsigned long long* cmax,
                              unsigned long long* err) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const bool ok = pm_corners(B, x, y, z, m, blk, p, [&](int64_t idx, double c) {
    const unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(c));
    if (bits) atomicMax(&cmax[idx], bits);
  });
  if (!ok) atomicOr(err, 1ULL);
}

// Pass 2: accH / accL += the exact integer split c 2^S = H 2^32 + L of every contribution,
// S from the cell's bound.
__global__ void k_pm_deposit(int64_t np, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ m,
                             const int64_t* __restrict__ blk, PmBlocks B,
                             const unsigned long long* __restrict__ cmax, unsigned long long* accH,
                             unsigned long long* accL) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  pm_corners(B, x, y, z, m, blk, p, [&](int64_t idx, double c) {
    const int S = pm_scale(cmax[idx], np);
    const double t = ldexp(c, S - 32);
    const double h = floor(t);
    const long long H = (long long)h;
    const unsigned long long L = (unsigned long long)ldexp(t - h, 32);
    atomicAdd(&accH[idx], (unsigned long long)H);
    atomicAdd(&accL[idx], L);
  });
}

// Pass 3: rho = (H 2^32 + L) 2^-S (or rho += that with accumulate).
__global__ void k_pm_finalize(int64_t n, const unsigned long long* __restrict__ accH,
                              const unsigned long long* __restrict__ accL,
                              const unsigned long long* __restrict__ cmax, int64_t np,
                              double* __restrict__ rho, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int S = pm_scale(cmax[i], np);
  const double v = ldexp((double)(long long)accH[i], 32 - S) + ldexp((double)accL[i], -S);
  rho[i] = accumulate ? rho[i] + v : v;
}

__global__ void k_zero_u64(int64_t n, unsigned long long* a, unsigned long long* b,
                           unsigned long long* c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] = 0ULL;
  if (b) b[i] = 0ULL;
  if (c) c[i] = 0ULL;
}

// g = -grad(phi) by the row rule delxinv (phi(i-1) - phi(i+1)), delxinv = 0.5 (nb / bs), the
// two ends of each row copying their neighbour; any of gx/gy/gz may be null.  With vx != null
// also the grid kick v += g dt (product and sum rounded separately).
// Grid: x = slice * 8 + (b mod 8), y = b / 8 - grp0, where a slice is 256 consecutive cells of
// a block.  Workgroups are dealt to the 8 XCDs round-robin by linear id (x fastest, gridDim.x a
// multiple of 8), so all slices of block b run on XCD (b mod 8) and a block's z-neighbour reads
// hit the L2 that its other slices filled.  Cell (i, j, k) comes from the slice offset r by two
// divisions by nbg done as multiply-high with m = ceil(2^32 / nbg), exact for r < 2^32 / nbg
// (nbg <= 128 and r < 2^21 here).
__global__ __launch_bounds__(256) void k_pm_gradient(PmBlocks B, uint32_t magic, int64_t grp0,
                                                     const double* __restrict__ phi,
                                                     double* __restrict__ gx, double* __restrict__ gy,
                                                     double* __restrict__ gz, double dt,
                                                     double* __restrict__ vx, double* __restrict__ vy,
                                                     double* __restrict__ vz) {
  const int64_t b = (grp0 + blockIdx.y) * 8 + (blockIdx.x & 7u);
  if (b >= B.nblk) return;
  const uint32_t r = (blockIdx.x >> 3) * 256u + threadIdx.x;
  if (r >= (uint32_t)B.nc) return;
  const int n = B.nbg;
  const uint32_t q = __umulhi(r, magic);        // r / n
  const uint32_t k = __umulhi(q, magic);        // r / n^2
  const int c[3] = {(int)(r - q * (uint32_t)n), (int)(q - k * (uint32_t)n), (int)k};
  const int stride[3] = {1, n, n * n};
  double delxinv = (double)B.nb / B.bs[b];
  delxinv = 0.5 * delxinv;
  const double* f = phi + b * B.nc;
  double g[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int ii = c[a] < 1 ? 1 : (c[a] > n - 2 ? n - 2 : c[a]);
    const int ctr = (int)r + (ii - c[a]) * stride[a];
    g[a] = delxinv * (f[ctr - stride[a]] - f[ctr + stride[a]]);
  }
  const int64_t t = b * B.nc + r;
  if (gx) gx[t] = g[0];
  if (gy) gy[t] = g[1];
  if (gz) gz[t] = g[2];
  if (vx) {
#pragma clang fp contract(off)
    vx[t] = vx[t] + g[0] * dt;
    vy[t] = vy[t] + g[1] * dt;
    vz[t] = vz[t] + g[2] * dt;
  }
}

// a_p = sum over the 8 corners of w g (corner order and roundings of the oracle).  A
// particle outside its guarded block gets NaN and sets bit 0 of *err.
__global__ void k_pm_interpolate(int64_t np, const double* __restrict__ x,
                                 const double* __restrict__ y, const double* __restrict__ z,
                                 const int64_t* __restrict__ blk, PmBlocks B,
                                 const double* __restrict__ gx, const double* __restrict__ gy,
                                 const double* __restrict__ gz, double* __restrict__ ax,
                                 double* __restrict__ ay, double* __restrict__ az,
                                 unsigned long long* err) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int64_t b;
  double del;
  int c0[3];
  double w[3][2];
  if (!pm_particle(B, blk, p, x[p], y[p], z[p], b, del, c0, w)) {
    atomicOr(err, 1ULL);
    ax[p] = ay[p] = az[p] = __longlong_as_double(0x7ff8000000000000LL);
    return;
  }
  const int64_t base = b * B.nc;
  double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
  for (int dk = 0; dk < 2; ++dk)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int di = 0; di < 2; ++di) {
#pragma clang fp contract(off)
        const int64_t idx = base + ((int64_t)(c0[2] + dk) * B.nbg + (c0[1] + dj)) * B.nbg + (c0[0] + di);
        const double wt = (w[0][di] * w[1][dj]) * w[2][dk];
        sx = sx + wt * gx[idx];
        sy = sy + wt * gy[idx];
        sz = sz + wt * gz[idx];
      }
  ax[p] = sx;
  ay[p] = sy;
  az[p] = sz;
}

}  // namespace tg
