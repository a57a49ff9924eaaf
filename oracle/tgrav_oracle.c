/*
 * tgrav oracle -- CPU FP64 restatement of Torch's softened direct-sum gravity field.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check results
 * against (or as the timed CPU baseline).  The product path (torch_amd/libtgrav.so)
 * never calls into it.
 *
 * What it restates (all FP64, as FLASH builds promote `real` to double,
 * support/Makefile.h.in:13):
 *
 *   oracle_eval_plummer   -- the direct-sum loop nest of
 *       src/flash/source/physics/Gravity/GravityMain/Poisson/BHTree/Couple_AMUSE/
 *           Gravity_getAccelAtPoint.F90:168-248   (acceleration, isolated branch)
 *           Gravity_getPotentialAtPoint.F90:152-227 (potential, isolated branch)
 *       targets outer, sources inner, one accumulator per output, with the
 *       Plummer pair kernel used by the AMUSE stellar side (src/torch_mainloop.py:130,
 *       epsilon_squared): eps2_ij = eps_s_j^2 + eps_t_i^2,
 *           a_i   =  G sum_j m_j d_ij (|d_ij|^2 + eps2_ij)^(-3/2),   d_ij = x_j - x_i
 *           phi_i = -G sum_j m_j      (|d_ij|^2 + eps2_ij)^(-1/2)
 *       (the reference's prefactor = -G*m times (x - x_c)/r^3 is the same sign).
 *       A pair with |d|^2 + eps2 == 0 contributes nothing (self-exclusion).
 *   oracle_eval_flash     -- the FLASH spline / linear softening kernels with a global
 *       softening radius (Gravity_getAccelAtPoint.F90:140-145, 206-241;
 *       Gravity_getPotentialAtPoint.F90:190-222), with or without the reference's
 *       potential quirks (wrong sign inside the radius, `pgpot = slope`), isolated or
 *       with the periodic image sum (Gravity_getAccelAtPoint.F90:250-312).
 *   oracle_bidirectional  -- the fused gas<->particle pass of
 *       src/flash/source/Particles/ParticlesMain/active/Sink/
 *           Particles_sinkAccelGasOnSinksAndSinksOnGas.F90:204-316
 *       (cells outer, particles inner, particle accelerations scattered in the
 *       inner loop, cell acceleration stored per cell), Plummer kernel.
 *   oracle_kick           -- v += dt * a  (interface.F90:1793-1795, 1432-1444;
 *       src/torch_mainloop.py:287,289 call it with dt/2).
 *   oracle_block_centres  -- implicit cell centres of 8^3 (nb^3) leaf blocks in the
 *       C order of .deprecated/bridge_multiples.py:566-571 (meshgrid 'ij', ravel 'C').
 *   oracle_pm_deposit / oracle_pm_gradient / oracle_pm_interpolate -- the particle-mesh
 *       pieces beside the kick on guarded FLASH leaf blocks: CIC mass deposit, the
 *       finite-difference acceleration row rule (Simulation/SimulationMain/Cube/
 *       Gravity_accelOneRow.F90:94-121) and CIC interpolation to particles
 *       (Particles_longRangeForce.F90:70-97).  FLASH's own mappers are not vendored: the
 *       deposit and interpolation restate the textbook cloud-in-cell scheme.
 *
 * Accumulation modes: ORACLE_SEQ (plain sequential FP64, the reference's order),
 * ORACLE_NEUMAIER (compensated FP64), ORACLE_LONGDOUBLE (x87 80-bit for the
 * pair math and the sums; used to measure the FP64 paths' own rounding).
 *
 * What pins each function (the reference holds no golden vectors for this path):
 *   oracle_eval_plummer    60-digit decimal brute-force KATs (tests/golden/plummer_kat.npz),
 *       closed forms (one source, far field, eps -> 0, symmetric pair, Plummer sphere),
 *       exact invariants (linearity, lattice translation, superposition, third law,
 *       a = -grad phi) and planted mistakes that must fail (tests/test_oracle.py,
 *       tests/test_oracle_properties.py).
 *   oracle_eval_flash      closed-form kernel limits and table values, continuity at the
 *       radius; the periodic sum against independently replicated, cutoff-filtered images
 *       (numpy) and symmetry points; planted mistakes (tests/test_oracle.py).
 *   oracle_bidirectional   equals the two directed sums of oracle_eval_plummer; the third
 *       law (tests/test_oracle.py, tests/test_oracle_properties.py).
 *   oracle_kick            exactness against numpy (tests/test_oracle.py).
 *   oracle_block_centres   numpy meshgrid('ij') / ravel('C') (tests/test_oracle.py).
 *   oracle_pm_*            a particle on a cell centre, quarter-cell splits, mass
 *       conservation, linearity, central differences exact on quadratics, CIC exact on
 *       linear fields, the deposit/interpolation adjoint identity, planted mistakes
 *       (tests/test_oracle_pm.py).
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <omp.h>

#define ORACLE_SEQ 0
#define ORACLE_NEUMAIER 1
#define ORACLE_LONGDOUBLE 2

typedef struct { double s, c; } nsum;

static inline void nadd(nsum* a, double v) {
  double t = a->s + v;
  if (fabs(a->s) >= fabs(v)) a->c += (a->s - t) + v;
  else a->c += (v - t) + a->s;
  a->s = t;
}

static inline double sq(double v) { return v * v; }

int oracle_version(void) { return 5; }

/* Plummer direct sum.  Any of se/te may be NULL (= 0).  Any of ax/ay/az/phi may be NULL. */
int oracle_eval_plummer(int64_t ns, const double* sx, const double* sy, const double* sz,
                        const double* sm, const double* se,
                        int64_t nt, const double* tx, const double* ty, const double* tz,
                        const double* te, double G,
                        double* ax, double* ay, double* az, double* phi,
                        int mode, int nthreads) {
  if (ns < 0 || nt < 0) return -1;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t i = 0; i < nt; ++i) {
    const double xi = tx[i], yi = ty[i], zi = tz[i];
    const double e2t = te ? te[i] * te[i] : 0.0;
    double gx = 0, gy = 0, gz = 0, gp = 0;
    if (mode == ORACLE_SEQ) {
      for (int64_t j = 0; j < ns; ++j) {
        const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
        const double e2 = (se ? se[j] * se[j] : 0.0) + e2t;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        if (r2 == 0.0) continue;
        const double rinv = 1.0 / sqrt(r2);
        const double mr = sm[j] * rinv;
        const double w = mr * rinv * rinv;
        gx += w * dx; gy += w * dy; gz += w * dz; gp += mr;
      }
    } else if (mode == ORACLE_NEUMAIER) {
      nsum sx_ = {0, 0}, sy_ = {0, 0}, sz_ = {0, 0}, sp_ = {0, 0};
      for (int64_t j = 0; j < ns; ++j) {
        const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
        const double e2 = (se ? se[j] * se[j] : 0.0) + e2t;
        const double r2 = dx * dx + dy * dy + dz * dz + e2;
        if (r2 == 0.0) continue;
        const double rinv = 1.0 / sqrt(r2);
        const double mr = sm[j] * rinv;
        const double w = mr * rinv * rinv;
        nadd(&sx_, w * dx); nadd(&sy_, w * dy); nadd(&sz_, w * dz); nadd(&sp_, mr);
      }
      gx = sx_.s + sx_.c; gy = sy_.s + sy_.c; gz = sz_.s + sz_.c; gp = sp_.s + sp_.c;
    } else {
      long double lx = 0, ly = 0, lz = 0, lp = 0;
      const long double e2tl = te ? (long double)te[i] * (long double)te[i] : 0.0L;
      for (int64_t j = 0; j < ns; ++j) {
        const long double dx = (long double)sx[j] - xi, dy = (long double)sy[j] - yi,
                          dz = (long double)sz[j] - zi;
        const long double e2 = (se ? (long double)se[j] * (long double)se[j] : 0.0L) + e2tl;
        const long double r2 = dx * dx + dy * dy + dz * dz + e2;
        if (r2 == 0.0L) continue;
        const long double rinv = 1.0L / sqrtl(r2);
        const long double mr = (long double)sm[j] * rinv;
        const long double w = mr * rinv * rinv;
        lx += w * dx; ly += w * dy; lz += w * dz; lp += mr;
      }
      const long double Gl = G;
      if (ax) ax[i] = (double)(Gl * lx);
      if (ay) ay[i] = (double)(Gl * ly);
      if (az) az[i] = (double)(Gl * lz);
      if (phi) phi[i] = (double)(-Gl * lp);
      continue;
    }
    if (ax) ax[i] = G * gx;
    if (ay) ay[i] = G * gy;
    if (az) az[i] = G * gz;
    if (phi) phi[i] = -G * gp;
  }
  return 0;
}

/* FLASH softening kernel at distance r from a cell (Gravity_getAccelAtPoint.F90:212-241,
 * Gravity_getPotentialAtPoint.F90:197-222).  type 1 = spline, 2 = linear.  Returns the force
 * kernel kv (a pair adds prefactor * kv * d, d = x_t - x_c) and the potential kernel pk (a pair
 * adds prefactor * pk).  quirks != 0 reproduces the reference potential as written: the spline
 * branch's pot_kernel (with its -q^4/10 term) multiplied by prefactor = -G m
 * (Gravity_getPotentialAtPoint.F90:203-211,225) and `pgpot = slope` in the linear branch
 * (:215-217).  quirks == 0 uses the standard M4-spline potential (Price & Monaghan 2007,
 * eq. A2 with h = rsoft/2) and the potential of the linear force, (r^2 - 3 rsoft^2)/(2 rsoft^3),
 * both continuous with -1/r at r = rsoft.  The reference leaves the spline force kernel
 * unassigned for q <= 1e-5 (a stale value from an earlier pair); both modes use its r -> 0
 * limit there, h^-3 (4/3 - 6/5 q^2 + 1/2 q^3), times d (so a coincident pair adds 0). */
static void flash_kernel(double r, double rsoft, double hinv, double h2inv, double slope,
                         int type, int quirks, double* kvp, double* pkp) {
  double kv, pk;
  if (r < rsoft) {
    if (type == 1) {
      const double q = r * hinv;
      if (q < 1.0) {
        if (q > 1.0e-5) kv = h2inv * (4.0 / 3.0 * q - 1.2 * q * q * q + 0.5 * q * q * q * q) / r;
        else kv = h2inv * hinv * (4.0 / 3.0 - 1.2 * q * q + 0.5 * q * q * q);
        if (quirks)
          pk = hinv * (2.0 / 3.0 * q * q - 0.1 * pow(q, 4) + 0.1 * pow(q, 5) - 7.0 / 5.0);
        else
          pk = -hinv * (2.0 / 3.0 * q * q - 0.3 * pow(q, 4) + 0.1 * pow(q, 5) - 7.0 / 5.0);
      } else {
        kv = h2inv * (8.0 / 3.0 * q - 3.0 * q * q + 1.2 * q * q * q - 1.0 / 6.0 * pow(q, 4) -
                      1.0 / (15.0 * q * q)) / r;
        pk = hinv * (4.0 / 3.0 * q * q - q * q * q + 0.3 * pow(q, 4) - 1.0 / 30.0 * pow(q, 5) -
                     8.0 / 5.0 + 1.0 / (15.0 * q));
        if (!quirks) pk = -pk;
      }
    } else {
      kv = slope;
      pk = quirks ? slope : -(r * r - 3.0 * rsoft * rsoft) / (2.0 * rsoft * rsoft * rsoft);
    }
  } else {
    kv = 1.0 / (r * r * r);
    pk = 1.0 / r;
  }
  *kvp = kv;
  *pkp = pk;
}

/* FLASH direct sum over cells with the global softening radius `rsoft`
 * (Gravity_getAccelAtPoint.F90:140-145, 168-248; Gravity_getPotentialAtPoint.F90:137-227).
 * Sum order and sign follow the reference: d = x_t - x_c, g += prefactor * k * d with
 * prefactor = -G m; radius = sqrt(dx*dx + dy*dy + dz*dz) summed left to right.
 *
 * nrep > 0 selects the periodic branch (Gravity_getAccelAtPoint.F90:250-312,
 * Gravity_getPotentialAtPoint.F90:229-283; the reference fixes nrep_pbc = 6, :60): for every
 * cell the images d + (nx Lx, ny Ly, nz Lz), n in [-nrep, nrep]^3 (nz fastest), whose radius is
 * below maxradius_pbc = nrep * min(Lx, Ly, Lz) (:125) are summed with the same kernel,
 * dx = dx_inside + nx*Lx and radius = sqrt(x2 + y2 + z2) as written there.  The periodic
 * potential of the reference *assigns* pgpot per image (:262-274), so with quirks != 0 a cell
 * contributes the kernel of its last image inside the cutoff; with quirks == 0 the images'
 * kernels are summed. */
int oracle_eval_flash(int64_t ns, const double* sx, const double* sy, const double* sz,
                      const double* sm, int64_t nt, const double* tx, const double* ty,
                      const double* tz, double G, double rsoft, int type, int quirks, int nrep,
                      double Lx, double Ly, double Lz,
                      double* ax, double* ay, double* az, double* phi, int nthreads) {
  if (ns < 0 || nt < 0 || !(rsoft > 0) || (type != 1 && type != 2) || nrep < 0) return -1;
  if (nrep > 0 && !(Lx > 0 && Ly > 0 && Lz > 0)) return -1;
  if (nthreads < 1) nthreads = 1;
  const double hinv = 2.0 / rsoft, h2inv = hinv * hinv, slope = 1.0 / (rsoft * rsoft * rsoft);
  double lmin = Lx < Ly ? Lx : Ly;
  lmin = lmin < Lz ? lmin : Lz;
  const double rmax = (double)nrep * lmin;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t i = 0; i < nt; ++i) {
    double gx = 0, gy = 0, gz = 0, gp = 0;
    for (int64_t j = 0; j < ns; ++j) {
      const double pref = -G * sm[j];
      double kv, pk;
      if (nrep == 0) {
        const double dx = tx[i] - sx[j], dy = ty[i] - sy[j], dz = tz[i] - sz[j];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        flash_kernel(r, rsoft, hinv, h2inv, slope, type, quirks, &kv, &pk);
        gx += pref * (kv * dx); gy += pref * (kv * dy); gz += pref * (kv * dz);
        gp += pref * pk;
      } else {
        const double dxi = tx[i] - sx[j], dyi = ty[i] - sy[j], dzi = tz[i] - sz[j];
        double px = 0, py = 0, pz = 0, pp = 0;
        for (int nx = -nrep; nx <= nrep; ++nx) {
          const double dx = dxi + nx * Lx, x2 = dx * dx;
          for (int ny = -nrep; ny <= nrep; ++ny) {
            const double dy = dyi + ny * Ly, y2 = dy * dy;
            for (int nz = -nrep; nz <= nrep; ++nz) {
              const double dz = dzi + nz * Lz, z2 = dz * dz;
              const double r = sqrt(x2 + y2 + z2);
              if (r < rmax) {
                flash_kernel(r, rsoft, hinv, h2inv, slope, type, quirks, &kv, &pk);
                px += kv * dx; py += kv * dy; pz += kv * dz;
                if (quirks) pp = pk;
                else pp += pk;
              }
            }
          }
        }
        gx += pref * px; gy += pref * py; gz += pref * pz;
        gp += pref * pp;
      }
    }
    if (ax) ax[i] = gx;
    if (ay) ay[i] = gy;
    if (az) az[i] = gz;
    if (phi) phi[i] = gp;
  }
  return 0;
}

/* Fused bidirectional pass (Particles_sinkAccelGasOnSinksAndSinksOnGas.F90:204-316), Plummer
 * kernel with the symmetric eps2 = eps_c^2 + eps_p^2.  Cells outer, particles inner; the
 * particle accelerations are scattered in the inner loop exactly like
 * particles_global(accxProp,p) there (:296-301), so this restatement is serial.
 * Outputs: (pax,pay,paz,pphi) on particles from cells, (cax,cay,caz,cphi) on cells from
 * particles.  Any output may be NULL. */
int oracle_bidirectional(int64_t nc, const double* cx, const double* cy, const double* cz,
                         const double* cm, const double* ce,
                         int64_t np, const double* px, const double* py, const double* pz,
                         const double* pm, const double* pe, double G,
                         double* pax, double* pay, double* paz, double* pphi,
                         double* cax, double* cay, double* caz, double* cphi) {
  if (nc < 0 || np < 0) return -1;
  for (int64_t p = 0; p < np; ++p) {
    if (pax) pax[p] = 0; if (pay) pay[p] = 0; if (paz) paz[p] = 0; if (pphi) pphi[p] = 0;
  }
  for (int64_t c = 0; c < nc; ++c) {
    double gx = 0, gy = 0, gz = 0, gp = 0;
    const double e2c = ce ? ce[c] * ce[c] : 0.0;
    for (int64_t p = 0; p < np; ++p) {
      const double dx = px[p] - cx[c], dy = py[p] - cy[c], dz = pz[p] - cz[c];
      const double r2 = dx * dx + dy * dy + dz * dz + e2c + (pe ? pe[p] * pe[p] : 0.0);
      if (r2 == 0.0) continue;
      const double rinv = 1.0 / sqrt(r2);
      const double r3 = rinv * rinv * rinv;
      /* cell -> particle: d points from the particle to the cell, i.e. -d */
      const double fc = G * cm[c] * r3;
      if (pax) pax[p] -= fc * dx;
      if (pay) pay[p] -= fc * dy;
      if (paz) paz[p] -= fc * dz;
      if (pphi) pphi[p] -= G * cm[c] * rinv;
      const double wp = pm[p] * r3;
      gx += wp * dx; gy += wp * dy; gz += wp * dz; gp += pm[p] * rinv;
    }
    if (cax) cax[c] = G * gx;
    if (cay) cay[c] = G * gy;
    if (caz) caz[c] = G * gz;
    if (cphi) cphi[c] = -G * gp;
  }
  return 0;
}

/* v += dt * a, FP64 (the bridge's half kick uses dt/2). */
int oracle_kick(int64_t n, double dt, const double* ax, const double* ay, const double* az,
                double* vx, double* vy, double* vz) {
  if (n < 0) return -1;
  for (int64_t i = 0; i < n; ++i) {
    vx[i] += dt * ax[i];
    vy[i] += dt * ay[i];
    vz[i] += dt * az[i];
  }
  return 0;
}

/* Cell centres of nblk cubic leaf blocks of nb^3 cells: block b has lower corner
 * (ox[b],oy[b],oz[b]) and cell width dx[b]; cell (i,j,k) has centre o + (idx + 1/2) dx and
 * flat index b*nb^3 + (i*nb + j)*nb + k (x slowest, z fastest). */
int oracle_block_centres(int64_t nblk, int nb, const double* ox, const double* oy,
                         const double* oz, const double* dx, double* x, double* y, double* z) {
  if (nblk < 0 || nb <= 0) return -1;
  const int64_t nc = (int64_t)nb * nb * nb;
  for (int64_t b = 0; b < nblk; ++b)
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < nb; ++k) {
          const int64_t t = b * nc + ((int64_t)i * nb + j) * nb + k;
          x[t] = ox[b] + (i + 0.5) * dx[b];
          y[t] = oy[b] + (j + 0.5) * dx[b];
          z[t] = oz[b] + (k + 0.5) * dx[b];
        }
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * Particle-mesh route on FLASH leaf blocks (the pieces the bridge kick sits beside).
 *
 * Block layout: one variable of nblk blocks of nbg^3 cells, nbg = nb + 2 ng (nb interior
 * cells per axis, ng guard layers), block b at offset b * nbg^3, cell (i, j, k) at
 * (k * nbg + j) * nbg + i -- i (x) fastest, the order of solnData(VAR, i, j, k) for one VAR.
 * Block b has interior lower corner (ox, oy, oz)[b] and physical size bs[b] (cubic); the
 * cell width is del = bs / nb (Grid_getDeltas) and guarded cell i has centre
 * ox + ((i - ng) + 0.5) del.
 */

/* Cloud-in-cell weights along one axis: the point's position u in cell-centre units of the
 * guarded block (u = (x - o) / del - 0.5 + ng), base cell i0 = floor(u), weights (1 - f, f)
 * on (i0, i0 + 1) with f = u - i0.  Returns 0 if the two cells lie inside [0, nbg). */
static int cic_axis(double x, double o, double del, int ng, int nbg, int* i0, double* w0,
                    double* w1) {
  const double u = (x - o) / del - 0.5 + (double)ng;
  const double fl = floor(u);
  if (!(fl >= 0.0 && fl + 1.0 < (double)nbg)) return -1;
  *i0 = (int)fl;
  *w1 = u - fl;
  *w0 = 1.0 - *w1;
  return 0;
}

/* CIC deposit of particle masses as density, rho += m w / del^3, onto the guarded block
 * blk[p] of each particle, particles in index order, corners with dk slowest, di fastest;
 * w = (wx * wy) * wz.  Restates the mapping that Particles_updateGridVar(MASS_PART_PROP,
 * PDEN_VAR) performs through FLASH's Grid_mapParticlesToMesh (called at
 * Gravity_potentialListOfBlocksPDENonly.F90:260-266 and
 * Particles_updateGridVarMassiveOnly.F90:40-58); that mapper belongs to FLASH 4.6.2, which
 * is not vendored, so this is the textbook cloud-in-cell assignment (Hockney & Eastwood,
 * Computer Simulation Using Particles, 1988, sec. 5-2-2).  Guard-cell contributions stay
 * in the guard cells (folding them into neighbouring blocks is FLASH's Grid communication).
 * Returns -2 if a particle's cloud leaves its guarded block, -1 on bad arguments. */
int oracle_pm_deposit(int64_t np, const double* x, const double* y, const double* z,
                      const double* m, const int64_t* blk, int64_t nblk, int nb, int ng,
                      const double* ox, const double* oy, const double* oz, const double* bs,
                      double* rho) {
  if (np < 0 || nblk < 0 || nb < 1 || ng < 1) return -1;
  const int nbg = nb + 2 * ng;
  const int64_t nc = (int64_t)nbg * nbg * nbg;
  for (int64_t p = 0; p < np; ++p) {
    const int64_t b = blk[p];
    if (b < 0 || b >= nblk) return -2;
    const double del = bs[b] / (double)nb;
    const double md = m[p] / (del * del * del);
    int i0, j0, k0;
    double wx[2], wy[2], wz[2];
    if (cic_axis(x[p], ox[b], del, ng, nbg, &i0, &wx[0], &wx[1]) ||
        cic_axis(y[p], oy[b], del, ng, nbg, &j0, &wy[0], &wy[1]) ||
        cic_axis(z[p], oz[b], del, ng, nbg, &k0, &wz[0], &wz[1]))
      return -2;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int64_t idx = b * nc + ((int64_t)(k0 + dk) * nbg + (j0 + dj)) * nbg + (i0 + di);
          rho[idx] += md * ((wx[di] * wy[dj]) * wz[dk]);
        }
  }
  return 0;
}

/* Finite-difference acceleration g = -grad(phi) on every row of every guarded block, the
 * row rule of Gravity_accelOneRow (src/flash/source/Simulation/SimulationMain/Cube/
 * Gravity_accelOneRow.F90:94-121, FLASH's Gravity_accelListOfBlocks applies it to each row):
 * delxinv = 0.5 * (nb / bs), g(i) = delxinv * (phi(i-1) - phi(i+1)) for the cells between
 * the two ends of the row, and each end copies its neighbour ("invalid data - must not be
 * used").  Any of gx/gy/gz may be NULL. */
int oracle_pm_gradient(int64_t nblk, int nb, int ng, const double* bs, const double* phi,
                       double* gx, double* gy, double* gz) {
  if (nblk < 0 || nb < 1 || ng < 1) return -1;
  const int n = nb + 2 * ng;
  const int64_t nc = (int64_t)n * n * n;
  for (int64_t b = 0; b < nblk; ++b) {
    double delxinv = (double)nb / bs[b];
    delxinv = 0.5 * delxinv;
    const double* f = phi + b * nc;
    double* g[3] = {gx ? gx + b * nc : NULL, gy ? gy + b * nc : NULL, gz ? gz + b * nc : NULL};
    const int64_t stride[3] = {1, n, (int64_t)n * n};
    for (int a = 0; a < 3; ++a) {
      if (!g[a]) continue;
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const int c[3] = {i, j, k};
            const int64_t idx = ((int64_t)k * n + j) * n + i;
            int ii = c[a];
            if (ii == 0) ii = 1;
            if (ii == n - 1) ii = n - 2;
            const int64_t ctr = idx + (int64_t)(ii - c[a]) * stride[a];
            g[a][idx] = delxinv * (f[ctr - stride[a]] - f[ctr + stride[a]]);
          }
    }
  }
  return 0;
}

/* CIC interpolation of a block field to particle positions, a_p = sum_corners w g(corner)
 * with the weights and corner order of oracle_pm_deposit: the mapping Grid_mapMeshToParticles
 * performs in Particles_longRangeForce (src/flash/source/Particles/ParticlesForces/longRange/
 * gravity/ParticleMesh/Particles_longRangeForce.F90:70-97) -- FLASH 4.6.2's mapper, not
 * vendored, restated as the textbook CIC interpolation (the adjoint of the deposit).
 * Returns -2 if a particle's stencil leaves its guarded block. */
int oracle_pm_interpolate(int64_t np, const double* x, const double* y, const double* z,
                          const int64_t* blk, int64_t nblk, int nb, int ng, const double* ox,
                          const double* oy, const double* oz, const double* bs,
                          const double* gx, const double* gy, const double* gz, double* ax,
                          double* ay, double* az) {
  if (np < 0 || nblk < 0 || nb < 1 || ng < 1) return -1;
  const int nbg = nb + 2 * ng;
  const int64_t nc = (int64_t)nbg * nbg * nbg;
  for (int64_t p = 0; p < np; ++p) {
    const int64_t b = blk[p];
    if (b < 0 || b >= nblk) return -2;
    const double del = bs[b] / (double)nb;
    int i0, j0, k0;
    double wx[2], wy[2], wz[2];
    if (cic_axis(x[p], ox[b], del, ng, nbg, &i0, &wx[0], &wx[1]) ||
        cic_axis(y[p], oy[b], del, ng, nbg, &j0, &wy[0], &wy[1]) ||
        cic_axis(z[p], oz[b], del, ng, nbg, &k0, &wz[0], &wz[1]))
      return -2;
    double sx = 0, sy = 0, sz = 0;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int64_t idx = b * nc + ((int64_t)(k0 + dk) * nbg + (j0 + dj)) * nbg + (i0 + di);
          const double w = (wx[di] * wy[dj]) * wz[dk];
          sx += w * gx[idx];
          sy += w * gy[idx];
          sz += w * gz[idx];
        }
    ax[p] = sx;
    ay[p] = sy;
    az[p] = sz;
  }
  return 0;
}
