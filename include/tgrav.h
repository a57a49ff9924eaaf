/*
 * tgrav -- MI355X (gfx950) softened direct-sum gravity field, C ABI.
 *
 * Drop-in replacement for the CPU field evaluation behind Torch's gas/star
 * gravity bridge.  Units are the caller's; the default constant is
 * G = 4.300917270e-3 pc (km/s)^2 / Msun.  All arrays are FP64 SoA of equal
 * length; ptr_kind says whether they are host or device (hipMalloc'd, e.g. a
 * PyTorch-ROCm tensor's data_ptr()) pointers.
 *
 * Pair kernel (Plummer): eps2_ij = eps_s_j^2 + eps_t_i^2,
 *   a_i   =  G sum_j m_j (x_j - x_i) (|x_j - x_i|^2 + eps2_ij)^(-3/2)
 *   phi_i = -G sum_j m_j             (|x_j - x_i|^2 + eps2_ij)^(-1/2)
 * A pair with |x_j - x_i|^2 + eps2_ij == 0 contributes nothing (self-exclusion).
 *
 * Conventions (mirroring the AMUSE legacy-function boundary):
 *   - every entry point returns 0 on success and a negative TGRAV_E* code on
 *     failure (AMUSE's ERROR_CODE, src/amuse/torch_amuse_flash/interface.py:1436-1440);
 *     tgrav_last_error() then describes the failure (thread-local string);
 *   - the caller owns all input/output memory; the context owns its scratch;
 *   - a context is bound to one device and one HIP stream and is not thread-safe
 *     (like the single-threaded AMUSE worker); separate contexts may run concurrently;
 *   - with TGRAV_HOST pointers a call is synchronous (outputs are valid on return);
 *     with TGRAV_DEVICE pointers work is enqueued on the context's stream and the
 *     caller synchronises (tgrav_synchronize, or stream order).
 */
#ifndef TGRAV_H
#define TGRAV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TGRAV_PLAIN_FP64 0   /* FP64 throughout; matches the FP64 oracle to <= 1e-12 */
#define TGRAV_MIXED 1        /* FP64 displacements, FP32 bulk, compensated sums, FP64 kick */

#define TGRAV_HOST 0
#define TGRAV_DEVICE 1

#define TGRAV_SOFT_PLUMMER 0       /* eps2_ij = eps_s^2 + eps_t^2 (the default) */
#define TGRAV_SOFT_FLASH_SPLINE 1  /* FLASH sink_softening_type_gas = "spline" */
#define TGRAV_SOFT_FLASH_LINEAR 2  /* FLASH sink_softening_type_gas = "linear" */
#define TGRAV_FLASH_QUIRKS 1       /* potential exactly as the reference computes it */

#define TGRAV_OK 0
#define TGRAV_EINVAL -1      /* bad argument: n < 0, NULL required pointer, bad mode */
#define TGRAV_ENOSOURCES -2  /* evaluation before tgrav_set_sources */
#define TGRAV_EHIP -3        /* a HIP runtime call or kernel launch failed */
#define TGRAV_ENONFINITE -4  /* validation found NaN/Inf in an input */
#define TGRAV_ENOMEM -5      /* device allocation failed */
#define TGRAV_ELIMIT -6      /* size beyond what the library supports */

typedef struct tgrav_ctx tgrav_ctx;
typedef struct tgrav_batch tgrav_batch;

/* One problem of a batched evaluation (device pointers).  eps arrays may be NULL
 * (= 0); any output may be NULL; v* may be NULL unless the batch kicks.
 * n_target_blocks > 0 gives the targets as that many 8^3 leaf blocks instead of tx/ty/tz/te:
 * lower corners (tbx, tby, tbz), cell widths tbd and softening tbe (NULL = 0) per block, and
 * n_targets must be 512 * n_target_blocks; outputs and velocities are then in the blocks' C
 * order (as tgrav_eval_blocks), and the plan evaluates them with the column kernel.
 * n_source_blocks > 0 gives the sources as that many 8^3 leaf blocks instead of sx/sy/sz/se:
 * lower corners (sbx, sby, sbz), cell widths sbd and softening sbe (NULL = 0) per block, and
 * the cell masses in sm, C order b*512 + (i*8 + j)*8 + k (as tgrav_set_source_blocks);
 * n_sources must be 512 * n_source_blocks, and the plan walks them block by block and column by
 * column (the per-block source loop of Gravity_getAccelAtPoint.F90:168-204). */
typedef struct tgrav_problem {
  int64_t n_sources;
  const double *sx, *sy, *sz, *sm, *se;
  int64_t n_targets;
  const double *tx, *ty, *tz, *te;
  double *ax, *ay, *az, *phi;
  double *vx, *vy, *vz;
  int64_t n_target_blocks;
  const double *tbx, *tby, *tbz, *tbd, *tbe;
  int64_t n_source_blocks;
  const double *sbx, *sby, *sbz, *sbd, *sbe;
} tgrav_problem;

/* ---- context ---------------------------------------------------------- */
int tgrav_version(void);
const char* tgrav_last_error(void);
/* Replaces the worker start-up of the AMUSE plugin (interface.F90 initialize_code) for
 * this path: binds a device, creates a default-priority stream.  tune_cache_path (may be
 * NULL) names the autotuner's JSON cache (torch_amd/tune_cache.json, written by
 * python -m torch_amd.autotune): evaluations, column-kernel launches and batch plans then
 * take the tuned launch shape for their device and size bucket, unless tgrav_set_config /
 * tgrav_set_batch_config set one explicitly. */
int tgrav_create(tgrav_ctx** out, int device, const char* tune_cache_path);
/* Number of entries in a tune cache file, or TGRAV_EINVAL if it cannot be read or parsed
 * (needs no device). */
int tgrav_tune_cache_entries(const char* path);
int tgrav_destroy(tgrav_ctx* ctx);
/* Run on an existing HIP stream of the same HIP runtime (e.g.
 * torch.cuda.current_stream().cuda_stream); NULL is HIP's null stream. */
int tgrav_set_stream(tgrav_ctx* ctx, void* hip_stream);
/* Back to the context's own (default-priority, non-blocking) stream, the initial state. */
int tgrav_use_own_stream(tgrav_ctx* ctx);
void* tgrav_get_stream(tgrav_ctx* ctx);
int tgrav_synchronize(tgrav_ctx* ctx);
/* Gravitational constant; the reference reads FLASH's `Newton`
 * (Gravity_getAccelAtPoint.F90:134). */
int tgrav_set_constant(tgrav_ctx* ctx, double G);
/* Validation of inputs for NaN/Inf (default on).  With device pointers it costs a
 * device->host round trip per call. */
int tgrav_set_validation(tgrav_ctx* ctx, int on);
/* Kernel configuration: variant index (see tgrav_variant_info) and the number of
 * source chunks (0 = automatic).  Used by the autotuner; -1 leaves a field as is. */
int tgrav_set_config(tgrav_ctx* ctx, int mode, int variant, int nsplit);
int tgrav_num_variants(int mode);
int tgrav_variant_info(int mode, int variant, int* targets_per_lane, int* threads,
                       int* tile_sources);

/* ---- the field -------------------------------------------------------- */
/* Source set: the stellar side (stars, for the forward direction of
 * get_gravity_at_point, src/torch_mainloop.py:128-140) or the gas cells as point
 * masses m = rho dV (for the reverse direction, Gravity_getAccelAtPoint.F90:168-204).
 * eps may be NULL (= 0).  The arrays are copied into the context (packed FP64).  Sources
 * that are 8^3 leaf blocks flattened in C order (each aligned 512 a tensor-product block
 * with one eps) are recognised on the device and evaluated as source blocks. */
int tgrav_set_sources(tgrav_ctx* ctx, int64_t n, const double* x, const double* y,
                      const double* z, const double* m, const double* eps, int ptr_kind);

/* Source set given as 8^3 leaf blocks (FLASH's NXB = 8): block b has lower corner
 * (ox,oy,oz)[b], cell width dx[b] and softening eps[b] (NULL = 0); m holds the nblk*512
 * cell masses rho dV in the C order b*512 + (i*8 + j)*8 + k -- the leaf-cell loop of
 * Gravity_getAccelAtPoint.F90:168-204 as a flat array.  Evaluations then walk the cells
 * column by column (dx, dy shared by the 8 cells of a column).  nb must be 8. */
int tgrav_set_source_blocks(tgrav_ctx* ctx, int64_t nblk, int nb, const double* ox,
                            const double* oy, const double* oz, const double* dx,
                            const double* m, const double* eps, int ptr_kind);

/* Acceleration and potential at n points from the source set.  Replaces the per-point
 * loops of get_accel_gas_on_particles (interface.F90:1130-1216) and
 * get_potential_at_point (interface.F90:1073-1127) -> Gravity_getAccelAtPoint
 * (Gravity_getAccelAtPoint.F90:32-419) / Gravity_getPotentialAtPoint
 * (Gravity_getPotentialAtPoint.F90:33-305), and the stellar-side AMUSE
 * get_gravity_at_point / get_potential_at_point(eps, x, y, z) (interface.py:221-243).
 * eps_t is the per-point softening (may be NULL = 0); any of ax/ay/az/phi may be NULL
 * (phi == NULL skips the potential).  Points that are 8^3 leaf-block meshgrids flattened in
 * C order (.deprecated/bridge_multiples.py:566-575: each aligned 8 share x and y, each
 * aligned 512 share eps) are recognised on the device and take the column kernel. */
int tgrav_eval(tgrav_ctx* ctx, int64_t n, const double* eps_t, const double* x,
               const double* y, const double* z, double* ax, double* ay, double* az,
               double* phi, int mode, int ptr_kind);

/* Evaluation with the bridge kick fused: v += dt * a in FP64 (the gas kick of
 * interface.F90:1793-1795 and the star kick of interface.F90:1432-1444; the bridge
 * passes dt/2, src/torch_mainloop.py:287,289).  ax..phi may be NULL. */
int tgrav_eval_kick(tgrav_ctx* ctx, int64_t n, const double* eps_t, const double* x,
                    const double* y, const double* z, double dt, double* vx, double* vy,
                    double* vz, double* ax, double* ay, double* az, double* phi, int mode,
                    int ptr_kind);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  tgrav_eval's direct sum
 * (generic kernels, same launch shape) with the source set passed in instead of set beforehand,
 * device pointers only, no validation; it calls no hipMalloc/hipFree.  The packed sources (40 B
 * each) and the chunk partial sums (32 B per target per source chunk) live in a fixed 64 MiB array
 * per GPU, held until the process ends: sources beyond what fits (~1.6 million) give TGRAV_ELIMIT,
 * and the source split is reduced until the partial sums fit.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (several streams or threads) race.  FLASH kernels and
 * periodic images are not available here (TGRAV_EINVAL). */
int tgrav_eval_static(tgrav_ctx* ctx, int64_t ns, const double* sx, const double* sy,
                      const double* sz, const double* sm, const double* se, int64_t nt,
                      const double* eps_t, const double* x, const double* y, const double* z,
                      double* ax, double* ay, double* az, double* phi, int mode);

/* Standalone FP64 kick v += dt * a (.deprecated/bridge_multiples.py:612-616). */
int tgrav_kick(tgrav_ctx* ctx, int64_t n, double dt, const double* ax, const double* ay,
               const double* az, double* vx, double* vy, double* vz, int ptr_kind);

/* Block-structured targets: nblk cubic leaf blocks of nb^3 cells, block b with lower
 * corner (ox,oy,oz)[b] and cell width dx[b]; cell (i,j,k) centre o + (idx+1/2) dx,
 * output index b*nb^3 + (i*nb + j)*nb + k (the meshgrid('ij')/ravel('C') order of
 * .deprecated/bridge_multiples.py:566-571, FLASH's get_1blk_cell_coords,
 * interface.F90:418-437).  eps_b is the per-block target softening (may be NULL).
 * vx/vy/vz may be NULL (no kick) or per-cell velocities kicked by dt. */
int tgrav_eval_blocks(tgrav_ctx* ctx, int64_t nblk, int nb, const double* ox,
                      const double* oy, const double* oz, const double* dx,
                      const double* eps_b, double* ax, double* ay, double* az, double* phi,
                      double dt, double* vx, double* vy, double* vz, int mode, int ptr_kind);

/* Fused bidirectional pass of Particles_sinkAccelGasOnSinksAndSinksOnGas
 * (Particles_sinkAccelGasOnSinksAndSinksOnGas.F90:47-411): gas cells (as point
 * masses, softening ce) and particles (softening pe), symmetric eps2 = ce^2 + pe^2;
 * returns the acceleration/potential of the cells from the particles and of the
 * particles from the cells in one pass over the pairs.  Outputs may be NULL.  Cells in
 * the C order of 8^3 leaf-block meshgrids (each aligned 8 share x and y, each aligned 512
 * share eps) are recognised on the device, without a host wait, and take the block-column
 * kernel of tgrav_bidirectional_blocks. */
int tgrav_bidirectional(tgrav_ctx* ctx, int64_t nc, const double* cx, const double* cy,
                        const double* cz, const double* cm, const double* ce, int64_t np,
                        const double* px, const double* py, const double* pz,
                        const double* pm, const double* pe, double* cax, double* cay,
                        double* caz, double* cphi, double* pax, double* pay, double* paz,
                        double* pphi, int mode, int ptr_kind);
/* Pair kernel of later tgrav_eval / tgrav_eval_kick / tgrav_eval_blocks calls.
 * TGRAV_SOFT_PLUMMER (default): the Plummer kernel above, per-point softening.
 * TGRAV_SOFT_FLASH_SPLINE / _LINEAR: FLASH's gas-gravity kernels with one global radius
 * rsoft (sink_softening_radius; Gravity_getAccelAtPoint.F90:88-100,140-145), per-point
 * softening ignored:  a_i = -G sum_j m_j k(r) (x_i - x_j),  phi_i = -G sum_j m_j p(r),
 * k = 1/r^3 and p = 1/r for r >= rsoft, the cubic spline (h = rsoft/2) or the constant
 * 1/rsoft^3 inside (:212-241).  flags = TGRAV_FLASH_QUIRKS takes the potential inside
 * rsoft as the reference computes it (Gravity_getPotentialAtPoint.F90:197-222: the spline
 * branch with the opposite sign, `pgpot = slope` in the linear one); without it the
 * potential is the one whose gradient is the force.  FP64 only (TGRAV_PLAIN_FP64); the
 * fused pass and batch plans stay Plummer. */
int tgrav_set_softening(tgrav_ctx* ctx, int kernel, double rsoft, int flags);
/* Periodic boundaries of the FLASH kernels (grav_boundary_type = "periodic",
 * Gravity_getAccelAtPoint.F90:101-131, 250-312): every cell also acts through its images
 * x_c - (nx lx, ny ly, nz lz), n in [-nrep, nrep]^3, that lie closer than
 * nrep * min(lx, ly, lz) to the point (the reference fixes nrep_pbc = 6, :60).  With
 * TGRAV_FLASH_QUIRKS the potential of a cell is that of its last image in the reference's
 * loop order (it assigns, not adds, per image, Gravity_getPotentialAtPoint.F90:262-274).
 * on = 0 restores isolated boundaries.  nrep in [1, 64].  on = 1 sorts the (target, cell)
 * pairs by separation and walks each wave's image rows together where nrep <= 10 and there are
 * at most 2^25 cells, in batches of targets of at most 2^25 pairs (it allocates ~40 B per pair
 * of a batch, up to 1.3 GB); on = 2 always walks each target's images on its own lane (the only
 * way above those limits).  Both give the reference's decisions. */
int tgrav_set_periodic(tgrav_ctx* ctx, int on, double lx, double ly, double lz, int nrep);

/* The fused pass with the gas cells handed over as 8^3 leaf blocks (nb must be 8): block b has
 * lower corner (ox, oy, oz)[b], cell width dx[b], softening ceps[b] (NULL = 0) and cell masses
 * cm[b*512 + (i*8 + j)*8 + k]; the cell outputs use the same C order.  Same pair set and
 * results as tgrav_bidirectional on the blocks' cell centres; each lane walks a column of 8
 * cells, so the pair arithmetic shared along the column is done once. */
int tgrav_bidirectional_blocks(tgrav_ctx* ctx, int64_t nblk, int nb, const double* ox,
                               const double* oy, const double* oz, const double* dx,
                               const double* cm, const double* ceps, int64_t np,
                               const double* px, const double* py, const double* pz,
                               const double* pm, const double* pe, double* cax, double* cay,
                               double* caz, double* cphi, double* pax, double* pay, double* paz,
                               double* pphi, int mode, int ptr_kind);
/* Block targets with nb == 8 use a column kernel (each lane owns one (i, j) column of 8
 * cells and shares dx, dy, dx^2 + dy^2 + eps^2 between them), and so do explicit arrays
 * recognised as blocks (tgrav_eval, tgrav_set_sources); on = 0 falls back to the generic
 * kernel (same results to rounding; for tests and comparisons).  Default on. */
int tgrav_set_block_kernel(tgrav_ctx* ctx, int on);
/* Waves per workgroup of the column kernels (1, 2 or 4, default 4; no counterpart in the reference, a
 * launch-shape knob like tgrav_set_config): the W waves of a workgroup hold the same leaf block
 * and split its sources, then add their partial sums in LDS in wave order, so the sources need
 * W times fewer chunks (and partial sums in memory) for the same number of waves.  Results are
 * deterministic for a given W and differ from another W's in rounding only.  Batch plans take
 * the value the context has when they are created.  tgrav_get_column_waves returns it. */
int tgrav_set_column_waves(tgrav_ctx* ctx, int waves);
int tgrav_get_column_waves(tgrav_ctx* ctx);
/* Fused-pass launch shape (variant, particle chunks; -1 keeps, 0 = automatic chunks). */
int tgrav_set_bidir_config(tgrav_ctx* ctx, int mode, int variant, int nsplit);
int tgrav_bidir_variant_info(int mode, int variant, int* targets_per_lane, int* threads);

/* ---- particle-mesh pieces beside the kick (FLASH leaf blocks) ----------- */
/* Block arrays hold one variable of nblk blocks of nbg^3 cells, nbg = nb + 2 ng <= 128 (nb interior
 * cells per axis, ng >= 1 guard layers); block b starts at b * nbg^3 and cell (i, j, k) is
 * at (k * nbg + j) * nbg + i, x fastest -- solnData(VAR, i, j, k) of one variable.  Block b
 * has interior lower corner (ox, oy, oz)[b] and physical size bsize[b] (Grid_getBlkPhysicalSize,
 * cubic); cells are bsize / nb wide and guarded cell i has centre o + (i - ng + 1/2) bsize/nb.
 * blk[p] is particle p's block (BLK_PART_PROP, 0-based).  A particle whose cloud leaves its
 * guarded block is an error (TGRAV_EINVAL) with validation on, and is skipped (deposit) or
 * gets NaN (interpolation) with it off. */

/* Cloud-in-cell mass deposit as density, rho = sum m w / (bsize/nb)^3 over the particles
 * (rho += ... with accumulate != 0): the mapping of Particles_updateGridVar(MASS_PART_PROP,
 * PDEN_VAR) through FLASH's Grid_mapParticlesToMesh (Gravity_potentialListOfBlocksPDENonly.F90:
 * 260-266).  Guard-cell contributions stay in the guard cells.  Deterministic: the sums are
 * formed in exact fixed point, independent of the order of the particles. */
int tgrav_pm_deposit(tgrav_ctx* ctx, int64_t np, const double* x, const double* y,
                     const double* z, const double* m, const int64_t* blk, int64_t nblk, int nb,
                     int ng, const double* ox, const double* oy, const double* oz,
                     const double* bsize, double* rho, int accumulate, int ptr_kind);
/* Finite-difference acceleration g = -grad(phi) on every cell of every guarded block, the
 * row rule of Gravity_accelOneRow (Simulation/SimulationMain/Cube/Gravity_accelOneRow.F90:
 * 94-121) as FLASH's Gravity_accelListOfBlocks applies it (interface.F90:1767-1769):
 * g(i) = 0.5 (nb / bsize) (phi(i-1) - phi(i+1)), the two end cells of each row copying their
 * neighbour.  gx/gy/gz may be NULL.  vx/vy/vz != NULL also apply the grid kick
 * v += g dt of interface.F90:1777-1830 in the same pass. */
int tgrav_pm_gradient(tgrav_ctx* ctx, int64_t nblk, int nb, int ng, const double* bsize,
                      const double* phi, double* gx, double* gy, double* gz, double dt,
                      double* vx, double* vy, double* vz, int ptr_kind);
/* Cloud-in-cell interpolation of a block vector field to the particles, a_p = sum w g: the
 * Grid_mapMeshToParticles step of Particles_longRangeForce (Particles_longRangeForce.F90:
 * 70-97). */
int tgrav_pm_interpolate(tgrav_ctx* ctx, int64_t np, const double* x, const double* y,
                         const double* z, const int64_t* blk, int64_t nblk, int nb, int ng,
                         const double* ox, const double* oy, const double* oz,
                         const double* bsize, const double* gx, const double* gy,
                         const double* gz, double* ax, double* ay, double* az, int ptr_kind);

/* ---- batched uneven evaluations --------------------------------------- */
/* Many independent problems in one persistent launch (device pointers only).  The plan
 * packs the sources, builds cost-sorted work lists (target tile x source chunk) and
 * owns the partial-sum scratch.  dt != 0 fuses the FP64 kick v += dt * a.  Explicit
 * targets in the C order of 8^3 leaf-block meshgrids (x, y shared by each aligned 8, eps by
 * each aligned 512) are evaluated like block targets: the plan tests that structure on the
 * device at every run (inside the graph too), so targets updated in place between runs are
 * always evaluated by a kernel that fits them.  Creation does not wait for the device. */
int tgrav_batch_create(tgrav_ctx* ctx, int32_t n_problems, const tgrav_problem* problems,
                       int mode, double dt, tgrav_batch** out);
/* Runs the plan.  use_graph != 0 captures the launch sequence (counter reset,
 * persistent evaluation, chunk reduction + kick) in a hipGraph on first use and
 * replays it afterwards. */
int tgrav_batch_run(tgrav_batch* b, int use_graph);
/* Destroy every plan before its context (tgrav_destroy): a plan uses the context's device and
 * stream. */
int tgrav_batch_destroy(tgrav_batch* b);
/* Work-list statistics: items, total interactions. */
int tgrav_batch_info(tgrav_batch* b, int64_t* n_items, double* interactions);
/* Batch plan shape for plans created afterwards: interactions per work item
 * (0 = total / (8 x CUs)) and persistent workgroups per CU (0 = occupancy limit).
 * The kernel variant is the one tgrav_set_config chose for the mode.  Set by the autotuner. */
int tgrav_set_batch_config(tgrav_ctx* ctx, double item_interactions, int wg_per_cu);
/* Plans created afterwards cut the problems of their column work list (targets that are 8^3 leaf
 * blocks, or explicit targets in their C order) into items of column_scale x the item size, and
 * those of their block-source work list (sources that are 8^3 leaf blocks) into items of
 * block_source_scale x it: every source chunk of a problem writes 32 B of partial sums per target,
 * and a list with many items balances with larger ones.  Scales in (0, 64]; both 0 return to the
 * tune cache's values (or 1).  Set by the autotuner.  No counterpart in the reference. */
int tgrav_set_batch_scales(tgrav_ctx* ctx, double column_scale, double block_source_scale);
/* Bytes of chunk partial sums a run of the plan writes (and its reduction reads back). */
int tgrav_batch_partials(tgrav_batch* b, int64_t* bytes);

/* ---- instrumentation (no counterpart in the reference) ------------------ */
/* on != 0: record HIP events on the context's stream around every evaluation kernel
 * (k_eval_*: the direct-sum kernel alone, not the chunk reduction), up to 256 launches;
 * also clears earlier records. */
int tgrav_set_timing(tgrav_ctx* ctx, int on);
/* Copies up to `max` records (milliseconds, targets, sources of each timed launch; waits
 * for them), clears the list and returns the number of records taken (>= 0). */
int tgrav_timing_read(tgrav_ctx* ctx, int max, double* ms, int64_t* n_targets,
                      int64_t* n_sources);

#ifdef __cplusplus
}
#endif

#endif /* TGRAV_H */
