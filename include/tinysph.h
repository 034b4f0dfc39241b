/*
 * tinysph.h -- C ABI of libtinysph_hip.so, the MI355X (gfx950) implementation of
 * the tinySPH per-step particle pipeline: cell-list build, pair loops and time
 * integration, run as hand-written HIP kernels on device-resident SoA data.
 *
 * Each entry point replaces one reference stage function that a case driver's
 * SPH_simulation::RUN() calls on a Particle_system (citations are to the
 * reference tree tomashalada/tinySPH_old).  Conventions:
 *   - every function returns 0 on success or a negative TSPH_E* code; the text
 *     of the last error is available from tsph_last_error();
 *   - one context owns one particle system on one GPU and one HIP stream; calls
 *     on a context are stream-ordered and not thread safe;
 *   - host arrays are in reference particle order (the order of Add_particle
 *     calls, src/SPH_particle_system.h:25-50), vectors interleaved as {x, y}
 *     like the reference's tdouble2 (src/SPH_defs.h:72-93).
 *   - stage calls are asynchronous; tsph_download / tsph_get_* synchronise.
 */
#ifndef TINYSPH_H
#define TINYSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSPH_ABI_VERSION 1

/* error codes */
#define TSPH_OK 0
#define TSPH_EINVAL -1   /* bad argument */
#define TSPH_EHIP -2     /* HIP runtime error */
#define TSPH_ESTATE -3   /* call out of order (e.g. forces before a cell list) */
#define TSPH_ECAPACITY -4/* particle count beyond the context capacity */
#define TSPH_EGRID -5    /* a particle lies outside the cell grid */

/* particle types, src/SPH_defs.h:44-65 */
#define TSPH_FLUID 0
#define TSPH_WALL 1
#define TSPH_INLET 6
#define TSPH_OUTLET 11

/* arithmetic of the pair loops */
#define TSPH_PREC_FP64 0  /* IEEE double with the reference's float-rounded kernel */
#define TSPH_PREC_MIXED 1 /* FP32 pair arithmetic on cell-relative FP32 positions,
                             FP64 integration of position and density */

/* pair-loop variants (which reference Compute_Forces is replaced) */
#define TSPH_FORCES_WCSPH 0 /* src/SPH_interaction.h:8-149 (fluid only; mDBC cases) */
#define TSPH_FORCES_DBC 1   /* src/SPH_interactionDBC.h:8-155 */
#define TSPH_FORCES_BT 2    /* src/SPH_interactionBT.h:8-192 */

/* wall treatment used by tsph_step_symplectic */
#define TSPH_WALL_DBC 0
#define TSPH_WALL_MDBC 1
#define TSPH_WALL_BT 2

/* SPH_constants + Simulation_data (src/SPH_data.h:30-46, 111-125).  The domain
 * is the padded one the drivers' PREP_SIMULATION_DATA produces
 * (e.g. cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:206-221).  Mass and
 * gravity are derived as in the Particle_system constructor
 * (src/SPH_particle_system.h:191-211). */
typedef struct tsph_params {
  double h, kap, cs, rho0, avisc, delta, dp;
  double x_0, x_m, y_0, y_m;
  int precision;   /* TSPH_PREC_* */
  int capacity;    /* maximum particle count held on the device */
} tsph_params;

/* Host view of a particle system in reference order.  Pointers the caller does
 * not need may be NULL (on upload: fields start as Add_particle sets them). */
typedef struct tsph_particles {
  int np;
  int32_t *type;        /* part_type */
  int32_t *index;       /* data.index (download only) */
  double *p, *rho, *drho;
  double *r, *v, *a;    /* 2*np each, {x,y} interleaved */
  double *r_o, *v_o, *rho_o;
  double *gnr;          /* mDBC ghost node per particle (2*np) */
  double *n;            /* BT wall normal per particle (2*np) */
} tsph_particles;

/* counters accumulated since the last tsph_reset_counters */
typedef struct tsph_counters {
  uint64_t pairs;        /* ordered (i,j), i active, |r_i-r_j| < 2h, over all pair passes */
  uint64_t candidates;   /* ordered (i,j), j != i in i's 9-cell stencil */
  uint64_t active;       /* active particles summed over pair passes */
  uint64_t pair_passes;  /* number of Compute_Forces calls */
  uint64_t density_out_of_range; /* rho > 1500 or < 0 seen by the EOS (src/SPH_pressure.h:76-80) */
  uint64_t outside_grid; /* particles that fell outside the cell grid */
  uint64_t removed;      /* particles removed out of the domain (np_out) */
  uint64_t appended;     /* particles appended by inlet buffer zones */
  uint64_t zone_rejected;/* inlet appends refused: capacity or the per-pass list limit */
  uint64_t zone_chained; /* appended particles already past the zone start (not re-examined) */
  uint64_t reinit_out_of_range; /* fluid densities outside [990, 1010] when reinitialised
                                   (src/SPH_density_reinitialization.h:133-134) */
  uint64_t scan_stalls;  /* cell-list scan blocks that outwaited their bounded wait for the
                            blocks before them and summed the counts themselves (results
                            unchanged; expected 0) */
} tsph_counters;

typedef struct tsph_ctx tsph_ctx;

const char *tsph_last_error(void);
int tsph_abi_version(void);

/* Particle_system constructor + Divide_To_Cells, src/SPH_particle_system.h:191-211,
 * src/SPH_find_pairs.h:5-42 */
int tsph_create(const tsph_params *prm, tsph_ctx **out);
int tsph_destroy(tsph_ctx *ctx);

/* grid size chosen by Divide_To_Cells (src/SPH_find_pairs.h:19-20) */
int tsph_grid(tsph_ctx *ctx, int *ncx, int *ncy);
int tsph_mass(tsph_ctx *ctx, double *m);

/* Replaces the host std::vector fields of Particle_system (a whole upload
 * stands for the sequence of Add_particle calls that built them). */
int tsph_upload(tsph_ctx *ctx, const tsph_particles *host);
/* Copies the current state back in reference order; host->np is set. */
int tsph_download(tsph_ctx *ctx, tsph_particles *host);
int tsph_get_np(tsph_ctx *ctx, int *np);

/* Clear_cells, src/SPH_find_pairs.h:74-92 */
int tsph_clear_cells(tsph_ctx *ctx);
/* Particle_To_Cell, src/SPH_find_pairs.h:45-72: cell hash, counting sort by
 * cell with reference in-cell order, cell start offsets, SoA reorder. */
int tsph_particle_to_cell(tsph_ctx *ctx);
/* Cell lists as CSR in reference particle indices: start[ncx*ncy+1], items[np]. */
int tsph_get_cell_lists(tsph_ctx *ctx, int32_t *start, int32_t *items);

/* Density_to_pressure, src/SPH_pressure.h:65-89 */
int tsph_density_to_pressure(tsph_ctx *ctx);
/* Compute_Forces, variant TSPH_FORCES_* */
int tsph_compute_forces(tsph_ctx *ctx, int variant);
/* mDBC_compute_density_mdbcGeo, src/SPH_mDBC.h:9-50 */
int tsph_mdbc_density(tsph_ctx *ctx);
/* BT_compute_density_BTGeo, src/SPH_BT.h:9-50 */
int tsph_bt_density(tsph_ctx *ctx);
/* Integrate_SymplecticPredictor / Corrector, src/SPH_integration.h:266-309 */
int tsph_symplectic_predictor(tsph_ctx *ctx, double dt);
int tsph_symplectic_corrector(tsph_ctx *ctx, double dt);
/* RemoveParticlesOutOfDomain, src/SPH_btdebug.h:4-31 (takes effect at the
 * next tsph_particle_to_cell, which the drivers always call next) */
int tsph_remove_out_of_domain(tsph_ctx *ctx);

/* One or more whole symplectic steps in the drivers' stage order
 * (cases/stillWater_DBC/SPH_simulationSymplectic.h:141-198,
 *  cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:59-116,
 *  cases/damBreakLong_BT/SPH_simulationSymplectic.h:66-122), replayed from a
 * captured HIP graph.  wall = TSPH_WALL_*. */
int tsph_step_symplectic(tsph_ctx *ctx, int wall, double dt, int nsteps);

/* Water_Elevation / Water_Elevation_inverse, src/SPH_water_level.h:87-150
 * (NaN where the reference leaves the level unset) */
int tsph_water_elevation(tsph_ctx *ctx, double x, double y0, int inverse, double *level);
/* Pressure_in_point -> Kernel_velocity_approximation_TEST, src/SPH_measure_pressure.h:6-15,
 * src/testing.h:6-148: out4 = {rho, p, vx, vy} */
int tsph_interpolate_point(tsph_ctx *ctx, double x, double y, double *out4);

int tsph_get_counters(tsph_ctx *ctx, tsph_counters *out);
int tsph_reset_counters(tsph_ctx *ctx);
/* Pair-loop tile: one wavefront handles `rows` consecutive cells of a grid
 * column (1..24, default 7). */
int tsph_set_tile_rows(tsph_ctx *ctx, int rows);
int tsph_get_tile_rows(tsph_ctx *ctx, int *rows);
/* Pair-loop waves launched per compute unit (1..32; default 12 on the mixed
 * path, 8 on the FP64 path); the waves walk the list of non-empty tiles. */
int tsph_set_waves_per_cu(tsph_ctx *ctx, int waves);
/* Polls a cell-list scan block makes for the published sums of an earlier
 * block before it forms its prefix from the cell counts itself (counted in
 * tsph_counters.scan_stalls; results are the same either way).  Default 2^20;
 * -1: never poll, always sum the counts (slow: for testing that path).  No
 * counterpart in the reference. */
int tsph_set_scan_wait(tsph_ctx *ctx, int polls);
/* wait for all work queued on the context */
int tsph_synchronize(tsph_ctx *ctx);

/* Pair-kernel timing with HIP events on the context stream.  While enabled
 * (max_launches > 0) every pair-loop launch, up to max_launches, is bracketed
 * by its own event pair and tsph_step_symplectic launches eagerly instead of
 * replaying its graph.  tsph_pair_timing_read synchronises, returns the summed
 * kernel time and the number of timed launches, and restarts the record. */
int tsph_pair_timing(tsph_ctx *ctx, int max_launches);
int tsph_pair_timing_read(tsph_ctx *ctx, double *total_ms, int *launches);

/* ---- Riemann-ALE-SPH (first order), the pipeline of the ALE dam-break driver
 * cases/ALESPH_dambreakLong/SPH_simulationINT.h:84-132 ---- */

/* ALE volume variables (SPH_special, src/SPH_data.h:54-67), reference order;
 * vector fields interleaved {x, y} */
typedef struct tsph_ale_fields {
  int np;
  double *omega, *omegarho, *omegav;
  double *omega_o, *omegarho_o, *omegav_o;
  double *domega, *domegarho, *domegav;
  double *gradrho, *gradvx, *gradvy;  /* MUSCL gradients (src/SPH_data.h:68-70) */
} tsph_ale_fields;

/* Allocates the ALE variables and sets them as ALESPH_Add_particle does
 * (src/SPH_particle_system.h:51-108): omega = dp^2, omega*rho, omega*v. */
int tsph_ale_enable(tsph_ctx *ctx);
/* tsph_ale_enable, then overwrites the fields given (non-NULL) with host values
 * in reference order; host->np must equal the live particle count. */
int tsph_ale_upload(tsph_ctx *ctx, const tsph_ale_fields *host);
int tsph_ale_download(tsph_ctx *ctx, tsph_ale_fields *host);
/* Compute_Forces of src/RALESPH_interaction.h:12-287 */
int tsph_rale_forces(tsph_ctx *ctx);
/* Integrate_LeapFrog_partOne / partTwo, src/RALESPH_integration.h:10-107 */
int tsph_rale_leapfrog_one(tsph_ctx *ctx, double dt);
int tsph_rale_leapfrog_two(tsph_ctx *ctx, double dt, int step);
/* nsteps whole steps numbered from `step` (the driver's step counter, from 1) */
int tsph_step_rale(tsph_ctx *ctx, double dt, int step, int nsteps);

/* ---- ALE-SPH with MUSCL reconstruction, the pipeline of the ALE dam-break
 * driver cases/ALESPH_dambreakLong/SPH_simulation.h:78-125 (shares the ALE
 * variables of tsph_ale_enable) ---- */
#define TSPH_MUSCL_NONE 0   /* first order: the gradients stay zero, as the driver runs */
#define TSPH_MUSCL_LINEAR 1 /* Kernel_gradient_approx after Compute_Density, weight 1 */
#define TSPH_MUSCL_MINMOD 2 /* as LINEAR, each reconstruction weighted by the minmod
                               limiter of src/ALESPH_riemann_solver.h:180-191 */
/* Compute_Density, src/ALESPH_density_and_omega_bt.h:9-281 (limited: minmod weights) */
int tsph_ale_compute_density(tsph_ctx *ctx, int limited);
/* Kernel_gradient_approx, src/ALESPH_kernel_gradient.h:10-190 */
int tsph_ale_kernel_gradient(tsph_ctx *ctx);
/* Integrate_density_compute_pressure, src/ALESPH_integration.h:86-135 */
int tsph_ale_density_pressure(tsph_ctx *ctx, double dt);
/* Compute_Acceleration_BT, src/ALESPH_acceleration_bt.h:12-295 */
int tsph_ale_compute_acceleration(tsph_ctx *ctx, int limited);
/* Integrate_LeapFrog_partOne_ALESPH / partTwo_ALESPH, src/ALESPH_integration.h:9-83 */
int tsph_ale_leapfrog_one(tsph_ctx *ctx, double dt);
int tsph_ale_leapfrog_two(tsph_ctx *ctx, double dt, int step);
/* nsteps whole ALE-SPH steps numbered from `step`; muscl = TSPH_MUSCL_* */
int tsph_step_ale(tsph_ctx *ctx, double dt, int step, int nsteps, int muscl);

/* ---- buffer zones (src/SPH_buffer_zones.h, src/SPH_inlet_outlet_data.h) ---- */
/* BufferZone, src/SPH_inlet_outlet_data.h:18-46: the region as DRAW_Create*
 * passes it (src/SPH_create.h:439-542), the buffer velocity, the orientation
 * {+-1, 0} (zone normal along x) or {0, +-1} (along y), and the particle type
 * of the zone's particles. */
typedef struct tsph_zone {
  double x0, xm, y0, ym;
  double vx, vy;
  int ox, oy;
  int btype;
} tsph_zone;
#define TSPH_MAX_ZONES 8
/* particles.zones.push_back(BufferZone(...)); *index = the zone's index */
int tsph_bz_add_zone(tsph_ctx *ctx, const tsph_zone *zone, int *index);
/* BZ_Integrate_inlet, src/SPH_buffer_zones.h:10-107: zone particles past the
 * zone start become fluid and a new zone particle is appended one buffer length
 * behind each, in particle order (at most 4096 per call); the context capacity
 * must leave room for them.  Takes effect for pair loops at the next
 * tsph_particle_to_cell. */
int tsph_bz_integrate_inlet(tsph_ctx *ctx, int zone, double dt);
/* BZ_Integrate_outlet, src/SPH_buffer_zones.h:111-206: fluid past the zone start
 * becomes the zone type (a = 0, drho = 0); zone particles past the zone end or
 * below dp/2 are erased, with the reference's in-loop skip of the next particle. */
int tsph_bz_integrate_outlet(tsph_ctx *ctx, int zone, double dt);
/* Kahan-compensated FP64 position and density updates in the symplectic
 * corrector (on != 0), where the reference's corrector
 * (src/SPH_integration.h:287-309) rounds each update once: the rounding error
 * of every step is carried to the next, so the drift of r and rho over long
 * runs stays at the rounding level of one step instead of growing with the
 * step count.  Results differ from the uncompensated path at that level.
 * Adds 24 B per particle (two generations), moved with the particles; the
 * terms start at zero (here and at every tsph_upload).  on = 0 frees them. */
int tsph_set_compensated(tsph_ctx *ctx, int on);
/* Relabels the reference positions of the live particles 0..np-1, keeping
 * their order.  No counterpart in the reference, where positions are vector
 * indices that every erase renumbers (src/SPH_particle_system.h:125-145).
 * Inlet appends take fresh positions and the library relabels by itself
 * before they run out (tsph_bz_integrate_inlet, tsph_step_inflow); results
 * are unchanged by a relabel. */
int tsph_renumber(tsph_ctx *ctx);
/* nsteps whole steps of the inflow driver
 * (cases/inflowSlab_mDBC/SPH_simulationSymplectic.h:57-123): the symplectic step
 * with BZ_Integrate_inlet / _outlet after the predictor (zone -1: none). */
int tsph_step_inflow(tsph_ctx *ctx, int wall, double dt, int nsteps, int inlet_zone,
                     int outlet_zone);

/* ---- other WCSPH integrators and density reinitialisation ----
 * The Verlet drivers keep a second time layer (v_oo, rho_oo) and the
 * reinitialisation a per-particle buffer (SPH_special::rhoTemp); they exist on
 * the device once tsph_history_enable has run, which sets them as
 * Add_particle does (v_oo = v, rho_oo = rho, rhoTemp = rho). */
typedef struct tsph_history {
  int np;
  double *v_oo;      /* 2*np */
  double *rho_oo;
  double *rho_temp;  /* rhoTemp */
} tsph_history;
int tsph_history_enable(tsph_ctx *ctx);
/* in reference order; NULL arrays keep the device values */
int tsph_history_upload(tsph_ctx *ctx, const tsph_history *host);
int tsph_history_download(tsph_ctx *ctx, tsph_history *host);
/* Integrate_Verlet / Integrate_Euler, src/SPH_integration.h:103-158 */
int tsph_verlet(tsph_ctx *ctx, double dt);
int tsph_euler(tsph_ctx *ctx, double dt);
/* Integrate_LeapFrog_partOne/partTwo_withDensity, src/SPH_integration.h:161-230 */
int tsph_leapfrog_one(tsph_ctx *ctx, double dt);
int tsph_leapfrog_two(tsph_ctx *ctx, double dt, int step);
/* Density_reinitialization_sumation / _overwrite,
 * src/SPH_density_reinitialization.h:8-139 (the sum needs a cell list) */
int tsph_density_reinit_sum(tsph_ctx *ctx);
int tsph_density_reinit_overwrite(tsph_ctx *ctx);
/* nsteps whole steps of the Verlet driver, steps step .. step+nsteps-1
 * (cases/damBreakLong_mDBC/SPH_simulationVerlet.h:64-116 and the DBC / BT
 * variants): wall density, cell list, EOS, pair loop, then Euler on step 1 and
 * every 20th step, Verlet otherwise, then removal */
int tsph_step_verlet(tsph_ctx *ctx, int wall, double dt, int step, int nsteps);
/* nsteps whole steps of the leap-frog driver
 * (cases/damBreakLong_mDBC/SPH_simulationLeapFrog.h and the DBC / BT variants):
 * part one (from step 2), cell list, wall density, EOS, pair loop, part two,
 * removal */
int tsph_step_leapfrog(tsph_ctx *ctx, int wall, double dt, int step, int nsteps);
/* nsteps symplectic steps, steps step .. step+nsteps-1, each followed on every
 * `interval`-th step by the density reinitialisation
 * (cases/damBreakLong_mDBC/SPH_simulationSymplectic_reinit.h, interval 20) */
int tsph_step_symplectic_reinit(tsph_ctx *ctx, int wall, double dt, int step, int nsteps,
                                int interval);

/* ---- output, snapshot and resume ----
 * write_to_ASCII_VTK, src/SPH_data_to_vtk.h:4-51: the particles in reference
 * order as legacy ASCII VTK polydata (type, pressure, density, velocity), the
 * same text as the reference's writer.  The device-to-host copies are queued on
 * the context's stream at the call (they see the stages queued before it); a
 * host thread formats and writes the file while later stages run.
 * tsph_io_wait joins the pending writes and reports the first failure;
 * tsph_destroy joins them too. */
int tsph_write_vtk(tsph_ctx *ctx, const char *path);
int tsph_io_wait(tsph_ctx *ctx);
/* GenerateInterpol, src/SPH_dataInterpolation.h:13-91: Pressure_in_point's
 * interpolation (src/testing.h:6-148) on the grid x0 + i dh, y0 + j dh with
 * dh = h/4, ncx = (int)((xm-x0)/h)*4, ncy = (int)((ym-y0)/h)*4, value of point
 * (i, j) at j + i*ncy.  With rho, p and v all NULL only the sizes are returned. */
int tsph_interpolate_grid(tsph_ctx *ctx, double x0, double y0, double xm, double ym, int *ncx,
                          int *ncy, double *rho, double *p, double *v);
/* tsph_interpolate_grid without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * grid values are staged in one device array per GPU, allocated when the
 * library is loaded and held until the process ends, instead of a buffer
 * allocated per call; the caller passes the host outputs as for
 * tsph_interpolate_grid and gets the same values.  At most
 * TSPH_GRID_SCRATCH_POINTS points (ncx*ncy) per call (a larger grid returns
 * TSPH_ECAPACITY, ncx and ncy untouched).  Calls that overlap on one GPU (from
 * several threads or contexts) race on that array. */
#define TSPH_GRID_SCRATCH_POINTS (1 << 19)
int tsph_interpolate_grid_noalloc(tsph_ctx *ctx, double x0, double y0, double xm, double ym,
                                  int *ncx, int *ncy, double *rho, double *p, double *v);
/* GenerateInterpol's file: write_INTERPOLATED_mesh_to_ASCII_VTK,
 * src/SPH_grid_to_vtk.h:139-243 */
int tsph_write_interpolated_vtk(tsph_ctx *ctx, double x0, double y0, double xm, double ym,
                                const char *path);
/* The two writers' formatting on host arrays (no device needed). */
int tsph_format_vtk(const tsph_particles *host, const char *path);
int tsph_format_interpolated_vtk(int ncx, int ncy, double dh, double x0, double y0,
                                 const double *rho, const double *p, const double *v,
                                 const char *path);
/* write_mesh_to_ASCII_VTK, src/SPH_grid_to_vtk.h:6-67: the cell grid of the
 * last Particle_To_Cell with each cell's particle count; x_0, y_0 are the
 * driver's (padded) domain origin, Simulation_data's x_0 and y_0. */
int tsph_write_grid_vtk(tsph_ctx *ctx, double x_0, double y_0, const char *path);
/* The same file from host counts (count of cell y + x*ncy). */
int tsph_format_grid_vtk(int ncx, int ncy, double kh, double of, double x_0, double y_0,
                         const int32_t *counts, const char *path);
/* Output_hydrostatic_pressure, src/SPH_measure_pressure.h:17-34: per fluid
 * particle in reference order, "p/(1000*9.81*0.5) y/0.5". */
int tsph_write_hydrostatic_pressure(tsph_ctx *ctx, const char *path);
int tsph_format_hydrostatic_pressure(const tsph_particles *host, const char *path);
/* Binary FP64 snapshot of every per-particle field in reference order (with
 * the ALE variables and the history fields when present), and its reload into
 * a context created with the same tsph_params.  After a reload, call
 * tsph_particle_to_cell; the run then continues as it would have.  No
 * counterpart in the reference. */
int tsph_save_state(tsph_ctx *ctx, const char *path);
int tsph_load_state(tsph_ctx *ctx, const char *path);

/* ---- batched ensembles ----
 * Many independent cases (each with its own constants, domain, time step and
 * particle count) held in one context and advanced together: every stage is
 * one launch over all cases.  The cell grids of the cases are laid end to end
 * in one cell range (a segmented cell list) and the pair loops walk one list
 * of tiles over all grids, so uneven cases share the persistent waves.  Each
 * case keeps the semantics of a single-case context run on its own, including
 * its reference particle order and the removal rule.  The per-case stage
 * calls tsph_particle_to_cell, tsph_density_to_pressure, tsph_compute_forces,
 * tsph_mdbc_density, tsph_bt_density and tsph_remove_out_of_domain act on all
 * cases; entry points that take one case's state or one time step (upload,
 * download, predictor/corrector, ALE, buffer zones, probes) are refused.  No
 * counterpart in the reference, which runs one case per executable. */
#define TSPH_BATCH_MAX_CASES 255
/* prm[q], dt[q]: constants, padded domain, capacity and time step of case q */
int tsph_batch_create(const tsph_params *prm, const double *dt, int ncases, tsph_ctx **out);
/* *ncases = the case count (0 for a single-case context) */
int tsph_batch_cases(tsph_ctx *ctx, int *ncases);
/* grid (src/SPH_find_pairs.h:19-20) and particle mass of case icase; NULL outputs skipped */
int tsph_batch_grid(tsph_ctx *ctx, int icase, int *ncx, int *ncy, double *mass);
/* host[q] for q < ncases, in each case's reference order (as tsph_upload) */
int tsph_batch_upload(tsph_ctx *ctx, const tsph_particles *host);
/* np[q] = live particles of case q (ncases entries) */
int tsph_batch_get_np(tsph_ctx *ctx, int *np);
/* case icase's state in its reference order (as tsph_download) */
int tsph_batch_download(tsph_ctx *ctx, int icase, tsph_particles *host);
/* nsteps symplectic steps of every case, each by its own dt, replayed from a
 * captured graph (the stage order of tsph_step_symplectic) */
int tsph_batch_step(tsph_ctx *ctx, int wall, int nsteps);

/* ---- autotuner ----
 * Times the pair pass `variant` on the current state for each tile walk
 * (pooled or static, see tsph_set_balanced), tile depth (cells per tile: 5,
 * 6, 7, 8, 10, 12, 14, 16, 20, 24) and persistent-wave count (waves per CU:
 * 4, 8, 12, 16), three passes each, and keeps the fastest.  The 64 lanes of a
 * wavefront take the particles of its tile in turn, so the depth sets the
 * particles per lane (thread): the depths above span about one to three at
 * the configured resolutions, and *particles_per_lane returns the chosen
 * tiling's mean (particles / (64 x non-empty tiles)).  Needs a cell list;
 * leaves a, drho as one pair pass with the chosen setting leaves them.
 * Outputs may be NULL.  No counterpart in the reference. */
int tsph_autotune(tsph_ctx *ctx, int variant, int *tile_rows, int *waves_per_cu, int *balanced,
                  double *best_ms, double *particles_per_lane);
/* Tile walk of the pair loops for the context's precision: 1 = pooled (the
 * default: workgroups of up to 4 wavefronts, each with a chunk of the tile
 * list of about equal cost, cut by the time each tile took in the previous
 * pass, whose tiles its wavefronts take from a counter in LDS, the costliest
 * first), 0 = static striding, one wavefront per workgroup.  Grids with fewer
 * than 16 tiles per persistent wave, and the ALE passes, always stride.
 * Results do not depend on it.  No counterpart in the reference. */
int tsph_set_balanced(tsph_ctx *ctx, int balanced);
/* Per-cell order of the cell-list build (Particle_To_Cell's push_back in
 * ascending particle index, src/SPH_find_pairs.h:45-92): 1 = ranked by the
 * pass that moves the particles (each particle counts those of its cell with
 * a lower index: one launch fewer), 2 = a separate ordering pass over the
 * cells (fewer bytes moved), 0 = by size (the default: 1 up to 2,000,000
 * particles, 2 above).  Results do not depend on it.  No counterpart in the
 * reference. */
int tsph_set_cell_order(tsph_ctx *ctx, int mode);

/* ---- case setup (host only; no GPU needed) ----
 * Presets built with the reference's geometry routines (src/SPH_create.h,
 * src/SPH_create_bt.h) in the order of the reference case geometry files. */
#define TSPH_CASE_STILL_WATER_DBC 1 /* cases/stillWater_DBC/geometry.h, 1.0 m tank, 0.5 m water */
#define TSPH_CASE_DAM_BREAK_MDBC 2  /* cases/damBreakLong_mDBC/geometry.h, 3.22x1.8 m tank,
                                       1.2x0.6 m column, fluid positions jittered */
#define TSPH_CASE_DAM_BREAK_BT 3    /* cases/damBreakLong_BT/geometry.h, 1.61x1.0 m tank,
                                       0.6x0.3 m column */
#define TSPH_CASE_ALE_DAM_BREAK 4   /* cases/ALESPH_dambreakLong/geometry.h (BT walls, ALE
                                       particles) in the 3.22x1.8 m tank with the 1.2x0.6 m
                                       column, dx = 2 mm */
#define TSPH_CASE_INFLOW_SLAB 5     /* cases/inflowSlab_mDBC/geometry.h + constants.h: mDBC
                                       step, inlet (zone 0, type 7) and empty outlet (zone 1,
                                       type 12), dx = 5 mm */
#define TSPH_CASE_INFLOW_CHANNEL 6  /* the inflow driver on an 8 x 0.5 m channel: 0.125 m of
                                       water moving at 1 m/s over an mDBC floor, inlet at
                                       x = 0 (zone 0), outlet at x = 7.9 (zone 1), dx = 0.5 mm
                                       (about 4M particles) */
#define TSPH_CASE_DAM_BREAK_MDBC_REF 7 /* cases/damBreakLong_mDBC/{constants,geometry}.h as
                                       they stand: 1.61x1.0 m tank, 0.6x0.3 m column, dx = 5 mm,
                                       c0 = 34.3 m/s, dt = 5e-5 s scaled with dx; the tank of
                                       the reference's water-level gauges and wall sensors */

/* time-stepping scheme of a preset */
#define TSPH_SCHEME_SYMPLECTIC 0 /* WCSPH, tsph_step_symplectic */
#define TSPH_SCHEME_RALE 1       /* Riemann-ALE-SPH, tsph_step_rale */

typedef struct tsph_case {
  tsph_params prm;           /* constants and padded domain (capacity = np) */
  double dt;                 /* time step of the preset */
  int wall;                  /* TSPH_WALL_* */
  int steps;                 /* step count of the preset */
  int scheme;                /* TSPH_SCHEME_* */
  int n_fluid, n_wall;
  tsph_particles particles;  /* arrays owned by the case */
  int n_zones;               /* buffer zones of the preset, in index order */
  tsph_zone zones[4];
} tsph_case;

/* dx <= 0 selects the preset's spacing; jitter is the uniform position
 * perturbation of fluid particles in units of dx (negative: preset default);
 * seed drives the jitter (SplitMix64). */
int tsph_case_build(int preset, double dx, double jitter, uint64_t seed, tsph_case **out);
int tsph_case_free(tsph_case *c);

/* The batched-ensemble workload: ncases tanks drawn from `seed` (SplitMix64),
 * each with mDBC walls and the dam-break constants (h = sqrt(2) dx, kap = 2.1,
 * c0 = 10 sqrt(g depth), dt = 0.2 h / c0, 500 steps): tank width W ~ U[0.5, 3] m,
 * tank height = 2 W, fill ~ U[0.3, 0.8] (water depth = fill * 2 W), dx ~ U[2, 10] mm,
 * and with probability 1/2 each still water over the whole width or a dam-break
 * column W/2 wide; dx is then rescaled so that the tank holds between nmin and
 * nmax fluid particles.  Fills cases[0..ncases-1]; free each with tsph_case_free. */
int tsph_ensemble_build(int ncases, uint64_t seed, int nmin, int nmax, tsph_case **cases);

#ifdef __cplusplus
}
#endif
#endif /* TINYSPH_H */
