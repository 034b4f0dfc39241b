// Internal declarations shared by the translation units of libtinysph_hip.so.
// Device data layout (HBM), one SoA generation per field:
//   r, v, a, r_o, v_o, gnr, n : double2   (16 B per particle each)
//   rho, drho, p, rho_o       : double    ( 8 B)
//   type                      : uint8     (reference part_type; 255 = removed)
//   seq                       : uint32    (position in reference order; gaps allowed)
//   index                     : int32     (reference data.index)
// After every Particle_To_Cell the arrays are permuted into (cell, seq) order,
// so a cell's particles are contiguous and ordered exactly as the reference
// iterates them; `cellof[k]` holds the cell of slot k and `cell_start` the CSR
// offsets.  The particle count lives on the device (`scal[S_NP]`) so that a
// whole step can be captured in a HIP graph; the host keeps an upper bound.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tinysph.h"

namespace tsph {

constexpr uint8_t kRemoved = 255;

constexpr int kMaxTileRows = 24;  // upper bound of tile_rows
// Pooled walk's cost model.  Every tile pass records the time each tile took
// its wavefront (tq_mark, tsph_tiles.h; 10 ns ticks of the device's constant
// clock, averaged with the tile's previous record) in one of kCostTables
// tables, at scal[S_NSCAL + kCostTables t + table] for global tile t; the
// next pooled pass of the same table cuts its chunks by these costs
// (cut_chunks_scan in a cell-list build, cut_chunks_block for a list reused
// without one).  A recorded time carries what a model of the tile's
// work would not: what else ran on its compute unit, so that the chunks of a
// compute unit that ran slow shrink in the next pass.  Table 1 holds the mixed
// pair passes that apply the symplectic corrector as well, table 0 every
// other tile pass.  A tile without a record (the first pass after a change of
// tile layout, or a tile newly occupied) is estimated from its own particles:
// own particles x own + tile, the least-squares fit of the measured times of
// the 740k-particle dam break's 13,147 tiles (near-empty tiles 4.4 us mixed,
// 3.3 us FP64; full 14.2 us and 32 us).
struct CostEstimate {
  int own, tile;
};
constexpr CostEstimate kCostMixed{19, 425}, kCostFp64{54, 300};
constexpr int kCostTables = 2;
inline CostEstimate cost_estimate(bool mixed) { return mixed ? kCostMixed : kCostFp64; }

// device scalar slots
// (the wall count alternates between S_NWALL and S_NWALL_B with the cell-offset
// buffers: a build fills the other one, see cell_start_next); the per-tile
// costs follow them (kCostTables (ncells_tot + 8) slots)
enum { S_NP = 0, S_MAXSEQ = 1, S_NTILES = 2, S_NWALL = 3, S_BZCNT = 4, S_NWALL_B = 5, S_NSCAL = 32 };
// device counter slots (uint64)
enum {
  C_PAIRS = 0, C_CAND = 1, C_ACTIVE = 2, C_PASSES = 3, C_BADRHO = 4, C_OUTGRID = 5,
  C_REMOVED = 6, C_APPENDED = 7, C_BZREJECT = 8, C_BZCHAIN = 9, C_REINIT = 10, C_SCANSTALL = 11,
  C_NCOUNT = 12
};
// The pair loops' counts (pairs, candidates, active) go to per-wavefront
// slots after the counters above, 4 words per wavefront, and are summed when
// read: one device-wide atomic per wavefront on three shared words, all at the
// end of a launch, serialises there and delays the launch's last tiles.
constexpr int kWaveCtr0 = 16;


// buffer zone (BufferZone, src/SPH_inlet_outlet_data.h:18-46)
struct Zone {
  double x0, xm, y0, ym, bl, bh, vx, vy;
  int ox, oy, btype;
};
constexpr int kBzMax = 4096;  // zone particles an inlet pass may convert

struct Soa {
  uint8_t *cid = nullptr;  // case of each particle (batched ensembles only)
  double2 *r = nullptr, *v = nullptr, *a = nullptr, *r_o = nullptr, *v_o = nullptr;
  double2 *gnr = nullptr, *n = nullptr;
  double *rho = nullptr, *drho = nullptr, *p = nullptr, *rho_o = nullptr;
  uint8_t *type = nullptr;
  uint32_t *seq = nullptr;
  int32_t *index = nullptr;
  // ALE-SPH volume variables (allocated for ALE runs only)
  double *omega = nullptr, *omegarho = nullptr, *omega_o = nullptr, *omegarho_o = nullptr;
  double *domega = nullptr, *domegarho = nullptr;
  double2 *omegav = nullptr, *omegav_o = nullptr, *domegav = nullptr;
  // ALE-SPH MUSCL gradients of rho, vx, vy (allocated with the volume variables)
  double2 *gradrho = nullptr, *gradvx = nullptr, *gradvy = nullptr;
  // Verlet second time layer and the density-reinitialisation buffer (rhoTemp)
  // (allocated by tsph_history_enable)
  double2 *v_oo = nullptr;
  double *rho_oo = nullptr, *rho_tmp = nullptr;
  // Kahan compensation terms of the symplectic corrector's position and
  // density updates (allocated by tsph_set_compensated)
  double2 *rc = nullptr;
  double *rhoc = nullptr;
};

// FP32 pair records of the mixed path, rebuilt after every reorder.
struct Mixed {
  float4 *rv = nullptr;   // {x - cell origin x, y - cell origin y, vx, vy}
  float2 *rp = nullptr;   // {rho - rho0, p}
  float2 *nrm = nullptr;  // BT wall normal
};

// constants of the physics, precomputed on the host exactly as the reference does
struct Consts {
  double h, kap, cs, rho0, avisc, delta, dp, m, gx, gy;
  double x0, xm, y0, ym;
  double kh, of;
  double awen, bwen;  // float-rounded Wendland constants, src/SPH_kernel.h:56-57
  int ncx, ncy, ncells;
  // batched ensembles: the case's first global cell id and its time step
  // (single cases: 0 and unused)
  int cell_base;
  double dt;
};

// constants of particle i: its case's entry of the batch table, or k0
__device__ __forceinline__ Consts case_consts(const Consts &k0, const Consts *__restrict__ cases,
                                              const uint8_t *__restrict__ cid, int i) {
  return cases ? cases[cid[i]] : k0;
}

// host-side flags that describe what the device buffers hold
struct HostState {
  bool cells_valid, tiles_valid, records_valid, marks_valid, olds_stale;
  int list_groups, list_table, list_cost_table;
};

// one captured step.  A step may rebuild the cell list an odd number of times
// (the Riemann-ALE step does it once), which swaps the cur/alt particle buffers:
// graphs are therefore kept per buffer parity, and a replay applies the swap
// and the host flags that the captured step left behind.
struct Graph {
  hipGraphExec_t exec = nullptr;
  int wall = -1;
  double dt = 0;
  int n_ub = -1;
  bool flips = false;  // the step swaps cur and alt an odd number of times
  HostState post{};    // host flags after the step
  int steps = 1;       // steps captured in the graph
};
// steps captured in one graph of a run of whole symplectic steps: each graph
// launch costs about 8 us of idle time between the steps it separates
constexpr int kGraphSteps = 8;

}  // namespace tsph

struct tsph_ctx {
  tsph_params prm;
  tsph::Consts k;
  int capacity = 0;
  int n_ub = 0;  // host upper bound of the device particle count
  int ncells_tot = 0;            // cells of all cases (the single case: k.ncells)
  // batched ensembles: per-case constants on the device (nullptr for one case)
  int ncases = 1;
  tsph::Consts *cases = nullptr;
  std::vector<tsph::Consts> host_cases;
  int *tile_base = nullptr;      // ncases+1 first global tile id per case (current tile_rows)
  std::vector<int> host_tile_base;  // host copy of tile_base
  std::vector<int> seq_base;     // first reference position per case (ncases+1)
  std::vector<int> case_cap;     // particle slots reserved per case
  bool has_gnr = false, has_n = false;
  bool cells_valid = false;
  int tile_rows = 16;  // cells per pair-loop tile (one grid-column run)
  int waves_per_cu = 8;  // persistent pair-loop waves per compute unit
  int num_cu = 256;      // compute units of the device
  int max_lds_block = 65536;  // LDS bytes a workgroup may allocate
  int lds_cu = 163840;        // LDS bytes of one compute unit
  // pair-loop tile walk per precision (FP64, mixed): pooled (work-balanced
  // chunks per workgroup, tiles taken from an LDS counter by the workgroup's
  // waves) or the plain static walk; defaults set at creation
  bool balanced[2] = {false, false};
  int cell_order = 0;  // tsph_set_cell_order: 0 by size, 1 ranked, 2 separate pass
  int *tile_list = nullptr;  // non-empty tiles of the current cell list, in tile order
  bool tiles_valid = false;  // tile_list matches the cell list and tile_rows
  int *list_part = nullptr;  // per list-build block: tiles, work
  int list_part_cap = 0;     // list-build blocks list_part holds
  int *tile_bound = nullptr; // pooled walk: first list entry of each workgroup's chunk
  int wave_ctr_cap = 0;      // wavefronts with count slots in `counters`
  int bound_cap = 0;         // chunks tile_bound holds
  int list_groups = 0;       // chunks the current tile_bound was cut for (0: none or stale)
  int list_table = 0;        // the cost table they were cut by
  int *list_cost = nullptr;  // costs of the listed tiles by list position (pooled cuts)
  int list_cost_table = -1;  // the cost table list_cost was gathered from (-1: none)
  // a pooled pair pass of `cut_groups` workgroups and cost table `cut_table`
  // follows the next cell build (0: none): the build cuts its chunks
  int cut_groups = 0, cut_table = 0;
  int *wall_list = nullptr;  // slots of the wall particles of the current cell list
  int n_wall = 0;            // wall particles (walls are never created or removed)
  bool records_valid = false;  // mixed-path FP32 records match the FP64 state
  bool marks_valid = false;    // removal marks match the current positions
  // r_o, v_o, rho_o were not moved by the last cell build (the build that
  // follows a symplectic corrector: the next predictor rewrites them); until
  // rebuilt they are alt's values through perm (ensure_olds)
  bool olds_stale = false;
  // device address of the allocation-free interpolation grid's static scratch
  // array on this context's device (looked up at creation)
  double *grid_scratch = nullptr;
  bool has_ale = false;        // ALE volume variables allocated
  uint8_t *ale_flag = nullptr; // per particle: its own ALE gamma sum is non-zero
  hipStream_t stream = nullptr;
  tsph::Soa cur, alt;
  tsph::Mixed mx;
  int *scal = nullptr;                 // S_* slots, then the per-tile costs
  unsigned long long *counters = nullptr;
  int *cell_count = nullptr;           // ncells+1 (last = removed bucket)
  int *cell_start = nullptr;           // ncells+2
  // Every build writes its cell offsets and wall count into the buffers the
  // previous build did not (cell_start_next, the other wall-count slot) and
  // then swaps, so that the wall-density pass that reads the previous cell
  // list can run while the next build sorts (record_step).
  int *cell_start_next = nullptr;      // ncells+2
  int nwall_idx = 3;                   // S_NWALL or S_NWALL_B: the current wall count
  int *cellof = nullptr;               // cell of each slot (sorted order)
  int *hrank = nullptr;                // rank inside the cell from the hash pass
  int *hcell = nullptr;                // cell from the hash pass (pre-sort order)
  int *perm = nullptr;                 // sorted position -> old slot
  // the scatter's placement (cell, then arrival order) and the seq of each
  // placed particle, which the reorder ranks within the cell
  int *arrival = nullptr;              // arrival position -> old slot
  uint32_t *arrival_seq = nullptr;     // arrival position -> seq
  uint32_t *arrival_span = nullptr;    // arrival rank | cell count << 16
  // cell-list scan: one published word per block (zero: not yet published),
  // and the polls a block makes for an earlier block's word (-1: none)
  unsigned long long *scan_state = nullptr;
  // per scan block: the listed tiles' summed cost and first list position
  int *scan_cost = nullptr, *scan_tile0 = nullptr;
  int scan_wait = 1 << 20;
  uint8_t *seqstate = nullptr;         // removal: per-seq state
  double *probe_out = nullptr;         // 4 doubles for probe results
  // pair-kernel timing: one event pair per pair launch while enabled
  bool time_pairs = false;
  std::vector<hipEvent_t> tev;         // 2 per recorded launch
  int tev_used = 0;
  float last_pair_ms = 0.f;
  int parity = 0;             // flips whenever cur and alt are swapped
  // symplectic step, by parity at the step's start: one step ([0], [1]) or
  // kGraphSteps steps ([2], [3])
  tsph::Graph graph[4];
  tsph::Graph rale_graph[2];  // Riemann-ALE step, by parity at the step's start
  tsph::Graph ale_graph[2];   // ALE-SPH MUSCL step, by parity at the step's start
  tsph::Graph inflow_graph[2];  // symplectic step with buffer zones
  tsph::Graph verlet_graph[2];  // Verlet-driver step (Verlet integrator)
  tsph::Graph euler_graph[2];   // Verlet-driver step on its Euler steps
  tsph::Graph lf_graph[2];      // leap-frog-driver step (after the first)
  bool has_history = false;     // v_oo, rho_oo, rhoTemp allocated
  // asynchronous output (tsph_write_vtk): writer threads and their first error
  std::vector<std::thread> writers;
  std::mutex io_mu;
  std::string io_error;
  tsph::Zone zones[TSPH_MAX_ZONES];
  int n_zones = 0;
  int *bz_list = nullptr;     // 2 * kBzMax ints: {seq, slot} of the crossing zone particles
  int seq_cap = 0;            // size of seqstate (reference positions ever handed out)
  int seq_ub = 0;             // host upper bound of scal[S_MAXSEQ]
};

namespace tsph {

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define TSPH_HIP(call)                                                  \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) return tsph::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

inline int grid_for(int n, int block) { return n <= 0 ? 1 : (n + block - 1) / block; }

// words of the counters buffer: the counters, then the per-wavefront slots
inline size_t counter_words(const tsph_ctx *c) { return kWaveCtr0 + 4 * (size_t)c->wave_ctr_cap; }

// pair-loop tiles of the cell grid(s) at the current tile_rows
inline int total_tiles(const tsph_ctx *c) {
  if (!c->cases) return c->k.ncx * ((c->k.ncy + c->tile_rows - 1) / c->tile_rows);
  return c->host_tile_base[c->ncases];
}

// whether a pair launch of `waves` persistent waves takes the pooled walk:
// the precision's setting, for grids of at least 16 tiles per wave (below that
// the waves have a few tiles each and little to balance: a 183k-particle ALE
// dam break with 8-10 tiles per wave ran 0.35 ms per step pooled, 0.28 static)
inline bool walk_balanced(const tsph_ctx *c, bool fp32, int waves) {
  return c->balanced[fp32 ? 1 : 0] && total_tiles(c) >= 16 * waves;
}

// launch geometry of a tile kernel: `groups` workgroups (a multiple of 8) of
// block.x / 64 wavefronts, each wavefront with lds_wave bytes of LDS
struct TileGrid {
  dim3 grid, block;
  size_t lds = 0;    // dynamic LDS per workgroup
  int lds_wave = 0;
  int groups = 0;
  bool pooled = false;
};

// tsph_cells.hip
// hash, count, scan, scatter, order, reorder; eos: the reorder also runs
// Density_to_pressure (and writes the mixed records) for the moved particles
// predict: the symplectic predictor (step dt; per case in a batch) is applied
// on the way, in place of a launch_predictor immediately before.  ad_skip: a
// mask of particle types (bit t: type t) whose a, drho are left behind, for a
// build that a pair pass overwriting exactly those follows (forces_active_types)
// decide: RemoveParticlesOutOfDomain's decision is taken by the build, in
// place of a launch_remove immediately before (the removal states must be
// current: a corrector precedes)
// skip_olds: r_o, v_o, rho_o are not moved (olds_stale afterwards), for a build
// after which they are dead until the next predictor rewrites them
// walls (TSPH_WALL_MDBC or TSPH_WALL_BT; 0: none): the wall densities of that
// treatment are computed first, from the previous cell list, in the same
// launch as the build's hash (as launch_mdbc / launch_bt_density with
// count_density)
int launch_cell_list(tsph_ctx *c, bool eos = false, bool predict = false, double dt = 0.,
                     unsigned ad_skip = 0, bool decide = false, bool skip_olds = false,
                     int walls = 0, bool count_density = false);
// r_o, v_o, rho_o of the current generation, gathered from the previous one
// through the last build's permutation if that build left them behind; every
// launch that reads or moves them calls this first
int ensure_olds(tsph_ctx *c);
// types whose a, drho a pair pass of `variant` writes (its active particles)
inline unsigned forces_active_types(int variant) {
  return variant == TSPH_FORCES_DBC ? (1u << TSPH_FLUID) | (1u << TSPH_WALL) : 1u << TSPH_FLUID;
}
// tsph_pairs.hip
// list of the non-empty pair-loop tiles; groups > 0: also its chunk bounds
// for a pooled walk of that many workgroups, cut by the tile costs recorded
// in cost table `table`
int ensure_tile_list(tsph_ctx *c, int groups, int table = 0);
// geometry of a tile-kernel launch at the context's waves per CU (pooled when
// may_pool and walk_balanced hold for the precision)
TileGrid tile_grid(tsph_ctx *c, bool fp32, size_t lds_wave, bool may_pool = true);
// clears the per-tile costs (the tile layout changed)
int clear_tile_costs(tsph_ctx *c);
// cuts the current tile list into `groups` chunks by cost table `table` on
// stream s (gathering the listed tiles' costs first unless the list build did)
int launch_cut(tsph_ctx *c, hipStream_t s, int groups, int table);
// workgroups of a pair pass of `variant` if it takes the pooled walk, else 0
int pass_groups(tsph_ctx *c, int variant);
// corr (mixed path only): the pass also applies the symplectic corrector to
// every particle, with time step corr_dt (batched contexts: each case's own)
int launch_forces(tsph_ctx *c, int variant, bool corr = false, double corr_dt = 0.);
int launch_mixed_records(tsph_ctx *c);
// tsph_stages.hip
int launch_eos(tsph_ctx *c);
int launch_predictor(tsph_ctx *c, double dt);
int launch_corrector(tsph_ctx *c, double dt);
// count_density: also take the out-of-range density count of an EOS pass
// right after (in place of that pass, where nothing reads its pressures)
int launch_mdbc(tsph_ctx *c, bool count_density = false);
int launch_bt_density(tsph_ctx *c, bool count_density = false);
int launch_remove(tsph_ctx *c);
int run_water_elevation(tsph_ctx *c, double x, double y0, int inverse, double *out);
// soa: out holds rho[n], p[n], then v[2n] (n = ncx ncy); else 4 values per point
int launch_interp_grid(tsph_ctx *c, double x0, double y0, double dh, int ncx, int ncy, double *out,
                       bool soa = false);
// device address of the allocation-free interpolation grid's scratch on the
// current device (looked up once per device)
int grid_scratch_ptr(double **out);
// tsph_ale.hip
int launch_ale_init(tsph_ctx *c);
int launch_rale_forces(tsph_ctx *c);
int launch_rale_leapfrog_one(tsph_ctx *c, double dt);
int launch_rale_leapfrog_two(tsph_ctx *c, double dt, int step);
int run_interpolate_point(tsph_ctx *c, double x, double y, double *out4);
// tsph_alemuscl.hip (muscl: 0 first order, 1 alpha = 1, 2 minmod-limited)
int launch_ale_density(tsph_ctx *c, int muscl);
int launch_ale_acceleration(tsph_ctx *c, int muscl);
int launch_ale_gradient(tsph_ctx *c);
int launch_ale_density_pressure(tsph_ctx *c, double dt);
int launch_ale_leapfrog_one(tsph_ctx *c, double dt);
// tsph_integrators.hip
int launch_history_init(tsph_ctx *c);
int launch_verlet(tsph_ctx *c, double dt, bool euler);
int launch_leapfrog_one(tsph_ctx *c, double dt);
int launch_leapfrog_two(tsph_ctx *c, double dt, int step);
int launch_reinit_sum(tsph_ctx *c);
int launch_reinit_overwrite(tsph_ctx *c);
int launch_ale_leapfrog_two(tsph_ctx *c, double dt, int step);
// tsph_zones.hip
int launch_bz_inlet(tsph_ctx *c, int zone, double dt);
int launch_bz_outlet(tsph_ctx *c, int zone, double dt);

}  // namespace tsph
