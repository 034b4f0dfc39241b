// C ABI of libtinysph_hip.so (declared in include/tinysph.h): context
// lifetime, host <-> device transfer in reference order, stage entry points
// and the graph-replayed symplectic step.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "tsph_internal.h"

namespace tsph {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_err = buf;
  return TSPH_EHIP;
}

static int fail(int code, const char *msg) {
  g_err = msg;
  return code;
}

template <typename T>
static int dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  TSPH_HIP(hipMalloc((void **)p, count * sizeof(T)));
  return 0;
}

static int alloc_soa(Soa &s, int cap, bool gnr, bool n) {
  int e = 0;
  if ((e = dalloc(&s.r, cap)) || (e = dalloc(&s.v, cap)) || (e = dalloc(&s.a, cap)) ||
      (e = dalloc(&s.r_o, cap)) || (e = dalloc(&s.v_o, cap)) || (e = dalloc(&s.rho, cap)) ||
      (e = dalloc(&s.drho, cap)) || (e = dalloc(&s.p, cap)) || (e = dalloc(&s.rho_o, cap)) ||
      (e = dalloc(&s.type, cap)) || (e = dalloc(&s.seq, cap)) || (e = dalloc(&s.index, cap)))
    return e;
  if (gnr && (e = dalloc(&s.gnr, cap))) return e;
  if (n && (e = dalloc(&s.n, cap))) return e;
  return 0;
}

static void free_soa(Soa &s) {
  void *ps[] = {s.cid, s.r, s.v, s.a, s.r_o, s.v_o, s.gnr, s.n, s.rho, s.drho, s.p, s.rho_o,
                s.type, s.seq, s.index, s.omega, s.omegarho, s.omega_o, s.omegarho_o,
                s.domega, s.domegarho, s.omegav, s.omegav_o, s.domegav, s.gradrho, s.gradvx,
                s.gradvy, s.v_oo, s.rho_oo, s.rho_tmp, s.rc, s.rhoc};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  s = Soa{};
}

static void drop_graph(tsph_ctx *c) {
  for (Graph *g : {&c->graph[0], &c->graph[1], &c->graph[2], &c->graph[3], &c->rale_graph[0],
                   &c->rale_graph[1],
                   &c->ale_graph[0], &c->ale_graph[1], &c->inflow_graph[0], &c->inflow_graph[1],
                   &c->verlet_graph[0], &c->verlet_graph[1], &c->euler_graph[0],
                   &c->euler_graph[1], &c->lf_graph[0], &c->lf_graph[1]}) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    *g = Graph{};
  }
}

// default pair-loop launch: 7-cell tiles in the pooled walk (chunks cut by
// the recorded tile costs); the mixed path's shorter records hide more
// latency with 12 waves per CU, the FP64 path's larger LDS records allow 8
// (740k-particle dam break, pair pass: mixed 0.070 ms pooled against 0.088
// ms static, FP64 0.248 ms against 0.277 ms)
static void set_walk_defaults(tsph_ctx *c) {
  const bool mixed = c->prm.precision == TSPH_PREC_MIXED;
  c->waves_per_cu = mixed ? 12 : 8;
  c->balanced[0] = true;
  c->balanced[1] = true;
}

static HostState host_state(const tsph_ctx *c) {
  return {c->cells_valid, c->tiles_valid, c->records_valid, c->marks_valid, c->olds_stale,
          c->list_groups, c->list_table, c->list_cost_table};
}

static void set_host_state(tsph_ctx *c, const HostState &h) {
  c->cells_valid = h.cells_valid;
  c->tiles_valid = h.tiles_valid;
  c->list_groups = h.list_groups;
  c->list_table = h.list_table;
  c->list_cost_table = h.list_cost_table;
  c->records_valid = h.records_valid;
  c->marks_valid = h.marks_valid;
  c->olds_stale = h.olds_stale;
}

// Runs nsteps steps from graphs: the graph of the current buffer parity is
// captured from `record` (one step's launches) unless it already holds a step
// for the same tag, dt and launch shape.  Recording mutates the host state as
// launching would (buffer swaps, flags); a replay applies the same changes.
// Steps other than the symplectic ones (record_step) read or move r_o, v_o,
// rho_o in their first launches, so a symplectic step before them must have
// its left-behind olds gathered first, eagerly: a gather captured into their
// graph would repeat on every replay.
static bool step_reads_olds(const Graph *slots, const tsph_ctx *c) {
  return slots != c->graph && slots != c->inflow_graph;
}

// group > 1: slots has 4 entries, and runs of `group` steps replay one graph
// of that many steps (slots[2 + parity]), the rest single-step graphs
template <typename R>
static int replay_step(tsph_ctx *c, Graph *slots, int tag, double dt, int nsteps, R record,
                       int group = 1) {
  if (step_reads_olds(slots, c))
    if (int e = ensure_olds(c)) return e;
  for (int s = 0; s < nsteps;) {
    const int k = group > 1 && nsteps - s >= group ? group : 1;
    Graph &g = slots[(k > 1 ? 2 : 0) + c->parity];
    if (!g.exec || g.wall != tag || g.dt != dt || g.n_ub != c->n_ub || g.steps != k) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = Graph{};
      const Soa cur0 = c->cur, alt0 = c->alt;
      int *const cs0 = c->cell_start, *const csn0 = c->cell_start_next;
      const int nwall0 = c->nwall_idx;
      const int parity0 = c->parity;
      const HostState h0 = host_state(c);
      auto undo = [&] {  // nothing was launched: restore the host view
        c->cur = cur0;
        c->alt = alt0;
        c->cell_start = cs0;
        c->cell_start_next = csn0;
        c->nwall_idx = nwall0;
        c->parity = parity0;
        set_host_state(c, h0);
      };
      hipGraph_t graph = nullptr;
      TSPH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int e = 0;
      for (int j = 0; j < k && !e; ++j) e = record();
      const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
      if (e || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        undo();
        if (e) return e;
        TSPH_HIP(ce);
      }
      const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        g = Graph{};
        undo();
        TSPH_HIP(ie);
      }
      g.wall = tag;
      g.dt = dt;
      g.n_ub = c->n_ub;
      g.flips = c->parity != parity0;
      g.post = host_state(c);
      g.steps = k;
    } else {
      if (g.flips) {  // what an odd number of cell builds leaves (launch_cell_list)
        std::swap(c->cur, c->alt);
        std::swap(c->cell_start, c->cell_start_next);
        c->nwall_idx = c->nwall_idx == S_NWALL ? S_NWALL_B : S_NWALL;
        c->parity ^= 1;
      }
      set_host_state(c, g.post);
    }
    TSPH_HIP(hipGraphLaunch(g.exec, c->stream));
    s += k;
  }
  return 0;
}

// Allocates the wall-extra fields (gnr / n) on first use.
static int ensure_extras(tsph_ctx *c, bool gnr, bool n) {
  int e = 0;
  if (gnr && !c->has_gnr) {
    if ((e = dalloc(&c->cur.gnr, c->capacity)) || (e = dalloc(&c->alt.gnr, c->capacity))) return e;
    c->has_gnr = true;
  }
  if (n && !c->has_n) {
    if ((e = dalloc(&c->cur.n, c->capacity)) || (e = dalloc(&c->alt.n, c->capacity))) return e;
    if (c->prm.precision == TSPH_PREC_MIXED && (e = dalloc(&c->mx.nrm, c->capacity))) return e;
    c->has_n = true;
  }
  return 0;
}

static int alloc_ale(Soa &s, int cap) {
  int e = 0;
  if ((e = dalloc(&s.omega, cap)) || (e = dalloc(&s.omegarho, cap)) || (e = dalloc(&s.omega_o, cap)) ||
      (e = dalloc(&s.omegarho_o, cap)) || (e = dalloc(&s.domega, cap)) ||
      (e = dalloc(&s.domegarho, cap)) || (e = dalloc(&s.omegav, cap)) ||
      (e = dalloc(&s.omegav_o, cap)) || (e = dalloc(&s.domegav, cap)) ||
      (e = dalloc(&s.gradrho, cap)) || (e = dalloc(&s.gradvx, cap)) || (e = dalloc(&s.gradvy, cap)))
    return e;
  return 0;
}

// one Riemann-ALE step, cases/ALESPH_dambreakLong/SPH_simulationINT.h:84-132
static int record_rale_step(tsph_ctx *c, double dt, int step) {
  int e = 0;
  if (step > 1 && (e = launch_rale_leapfrog_one(c, dt))) return e;
  if ((e = launch_cell_list(c, true)) || (e = launch_rale_forces(c))) return e;
  if ((e = launch_rale_leapfrog_two(c, dt, step)) || (e = launch_remove(c))) return e;
  return 0;
}

// one ALE-SPH step, cases/ALESPH_dambreakLong/SPH_simulation.h:78-125; with
// muscl > 0 the gradients are computed where the driver has the (commented)
// Kernel_gradient_approx call, right after Compute_Density (:103)
static int record_ale_step(tsph_ctx *c, double dt, int step, int muscl) {
  int e = 0;
  if (step > 1 && (e = launch_ale_leapfrog_one(c, dt))) return e;
  if ((e = launch_cell_list(c)) || (e = launch_ale_density(c, muscl))) return e;
  if (muscl && (e = launch_ale_gradient(c))) return e;
  if ((e = launch_ale_density_pressure(c, dt)) || (e = launch_ale_acceleration(c, muscl))) return e;
  if ((e = launch_ale_leapfrog_two(c, dt, step)) || (e = launch_remove(c))) return e;
  return 0;
}

// one symplectic step in the drivers' stage order; inlet / outlet >= 0 add the
// buffer zones of the inflow driver after the predictor
// (cases/inflowSlab_mDBC/SPH_simulationSymplectic.h:71-76)
static int record_step(tsph_ctx *c, int wall, double dt, int inlet = -1, int outlet = -1) {
  const int variant = wall == TSPH_WALL_DBC    ? TSPH_FORCES_DBC
                      : wall == TSPH_WALL_MDBC ? TSPH_FORCES_WCSPH
                                               : TSPH_FORCES_BT;
  int e = 0;
  const bool zones = inlet >= 0 || outlet >= 0;
  // each cell build is followed by a pair pass that rewrites a, drho of its
  // active particles, so the builds need not move those
  const unsigned active = forces_active_types(variant);
  // wall density + EOS (cases/damBreakLong_mDBC/SPH_simulationSymplectic.h:65-66,
  // cases/damBreakLong_BT/SPH_simulationSymplectic.h:72-73); DBC has none.
  // Without buffer zones nothing reads the pressures (or the mixed records)
  // this EOS writes before the cell build rewrites them all, so only its
  // out-of-range density count is taken, by the wall-density pass.  Without
  // buffer zones the pass runs in workgroups of the first build's hash launch,
  // beside the hash (on a second stream the fork and the join in the step
  // graph left about 25 us of idle time per step, more than the pass takes).
  // With buffer zones the pass runs in a launch of its own, before the
  // predictor (the zone passes act on predicted positions, so the build cannot
  // apply the predictor); its EOS is still only the out-of-range density
  // count, taken by the pass's sweep: the zone passes read no pressure, and the
  // build's EOS rewrites every pressure that EOS would write (inlet and outlet
  // particles keep theirs either way).
  const bool walls_in_build = !zones && (wall == TSPH_WALL_MDBC || wall == TSPH_WALL_BT);
  if (zones && (wall == TSPH_WALL_MDBC || wall == TSPH_WALL_BT)) {
    if ((e = wall == TSPH_WALL_MDBC ? launch_mdbc(c, true) : launch_bt_density(c, true)))
      return e;
  }
  // each build cuts the chunks of the pooled pair pass that follows it
  const bool mixed = c->prm.precision == TSPH_PREC_MIXED;
  const int groups = pass_groups(c, variant);
  c->cut_groups = groups;
  c->cut_table = mixed ? 1 : 0;
  if (zones) {
    // the buffer zones act on the predicted positions
    if ((e = launch_predictor(c, dt))) return e;
    if (inlet >= 0 && (e = launch_bz_inlet(c, inlet, dt))) return e;
    if (outlet >= 0 && (e = launch_bz_outlet(c, outlet, dt))) return e;
    if ((e = launch_cell_list(c, true, false, 0., active))) return e;
  } else if ((e = launch_cell_list(c, true, true, dt, active, false, false,
                                   walls_in_build ? wall : 0, true))) {
    return e;  // (the predictor is applied by the build)
  }
  // the corrector writes the removal states; the build takes the removal
  // decision (RemoveParticlesOutOfDomain) on the way.  The mixed pair pass
  // applies the corrector itself (it never reads the FP64 state it updates).
  if (c->prm.precision == TSPH_PREC_MIXED) {
    if ((e = launch_forces(c, variant, true, dt))) return e;
  } else if ((e = launch_forces(c, variant)) || (e = launch_corrector(c, dt))) {
    return e;
  }
  // r_o, v_o, rho_o are dead after the corrector until the next predictor
  // rewrites them: this build leaves them behind
  c->cut_groups = groups;
  c->cut_table = 0;
  if ((e = launch_cell_list(c, true, false, 0., active, true, true)) ||
      (e = launch_forces(c, variant)))
    return e;
  return 0;
}

static int forces_variant_of(int wall) {
  return wall == TSPH_WALL_DBC ? TSPH_FORCES_DBC : wall == TSPH_WALL_MDBC ? TSPH_FORCES_WCSPH
                                                                         : TSPH_FORCES_BT;
}

// one Verlet-driver step, cases/damBreakLong_mDBC/SPH_simulationVerlet.h:64-116
// (mDBC: no EOS before the cell list), cases/damBreakLong_BT/SPH_simulationVerlet.h
// (BT: wall density and EOS), cases/stillWater_DBC/SPH_simulationVerlet.h (DBC)
static int record_verlet_step(tsph_ctx *c, int wall, double dt, bool euler) {
  int e = 0;
  if (wall == TSPH_WALL_MDBC) {
    if ((e = launch_mdbc(c))) return e;
  } else if (wall == TSPH_WALL_BT) {
    if ((e = launch_bt_density(c)) || (e = launch_eos(c))) return e;
  }
  if ((e = launch_cell_list(c, true)) || (e = launch_forces(c, forces_variant_of(wall)))) return e;
  if ((e = launch_verlet(c, dt, euler)) || (e = launch_remove(c))) return e;
  return 0;
}

// one leap-frog-driver step, cases/damBreakLong_mDBC/SPH_simulationLeapFrog.h
// (the wall density follows the cell list), cases/stillWater_DBC/SPH_simulationLeapFrog.h
static int record_leapfrog_step(tsph_ctx *c, int wall, double dt, int step) {
  int e = 0;
  if (step > 1 && (e = launch_leapfrog_one(c, dt))) return e;
  if (wall == TSPH_WALL_DBC) {
    if ((e = launch_cell_list(c, true))) return e;
  } else {
    if ((e = launch_cell_list(c, false))) return e;
    if ((e = wall == TSPH_WALL_MDBC ? launch_mdbc(c) : launch_bt_density(c)) || (e = launch_eos(c)))
      return e;
  }
  if ((e = launch_forces(c, forces_variant_of(wall)))) return e;
  if ((e = launch_leapfrog_two(c, dt, step)) || (e = launch_remove(c))) return e;
  return 0;
}

}  // namespace tsph

using namespace tsph;

extern "C" {

const char *tsph_last_error(void) { return g_err.c_str(); }
int tsph_abi_version(void) { return TSPH_ABI_VERSION; }

}  // extern "C"

namespace tsph {

// Particle_system constructor + Divide_To_Cells for one case's constants
static int make_consts(const tsph_params *prm, Consts &k) {
  if (!(prm->h > 0) || !(prm->kap > 0) || !(prm->dp > 0) || prm->capacity <= 0 ||
      !(prm->x_m > prm->x_0) || !(prm->y_m > prm->y_0))
    return fail(TSPH_EINVAL, "invalid parameters");
  if (prm->precision != TSPH_PREC_FP64 && prm->precision != TSPH_PREC_MIXED)
    return fail(TSPH_EINVAL, "unknown precision");
  k = Consts{};
  k.h = prm->h; k.kap = prm->kap; k.cs = prm->cs; k.rho0 = prm->rho0; k.avisc = prm->avisc;
  k.delta = prm->delta; k.dp = prm->dp;
  k.m = prm->rho0 / (1 / prm->dp * 1 / prm->dp);  // src/SPH_particle_system.h:207
  k.gx = 0.; k.gy = 9.81;                          // src/SPH_particle_system.h:205
  k.x0 = prm->x_0; k.xm = prm->x_m; k.y0 = prm->y_0; k.ym = prm->y_m;
  k.kh = prm->kap * prm->h;                        // src/SPH_find_pairs.h:11
  k.of = prm->h * 0.273;                           // src/SPH_find_pairs.h:13
  k.awen = (double)(float)(0.557 / (prm->h * prm->h));
  k.bwen = (double)(float)(-2.7852 / (prm->h * prm->h * prm->h));
  k.ncx = (int)((k.xm - k.x0 + k.of) / k.kh) + 1;  // src/SPH_find_pairs.h:19-20
  k.ncy = (int)((k.ym - k.y0 + k.of) / k.kh) + 1;
  const long long nc = (long long)k.ncx * k.ncy;
  if (nc <= 0 || nc > (1ll << 30)) return fail(TSPH_EINVAL, "cell grid too large");
  k.ncells = (int)nc;
  return 0;
}

// device buffers of a context holding `cap` particles over `ncells_tot` cells
static int alloc_context(tsph_ctx *c) {
  int e = 0;
  // the stream at the default priority
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(TSPH_EHIP, "hipStreamCreate failed");
  const int cap = c->capacity, nct = c->ncells_tot;
  c->seq_cap = 2 * cap + kBzMax + c->ncases;
  if ((e = alloc_soa(c->cur, cap, false, false)) || (e = alloc_soa(c->alt, cap, false, false)))
    return e;
  if ((e = dalloc(&c->scal, S_NSCAL + (size_t)kCostTables * (nct + 8))) ||
      (e = dalloc(&c->cell_count, nct + 1)) || (e = dalloc(&c->cell_start, nct + 2)) ||
      (e = dalloc(&c->cell_start_next, nct + 2)) ||
      (e = dalloc(&c->cellof, cap)) || (e = dalloc(&c->hrank, cap)) || (e = dalloc(&c->hcell, cap)) ||
      (e = dalloc(&c->perm, cap)) || (e = dalloc(&c->arrival, cap)) ||
      (e = dalloc(&c->arrival_seq, cap)) || (e = dalloc(&c->arrival_span, cap)) ||
      (e = dalloc(&c->seqstate, (size_t)c->seq_cap)) || (e = dalloc(&c->probe_out, 4)) ||
      (e = dalloc(&c->bz_list, 2 * kBzMax)) ||
      (e = dalloc(&c->tile_list, nct + 8)) || (e = dalloc(&c->list_cost, nct + 8)) ||
      (e = dalloc(&c->wall_list, cap)) ||
      (e = dalloc(&c->scan_state, (nct + 1) / 1024 + 2)) ||
      (e = dalloc(&c->scan_cost, (nct + 1) / 1024 + 2)) ||
      (e = dalloc(&c->scan_tile0, (nct + 1) / 1024 + 2)))
    return e;
  // the per-cell counters and removal states start cleared; later builds clear
  // them as they consume them (the single-pass scan's words likewise)
  if (hipMemsetAsync(c->cell_count, 0, sizeof(int) * (nct + 1), c->stream) != hipSuccess ||
      hipMemsetAsync(c->scan_state, 0, sizeof(unsigned long long) * ((nct + 1) / 1024 + 2),
                     c->stream) != hipSuccess ||
      hipMemsetAsync(c->seqstate, 0, (size_t)c->seq_cap, c->stream) != hipSuccess)
    return fail(TSPH_EHIP, "initialisation failed");
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
      if (prop.multiProcessorCount > 0) c->num_cu = prop.multiProcessorCount;
      if (prop.sharedMemPerBlock > 0) c->max_lds_block = (int)prop.sharedMemPerBlock;
      if (prop.maxSharedMemoryPerMultiProcessor > 0)
        c->lds_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
    }
  }
  // the tile-list build: one block per 256 tiles (tiles <= cells); pooled
  // walks: one chunk per workgroup, at most one workgroup per wave
  c->list_part_cap = (nct + 8) / 256 + 2;
  c->bound_cap = 8 * ((c->num_cu * 32 + 7) / 8);
  // tile launches have at most num_cu * 32 + 31 wavefronts (tile_grid)
  c->wave_ctr_cap = c->num_cu * 32 + 32;
  if ((e = dalloc(&c->list_part, 2 * c->list_part_cap)) || (e = dalloc(&c->tile_bound, c->bound_cap + 1)) ||
      (e = dalloc(&c->counters, counter_words(c))))
    return e;
  if (c->prm.precision == TSPH_PREC_MIXED) {
    if ((e = dalloc(&c->mx.rv, cap)) || (e = dalloc(&c->mx.rp, cap))) return e;
  }
  if (hipMemsetAsync(c->scal, 0, (S_NSCAL + (size_t)kCostTables * (nct + 8)) * sizeof(int),
                     c->stream) != hipSuccess ||
      hipMemsetAsync(c->counters, 0, counter_words(c) * sizeof(unsigned long long), c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(TSPH_EHIP, "initialisation failed");
  return 0;
}

// entry points that act on one case's state are refused on batched contexts
static int single_only(const tsph_ctx *c) {
  return c->cases ? fail(TSPH_ESTATE, "not available on a batched context") : 0;
}

// first global pair-loop tile of each case of a batched context at the
// current tile_rows (host and device copies; no-op for a single case)
static int update_tile_base(tsph_ctx *c) {
  if (!c->cases) return 0;
  std::vector<int> &tb = c->host_tile_base;
  tb.assign(c->ncases + 1, 0);
  for (int q = 0; q < c->ncases; ++q) {
    const Consts &k = c->host_cases[q];
    tb[q + 1] = tb[q] + k.ncx * ((k.ncy + c->tile_rows - 1) / c->tile_rows);
  }
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(c->tile_base, tb.data(), sizeof(int) * (c->ncases + 1), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace tsph

extern "C" {

int tsph_create(const tsph_params *prm, tsph_ctx **out) {
  if (!prm || !out) return fail(TSPH_EINVAL, "null argument");
  *out = nullptr;
  Consts k;
  if (int e = make_consts(prm, k)) return e;
  tsph_ctx *c = new tsph_ctx();
  c->prm = *prm;
  c->capacity = prm->capacity;
  c->tile_rows = 7;
  set_walk_defaults(c);
  c->k = k;
  c->ncells_tot = k.ncells;
  if (int e = alloc_context(c)) {
    tsph_destroy(c);
    return e;
  }
  // set-up of tsph_interpolate_grid_noalloc: the scratch array's address on
  // this device (the device current at creation, where the context runs)
  if (int e = grid_scratch_ptr(&c->grid_scratch)) {
    tsph_destroy(c);
    return e;
  }
  *out = c;
  return 0;
}

int tsph_destroy(tsph_ctx *c) {
  if (!c) return 0;
  (void)tsph_io_wait(c);  // writer threads hold copies of this context's state
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  free_soa(c->cur);
  free_soa(c->alt);
  void *ps[] = {c->scal, c->counters, c->cell_count, c->cell_start, c->cell_start_next, c->cellof, c->hrank,
                c->hcell, c->perm, c->arrival, c->arrival_seq, c->arrival_span, c->scan_state, c->scan_cost, c->scan_tile0, c->seqstate, c->probe_out, c->mx.rv,
                c->mx.rp, c->mx.nrm, c->tile_list, c->list_cost, c->wall_list, c->bz_list,
                c->cases, c->tile_base, c->list_part, c->tile_bound};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  for (hipEvent_t ev : c->tev) (void)hipEventDestroy(ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int tsph_grid(tsph_ctx *c, int *ncx, int *ncy) {
  if (!c || !ncx || !ncy) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  *ncx = c->k.ncx;
  *ncy = c->k.ncy;
  return 0;
}

int tsph_mass(tsph_ctx *c, double *m) {
  if (!c || !m) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  *m = c->k.m;
  return 0;
}

}  // extern "C"

namespace tsph {

// Writes n particles (reference order, fields as in tsph_particles) into the
// device state.  seq / cid / index: per-particle reference position, case and
// reference index (nullptr: i, none, h->index or i).  maxseq: first unused
// reference position.
static int upload_impl(tsph_ctx *c, const tsph_particles *h, const uint32_t *seq_in,
                       const uint8_t *cid, int maxseq) {
  const int n = h->np;
  if (n < 0) return fail(TSPH_EINVAL, "negative particle count");
  if (n > c->capacity) return fail(TSPH_ECAPACITY, "particle count exceeds capacity");
  if (!h->type || !h->rho || !h->r) return fail(TSPH_EINVAL, "type, rho and r are required");
  int e = 0;
  if ((e = ensure_extras(c, h->gnr != nullptr, h->n != nullptr))) return e;
  TSPH_HIP(hipStreamSynchronize(c->stream));
  drop_graph(c);
  std::vector<uint8_t> t(n);
  std::vector<uint32_t> seq(n);
  std::vector<int32_t> idx(n);
  for (int i = 0; i < n; ++i) {
    if (h->type[i] < 0 || h->type[i] > 254) return fail(TSPH_EINVAL, "particle type out of range");
    t[i] = (uint8_t)h->type[i];
    seq[i] = seq_in ? seq_in[i] : (uint32_t)i;
    idx[i] = h->index ? h->index[i] : i;
  }
  std::vector<double> zeros(2 * (size_t)n, 0.0);
  const size_t s1 = sizeof(double) * n, s2 = 2 * s1;
  Soa &d = c->cur;
  auto cp = [&](void *dst, const void *src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  const double *v = h->v ? h->v : zeros.data();
  const double *p = h->p ? h->p : zeros.data();
  if (n > 0) {
    TSPH_HIP(cp(d.r, h->r, s2));
    TSPH_HIP(cp(d.v, v, s2));
    TSPH_HIP(cp(d.a, h->a ? h->a : zeros.data(), s2));
    TSPH_HIP(cp(d.r_o, h->r_o ? h->r_o : h->r, s2));
    TSPH_HIP(cp(d.v_o, h->v_o ? h->v_o : v, s2));
    TSPH_HIP(cp(d.rho, h->rho, s1));
    TSPH_HIP(cp(d.drho, h->drho ? h->drho : zeros.data(), s1));
    TSPH_HIP(cp(d.p, p, s1));
    TSPH_HIP(cp(d.rho_o, h->rho_o ? h->rho_o : h->rho, s1));
    TSPH_HIP(cp(d.type, t.data(), n));
    TSPH_HIP(cp(d.seq, seq.data(), 4 * (size_t)n));
    TSPH_HIP(cp(d.index, idx.data(), 4 * (size_t)n));
    if (cid) TSPH_HIP(cp(d.cid, cid, n));
    if (h->gnr) TSPH_HIP(cp(d.gnr, h->gnr, s2));
    else if (c->has_gnr) TSPH_HIP(cp(d.gnr, zeros.data(), s2));
    if (h->n) TSPH_HIP(cp(d.n, h->n, s2));
    else if (c->has_n) TSPH_HIP(cp(d.n, zeros.data(), s2));
  }
  if (d.rc) {  // compensation starts from zero
    TSPH_HIP(hipMemsetAsync(d.rc, 0, sizeof(double2) * (size_t)c->capacity, c->stream));
    TSPH_HIP(hipMemsetAsync(d.rhoc, 0, sizeof(double) * (size_t)c->capacity, c->stream));
  }
  int sc[S_NSCAL] = {};
  sc[S_NP] = n;
  sc[S_MAXSEQ] = maxseq;
  c->seq_ub = maxseq;
  TSPH_HIP(hipMemcpy(c->scal, sc, sizeof sc, hipMemcpyHostToDevice));
  TSPH_HIP(hipMemsetAsync(c->seqstate, 0, (size_t)c->seq_cap, c->stream));
  TSPH_HIP(hipStreamSynchronize(c->stream));
  c->n_ub = n;
  c->n_wall = 0;
  for (int i = 0; i < n; ++i) c->n_wall += t[i] == TSPH_WALL;
  c->olds_stale = false;
  c->cells_valid = false;
  c->tiles_valid = false;
  c->records_valid = false;
  c->marks_valid = false;
  return 0;
}

}  // namespace tsph

extern "C" {

int tsph_upload(tsph_ctx *c, const tsph_particles *h) {
  if (!c || !h) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  return upload_impl(c, h, nullptr, nullptr, h->np);
}

int tsph_get_np(tsph_ctx *c, int *np) {
  if (!c || !np) return fail(TSPH_EINVAL, "null argument");
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(np, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"

namespace tsph {

// order of the live slots in reference order (by seq), skipping removed ones;
// icase >= 0 keeps the slots of that case of a batched context only
static int live_order(tsph_ctx *c, int n, int icase, std::vector<int> &order,
                      std::vector<uint32_t> &seq, std::vector<uint8_t> &type) {
  seq.resize(n);
  type.resize(n);
  std::vector<uint8_t> cid;
  if (n > 0) {
    TSPH_HIP(hipMemcpy(seq.data(), c->cur.seq, 4 * (size_t)n, hipMemcpyDeviceToHost));
    TSPH_HIP(hipMemcpy(type.data(), c->cur.type, n, hipMemcpyDeviceToHost));
    if (icase >= 0) {
      cid.resize(n);
      TSPH_HIP(hipMemcpy(cid.data(), c->cur.cid, n, hipMemcpyDeviceToHost));
    }
  }
  order.clear();
  for (int i = 0; i < n; ++i)
    if (type[i] != kRemoved && (icase < 0 || cid[i] == icase)) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return seq[a] < seq[b]; });
  return 0;
}

// Reference positions (seq) are only ever compared, never counted, so they may
// be relabelled between stages as long as the live particles keep their order
// (in the reference they are vector indices, renumbered by every erase,
// src/SPH_particle_system.h:125-145).  Inlet appends take fresh positions from
// S_MAXSEQ upwards and nothing hands positions back, so a long inflow run
// would run out of removal states (seq_cap); this relabels the live particles
// 0..m-1 in order, carrying each one's removal state along, gives the removed
// slots still in the range the positions after them (state 0: absent), and
// resets S_MAXSEQ.  Cell order, tiles and records stay valid: the in-cell
// order by seq is unchanged.
static int renumber_seq(tsph_ctx *c) {
  int n = 0, e = 0;
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(&n, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> order;
  std::vector<uint32_t> seq;
  std::vector<uint8_t> type;
  if ((e = live_order(c, n, -1, order, seq, type))) return e;
  std::vector<uint8_t> st((size_t)c->seq_cap), nst((size_t)c->seq_cap, 0);
  TSPH_HIP(hipMemcpy(st.data(), c->seqstate, (size_t)c->seq_cap, hipMemcpyDeviceToHost));
  std::vector<uint32_t> nseq((size_t)std::max(n, 1));
  uint32_t next = 0;
  for (int i : order) {
    nst[next] = st[seq[i]];
    nseq[i] = next++;
  }
  for (int i = 0; i < n; ++i)
    if (type[i] == kRemoved) nseq[i] = next++;
  if (n > 0) TSPH_HIP(hipMemcpy(c->cur.seq, nseq.data(), 4 * (size_t)n, hipMemcpyHostToDevice));
  TSPH_HIP(hipMemcpy(c->seqstate, nst.data(), (size_t)c->seq_cap, hipMemcpyHostToDevice));
  const int ms = (int)next;
  TSPH_HIP(hipMemcpy(&c->scal[S_MAXSEQ], &ms, sizeof(int), hipMemcpyHostToDevice));
  c->seq_ub = ms;
  return 0;
}

// Inlet passes (each hands out up to kBzMax positions) that fit before the
// reference positions run out; reads the device's S_MAXSEQ, and relabels, only
// when the host's bound allows less than one.  After a relabel at least one
// fits: S_MAXSEQ <= capacity and seq_cap = 2 capacity + kBzMax + ncases.
static int seq_room(tsph_ctx *c, int *passes) {
  auto room = [c] { return (c->seq_cap - c->seq_ub) / kBzMax; };
  if (room() < 1) {
    int ms = 0;
    TSPH_HIP(hipStreamSynchronize(c->stream));
    TSPH_HIP(hipMemcpy(&ms, &c->scal[S_MAXSEQ], sizeof(int), hipMemcpyDeviceToHost));
    c->seq_ub = ms;
    if (room() < 1)
      if (int e = renumber_seq(c)) return e;
  }
  *passes = room();
  return 0;
}

static int download_impl(tsph_ctx *c, tsph_particles *h, int icase) {
  int n = 0, e = 0;
  if ((e = ensure_olds(c))) return e;
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(&n, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> order;
  std::vector<uint32_t> seq;
  std::vector<uint8_t> type;
  if ((e = live_order(c, n, icase, order, seq, type))) return e;
  const int m = (int)order.size();
  h->np = m;
  std::vector<double> b1(n), b2(2 * (size_t)n);
  auto get1 = [&](double *dst, const double *src) -> int {
    if (!dst) return 0;
    TSPH_HIP(hipMemcpy(b1.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) dst[k] = b1[order[k]];
    return 0;
  };
  auto get2 = [&](double *dst, const double2 *src) -> int {
    if (!dst) return 0;
    if (!src) {
      std::fill(dst, dst + 2 * (size_t)m, 0.0);
      return 0;
    }
    TSPH_HIP(hipMemcpy(b2.data(), src, 2 * sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) {
      dst[2 * k] = b2[2 * order[k]];
      dst[2 * k + 1] = b2[2 * order[k] + 1];
    }
    return 0;
  };
  if (n == 0) return 0;
  const Soa &d = c->cur;
  if ((e = get1(h->p, d.p)) || (e = get1(h->rho, d.rho)) || (e = get1(h->drho, d.drho)) ||
      (e = get1(h->rho_o, d.rho_o)) || (e = get2(h->r, d.r)) || (e = get2(h->v, d.v)) ||
      (e = get2(h->a, d.a)) || (e = get2(h->r_o, d.r_o)) || (e = get2(h->v_o, d.v_o)) ||
      (e = get2(h->gnr, d.gnr)) || (e = get2(h->n, d.n)))
    return e;
  if (h->type)
    for (int k = 0; k < m; ++k) h->type[k] = type[order[k]];
  if (h->index) {
    std::vector<int32_t> idx(n);
    TSPH_HIP(hipMemcpy(idx.data(), d.index, 4 * (size_t)n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) h->index[k] = idx[order[k]];
  }
  return 0;
}

}  // namespace tsph

extern "C" {

int tsph_download(tsph_ctx *c, tsph_particles *h) {
  if (!c || !h) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  return download_impl(c, h, -1);
}

int tsph_clear_cells(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  c->cells_valid = false;  // the cell list is rebuilt whole by tsph_particle_to_cell
  return 0;
}

int tsph_particle_to_cell(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  return launch_cell_list(c);
}

int tsph_get_cell_lists(tsph_ctx *c, int32_t *start, int32_t *items) {
  if (!c || !start || !items) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  int n = 0, e = 0;
  if ((e = tsph_get_np(c, &n))) return e;
  std::vector<int> order;
  std::vector<uint32_t> seq;
  std::vector<uint8_t> type;
  if ((e = live_order(c, n, -1, order, seq, type))) return e;
  // reference index of slot = its rank in reference order
  std::vector<int> rank(n, -1);
  for (size_t k = 0; k < order.size(); ++k) rank[order[k]] = (int)k;
  TSPH_HIP(hipMemcpy(start, c->cell_start, sizeof(int32_t) * (c->ncells_tot + 1), hipMemcpyDeviceToHost));
  for (int k = 0; k < n; ++k) items[k] = rank[k];
  return 0;
}

int tsph_density_to_pressure(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  return launch_eos(c);
}

int tsph_compute_forces(tsph_ctx *c, int variant) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (variant < TSPH_FORCES_WCSPH || variant > TSPH_FORCES_BT) return fail(TSPH_EINVAL, "unknown variant");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  if (variant == TSPH_FORCES_BT && !c->has_n) return fail(TSPH_ESTATE, "BT forces need wall normals (n)");
  return launch_forces(c, variant);
}

int tsph_mdbc_density(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  if (!c->has_gnr) return fail(TSPH_ESTATE, "mDBC needs ghost nodes (gnr)");
  return launch_mdbc(c);
}

int tsph_bt_density(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return launch_bt_density(c);
}

int tsph_symplectic_predictor(tsph_ctx *c, double dt) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  return launch_predictor(c, dt);
}

int tsph_symplectic_corrector(tsph_ctx *c, double dt) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  return launch_corrector(c, dt);
}

int tsph_remove_out_of_domain(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  return launch_remove(c);
}

int tsph_step_symplectic(tsph_ctx *c, int wall, double dt, int nsteps) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  if (wall < TSPH_WALL_DBC || wall > TSPH_WALL_BT) return fail(TSPH_EINVAL, "unknown wall treatment");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  if (wall == TSPH_WALL_MDBC && !c->has_gnr) return fail(TSPH_ESTATE, "mDBC needs ghost nodes (gnr)");
  if (wall == TSPH_WALL_BT && !c->has_n) return fail(TSPH_ESTATE, "BT needs wall normals (n)");
  if (nsteps <= 0) return 0;
  if (c->time_pairs) {
    // eager launches so that every pair kernel is bracketed by its own events
    for (int s = 0; s < nsteps; ++s)
      if (int e = record_step(c, wall, dt)) return e;
    return 0;
  }
  return replay_step(c, c->graph, wall, dt, nsteps, [&] { return record_step(c, wall, dt); },
                     kGraphSteps);
}

int tsph_ale_enable(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  int e = 0;
  if (!c->has_ale) {
    if ((e = alloc_ale(c->cur, c->capacity)) || (e = alloc_ale(c->alt, c->capacity)) ||
        (e = dalloc(&c->ale_flag, c->capacity)))
      return e;
    c->has_ale = true;
  }
  drop_graph(c);
  return launch_ale_init(c);
}

int tsph_ale_upload(tsph_ctx *c, const tsph_ale_fields *h) {
  if (!c || !h) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  int e = 0, n = 0;
  if ((e = tsph_ale_enable(c)) || (e = tsph_get_np(c, &n))) return e;
  std::vector<int> order;
  std::vector<uint32_t> seq;
  std::vector<uint8_t> type;
  if ((e = live_order(c, n, -1, order, seq, type))) return e;
  const int m = (int)order.size();
  if (h->np != m) return fail(TSPH_EINVAL, "ALE field count differs from the particle count");
  std::vector<double> b(2 * (size_t)n + 2, 0.0);
  // host entry k (reference order) goes to the device slot of the k-th live particle
  auto put1 = [&](double *dst, const double *src) -> int {
    if (!src) return 0;
    if (n) TSPH_HIP(hipMemcpy(b.data(), dst, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) b[order[k]] = src[k];
    if (n) TSPH_HIP(hipMemcpy(dst, b.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    return 0;
  };
  auto put2 = [&](double2 *dst, const double *src) -> int {
    if (!src) return 0;
    if (n) TSPH_HIP(hipMemcpy(b.data(), dst, 2 * sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) {
      b[2 * order[k]] = src[2 * k];
      b[2 * order[k] + 1] = src[2 * k + 1];
    }
    if (n) TSPH_HIP(hipMemcpy(dst, b.data(), 2 * sizeof(double) * n, hipMemcpyHostToDevice));
    return 0;
  };
  const Soa &d = c->cur;
  if ((e = put1(d.omega, h->omega)) || (e = put1(d.omegarho, h->omegarho)) ||
      (e = put2(d.omegav, h->omegav)) || (e = put1(d.omega_o, h->omega_o)) ||
      (e = put1(d.omegarho_o, h->omegarho_o)) || (e = put2(d.omegav_o, h->omegav_o)) ||
      (e = put1(d.domega, h->domega)) || (e = put1(d.domegarho, h->domegarho)) ||
      (e = put2(d.domegav, h->domegav)) || (e = put2(d.gradrho, h->gradrho)) ||
      (e = put2(d.gradvx, h->gradvx)) || (e = put2(d.gradvy, h->gradvy)))
    return e;
  return 0;
}

int tsph_ale_download(tsph_ctx *c, tsph_ale_fields *h) {
  if (!c || !h) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  if (!c->has_ale) return fail(TSPH_ESTATE, "ALE variables not enabled (tsph_ale_enable)");
  int n = 0, e = 0;
  if ((e = tsph_get_np(c, &n))) return e;
  std::vector<int> order;
  std::vector<uint32_t> seq;
  std::vector<uint8_t> type;
  if ((e = live_order(c, n, -1, order, seq, type))) return e;
  const int m = (int)order.size();
  std::vector<double> b(2 * (size_t)n + 2);
  auto get1 = [&](double *dst, const double *src) -> int {
    if (!dst) return 0;
    if (n) TSPH_HIP(hipMemcpy(b.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) dst[k] = b[order[k]];
    return 0;
  };
  auto get2 = [&](double *dst, const double2 *src) -> int {
    if (!dst) return 0;
    if (n) TSPH_HIP(hipMemcpy(b.data(), src, 2 * sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) {
      dst[2 * k] = b[2 * order[k]];
      dst[2 * k + 1] = b[2 * order[k] + 1];
    }
    return 0;
  };
  const Soa &d = c->cur;
  if ((e = get1(h->omega, d.omega)) || (e = get1(h->omegarho, d.omegarho)) ||
      (e = get2(h->omegav, d.omegav)) || (e = get1(h->omega_o, d.omega_o)) ||
      (e = get1(h->omegarho_o, d.omegarho_o)) || (e = get2(h->omegav_o, d.omegav_o)) ||
      (e = get1(h->domega, d.domega)) || (e = get1(h->domegarho, d.domegarho)) ||
      (e = get2(h->domegav, d.domegav)) || (e = get2(h->gradrho, d.gradrho)) ||
      (e = get2(h->gradvx, d.gradvx)) || (e = get2(h->gradvy, d.gradvy)))
    return e;
  h->np = m;
  return 0;
}

static int rale_ready(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (!c->has_ale) return fail(TSPH_ESTATE, "ALE variables not enabled (tsph_ale_enable)");
  return 0;
}

int tsph_rale_forces(tsph_ctx *c) {
  if (int e = rale_ready(c)) return e;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return launch_rale_forces(c);
}

int tsph_rale_leapfrog_one(tsph_ctx *c, double dt) {
  if (int e = rale_ready(c)) return e;
  return launch_rale_leapfrog_one(c, dt);
}

int tsph_rale_leapfrog_two(tsph_ctx *c, double dt, int step) {
  if (int e = rale_ready(c)) return e;
  return launch_rale_leapfrog_two(c, dt, step);
}

int tsph_step_rale(tsph_ctx *c, double dt, int step, int nsteps) {
  if (int e = rale_ready(c)) return e;
  if (step < 1 || nsteps < 0) return fail(TSPH_EINVAL, "steps are numbered from 1");
  int s = step;
  const int last = step + nsteps;
  // the first step differs (src/RALESPH_integration.h:42-70): launched eagerly
  for (; s < last && (s == 1 || c->time_pairs); ++s)
    if (int e = record_rale_step(c, dt, s)) return e;
  if (s >= last) return 0;
  return replay_step(c, c->rale_graph, 100, dt, last - s, [&] { return record_rale_step(c, dt, 2); });
}

int tsph_ale_compute_density(tsph_ctx *c, int limited) {
  if (int e = rale_ready(c)) return e;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return launch_ale_density(c, limited ? 2 : 1);
}

int tsph_ale_kernel_gradient(tsph_ctx *c) {
  if (int e = rale_ready(c)) return e;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return launch_ale_gradient(c);
}

int tsph_ale_density_pressure(tsph_ctx *c, double dt) {
  if (int e = rale_ready(c)) return e;
  return launch_ale_density_pressure(c, dt);
}

int tsph_ale_compute_acceleration(tsph_ctx *c, int limited) {
  if (int e = rale_ready(c)) return e;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return launch_ale_acceleration(c, limited ? 2 : 1);
}

int tsph_ale_leapfrog_one(tsph_ctx *c, double dt) {
  if (int e = rale_ready(c)) return e;
  return launch_ale_leapfrog_one(c, dt);
}

int tsph_ale_leapfrog_two(tsph_ctx *c, double dt, int step) {
  if (int e = rale_ready(c)) return e;
  return launch_ale_leapfrog_two(c, dt, step);
}

int tsph_step_ale(tsph_ctx *c, double dt, int step, int nsteps, int muscl) {
  if (int e = rale_ready(c)) return e;
  if (step < 1 || nsteps < 0) return fail(TSPH_EINVAL, "steps are numbered from 1");
  if (muscl < TSPH_MUSCL_NONE || muscl > TSPH_MUSCL_MINMOD) return fail(TSPH_EINVAL, "unknown MUSCL mode");
  int s = step;
  const int last = step + nsteps;
  // the first step differs (src/ALESPH_integration.h:37-58): launched eagerly
  for (; s < last && (s == 1 || c->time_pairs); ++s)
    if (int e = record_ale_step(c, dt, s, muscl)) return e;
  if (s >= last) return 0;
  return replay_step(c, c->ale_graph, 200 + muscl, dt, last - s,
                     [&] { return record_ale_step(c, dt, 2, muscl); });
}

int tsph_bz_add_zone(tsph_ctx *c, const tsph_zone *z, int *index) {
  if (!c || !z) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  if (c->n_zones >= TSPH_MAX_ZONES) return fail(TSPH_ECAPACITY, "too many buffer zones");
  if ((z->ox != 0) == (z->oy != 0)) return fail(TSPH_EINVAL, "zone orientation must be along x or y");
  Zone &q = c->zones[c->n_zones];
  q.x0 = z->x0; q.xm = z->xm; q.y0 = z->y0; q.ym = z->ym; q.vx = z->vx; q.vy = z->vy;
  q.ox = z->ox; q.oy = z->oy; q.btype = z->btype;
  // src/SPH_inlet_outlet_data.h:37-42
  if (q.ox != 0) { q.bl = std::fabs(q.xm - q.x0); q.bh = std::fabs(q.ym - q.y0); }
  else { q.bl = std::fabs(q.ym - q.y0); q.bh = std::fabs(q.xm - q.x0); }
  if (index) *index = c->n_zones;
  c->n_zones++;
  drop_graph(c);
  return 0;
}

static int zone_ok(tsph_ctx *c, int zone) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (zone < 0 || zone >= c->n_zones) return fail(TSPH_EINVAL, "no such buffer zone");
  return 0;
}

int tsph_bz_integrate_inlet(tsph_ctx *c, int zone, double dt) {
  if (int e = zone_ok(c, zone)) return e;
  int passes = 0;
  if (int e = seq_room(c, &passes)) return e;
  c->seq_ub += kBzMax;
  return launch_bz_inlet(c, zone, dt);
}

int tsph_renumber(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  return renumber_seq(c);
}

int tsph_bz_integrate_outlet(tsph_ctx *c, int zone, double dt) {
  if (int e = zone_ok(c, zone)) return e;
  return launch_bz_outlet(c, zone, dt);
}

int tsph_step_inflow(tsph_ctx *c, int wall, double dt, int nsteps, int inlet, int outlet) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  if (wall < TSPH_WALL_DBC || wall > TSPH_WALL_BT || nsteps < 0)
    return fail(TSPH_EINVAL, "invalid wall treatment or step count");
  if (inlet >= 0 && (inlet >= c->n_zones)) return fail(TSPH_EINVAL, "no such inlet zone");
  if (outlet >= 0 && (outlet >= c->n_zones)) return fail(TSPH_EINVAL, "no such outlet zone");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  if (wall == TSPH_WALL_MDBC && !c->has_gnr) return fail(TSPH_ESTATE, "mDBC needs ghost nodes (gnr)");
  if (wall == TSPH_WALL_BT && !c->has_n) return fail(TSPH_ESTATE, "BT needs wall normals (n)");
  if (inlet >= 0) c->n_ub = c->capacity;
  // one inlet pass per step: replay in runs that cannot exhaust the reference
  // positions, relabelling them in between when needed
  for (int done = 0; done < nsteps;) {
    int k = nsteps - done;
    if (inlet >= 0) {
      int passes = 0;
      if (int e = seq_room(c, &passes)) return e;
      k = std::min(k, passes);
      c->seq_ub += k * kBzMax;
    }
    if (c->time_pairs) {
      for (int s = 0; s < k; ++s)
        if (int e = record_step(c, wall, dt, inlet, outlet)) return e;
    } else if (int e = replay_step(c, c->inflow_graph,
                                   300 + 16 * wall + 4 * (inlet + 1) + (outlet + 1), dt, k,
                                   [&] { return record_step(c, wall, dt, inlet, outlet); })) {
      return e;
    }
    done += k;
  }
  return 0;
}

int tsph_water_elevation(tsph_ctx *c, double x, double y0, int inverse, double *level) {
  if (!c || !level) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return run_water_elevation(c, x, y0, inverse, level);
}

int tsph_interpolate_point(tsph_ctx *c, double x, double y, double *out4) {
  if (!c || !out4) return fail(TSPH_EINVAL, "null argument");
  if (int e_ = single_only(c)) return e_;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return run_interpolate_point(c, x, y, out4);
}

int tsph_get_counters(tsph_ctx *c, tsph_counters *out) {
  if (!c || !out) return fail(TSPH_EINVAL, "null argument");
  std::vector<unsigned long long> v(counter_words(c));
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(v.data(), c->counters, v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int w = 0; w < c->wave_ctr_cap; ++w)
    for (int q = 0; q < 3; ++q) v[C_PAIRS + q] += v[kWaveCtr0 + 4 * w + q];  // C_PAIRS, C_CAND, C_ACTIVE
  out->pairs = v[C_PAIRS];
  out->candidates = v[C_CAND];
  out->active = v[C_ACTIVE];
  out->pair_passes = v[C_PASSES];
  out->density_out_of_range = v[C_BADRHO];
  out->outside_grid = v[C_OUTGRID];
  out->removed = v[C_REMOVED];
  out->appended = v[C_APPENDED];
  out->zone_rejected = v[C_BZREJECT];
  out->zone_chained = v[C_BZCHAIN];
  out->reinit_out_of_range = v[C_REINIT];
  out->scan_stalls = v[C_SCANSTALL];
  return 0;
}

int tsph_reset_counters(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  // on the context's stream: a hipMemset on the null stream is not ordered
  // with the (non-blocking) context stream and could land after the next
  // launch's first counts
  TSPH_HIP(hipMemsetAsync(c->counters, 0, counter_words(c) * sizeof(unsigned long long), c->stream));
  TSPH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int tsph_set_tile_rows(tsph_ctx *c, int rows) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (rows < 1 || rows > kMaxTileRows) return fail(TSPH_EINVAL, "tile rows must be 1..24");
  if (rows != c->tile_rows) {
    drop_graph(c);
    c->tiles_valid = false;
    c->tile_rows = rows;
    if (int e = clear_tile_costs(c)) return e;  // costs of the old tile layout
  }
  return update_tile_base(c);
}

int tsph_set_waves_per_cu(tsph_ctx *c, int waves) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (waves < 1 || waves > 32) return fail(TSPH_EINVAL, "waves per CU must be 1..32");
  if (waves != c->waves_per_cu) drop_graph(c);
  c->waves_per_cu = waves;
  return 0;
}

int tsph_set_scan_wait(tsph_ctx *c, int polls) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (polls < -1) return fail(TSPH_EINVAL, "scan polls must be -1 or more");
  if (polls != c->scan_wait) drop_graph(c);
  c->scan_wait = polls;
  return 0;
}

int tsph_get_tile_rows(tsph_ctx *c, int *rows) {
  if (!c || !rows) return fail(TSPH_EINVAL, "null argument");
  *rows = c->tile_rows;
  return 0;
}

int tsph_synchronize(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  TSPH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int tsph_pair_timing(tsph_ctx *c, int max_launches) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  TSPH_HIP(hipStreamSynchronize(c->stream));
  if (max_launches <= 0) {
    c->time_pairs = false;
    c->tev_used = 0;
    return 0;
  }
  while ((int)c->tev.size() < 2 * max_launches) {
    hipEvent_t ev;
    TSPH_HIP(hipEventCreate(&ev));
    c->tev.push_back(ev);
  }
  c->tev_used = 0;
  c->time_pairs = true;
  return 0;
}

int tsph_pair_timing_read(tsph_ctx *c, double *total_ms, int *launches) {
  if (!c || !total_ms || !launches) return fail(TSPH_EINVAL, "null argument");
  TSPH_HIP(hipStreamSynchronize(c->stream));
  double tot = 0;
  for (int k = 0; k + 1 < c->tev_used; k += 2) {
    float ms = 0.f;
    TSPH_HIP(hipEventElapsedTime(&ms, c->tev[k], c->tev[k + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = c->tev_used / 2;
  c->tev_used = 0;
  return 0;
}


// ---- other WCSPH integrators and density reinitialisation ---------------------

int tsph_set_compensated(tsph_ctx *c, int on) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  TSPH_HIP(hipStreamSynchronize(c->stream));
  drop_graph(c);
  if (!on) {
    for (Soa *d : {&c->cur, &c->alt}) {
      if (d->rc) (void)hipFree(d->rc);
      if (d->rhoc) (void)hipFree(d->rhoc);
      d->rc = nullptr;
      d->rhoc = nullptr;
    }
    return 0;
  }
  if (!c->cur.rc) {
    for (Soa *d : {&c->cur, &c->alt})
      if (int e = dalloc(&d->rc, c->capacity)) return e;
    for (Soa *d : {&c->cur, &c->alt})
      if (int e = dalloc(&d->rhoc, c->capacity)) return e;
    TSPH_HIP(hipMemsetAsync(c->cur.rc, 0, sizeof(double2) * (size_t)c->capacity, c->stream));
    TSPH_HIP(hipMemsetAsync(c->cur.rhoc, 0, sizeof(double) * (size_t)c->capacity, c->stream));
    TSPH_HIP(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int tsph_history_enable(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  int e = 0;
  if (!c->has_history) {
    for (Soa *d : {&c->cur, &c->alt})
      if ((e = dalloc(&d->v_oo, c->capacity)) || (e = dalloc(&d->rho_oo, c->capacity)) ||
          (e = dalloc(&d->rho_tmp, c->capacity)))
        return e;
    c->has_history = true;
  }
  drop_graph(c);
  return launch_history_init(c);
}

static int history_ready(tsph_ctx *c) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  if (!c->has_history) return fail(TSPH_ESTATE, "history fields not enabled (tsph_history_enable)");
  return 0;
}

int tsph_history_upload(tsph_ctx *c, const tsph_history *h) {
  if (int e = history_ready(c)) return e;
  if (!h) return fail(TSPH_EINVAL, "null argument");
  int e = 0, n = 0;
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(&n, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> order;
  std::vector<uint32_t> seq;
  std::vector<uint8_t> type;
  if ((e = live_order(c, n, -1, order, seq, type))) return e;
  const int m = (int)order.size();
  if (h->np != m) return fail(TSPH_EINVAL, "history field count differs from the particle count");
  std::vector<double> b(2 * (size_t)n + 2, 0.0);
  auto put = [&](void *dst, const double *src, int w) -> int {
    if (!src || !n) return 0;
    TSPH_HIP(hipMemcpy(b.data(), dst, w * sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k)
      for (int u = 0; u < w; ++u) b[(size_t)w * order[k] + u] = src[(size_t)w * k + u];
    TSPH_HIP(hipMemcpy(dst, b.data(), w * sizeof(double) * n, hipMemcpyHostToDevice));
    return 0;
  };
  const Soa &d = c->cur;
  if ((e = put(d.v_oo, h->v_oo, 2)) || (e = put(d.rho_oo, h->rho_oo, 1)) ||
      (e = put(d.rho_tmp, h->rho_temp, 1)))
    return e;
  return 0;
}

int tsph_history_download(tsph_ctx *c, tsph_history *h) {
  if (int e = history_ready(c)) return e;
  if (!h) return fail(TSPH_EINVAL, "null argument");
  int e = 0, n = 0;
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(&n, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> order;
  std::vector<uint32_t> seq;
  std::vector<uint8_t> type;
  if ((e = live_order(c, n, -1, order, seq, type))) return e;
  const int m = (int)order.size();
  h->np = m;
  std::vector<double> b(2 * (size_t)n + 2);
  auto get = [&](double *dst, const void *src, int w) -> int {
    if (!dst || !n) return 0;
    TSPH_HIP(hipMemcpy(b.data(), src, w * sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k)
      for (int u = 0; u < w; ++u) dst[(size_t)w * k + u] = b[(size_t)w * order[k] + u];
    return 0;
  };
  const Soa &d = c->cur;
  if ((e = get(h->v_oo, d.v_oo, 2)) || (e = get(h->rho_oo, d.rho_oo, 1)) ||
      (e = get(h->rho_temp, d.rho_tmp, 1)))
    return e;
  return 0;
}

int tsph_verlet(tsph_ctx *c, double dt) {
  if (int e = history_ready(c)) return e;
  return launch_verlet(c, dt, false);
}

int tsph_euler(tsph_ctx *c, double dt) {
  if (int e = history_ready(c)) return e;
  return launch_verlet(c, dt, true);
}

int tsph_leapfrog_one(tsph_ctx *c, double dt) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  return launch_leapfrog_one(c, dt);
}

int tsph_leapfrog_two(tsph_ctx *c, double dt, int step) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  return launch_leapfrog_two(c, dt, step);
}

int tsph_density_reinit_sum(tsph_ctx *c) {
  if (int e = history_ready(c)) return e;
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  return launch_reinit_sum(c);
}

int tsph_density_reinit_overwrite(tsph_ctx *c) {
  if (int e = history_ready(c)) return e;
  return launch_reinit_overwrite(c);
}

static int step_checks(tsph_ctx *c, int wall, int step, int nsteps) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  if (wall < TSPH_WALL_DBC || wall > TSPH_WALL_BT) return fail(TSPH_EINVAL, "unknown wall treatment");
  if (step < 1 || nsteps < 0) return fail(TSPH_EINVAL, "steps are numbered from 1");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  if (wall == TSPH_WALL_MDBC && !c->has_gnr) return fail(TSPH_ESTATE, "mDBC needs ghost nodes (gnr)");
  if (wall == TSPH_WALL_BT && !c->has_n) return fail(TSPH_ESTATE, "BT needs wall normals (n)");
  return 0;
}

int tsph_step_verlet(tsph_ctx *c, int wall, double dt, int step, int nsteps) {
  if (int e = step_checks(c, wall, step, nsteps)) return e;
  if (!c->has_history) return fail(TSPH_ESTATE, "history fields not enabled (tsph_history_enable)");
  for (int s = step; s < step + nsteps; ++s) {
    // src: `if(step > 1 && step%20 != 0) Verlet else Euler`
    const bool euler = !(s > 1 && s % 20 != 0);
    if (c->time_pairs) {
      if (int e = record_verlet_step(c, wall, dt, euler)) return e;
      continue;
    }
    if (int e = replay_step(c, euler ? c->euler_graph : c->verlet_graph, 500 + 4 * wall + euler, dt,
                            1, [&] { return record_verlet_step(c, wall, dt, euler); }))
      return e;
  }
  return 0;
}

int tsph_step_leapfrog(tsph_ctx *c, int wall, double dt, int step, int nsteps) {
  if (int e = step_checks(c, wall, step, nsteps)) return e;
  int s = step;
  const int last = step + nsteps;
  // the first step differs (no part one; part two's first-step branch): launched eagerly
  for (; s < last && (s == 1 || c->time_pairs); ++s)
    if (int e = record_leapfrog_step(c, wall, dt, s)) return e;
  if (s >= last) return 0;
  return replay_step(c, c->lf_graph, 520 + wall, dt, last - s,
                     [&] { return record_leapfrog_step(c, wall, dt, 2); });
}

int tsph_step_symplectic_reinit(tsph_ctx *c, int wall, double dt, int step, int nsteps,
                                int interval) {
  if (int e = step_checks(c, wall, step, nsteps)) return e;
  if (!c->has_history) return fail(TSPH_ESTATE, "history fields not enabled (tsph_history_enable)");
  if (interval < 1) return fail(TSPH_EINVAL, "reinitialisation interval must be positive");
  for (int s = step; s < step + nsteps; ++s) {
    int e = c->time_pairs ? record_step(c, wall, dt)
                          : replay_step(c, c->graph, wall, dt, 1, [&] { return record_step(c, wall, dt); });
    if (e) return e;
    if (s % interval == 0 && ((e = launch_reinit_sum(c)) || (e = launch_reinit_overwrite(c)))) return e;
  }
  return 0;
}

// ---- binary snapshot and resume -------------------------------------------------
// File: "TSPHSNP1", int32 version 1, np, flags (1 gnr, 2 n, 4 ALE variables,
// 8 history fields), 0; then in reference order int32 type[np], index[np];
// f64 r, v, a, r_o, v_o [2np each], p, rho, drho, rho_o [np each]; gnr, n
// [2np each, if flagged]; ALE omega, omegarho, omega_o, omegarho_o, domega,
// domegarho [np each], omegav, omegav_o, domegav, gradrho, gradvx, gradvy [2np
// each]; history v_oo [2np], rho_oo, rhoTemp [np each].  Little endian.

static const char kSnapMagic[8] = {'T', 'S', 'P', 'H', 'S', 'N', 'P', '1'};

int tsph_save_state(tsph_ctx *c, const char *path) {
  if (!c || !path) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  int n = 0, e = 0;
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(&n, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost));
  const size_t N = (size_t)std::max(n, 1);
  std::vector<int32_t> type(N), index(N);
  std::vector<double> s1[4], s2[7];
  for (auto &x : s1) x.resize(N);
  for (auto &x : s2) x.resize(2 * N);
  tsph_particles h{};
  h.type = type.data(); h.index = index.data();
  h.p = s1[0].data(); h.rho = s1[1].data(); h.drho = s1[2].data(); h.rho_o = s1[3].data();
  h.r = s2[0].data(); h.v = s2[1].data(); h.a = s2[2].data(); h.r_o = s2[3].data();
  h.v_o = s2[4].data();
  h.gnr = c->has_gnr ? s2[5].data() : nullptr;
  h.n = c->has_n ? s2[6].data() : nullptr;
  if ((e = download_impl(c, &h, -1))) return e;
  const int m = h.np;
  std::vector<double> a1[6], a2[6], h1[2], h2;
  if (c->has_ale) {
    for (auto &x : a1) x.resize(N);
    for (auto &x : a2) x.resize(2 * N);
    tsph_ale_fields f{};
    f.omega = a1[0].data(); f.omegarho = a1[1].data(); f.omega_o = a1[2].data();
    f.omegarho_o = a1[3].data(); f.domega = a1[4].data(); f.domegarho = a1[5].data();
    f.omegav = a2[0].data(); f.omegav_o = a2[1].data(); f.domegav = a2[2].data();
    f.gradrho = a2[3].data(); f.gradvx = a2[4].data(); f.gradvy = a2[5].data();
    if ((e = tsph_ale_download(c, &f))) return e;
  }
  if (c->has_history) {
    for (auto &x : h1) x.resize(N);
    h2.resize(2 * N);
    tsph_history hh{};
    hh.v_oo = h2.data(); hh.rho_oo = h1[0].data(); hh.rho_temp = h1[1].data();
    if ((e = tsph_history_download(c, &hh))) return e;
  }
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return fail(TSPH_EINVAL, (std::string("cannot write ") + path).c_str());
  const int32_t head[4] = {1, m, (c->has_gnr ? 1 : 0) | (c->has_n ? 2 : 0) | (c->has_ale ? 4 : 0) |
                                     (c->has_history ? 8 : 0), 0};
  bool ok = std::fwrite(kSnapMagic, 1, 8, fp) == 8 && std::fwrite(head, 4, 4, fp) == 4;
  auto w = [&](const void *src, size_t bytes) {
    if (ok && bytes) ok = std::fwrite(src, 1, bytes, fp) == bytes;
  };
  const size_t b1 = 8 * (size_t)m, b2 = 16 * (size_t)m;
  w(type.data(), 4 * (size_t)m);
  w(index.data(), 4 * (size_t)m);
  for (int k = 0; k < 5; ++k) w(s2[k].data(), b2);
  for (int k = 0; k < 4; ++k) w(s1[k].data(), b1);
  if (c->has_gnr) w(s2[5].data(), b2);
  if (c->has_n) w(s2[6].data(), b2);
  if (c->has_ale) {
    for (int k = 0; k < 6; ++k) w(a1[k].data(), b1);
    for (int k = 0; k < 6; ++k) w(a2[k].data(), b2);
  }
  if (c->has_history) {
    w(h2.data(), b2);
    w(h1[0].data(), b1);
    w(h1[1].data(), b1);
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) return fail(TSPH_EINVAL, (std::string("write failed: ") + path).c_str());
  return 0;
}

int tsph_load_state(tsph_ctx *c, const char *path) {
  if (!c || !path) return fail(TSPH_EINVAL, "null argument");
  if (int e = single_only(c)) return e;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return fail(TSPH_EINVAL, (std::string("cannot read ") + path).c_str());
  char magic[8];
  int32_t head[4];
  bool ok = std::fread(magic, 1, 8, fp) == 8 && std::fread(head, 4, 4, fp) == 4 &&
            std::memcmp(magic, kSnapMagic, 8) == 0 && head[0] == 1 && head[1] >= 0;
  const int m = ok ? head[1] : 0;
  const int flags = ok ? head[2] : 0;
  const size_t N = (size_t)std::max(m, 1);
  std::vector<int32_t> type(N), index(N);
  std::vector<double> s1[4], s2[7], a1[6], a2[6], h1[2], h2;
  auto r = [&](void *dst, size_t bytes) {
    if (ok && bytes) ok = std::fread(dst, 1, bytes, fp) == bytes;
  };
  const size_t b1 = 8 * (size_t)m, b2 = 16 * (size_t)m;
  for (auto &x : s1) x.resize(N);
  for (auto &x : s2) x.resize(2 * N);
  r(type.data(), 4 * (size_t)m);
  r(index.data(), 4 * (size_t)m);
  for (int k = 0; k < 5; ++k) r(s2[k].data(), b2);
  for (int k = 0; k < 4; ++k) r(s1[k].data(), b1);
  if (flags & 1) r(s2[5].data(), b2);
  if (flags & 2) r(s2[6].data(), b2);
  if (flags & 4) {
    for (auto &x : a1) x.resize(N);
    for (auto &x : a2) x.resize(2 * N);
    for (int k = 0; k < 6; ++k) r(a1[k].data(), b1);
    for (int k = 0; k < 6; ++k) r(a2[k].data(), b2);
  }
  if (flags & 8) {
    for (auto &x : h1) x.resize(N);
    h2.resize(2 * N);
    r(h2.data(), b2);
    r(h1[0].data(), b1);
    r(h1[1].data(), b1);
  }
  std::fclose(fp);
  if (!ok) return fail(TSPH_EINVAL, (std::string("not a complete snapshot: ") + path).c_str());
  tsph_particles h{};
  h.np = m;
  h.type = type.data(); h.index = index.data();
  h.p = s1[0].data(); h.rho = s1[1].data(); h.drho = s1[2].data(); h.rho_o = s1[3].data();
  h.r = s2[0].data(); h.v = s2[1].data(); h.a = s2[2].data(); h.r_o = s2[3].data();
  h.v_o = s2[4].data();
  h.gnr = (flags & 1) ? s2[5].data() : nullptr;
  h.n = (flags & 2) ? s2[6].data() : nullptr;
  int e = 0;
  if ((e = tsph_upload(c, &h))) return e;
  if (flags & 4) {
    tsph_ale_fields f{};
    f.np = m;
    f.omega = a1[0].data(); f.omegarho = a1[1].data(); f.omega_o = a1[2].data();
    f.omegarho_o = a1[3].data(); f.domega = a1[4].data(); f.domegarho = a1[5].data();
    f.omegav = a2[0].data(); f.omegav_o = a2[1].data(); f.domegav = a2[2].data();
    f.gradrho = a2[3].data(); f.gradvx = a2[4].data(); f.gradvy = a2[5].data();
    if ((e = tsph_ale_upload(c, &f))) return e;
  }
  if (flags & 8) {
    if ((e = tsph_history_enable(c))) return e;
    tsph_history hh{};
    hh.np = m;
    hh.v_oo = h2.data(); hh.rho_oo = h1[0].data(); hh.rho_temp = h1[1].data();
    if ((e = tsph_history_upload(c, &hh))) return e;
  }
  return 0;
}

// ---- batched ensembles -------------------------------------------------------

int tsph_batch_create(const tsph_params *prm, const double *dt, int ncases, tsph_ctx **out) {
  if (!prm || !dt || !out) return fail(TSPH_EINVAL, "null argument");
  *out = nullptr;
  if (ncases < 1 || ncases > TSPH_BATCH_MAX_CASES)
    return fail(TSPH_EINVAL, "case count must be 1..255");
  std::vector<Consts> ks(ncases);
  std::vector<int> caps(ncases);
  long long cells = 0, cap = 0;
  for (int q = 0; q < ncases; ++q) {
    if (prm[q].precision != prm[0].precision) return fail(TSPH_EINVAL, "cases differ in precision");
    if (!(dt[q] > 0.)) return fail(TSPH_EINVAL, "time steps must be positive");
    if (prm[q].capacity < 0) return fail(TSPH_EINVAL, "negative capacity");
    if (int e = make_consts(&prm[q], ks[q])) return e;
    ks[q].cell_base = (int)cells;
    ks[q].dt = dt[q];
    cells += ks[q].ncells;
    cap += prm[q].capacity;
    caps[q] = prm[q].capacity;
    if (cells > (1ll << 30)) return fail(TSPH_EINVAL, "cell grids too large");
    if (cap > (1ll << 30)) return fail(TSPH_ECAPACITY, "capacity too large");
  }
  tsph_ctx *c = new tsph_ctx();
  c->prm = prm[0];
  c->prm.capacity = (int)cap;
  c->capacity = (int)cap;
  c->tile_rows = 7;
  set_walk_defaults(c);
  c->k = ks[0];
  c->ncells_tot = (int)cells;
  c->ncases = ncases;
  c->host_cases = ks;
  c->case_cap = caps;
  int e = 0;
  if ((e = alloc_context(c)) || (e = dalloc(&c->cases, ncases)) ||
      (e = dalloc(&c->tile_base, ncases + 1)) || (e = dalloc(&c->cur.cid, c->capacity)) ||
      (e = dalloc(&c->alt.cid, c->capacity))) {
    tsph_destroy(c);
    return e;
  }
  if (hipMemcpy(c->cases, ks.data(), sizeof(Consts) * ncases, hipMemcpyHostToDevice) != hipSuccess) {
    tsph_destroy(c);
    return fail(TSPH_EHIP, "initialisation failed");
  }
  if ((e = update_tile_base(c))) {
    tsph_destroy(c);
    return e;
  }
  *out = c;
  return 0;
}

int tsph_batch_cases(tsph_ctx *c, int *ncases) {
  if (!c || !ncases) return fail(TSPH_EINVAL, "null argument");
  *ncases = c->cases ? c->ncases : 0;
  return 0;
}

int tsph_batch_grid(tsph_ctx *c, int icase, int *ncx, int *ncy, double *mass) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (!c->cases) return fail(TSPH_ESTATE, "not a batched context");
  if (icase < 0 || icase >= c->ncases) return fail(TSPH_EINVAL, "no such case");
  const Consts &k = c->host_cases[icase];
  if (ncx) *ncx = k.ncx;
  if (ncy) *ncy = k.ncy;
  if (mass) *mass = k.m;
  return 0;
}

int tsph_batch_upload(tsph_ctx *c, const tsph_particles *h) {
  if (!c || !h) return fail(TSPH_EINVAL, "null argument");
  if (!c->cases) return fail(TSPH_ESTATE, "not a batched context");
  const int nc = c->ncases;
  long long tot = 0;
  bool any_gnr = false, any_n = false;
  for (int q = 0; q < nc; ++q) {
    if (h[q].np < 0 || h[q].np > c->case_cap[q])
      return fail(TSPH_ECAPACITY, "case particle count outside 0..its capacity");
    if (h[q].np > 0 && (!h[q].type || !h[q].rho || !h[q].r))
      return fail(TSPH_EINVAL, "type, rho and r are required");
    tot += h[q].np;
    any_gnr |= h[q].gnr != nullptr;
    any_n |= h[q].n != nullptr;
  }
  // the cases one after another.  Each case's reference positions start one
  // past the previous case's last; the position between them holds the state
  // "kept", so a run of removal candidates never continues across cases.
  const int n = (int)tot;
  std::vector<int32_t> type(n), index(n);
  std::vector<uint32_t> seq(n);
  std::vector<uint8_t> cid(n);
  std::vector<double> p(n), rho(n), drho(n), rho_o(n);
  std::vector<double> r(2 * (size_t)n), v(2 * (size_t)n), a(2 * (size_t)n), r_o(2 * (size_t)n),
      v_o(2 * (size_t)n), gnr(any_gnr ? 2 * (size_t)n : 0), nrm(any_n ? 2 * (size_t)n : 0);
  c->seq_base.assign(nc + 1, 0);
  int o = 0;
  for (int q = 0; q < nc; ++q) {
    const tsph_particles &x = h[q];
    const int m = x.np;
    for (int i = 0; i < m; ++i, ++o) {
      type[o] = x.type[i];
      index[o] = x.index ? x.index[i] : i;
      seq[o] = (uint32_t)(c->seq_base[q] + i);
      cid[o] = (uint8_t)q;
      rho[o] = x.rho[i];
      p[o] = x.p ? x.p[i] : 0.;
      drho[o] = x.drho ? x.drho[i] : 0.;
      rho_o[o] = x.rho_o ? x.rho_o[i] : x.rho[i];
      for (int u = 0; u < 2; ++u) {
        const size_t g = 2 * (size_t)o + u, l = 2 * (size_t)i + u;
        r[g] = x.r[l];
        v[g] = x.v ? x.v[l] : 0.;
        a[g] = x.a ? x.a[l] : 0.;
        r_o[g] = x.r_o ? x.r_o[l] : x.r[l];
        v_o[g] = x.v_o ? x.v_o[l] : v[g];
        if (any_gnr) gnr[g] = x.gnr ? x.gnr[l] : 0.;
        if (any_n) nrm[g] = x.n ? x.n[l] : 0.;
      }
    }
    c->seq_base[q + 1] = c->seq_base[q] + m + 1;
  }
  tsph_particles all{};
  all.np = n;
  all.type = type.data();
  all.index = index.data();
  all.p = p.data();
  all.rho = rho.data();
  all.drho = drho.data();
  all.r = r.data();
  all.v = v.data();
  all.a = a.data();
  all.r_o = r_o.data();
  all.v_o = v_o.data();
  all.rho_o = rho_o.data();
  all.gnr = any_gnr ? gnr.data() : nullptr;
  all.n = any_n ? nrm.data() : nullptr;
  if (int e = upload_impl(c, &all, seq.data(), cid.data(), c->seq_base[nc])) return e;
  const uint8_t kept = 1;
  for (int q = 1; q <= nc; ++q)
    TSPH_HIP(hipMemcpy(c->seqstate + c->seq_base[q] - 1, &kept, 1, hipMemcpyHostToDevice));
  return 0;
}

int tsph_batch_get_np(tsph_ctx *c, int *np) {
  if (!c || !np) return fail(TSPH_EINVAL, "null argument");
  if (!c->cases) return fail(TSPH_ESTATE, "not a batched context");
  int n = 0;
  TSPH_HIP(hipStreamSynchronize(c->stream));
  TSPH_HIP(hipMemcpy(&n, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost));
  std::vector<uint8_t> type(n), cid(n);
  if (n > 0) {
    TSPH_HIP(hipMemcpy(type.data(), c->cur.type, n, hipMemcpyDeviceToHost));
    TSPH_HIP(hipMemcpy(cid.data(), c->cur.cid, n, hipMemcpyDeviceToHost));
  }
  std::fill(np, np + c->ncases, 0);
  for (int i = 0; i < n; ++i)
    if (type[i] != kRemoved) np[cid[i]]++;
  return 0;
}

int tsph_batch_download(tsph_ctx *c, int icase, tsph_particles *h) {
  if (!c || !h) return fail(TSPH_EINVAL, "null argument");
  if (!c->cases) return fail(TSPH_ESTATE, "not a batched context");
  if (icase < 0 || icase >= c->ncases) return fail(TSPH_EINVAL, "no such case");
  return download_impl(c, h, icase);
}

int tsph_batch_step(tsph_ctx *c, int wall, int nsteps) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (!c->cases) return fail(TSPH_ESTATE, "not a batched context");
  if (wall < TSPH_WALL_DBC || wall > TSPH_WALL_BT) return fail(TSPH_EINVAL, "unknown wall treatment");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  if (wall == TSPH_WALL_MDBC && !c->has_gnr) return fail(TSPH_ESTATE, "mDBC needs ghost nodes (gnr)");
  if (wall == TSPH_WALL_BT && !c->has_n) return fail(TSPH_ESTATE, "BT needs wall normals (n)");
  if (nsteps <= 0) return 0;
  // each case advances by its own time step (its entry of the case table)
  if (c->time_pairs) {
    for (int s = 0; s < nsteps; ++s)
      if (int e = record_step(c, wall, 0.)) return e;
    return 0;
  }
  return replay_step(c, c->graph, 400 + wall, 0., nsteps, [&] { return record_step(c, wall, 0.); },
                     kGraphSteps);
}

// ---- autotuner -----------------------------------------------------------------

int tsph_set_balanced(tsph_ctx *c, int on) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  bool &b = c->balanced[c->prm.precision == TSPH_PREC_MIXED ? 1 : 0];
  if (b != (on != 0)) drop_graph(c);
  b = on != 0;
  return 0;
}

int tsph_set_cell_order(tsph_ctx *c, int mode) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (mode < 0 || mode > 2) return fail(TSPH_EINVAL, "cell order must be 0, 1 or 2");
  if (c->cell_order != mode) drop_graph(c);
  c->cell_order = mode;
  return 0;
}

int tsph_autotune(tsph_ctx *c, int variant, int *tile_rows, int *waves_per_cu, int *balanced,
                  double *best_ms, double *particles_per_lane) {
  if (!c) return fail(TSPH_EINVAL, "null argument");
  if (variant < TSPH_FORCES_WCSPH || variant > TSPH_FORCES_BT) return fail(TSPH_EINVAL, "unknown variant");
  if (!c->cells_valid) return fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  if (variant == TSPH_FORCES_BT && !c->has_n) return fail(TSPH_ESTATE, "BT forces need wall normals (n)");
  if (c->time_pairs) return fail(TSPH_ESTATE, "pair timing is active");
  // depths from about one particle per lane to about three (8-9 particles
  // per non-empty cell at the configured resolutions)
  static const int kRows[] = {5, 6, 7, 8, 10, 12, 14, 16, 20, 24};
  static const int kWaves[] = {4, 8, 12, 16};
  constexpr int kReps = 3;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TSPH_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(TSPH_EHIP, "hipEventCreate failed");
  }
  int best_r = c->tile_rows, best_w = c->waves_per_cu, best_b = 0, e = 0;
  float best = 1e30f;
  // every candidate recomputes the same pair pass: its outputs a and drho are
  // overwritten from the unchanged state
  for (int bal = 0; bal < 2 && !e; ++bal) {
    if ((e = tsph_set_balanced(c, bal))) break;
    for (int r : kRows) {
      if ((e = tsph_set_tile_rows(c, r))) break;
      for (int w : kWaves) {
        if ((e = tsph_set_waves_per_cu(c, w)) || (e = launch_forces(c, variant))) break;  // warm-up
        // a pooled walk is timed with chunks cut by the tile costs the
        // warm-up recorded (its own chunks were cut by estimates)
        c->list_groups = 0;
        if ((e = launch_forces(c, variant))) break;
        // Every timed pass follows a cell-list build, as in a step: the build
        // re-sorts the sorted state (which leaves it as it is) and cuts a
        // pooled pass's chunks.  Passes repeated on one build favoured the
        // static walk, which then ran slower in whole steps (740k-particle dam
        // break: 0.68 against 0.63 ms per FP64 step).
        float ms = 0.f;
        for (int k = 0; k < kReps && !e; ++k) {
          c->cut_groups = pass_groups(c, variant);
          c->cut_table = 0;
          if ((e = launch_cell_list(c, false))) break;
          if (c->prm.precision == TSPH_PREC_MIXED) {
            if ((e = launch_mixed_records(c))) break;
            c->records_valid = true;
          }
          float rep = 0.f;
          if (hipEventRecord(e0, c->stream) != hipSuccess) { e = fail(TSPH_EHIP, "event record failed"); break; }
          if ((e = launch_forces(c, variant))) break;
          if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&rep, e0, e1) != hipSuccess) {
            e = fail(TSPH_EHIP, "event timing failed");
            break;
          }
          ms += rep;
        }
        if (e) break;
        if (ms < best) {
          best = ms;
          best_r = r;
          best_w = w;
          best_b = bal;
        }
      }
      if (e) break;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e) return e;
  // a last pass with the chosen setting, so that a and drho are what a pair
  // pass with it leaves (the mixed path rounds in a tile-relative frame)
  if ((e = tsph_set_balanced(c, best_b)) || (e = tsph_set_tile_rows(c, best_r)) ||
      (e = tsph_set_waves_per_cu(c, best_w)) || (e = launch_forces(c, variant)))
    return e;
  if (tile_rows) *tile_rows = best_r;
  if (waves_per_cu) *waves_per_cu = best_w;
  if (balanced) *balanced = best_b;
  if (best_ms) *best_ms = best / kReps;
  if (particles_per_lane) {
    // the chosen depth's particles per lane: every particle belongs to one
    // tile of the list, whose 64 lanes take its particles in turn
    int ntiles = 0, np = 0;
    TSPH_HIP(hipMemcpyAsync(&ntiles, c->scal + S_NTILES, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TSPH_HIP(hipMemcpyAsync(&np, c->scal + S_NP, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TSPH_HIP(hipStreamSynchronize(c->stream));
    *particles_per_lane = ntiles > 0 ? (double)np / (64.0 * ntiles) : 0.;
  }
  return 0;
}

}  // extern "C"
