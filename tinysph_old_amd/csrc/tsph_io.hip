// Output path (host code of the library):
//   * write_to_ASCII_VTK, src/SPH_data_to_vtk.h:4-51: the particles as legacy
//     ASCII VTK polydata (type, pressure, density, velocity).  tsph_write_vtk
//     queues the device-to-host copies on the context's stream into pinned
//     buffers and leaves the formatting and the file to a host thread, so the
//     following steps run while the file is written.
//   * GenerateInterpol, src/SPH_dataInterpolation.h:13-91, with
//     write_INTERPOLATED_mesh_to_ASCII_VTK, src/SPH_grid_to_vtk.h:139-243: the
//     interpolation grid is computed on the device (one lane per point).
// The text is produced with the same stream operations as the reference
// writers, so the files are byte-identical for the same values.
#include <algorithm>
#include <fstream>
#include <mutex>
#include <numeric>

#include "tsph_internal.h"

namespace tsph {

// Staging of tsph_interpolate_grid_noalloc: the grid values of one call, laid
// out as rho[n], p[n], v[2n] (n = ncx ncy <= TSPH_GRID_SCRATCH_POINTS),
// 16 MiB.  Each GPU holds its own copy at its own address until the process
// ends; every call on a GPU shares that GPU's copy.  In host code `scratch` is
// only a handle: the device address comes from hipGetSymbolAddress, once per
// device (grid_scratch_ptr), and only that address is passed to kernels and
// copies.
static __device__ __align__(256) double scratch[4 * (size_t)TSPH_GRID_SCRATCH_POINTS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

int grid_scratch_ptr(double **out) {
  int dev = 0;
  TSPH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index beyond the scratch table");
    return TSPH_EHIP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return TSPH_EHIP;
    }
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

namespace {

int io_fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

// the particle file of write_to_ASCII_VTK (np particles in reference order)
bool format_particles(const char *path, int np, const int32_t *type, const double *r,
                      const double *p, const double *rho, const double *v) {
  std::ofstream file(path);
  if (!file) return false;
  const unsigned n = (unsigned)np;
  file << "# vtk DataFile Version 3.0" << '\n' << "vtk output" << '\n' << "ASCII" << '\n'
       << "DATASET POLYDATA" << '\n';
  file << "POINTS " << n << " float" << '\n';
  for (int i = 0; i < np; ++i) file << r[2 * i] << " " << r[2 * i + 1] << " 0" << '\n';
  file << "POINT_DATA " << n << '\n' << "FIELDS FieldData 4" << '\n';
  file << "type 1 " << n << " int" << '\n';
  for (int i = 0; i < np; ++i) file << (int)type[i] << '\n';
  file << "pressure 1 " << n << " float" << '\n';
  for (int i = 0; i < np; ++i) file << p[i] << '\n';
  file << "density 1 " << n << " float" << '\n';
  for (int i = 0; i < np; ++i) file << rho[i] << '\n';
  file << "velocity 3 " << n << " float" << '\n';
  for (int i = 0; i < np; ++i) file << v[2 * i] << " " << v[2 * i + 1] << " 0" << '\n';
  file.close();
  return !file.fail();
}

// the grid file of write_INTERPOLATED_mesh_to_ASCII_VTK; values at
// POS(i, j) = j + i ncy (src/SPH_defs.h:36), points in row order
bool format_grid(const char *path, int ncx, int ncy, double dh, double x0, double y0,
                 const double *rho, const double *p, const double *v) {
  std::ofstream file(path);
  if (!file) return false;
  const int cells = ncx * ncy;
  file << "# vtk DataFile Version 3.0" << '\n';
  file << "vtk output" << '\n';
  file << "ASCII" << '\n';
  file << "DATASET STRUCTURED_GRID" << '\n';
  file << "DIMENSIONS " << ncx << " " << ncy << " 1" << '\n';
  file << "POINTS " << ncx * ncy << " float" << '\n';
  // float counters, as the reference's loop
  for (float j = 0; j < ncy; j++)
    for (float i = 0; i < ncx; i++) file << x0 + i * dh << " " << y0 + j * dh << " 0" << '\n';
  file << "POINT_DATA " << cells << '\n';
  file << "FIELD FieldData " << 3 << '\n';
  file << "DENSITY 1 " << cells << " float" << '\n';
  for (int j = 0; j < ncy; j++)
    for (int i = 0; i < ncx; i++) file << rho[j + i * ncy] << '\n';
  file << "PRESSURE 1 " << cells << " float" << '\n';
  for (int j = 0; j < ncy; j++)
    for (int i = 0; i < ncx; i++) file << p[j + i * ncy] << '\n';
  file << "VELOCITY 3 " << cells << " float" << '\n';
  for (int j = 0; j < ncy; j++)
    for (int i = 0; i < ncx; i++) {
      const int t = j + i * ncy;
      file << v[2 * t] << " " << v[2 * t + 1] << " 0" << '\n';
    }
  file.close();
  return !file.fail();
}

// the cell-grid file of write_mesh_to_ASCII_VTK, src/SPH_grid_to_vtk.h:6-67:
// (ncx+1) x (ncy+1) grid points at x_0 + i kh - of, y_0 + j kh - of, then per
// cell (in POS order, src/SPH_defs.h:36) its particle count and its 1-D index
bool format_cells(const char *path, int ncx, int ncy, double kh, double of, double x_0,
                  double y_0, const int32_t *count) {
  std::ofstream file(path);
  if (!file) return false;
  const int px = ncx + 1, py = ncy + 1;
  const int cells = ncx * ncy;
  file << "# vtk DataFile Version 3.0" << '\n' << "vtk output" << '\n' << "ASCII" << '\n'
       << "DATASET STRUCTURED_GRID" << '\n' << "DIMENSIONS " << px << " " << py << " 1" << '\n'
       << "POINTS " << px * py << " float" << '\n';
  for (int i = 0; i < px; i++)
    for (int j = 0; j < py; j++) file << x_0 + i * kh - of << " " << y_0 + j * kh - of << " 0" << '\n';
  file << "CELL_DATA " << cells << '\n' << "FIELD FieldData " << 2 << '\n';
  file << "NumberOfParticles 1 " << cells << " int" << '\n';
  for (int x = 0; x < ncx; x++)
    for (int y = 0; y < ncy; y++) file << (unsigned)count[y + x * ncy] << '\n';
  file << "Cell1DIndex 1 " << cells << " int" << '\n';
  for (int x = 0; x < ncx; x++)
    for (int y = 0; y < ncy; y++) file << y + x * ncy << '\n';
  file.close();
  return !file.fail();
}

// Output_hydrostatic_pressure, src/SPH_measure_pressure.h:17-34: one line per
// fluid particle in reference order, p / (rho0 g H) and y / H for the 0.5 m
// still-water column
bool format_hydrostatic(const char *path, int np, const int32_t *type, const double *r,
                        const double *p) {
  std::ofstream file(path);
  if (!file) return false;
  const double H_coef = 1. / 0.5;
  const double p_coef = 1. / (9.81 * 1000. * 0.5);
  for (int i = 0; i < np; i++) {
    if (type[i] != 0) continue;
    file << p[i] * p_coef << " " << r[2 * i + 1] * H_coef << '\n';
  }
  file.close();
  return !file.fail();
}

// one asynchronous particle file: pinned copies of the slot arrays, the event
// that marks them complete, and the target path
struct VtkJob {
  std::string path;
  int slots = 0;
  hipEvent_t done = nullptr;
  int *np = nullptr;
  uint32_t *seq = nullptr;
  uint8_t *type = nullptr;
  double2 *r = nullptr, *v = nullptr;
  double *p = nullptr, *rho = nullptr;
};

void free_job(VtkJob &j) {
  void *ps[] = {j.np, j.seq, j.type, j.r, j.v, j.p, j.rho};
  for (void *q : ps)
    if (q) (void)hipHostFree(q);
  if (j.done) (void)hipEventDestroy(j.done);
}

std::string write_job(VtkJob &j) {
  if (hipEventSynchronize(j.done) != hipSuccess) return "tsph_write_vtk: copy failed";
  const int n = std::min(*j.np, j.slots);
  // live slots in reference order (removed particles skipped)
  std::vector<int> order;
  order.reserve(n);
  for (int i = 0; i < n; ++i)
    if (j.type[i] != kRemoved) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return j.seq[a] < j.seq[b]; });
  const int m = (int)order.size();
  std::vector<int32_t> type(m);
  std::vector<double> r(2 * (size_t)m), v(2 * (size_t)m), p(m), rho(m);
  for (int k = 0; k < m; ++k) {
    const int i = order[k];
    type[k] = j.type[i];
    r[2 * k] = j.r[i].x; r[2 * k + 1] = j.r[i].y;
    v[2 * k] = j.v[i].x; v[2 * k + 1] = j.v[i].y;
    p[k] = j.p[i];
    rho[k] = j.rho[i];
  }
  if (!format_particles(j.path.c_str(), m, type.data(), r.data(), p.data(), rho.data(), v.data()))
    return "tsph_write_vtk: cannot write " + j.path;
  return std::string();
}

}  // namespace

}  // namespace tsph

using namespace tsph;

extern "C" {

int tsph_format_vtk(const tsph_particles *h, const char *path) {
  if (!h || !path || (h->np > 0 && (!h->type || !h->r || !h->p || !h->rho || !h->v)))
    return io_fail(TSPH_EINVAL, "tsph_format_vtk: null argument");
  if (!format_particles(path, h->np, h->type, h->r, h->p, h->rho, h->v))
    return io_fail(TSPH_EINVAL, std::string("tsph_format_vtk: cannot write ") + path);
  return 0;
}

int tsph_format_interpolated_vtk(int ncx, int ncy, double dh, double x0, double y0,
                                 const double *rho, const double *p, const double *v,
                                 const char *path) {
  if (!path || ncx < 0 || ncy < 0 || (ncx * ncy > 0 && (!rho || !p || !v)))
    return io_fail(TSPH_EINVAL, "tsph_format_interpolated_vtk: invalid argument");
  if (!format_grid(path, ncx, ncy, dh, x0, y0, rho, p, v))
    return io_fail(TSPH_EINVAL, std::string("tsph_format_interpolated_vtk: cannot write ") + path);
  return 0;
}

int tsph_format_grid_vtk(int ncx, int ncy, double kh, double of, double x_0, double y_0,
                         const int32_t *counts, const char *path) {
  if (!path || ncx < 0 || ncy < 0 || (ncx * ncy > 0 && !counts))
    return io_fail(TSPH_EINVAL, "tsph_format_grid_vtk: invalid argument");
  if (!format_cells(path, ncx, ncy, kh, of, x_0, y_0, counts))
    return io_fail(TSPH_EINVAL, std::string("tsph_format_grid_vtk: cannot write ") + path);
  return 0;
}

int tsph_write_grid_vtk(tsph_ctx *c, double x_0, double y_0, const char *path) {
  if (!c || !path) return io_fail(TSPH_EINVAL, "null argument");
  if (c->cases) return io_fail(TSPH_ESTATE, "not available on a batched context");
  if (!c->cells_valid) return io_fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  const int ncx = c->k.ncx, ncy = c->k.ncy, cells = ncx * ncy;
  std::vector<int32_t> start((size_t)cells + 1), count((size_t)std::max(cells, 1));
  if (hipMemcpyAsync(start.data(), c->cell_start, sizeof(int32_t) * ((size_t)cells + 1),
                     hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return io_fail(TSPH_EHIP, "tsph_write_grid_vtk: copy failed");
  for (int q = 0; q < cells; ++q) count[q] = start[q + 1] - start[q];
  // src/SPH_grid_to_vtk.h:16-18: kh = h kap, of = 0.273 h
  return tsph_format_grid_vtk(ncx, ncy, c->k.h * c->k.kap, c->k.h * 0.273, x_0, y_0, count.data(),
                              path);
}

int tsph_format_hydrostatic_pressure(const tsph_particles *h, const char *path) {
  if (!h || !path || (h->np > 0 && (!h->type || !h->r || !h->p)))
    return io_fail(TSPH_EINVAL, "tsph_format_hydrostatic_pressure: null argument");
  if (!format_hydrostatic(path, h->np, h->type, h->r, h->p))
    return io_fail(TSPH_EINVAL, std::string("tsph_format_hydrostatic_pressure: cannot write ") + path);
  return 0;
}

int tsph_write_hydrostatic_pressure(tsph_ctx *c, const char *path) {
  if (!c || !path) return io_fail(TSPH_EINVAL, "null argument");
  if (c->cases) return io_fail(TSPH_ESTATE, "not available on a batched context");
  int n = 0;
  if (int e = tsph_get_np(c, &n)) return e;
  std::vector<int32_t> type((size_t)std::max(n, 1));
  std::vector<double> r(2 * (size_t)std::max(n, 1)), p((size_t)std::max(n, 1));
  tsph_particles h{};
  h.type = type.data();
  h.r = r.data();
  h.p = p.data();
  if (int e = tsph_download(c, &h)) return e;
  return tsph_format_hydrostatic_pressure(&h, path);
}

int tsph_write_vtk(tsph_ctx *c, const char *path) {
  if (!c || !path) return io_fail(TSPH_EINVAL, "null argument");
  if (c->cases) return io_fail(TSPH_ESTATE, "not available on a batched context");
  VtkJob j;
  j.path = path;
  j.slots = c->n_ub;
  const size_t n = (size_t)std::max(j.slots, 1);
  bool ok = hipEventCreateWithFlags(&j.done, hipEventDisableTiming) == hipSuccess &&
            hipHostMalloc((void **)&j.np, sizeof(int)) == hipSuccess &&
            hipHostMalloc((void **)&j.seq, 4 * n) == hipSuccess &&
            hipHostMalloc((void **)&j.type, n) == hipSuccess &&
            hipHostMalloc((void **)&j.r, 16 * n) == hipSuccess &&
            hipHostMalloc((void **)&j.v, 16 * n) == hipSuccess &&
            hipHostMalloc((void **)&j.p, 8 * n) == hipSuccess &&
            hipHostMalloc((void **)&j.rho, 8 * n) == hipSuccess;
  if (ok) {
    const Soa &d = c->cur;
    const size_t s = (size_t)j.slots;
    hipStream_t st = c->stream;
    ok = hipMemcpyAsync(j.np, &c->scal[S_NP], sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
         (s == 0 ||
          (hipMemcpyAsync(j.seq, d.seq, 4 * s, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(j.type, d.type, s, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(j.r, d.r, 16 * s, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(j.v, d.v, 16 * s, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(j.p, d.p, 8 * s, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(j.rho, d.rho, 8 * s, hipMemcpyDeviceToHost, st) == hipSuccess)) &&
         hipEventRecord(j.done, st) == hipSuccess;
  }
  if (!ok) {
    (void)hipStreamSynchronize(c->stream);
    free_job(j);
    return io_fail(TSPH_EHIP, "tsph_write_vtk: copy setup failed");
  }
  c->writers.emplace_back([c, j]() mutable {
    const std::string err = write_job(j);
    free_job(j);
    if (!err.empty()) {
      std::lock_guard<std::mutex> lk(c->io_mu);
      if (c->io_error.empty()) c->io_error = err;
    }
  });
  return 0;
}

int tsph_io_wait(tsph_ctx *c) {
  if (!c) return io_fail(TSPH_EINVAL, "null argument");
  for (std::thread &t : c->writers)
    if (t.joinable()) t.join();
  c->writers.clear();
  std::lock_guard<std::mutex> lk(c->io_mu);
  if (!c->io_error.empty()) {
    const std::string e = c->io_error;
    c->io_error.clear();
    return io_fail(TSPH_EINVAL, e);
  }
  return 0;
}

int tsph_interpolate_grid(tsph_ctx *c, double x0, double y0, double xm, double ym, int *ncx,
                          int *ncy, double *rho, double *p, double *v) {
  if (!c || !ncx || !ncy) return io_fail(TSPH_EINVAL, "null argument");
  if (c->cases) return io_fail(TSPH_ESTATE, "not available on a batched context");
  // src/SPH_dataInterpolation.h:31-36: sf = 4, dh = h / sf
  const double h = c->k.h, sf = 4;
  const double dh = h / sf;
  const int nx = (int)((xm - x0) / h) * sf;
  const int ny = (int)((ym - y0) / h) * sf;
  if (nx < 0 || ny < 0 || (long long)nx * ny > (1ll << 28))
    return io_fail(TSPH_EINVAL, "interpolation grid too large");
  *ncx = nx;
  *ncy = ny;
  if (!rho && !p && !v) return 0;
  if (!c->cells_valid) return io_fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  const size_t cnt = (size_t)nx * ny;
  if (cnt == 0) return 0;
  double *dout = nullptr;
  TSPH_HIP(hipMalloc((void **)&dout, 4 * sizeof(double) * cnt));
  std::vector<double> out(4 * cnt);
  int e = launch_interp_grid(c, x0, y0, dh, nx, ny, dout);
  if (!e && (hipMemcpyAsync(out.data(), dout, 4 * sizeof(double) * cnt, hipMemcpyDeviceToHost,
                            c->stream) != hipSuccess ||
             hipStreamSynchronize(c->stream) != hipSuccess))
    e = io_fail(TSPH_EHIP, "tsph_interpolate_grid: copy failed");
  (void)hipFree(dout);
  if (e) return e;
  for (size_t t = 0; t < cnt; ++t) {
    if (rho) rho[t] = out[4 * t];
    if (p) p[t] = out[4 * t + 1];
    if (v) { v[2 * t] = out[4 * t + 2]; v[2 * t + 1] = out[4 * t + 3]; }
  }
  return 0;
}

int tsph_interpolate_grid_noalloc(tsph_ctx *c, double x0, double y0, double xm, double ym, int *ncx,
                                  int *ncy, double *rho, double *p, double *v) {
  if (!c || !ncx || !ncy) return io_fail(TSPH_EINVAL, "null argument");
  if (c->cases) return io_fail(TSPH_ESTATE, "not available on a batched context");
  // the grid of tsph_interpolate_grid (src/SPH_dataInterpolation.h:31-36)
  const double h = c->k.h, sf = 4;
  const double dh = h / sf;
  const int nx = (int)((xm - x0) / h) * sf;
  const int ny = (int)((ym - y0) / h) * sf;
  if (nx < 0 || ny < 0) return io_fail(TSPH_EINVAL, "interpolation grid too large");
  if ((long long)nx * ny > TSPH_GRID_SCRATCH_POINTS)
    return io_fail(TSPH_ECAPACITY, "interpolation grid of " + std::to_string((long long)nx * ny) +
                                       " points exceeds TSPH_GRID_SCRATCH_POINTS (" +
                                       std::to_string(TSPH_GRID_SCRATCH_POINTS) + ")");
  *ncx = nx;
  *ncy = ny;
  if (!rho && !p && !v) return 0;
  if (!c->cells_valid) return io_fail(TSPH_ESTATE, "no cell list: call tsph_particle_to_cell first");
  const size_t cnt = (size_t)nx * ny;
  if (cnt == 0) return 0;
  double *dout = c->grid_scratch;  // this context's device's copy of the array
  if (!dout) return io_fail(TSPH_ESTATE, "no scratch address for this context");
  if (int e = launch_interp_grid(c, x0, y0, dh, nx, ny, dout, true)) return e;
  const size_t b = sizeof(double) * cnt;
  if ((rho && hipMemcpyAsync(rho, dout, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
      (p && hipMemcpyAsync(p, dout + cnt, b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
      (v && hipMemcpyAsync(v, dout + 2 * cnt, 2 * b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return io_fail(TSPH_EHIP, "tsph_interpolate_grid_noalloc: copy failed");
  return 0;
}

int tsph_write_interpolated_vtk(tsph_ctx *c, double x0, double y0, double xm, double ym,
                                const char *path) {
  if (!path) return io_fail(TSPH_EINVAL, "null argument");
  int nx = 0, ny = 0;
  if (int e = tsph_interpolate_grid(c, x0, y0, xm, ym, &nx, &ny, nullptr, nullptr, nullptr)) return e;
  const size_t cnt = (size_t)nx * ny;
  std::vector<double> rho(cnt + 1), p(cnt + 1), v(2 * cnt + 2);
  if (int e = tsph_interpolate_grid(c, x0, y0, xm, ym, &nx, &ny, rho.data(), p.data(), v.data()))
    return e;
  return tsph_format_interpolated_vtk(nx, ny, c->k.h / 4, x0, y0, rho.data(), p.data(), v.data(),
                                      path);
}

}  // extern "C"
