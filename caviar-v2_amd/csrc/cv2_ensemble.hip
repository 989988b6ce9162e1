// cv2_ensemble.hip -- the batched-ensemble C ABI (include/cv2_md.h): many
// independent periodic LJ systems in one set of device arrays, advanced by
// one captured HIP graph per step.
//
// Layout: the atoms of system s are [atom_off[s], atom_off[s+1]) and keep
// that range for good (the cell sort permutes inside a system only); the
// cells of system s are [cell_off[s], cell_off[s] + nc^3).  Every kernel of
// a step covers all systems; the list build works only on the systems whose
// rebuild flag the update kernel raised (or on all of them under the
// rebuild-all policy), so a step is the same fixed launch sequence whatever
// the systems do, which is what lets one graph replay it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cv2_internal.h"
#include "cv2_md.h"

using namespace cv2;

struct cv2_ens {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool lj_set = false, nl_set = false, atoms = false;
  LJDev lj{};
  LJDev *lj_dev = nullptr;
  double nl_cutoff = 0.0, nl_coef = 1.0;
  // systems (host copies)
  int nsys = 0;
  int64_t n = 0;
  std::vector<int64_t> natoms, atom_off;
  std::vector<double> lo, hi;
  int ncell = 0;
  double thresh_sq = 0.0;
  bool iso = true;  // every system has one fixed-point unit on all axes
  bool per_atom = false;  // some system takes the per-atom list build
  // device state and tables
  DevState s{};
  int *d_sys = nullptr, *d_cell_sys = nullptr, *d_cell_off = nullptr, *d_atom_off = nullptr;
  BoxDev *d_box = nullptr;
  int *d_sflag = nullptr, *d_srebuilds = nullptr;
  int *d_act = nullptr;  // active-system list of a build (EnsDev::act)
  long long *eatom = nullptr;
  double *d_thermo = nullptr;
  int *h_flags = nullptr;
  // launch shapes, graph of one step
  cv2_ens_tuning tune{256, 128, 0, 0.0, 1.0};
  double cell_scale = 1.0;  // cell width of the current layout, in list cutoffs
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  size_t gnodes = 0;
  double graph_dt = 0.0;
  cv2_ens_tuning graph_tune{};
  bool forces_valid = false;
  int64_t steps = 0;

  EnsDev ens() const {
    EnsDev e{};
    e.nsys = nsys;
    e.sys = d_sys;
    e.cell_sys = d_cell_sys;
    e.cell_off = d_cell_off;
    e.atom_off = d_atom_off;
    e.box = d_box;
    e.sflag = d_sflag;
    e.srebuilds = d_srebuilds;
    e.act = d_act;
    return e;
  }
};

namespace {

int efail(cv2_ens *e, int code, const std::string &m) {
  if (e) e->err = m;
  return code;
}

#define EHIP(expr)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return efail(ens, CV2_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <class T>
hipError_t ralloc(T **p, size_t count) {
  if (*p) {
    hipError_t e = hipFree(*p);
    *p = nullptr;
    if (e != hipSuccess) return e;
  }
  return hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
}

template <class T>
void rfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void drop_graph(cv2_ens *ens) {
  if (ens->gexec) (void)hipGraphExecDestroy(ens->gexec);
  if (ens->graph) (void)hipGraphDestroy(ens->graph);
  ens->gexec = nullptr;
  ens->graph = nullptr;
  ens->gnodes = 0;
}

void free_all(cv2_ens *ens) {
  drop_graph(ens);
  DevState &s = ens->s;
  void *ptrs[] = {s.pos, s.pos_alt, s.pos_ref, s.q, s.q_alt, s.vel, s.vel_alt, s.acc, s.acc_alt, s.tag,
                  s.tag_alt, s.cell, s.slot, s.perm, s.order, s.cell_count, s.cell_start,
                  s.tile_sum, s.nbr, s.nnbr, s.flags, s.wstat};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  s = DevState{};
  rfree(ens->d_sys);
  rfree(ens->d_cell_sys);
  rfree(ens->d_cell_off);
  rfree(ens->d_atom_off);
  rfree(ens->d_box);
  rfree(ens->d_sflag);
  rfree(ens->d_srebuilds);
  rfree(ens->d_act);
  rfree(ens->eatom);
  rfree(ens->d_thermo);
  ens->atoms = false;
  ens->forces_valid = false;
}

// Box constants of one periodic system.  With at least 5 cells of width >=
// ce per axis the system takes the wave-per-cell build (coarse filter on
// the fixed-point coordinates); otherwise cells at least half the list cutoff
// wide for the per-atom build (+-2 cells).
// cell_scale >= 1: cells at least cell_scale * ce wide where that still
// leaves two cells per axis (the wave build's minimum), else ce wide.
BoxDev system_box(const double lo[3], const double hi[3], double ce, double rc, double cell_scale) {
  BoxDev b{};
  b.ce_sq = ce * ce;
  b.thresh_sq = 0.25 * (ce - rc) * (ce - rc);
  b.ghost_cut = ce;  // every image within the list cutoff exists
  bool fast = true;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = lo[d];
    b.hi[d] = hi[d];
    b.L[d] = hi[d] - lo[d];
    b.halfL[d] = 0.5 * b.L[d];
    b.bc[d] = 1;
    b.wrap[d] = 1;
    b.glo[d] = lo[d];
    // Two cells per axis suffice: with 2, the stencil meets a cell twice,
    // under two images L apart, and as L > 2 ce at most one image of an atom
    // lies within ce of a home atom.  Boxes with fewer than 5 cells on some
    // axis form the wave build's relative coordinates from the FP64
    // positions (pos_rel).
    fast = fast && std::floor(b.L[d] / (ce * (1.0 + 1e-12))) >= 2.0;
  }
  double scale = 1.0;
  if (fast && cell_scale > 1.0) {
    bool wide_ok = true;
    for (int d = 0; d < 3; ++d)
      wide_ok = wide_ok && std::floor(b.L[d] / (cell_scale * ce * (1.0 + 1e-12))) >= 2.0;
    if (wide_ok) scale = cell_scale;
  }
  double wmax = 0.0;
  for (int d = 0; d < 3; ++d) {
    int nc = (int)std::floor(b.L[d] / ((fast ? scale : 0.5) * ce * (1.0 + 1e-12)));
    nc = std::max(1, std::min(nc, 1024));
    b.nc[d] = nc;
    b.cinv[d] = nc / b.L[d];
    wmax = std::max(wmax, b.L[d] / nc);
  }
  // list pairs stay below ce + skin < 2 ce; the coarse filter forms
  // separations up to two cell widths
  set_qscale(b, fast ? 2.0 * std::max(wmax, ce) : 2.0 * ce);
  b.coarse = fast && coarse_filter_setup(b, ce, wmax) ? 1 : 0;
  if (fast && !b.coarse) {  // the filter cannot decide: per-atom build
    for (int d = 0; d < 3; ++d) {
      int nc = (int)std::floor(b.L[d] / (0.5 * ce * (1.0 + 1e-12)));
      nc = std::max(1, std::min(nc, 1024));
      b.nc[d] = nc;
      b.cinv[d] = nc / b.L[d];
    }
    set_qscale(b, 2.0 * ce);
  }
  b.ghost_ok = 1;
  return b;
}

int check_flags(cv2_ens *ens) {
  EHIP(hipMemcpyAsync(ens->h_flags, ens->s.flags, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost,
                      ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  const int *f = ens->h_flags;
  if (f[F_OVERFLOW])
    return efail(ens, CV2_ERR_CAPACITY, "neighbour list capacity " + std::to_string(ens->s.cap) +
                                            " exceeded (longest list " + std::to_string(f[F_MAXNBR]) + ")");
  if (f[F_FXOVF]) return efail(ens, CV2_ERR_FIXED_OVERFLOW, "fixed-point force range exceeded (r < sigma/2)");
  return CV2_OK;
}

// Force a build of every system and compute forces (state after upload).
int prime(cv2_ens *ens) {
  const EnsDev E = ens->ens();
  launch_fill_int(ens->d_sflag, ens->nsys, 1, ens->stream);
  launch_set_flag(ens->s.flags, F_REBUILD, 1, ens->stream);
  launch_build_ens(ens->s, E, ens->ncell, ens->tune.build_block, false, ens->per_atom,
                   ens->stream);
  launch_forces_ens(ens->s, E, ens->lj_dev, ens->tune.force_block, ens->iso, false, ens->eatom, ens->stream);
  ens->forces_valid = true;
  return check_flags(ens);
}

// Positions and velocities (AoS, upload order) into the device state; the
// atoms keep their system's index range.
int load_state(cv2_ens *ens, const double *pos, const double *vel) {
  const int64_t n = ens->n;
  DevState &s = ens->s;
  std::vector<double4> p4(n);
  std::vector<double> v3(3 * n, 0.0);
  std::vector<int> tag(n);
  for (int64_t i = 0; i < n; ++i) {
    p4[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0);
    for (int d = 0; d < 3; ++d) v3[d * n + i] = vel ? vel[3 * i + d] : 0.0;
    tag[i] = (int)i;
  }
  EHIP(hipMemcpyAsync(s.pos, p4.data(), sizeof(double4) * n, hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemcpyAsync(s.vel, v3.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemcpyAsync(s.tag, tag.data(), sizeof(int) * n, hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemsetAsync(s.acc, 0, sizeof(double) * 3 * n, ens->stream));
  EHIP(hipMemsetAsync(s.flags, 0, sizeof(int) * F_COUNT, ens->stream));
  EHIP(hipMemsetAsync(ens->d_srebuilds, 0, sizeof(int) * ens->nsys, ens->stream));
  EHIP(hipMemsetAsync(s.cell_count, 0, sizeof(int) * (ens->ncell + 1), ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  ens->steps = 0;
  return prime(ens);
}

// One step that closes the previous step and opens the next: the list build
// (whose first kernel takes over the coordinates the previous force pass
// wrote to q_alt) and the force pass with the integrator in its epilogue.
void middle_step(cv2_ens *ens, const EnsDev &E, double dt) {
  const cv2_ens_tuning &t = ens->tune;
  launch_build_ens(ens->s, E, ens->ncell, t.build_block, t.rebuild_all != 0, ens->per_atom,
                   ens->stream, ens->s.q_alt);
  launch_forces_ens(ens->s, E, ens->lj_dev, t.force_block, ens->iso, false, ens->eatom, ens->stream,
                    dt, ens->thresh_sq);
}

int ensure_graph(cv2_ens *ens, double dt) {
  const cv2_ens_tuning &t = ens->tune;
  if (ens->gexec && ens->graph_dt == dt && ens->graph_tune.force_block == t.force_block &&
      ens->graph_tune.build_block == t.build_block && ens->graph_tune.rebuild_all == t.rebuild_all)
    return CV2_OK;
  drop_graph(ens);
  const EnsDev E = ens->ens();
  EHIP(hipStreamBeginCapture(ens->stream, hipStreamCaptureModeThreadLocal));
  middle_step(ens, E, dt);
  EHIP(hipStreamEndCapture(ens->stream, &ens->graph));
  EHIP(hipGraphInstantiate(&ens->gexec, ens->graph, nullptr, nullptr, 0));
  EHIP(hipGraphGetNodes(ens->graph, nullptr, &ens->gnodes));
  ens->graph_dt = dt;
  ens->graph_tune = t;
  return CV2_OK;
}

int run(cv2_ens *ens, int64_t nsteps, double dt) {
  if (!ens->atoms || !ens->lj_set || !ens->nl_set)
    return efail(ens, CV2_ERR_STATE, "LJ parameters, neighbour settings and systems must be set first");
  if (nsteps <= 0) return CV2_OK;
  if (!ens->forces_valid) {
    int rc = prime(ens);
    if (rc) return rc;
  }
  const EnsDev E = ens->ens();
  const cv2_ens_tuning &t = ens->tune;
  // Step 1 opens eagerly (its coordinates also into q_alt, which the first
  // build takes over); steps 1..n-1 are the captured middle step (build,
  // forces with the next step's integrator pass); step n builds and computes
  // forces alone, and the final half kick closes it.
  launch_update_ens(ens->s, E, false, true, dt, ens->thresh_sq, ens->stream);
  EHIP(hipMemcpyAsync(ens->s.q_alt, ens->s.q, sizeof(int4) * ens->s.n, hipMemcpyDeviceToDevice,
                      ens->stream));
  if (nsteps > 1) {
    if (ens->use_graph) {
      int rc = ensure_graph(ens, dt);
      if (rc) return rc;
      for (int64_t k = 1; k < nsteps; ++k) EHIP(hipGraphLaunch(ens->gexec, ens->stream));
    } else {
      for (int64_t k = 1; k < nsteps; ++k) middle_step(ens, E, dt);
    }
  }
  launch_build_ens(ens->s, E, ens->ncell, t.build_block, t.rebuild_all != 0, ens->per_atom,
                   ens->stream, ens->s.q_alt);
  launch_forces_ens(ens->s, E, ens->lj_dev, t.force_block, ens->iso, false, ens->eatom, ens->stream);
  launch_update_ens(ens->s, E, true, false, dt, ens->thresh_sq, ens->stream);
  ens->steps += nsteps;
  return check_flags(ens);
}

bool valid_block(int b) { return b == 64 || b == 128 || b == 256; }

}  // namespace

extern "C" {

int cv2_ens_create(cv2_ens **out, int device) {
  if (!out) return CV2_ERR_BAD_ARG;
  *out = nullptr;
  cv2_ens *ens = new cv2_ens();
  ens->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ens;
    return CV2_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&ens->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&ens->lj_dev, sizeof(LJDev)) != hipSuccess ||
      hipHostMalloc((void **)&ens->h_flags, sizeof(int) * F_COUNT, hipHostMallocDefault) != hipSuccess) {
    cv2_ens_destroy(ens);
    return CV2_ERR_HIP;
  }
  *out = ens;
  return CV2_OK;
}

int cv2_ens_destroy(cv2_ens *ens) {
  if (!ens) return CV2_OK;
  (void)hipSetDevice(ens->device);
  if (ens->stream) (void)hipStreamSynchronize(ens->stream);
  free_all(ens);
  if (ens->lj_dev) (void)hipFree(ens->lj_dev);
  if (ens->h_flags) (void)hipHostFree(ens->h_flags);
  if (ens->stream) (void)hipStreamDestroy(ens->stream);
  delete ens;
  return CV2_OK;
}

const char *cv2_ens_last_error(const cv2_ens *ens) { return ens ? ens->err.c_str() : "null ensemble"; }

int cv2_ens_set_lj(cv2_ens *ens, double eps, double sigma, double cutoff, double mass) {
  if (!ens) return CV2_ERR_BAD_ARG;
  if (cutoff < 0.0) return efail(ens, CV2_ERR_BAD_ARG, "Force field cutoff must be non-negative.");
  std::string err;
  if (!build_lj_table(ens->lj, 1, &eps, &sigma, nullptr, cutoff, &mass, err))
    return efail(ens, CV2_ERR_BAD_ARG, err);
  (void)hipSetDevice(ens->device);
  EHIP(hipMemcpyAsync(ens->lj_dev, &ens->lj, sizeof(LJDev), hipMemcpyHostToDevice, ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  ens->lj_set = true;
  ens->forces_valid = false;
  return CV2_OK;
}

int cv2_ens_set_neighbor(cv2_ens *ens, double cutoff, double extra_coef) {
  if (!ens) return CV2_ERR_BAD_ARG;
  if (cutoff <= 0.0) return efail(ens, CV2_ERR_BAD_ARG, "cutoff have to non-negative.");
  if (extra_coef < 1.0) return efail(ens, CV2_ERR_BAD_ARG, "cutoff_extra_coef have to be larger than 1.0");
  if (ens->atoms) return efail(ens, CV2_ERR_STATE, "set the neighbour list before uploading systems");
  ens->nl_cutoff = cutoff;
  ens->nl_coef = extra_coef;
  ens->nl_set = true;
  return CV2_OK;
}

namespace {
// Cell grids of all systems for a cell width of cell_scale list cutoffs
// (system_box): boxes, cell offsets and system ids, the cell arrays and the
// per-block statistics.  The atoms keep their arrays; the caller reloads the
// state (which forces a build).
int layout_cells(cv2_ens *ens, double cell_scale) {
  const double ce = ens->nl_cutoff * ens->nl_coef;
  const int nsys = ens->nsys;
  std::vector<BoxDev> boxes(nsys);
  std::vector<int> coff(nsys, 0);
  int64_t ncell = 0;
  bool iso = true, per_atom = false;
  for (int s = 0; s < nsys; ++s) {
    boxes[s] = system_box(&ens->lo[3 * s], &ens->hi[3 * s], ce, ens->nl_cutoff, cell_scale);
    iso = iso && boxes[s].qiso;
    per_atom = per_atom || !boxes[s].coarse;
    coff[s] = (int)ncell;
    ncell += (int64_t)boxes[s].nc[0] * boxes[s].nc[1] * boxes[s].nc[2];
  }
  if (ncell > (1 << 28)) return efail(ens, CV2_ERR_BAD_ARG, "cell grids too large");
  drop_graph(ens);
  std::vector<int> csys(ncell);
  for (int s = 0; s < nsys; ++s) {
    const int nc = boxes[s].nc[0] * boxes[s].nc[1] * boxes[s].nc[2];
    for (int c = 0; c < nc; ++c) csys[coff[s] + c] = s;
  }
  DevState &st = ens->s;
  for (void **p : {(void **)&st.cell_count, (void **)&st.cell_start, (void **)&st.tile_sum,
                   (void **)&st.wstat, (void **)&ens->d_cell_sys, (void **)&ens->d_cell_off,
                   (void **)&ens->d_box}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  const int64_t n = ens->n;
  ens->iso = iso;
  ens->per_atom = per_atom;
  ens->ncell = (int)ncell;
  ens->cell_scale = cell_scale;
  ens->thresh_sq = boxes[0].thresh_sq;
  st.ncell = (int)ncell;
  EHIP(ralloc(&st.cell_count, ncell + 1));
  EHIP(ralloc(&st.cell_start, ncell + 1));
  EHIP(ralloc(&st.tile_sum, ncell / 1024 + 2));
  // per-block statistics: per-atom build blocks (>= 64 atoms), then
  // wave-per-cell build blocks (4 cells)
  EHIP(ralloc(&st.wstat, (n + 63) / 64 + (ncell + 3) / 4 + 16));
  EHIP(ralloc(&ens->d_cell_sys, ncell));
  EHIP(ralloc(&ens->d_cell_off, nsys));
  EHIP(ralloc(&ens->d_box, nsys));
  EHIP(hipMemcpyAsync(ens->d_cell_sys, csys.data(), sizeof(int) * ncell, hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemcpyAsync(ens->d_cell_off, coff.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemcpyAsync(ens->d_box, boxes.data(), sizeof(BoxDev) * nsys, hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemsetAsync(st.cell_count, 0, sizeof(int) * (ncell + 1), ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  return CV2_OK;
}
}  // namespace

int cv2_ens_upload(cv2_ens *ens, int32_t nsys, const int64_t *natoms, const double *lo,
                   const double *hi, const double *pos, const double *vel) {
  if (!ens || nsys < 1 || !natoms || !lo || !hi || !pos) return efail(ens, CV2_ERR_BAD_ARG, "bad systems");
  if (!ens->nl_set || !ens->lj_set) return efail(ens, CV2_ERR_STATE, "set LJ and neighbour parameters first");
  (void)hipSetDevice(ens->device);
  const double ce = ens->nl_cutoff * ens->nl_coef;
  std::vector<int64_t> aoff(nsys + 1, 0);
  double max_expect = 0.0;
  for (int s = 0; s < nsys; ++s) {
    if (natoms[s] < 1) return efail(ens, CV2_ERR_BAD_ARG, "every system needs atoms");
    double vol = 1.0;
    for (int d = 0; d < 3; ++d) {
      if (!(hi[3 * s + d] - lo[3 * s + d] > 2.0 * ce))
        return efail(ens, CV2_ERR_BAD_ARG, "periodic box must be wider than twice the list cutoff");
      vol *= hi[3 * s + d] - lo[3 * s + d];
    }
    aoff[s + 1] = aoff[s] + natoms[s];
    max_expect = std::max(max_expect, 4.18879020478639 * ce * ce * ce * (double)natoms[s] / vol);
  }
  const int64_t n = aoff[nsys];
  if (n > (int64_t)kIndexMask) return efail(ens, CV2_ERR_BAD_ARG, "too many atoms");
  free_all(ens);
  ens->nsys = nsys;
  ens->n = n;
  ens->natoms.assign(natoms, natoms + nsys);
  ens->atom_off = aoff;
  ens->lo.assign(lo, lo + 3 * nsys);
  ens->hi.assign(hi, hi + 3 * nsys);
  // per-atom system ids
  std::vector<int> sys(n), aoff32(nsys + 1);
  for (int s = 0; s < nsys; ++s)
    for (int64_t i = aoff[s]; i < aoff[s + 1]; ++i) sys[i] = s;
  for (int s = 0; s <= nsys; ++s) aoff32[s] = (int)aoff[s];
  DevState &st = ens->s;
  st.n = (int)n;
  st.ng = 0;
  int cap = (int)std::ceil(1.6 * max_expect + 64.0);
  st.cap = (cap + 15) / 16 * 16;
  st.stride = (int)((n + 63) / 64 * 64);
  st.write_q = true;
  EHIP(ralloc(&st.pos, n));
  EHIP(ralloc(&st.pos_alt, n));
  EHIP(ralloc(&st.pos_ref, n));
  EHIP(ralloc(&st.q, n));
  EHIP(ralloc(&st.q_alt, n));
  EHIP(ralloc(&st.vel, 3 * n));
  EHIP(ralloc(&st.vel_alt, 3 * n));
  EHIP(ralloc(&st.acc, 3 * n));
  EHIP(ralloc(&st.acc_alt, 3 * n));
  EHIP(ralloc(&st.tag, n));
  EHIP(ralloc(&st.tag_alt, n));
  EHIP(ralloc(&st.cell, n));
  EHIP(ralloc(&st.slot, n));
  EHIP(ralloc(&st.perm, n));
  EHIP(ralloc(&st.order, n));
  EHIP(ralloc(&st.nnbr, n));
  EHIP(ralloc(&st.nbr, (size_t)st.cap * st.stride));
  EHIP(ralloc(&st.flags, F_COUNT));
  EHIP(ralloc(&ens->d_sys, n));
  EHIP(ralloc(&ens->d_atom_off, nsys + 1));
  EHIP(ralloc(&ens->d_sflag, nsys));
  EHIP(ralloc(&ens->d_srebuilds, nsys));
  EHIP(ralloc(&ens->d_act, 4 * nsys + 8));  // systems and cells of a build (EnsDev::act)
  EHIP(ralloc(&ens->eatom, 2 * n));
  EHIP(ralloc(&ens->d_thermo, 3 * nsys));
  EHIP(hipMemcpyAsync(ens->d_sys, sys.data(), sizeof(int) * n, hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemcpyAsync(ens->d_atom_off, aoff32.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, ens->stream));
  EHIP(hipMemsetAsync(ens->eatom, 0, sizeof(long long) * 2 * n, ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  int rc = layout_cells(ens, ens->tune.cell_scale);
  if (rc) return rc;
  ens->atoms = true;
  return load_state(ens, pos, vel);
}

namespace {
// A new cell width: the systems' cell grids are laid out again and the
// current state reloaded into them (a build follows).
int change_cells(cv2_ens *ens, double cell_scale) {
  if (cell_scale == ens->cell_scale || !ens->atoms) {
    ens->cell_scale = ens->atoms ? ens->cell_scale : cell_scale;
    return CV2_OK;
  }
  const int64_t n = ens->n;
  std::vector<double> pos(3 * n), vel(3 * n);
  int rc = cv2_ens_download(ens, pos.data(), vel.data(), nullptr);
  if (!rc) rc = layout_cells(ens, cell_scale);
  if (!rc) rc = load_state(ens, pos.data(), vel.data());
  return rc;
}
}  // namespace

int cv2_ens_set_tuning(cv2_ens *ens, const cv2_ens_tuning *t) {
  if (!ens || !t) return CV2_ERR_BAD_ARG;
  if (!valid_block(t->force_block) || !valid_block(t->build_block) || (t->rebuild_all != 0 && t->rebuild_all != 1))
    return efail(ens, CV2_ERR_BAD_ARG, "block sizes must be 64, 128 or 256 and rebuild_all 0 or 1");
  if (!(t->cell_scale >= 1.0 && t->cell_scale <= 2.0))
    return efail(ens, CV2_ERR_BAD_ARG, "cell_scale must lie in [1, 2]");
  (void)hipSetDevice(ens->device);
  ens->tune.force_block = t->force_block;
  ens->tune.build_block = t->build_block;
  ens->tune.rebuild_all = t->rebuild_all;
  ens->tune.cell_scale = t->cell_scale;
  return change_cells(ens, t->cell_scale);
}

int cv2_ens_tune(cv2_ens *ens, double dt, int32_t trial_steps, cv2_ens_tuning *best) {
  if (!ens || trial_steps < 1) return CV2_ERR_BAD_ARG;
  if (!ens->atoms) return efail(ens, CV2_ERR_STATE, "upload systems first");
  (void)hipSetDevice(ens->device);
  const int64_t n = ens->n;
  std::vector<double> pos(3 * n), vel(3 * n);
  int rc = cv2_ens_download(ens, pos.data(), vel.data(), nullptr);
  if (rc) return rc;
  const int blocks[3] = {64, 128, 256};
  const double scale0 = ens->cell_scale;
  cv2_ens_tuning win{ens->tune.force_block, ens->tune.build_block, ens->tune.rebuild_all, 1e30, scale0};
  auto trial = [&](const cv2_ens_tuning &t) -> int {
    ens->tune = t;
    int r = load_state(ens, pos.data(), vel.data());
    if (!r) r = run(ens, 5, dt);  // warm-up (and graph capture)
    if (r) return r;
    const auto t0 = std::chrono::steady_clock::now();
    r = run(ens, trial_steps, dt);
    const auto t1 = std::chrono::steady_clock::now();
    if (r) return r;
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / trial_steps;
    if (ms < win.ms_per_step) {
      win = t;
      win.ms_per_step = ms;
    }
    return CV2_OK;
  };
  // launch shapes and rebuild policy on the current cells, then the cell
  // width with the best of them
  for (int fb : blocks)
    for (int bb : blocks)
      for (int pol = 0; pol < 2; ++pol)
        if ((rc = trial(cv2_ens_tuning{fb, bb, pol, 0.0, scale0}))) return rc;
  for (double sc : {1.0, 1.15, 1.3}) {
    if (sc == scale0) continue;
    if ((rc = layout_cells(ens, sc))) return rc;
    if ((rc = trial(cv2_ens_tuning{win.force_block, win.build_block, win.rebuild_all, 0.0, sc}))) return rc;
  }
  ens->tune = win;
  if (ens->cell_scale != win.cell_scale && (rc = layout_cells(ens, win.cell_scale))) return rc;
  rc = load_state(ens, pos.data(), vel.data());
  if (rc) return rc;
  if (best) *best = win;
  return CV2_OK;
}

int cv2_ens_run_vv(cv2_ens *ens, int64_t nsteps, double dt) {
  if (!ens) return CV2_ERR_BAD_ARG;
  (void)hipSetDevice(ens->device);
  return run(ens, nsteps, dt);
}

int cv2_ens_build_neighbors(cv2_ens *ens) {
  if (!ens) return CV2_ERR_BAD_ARG;
  if (!ens->atoms) return efail(ens, CV2_ERR_STATE, "upload systems first");
  (void)hipSetDevice(ens->device);
  return prime(ens);
}

int cv2_ens_thermo(cv2_ens *ens, double *ke, double *pe, double *virial) {
  if (!ens) return CV2_ERR_BAD_ARG;
  if (!ens->atoms) return efail(ens, CV2_ERR_STATE, "upload systems first");
  (void)hipSetDevice(ens->device);
  const EnsDev E = ens->ens();
  if (!ens->forces_valid) {
    int rc = prime(ens);
    if (rc) return rc;
  }
  // an energy pass at the current positions (accelerations are recomputed
  // identically), then fixed-order per-system sums
  launch_forces_ens(ens->s, E, ens->lj_dev, ens->tune.force_block, ens->iso, true, ens->eatom, ens->stream);
  launch_thermo_ens(ens->s, E, ens->lj_dev, ens->eatom, ens->d_thermo, ens->stream);
  std::vector<double> out(3 * ens->nsys);
  EHIP(hipMemcpyAsync(out.data(), ens->d_thermo, sizeof(double) * out.size(), hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  for (int s = 0; s < ens->nsys; ++s) {
    if (ke) ke[s] = out[3 * s];
    if (pe) pe[s] = out[3 * s + 1];
    if (virial) virial[s] = out[3 * s + 2];
  }
  return check_flags(ens);
}

int cv2_ens_download(cv2_ens *ens, double *pos, double *vel, double *acc) {
  if (!ens) return CV2_ERR_BAD_ARG;
  if (!ens->atoms) return efail(ens, CV2_ERR_STATE, "upload systems first");
  (void)hipSetDevice(ens->device);
  const int64_t n = ens->n;
  const DevState &s = ens->s;
  std::vector<double4> p4(n);
  std::vector<double> v3(3 * n), a3(3 * n);
  std::vector<int> tag(n);
  EHIP(hipMemcpyAsync(p4.data(), s.pos, sizeof(double4) * n, hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipMemcpyAsync(v3.data(), s.vel, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipMemcpyAsync(a3.data(), s.acc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipMemcpyAsync(tag.data(), s.tag, sizeof(int) * n, hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  for (int64_t i = 0; i < n; ++i) {
    const int64_t t = tag[i];
    if (pos) {
      pos[3 * t] = p4[i].x;
      pos[3 * t + 1] = p4[i].y;
      pos[3 * t + 2] = p4[i].z;
    }
    for (int d = 0; d < 3; ++d) {
      if (vel) vel[3 * t + d] = v3[d * n + i];
      if (acc) acc[3 * t + d] = a3[d * n + i];
    }
  }
  return CV2_OK;
}

int cv2_ens_get_pairs(cv2_ens *ens, int32_t sys, int64_t cap, int64_t *npairs, int64_t *pairs) {
  if (!ens || !npairs || (cap > 0 && !pairs)) return efail(ens, CV2_ERR_BAD_ARG, "bad pair buffer");
  if (!ens->atoms || sys < 0 || sys >= ens->nsys) return efail(ens, CV2_ERR_BAD_ARG, "no such system");
  (void)hipSetDevice(ens->device);
  const DevState &s = ens->s;
  const int64_t a0 = ens->atom_off[sys], a1 = ens->atom_off[sys + 1];
  std::vector<int> cnt(a1 - a0), tg(a1 - a0);
  std::vector<uint32_t> nb((size_t)s.cap * s.stride);
  EHIP(hipMemcpyAsync(cnt.data(), s.nnbr + a0, sizeof(int) * cnt.size(), hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipMemcpyAsync(tg.data(), s.tag + a0, sizeof(int) * tg.size(), hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipMemcpyAsync(nb.data(), s.nbr, sizeof(uint32_t) * nb.size(), hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  int64_t k = 0;
  for (int64_t i = a0; i < a1; ++i) {
    const int64_t ti = tg[i - a0] - a0;
    for (int m = 0; m < cnt[i - a0]; ++m) {
      const uint32_t e = nb[((size_t)(m >> 2) * s.stride + i) * 4 + (m & 3)];
      const int64_t j = e & kIndexMask;
      const uint32_t img = e >> 26;
      const int code = (int)(img & 3u) + 3 * (int)((img >> 2) & 3u) + 9 * (int)(img >> 4);
      if (j < a0 || j >= a1) return efail(ens, CV2_ERR_STATE, "list entry outside its system");
      // tags of the system's atoms: look up j's
      const int64_t tj = tg[j - a0] - a0;
      if (img == kNoImage && tj < ti) continue;  // half list: owned j > i
      if (k < cap) {
        pairs[3 * k] = ti;
        pairs[3 * k + 1] = tj;
        pairs[3 * k + 2] = code;
      }
      ++k;
    }
  }
  *npairs = k;
  return CV2_OK;
}

int cv2_ens_get_stats(cv2_ens *ens, cv2_ens_stats *out) {
  if (!ens || !out) return CV2_ERR_BAD_ARG;
  *out = cv2_ens_stats{};
  out->nsys = ens->nsys;
  out->natoms = ens->n;
  out->steps = ens->steps;
  out->capacity = ens->s.cap;
  out->graph_nodes = (int32_t)ens->gnodes;
  if (!ens->atoms) return CV2_OK;
  (void)hipSetDevice(ens->device);
  std::vector<int> cnt(ens->n), reb(ens->nsys);
  EHIP(hipMemcpyAsync(cnt.data(), ens->s.nnbr, sizeof(int) * ens->n, hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipMemcpyAsync(reb.data(), ens->d_srebuilds, sizeof(int) * ens->nsys, hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipMemcpyAsync(ens->h_flags, ens->s.flags, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost, ens->stream));
  EHIP(hipStreamSynchronize(ens->stream));
  for (int c : cnt) out->list_entries += c;
  for (int r : reb) out->system_builds += r;
  out->build_steps = ens->h_flags[F_REBUILDS];
  out->max_neighbors = ens->h_flags[F_MAXNBR];
  return CV2_OK;
}

int cv2_ens_use_graph(cv2_ens *ens, int enable) {
  if (!ens) return CV2_ERR_BAD_ARG;
  ens->use_graph = enable != 0;
  return CV2_OK;
}

}  // extern "C"
