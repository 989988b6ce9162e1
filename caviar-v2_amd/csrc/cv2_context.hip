// cv2_context.hip -- the C ABI (include/cv2_md.h): context, buffers, step
// sequencing.  All device work of a context is queued on its own stream;
// the host waits only where a result must come back (thermo, download,
// error flags at the end of a run).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cv2_internal.h"
#include "cv2_md.h"

using namespace cv2;

// ---- device scratch of cv2_halo_pack_noalloc ------------------------------
// One static array per GPU (each device holds its own copy until the process
// ends): per-atom direction codes and sources, per-tile counts and offsets,
// per-direction totals and bases, the 28 results copied back, and the packet
// staging area.  Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, taken once per device at
// cv2_create with that device current.
namespace {
constexpr size_t kNaMaxAtoms = CV2_NOALLOC_MAX_ATOMS;
constexpr size_t kNaMaxTiles = kNaMaxAtoms / 1024;  // halo tiles of 1024 upload indices
constexpr size_t kNaCode = 0;
constexpr size_t kNaSrc = kNaCode + 4 * kNaMaxAtoms;
constexpr size_t kNaTcnt = kNaSrc + 4 * kNaMaxAtoms;
constexpr size_t kNaToff = kNaTcnt + 4 * 27 * kNaMaxTiles;
constexpr size_t kNaDirN = kNaToff + 4 * 27 * kNaMaxTiles;
constexpr size_t kNaBase = kNaDirN + 256;
constexpr size_t kNaRes = kNaBase + 256;
constexpr size_t kNaBuf = kNaRes + 256;
constexpr size_t kNaBytes = kNaBuf + 8 * (size_t)CV2_NOALLOC_MAX_DOUBLES;
static_assert(kNaBuf % 256 == 0, "scratch sections are 256-byte aligned");

__device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
unsigned char *g_na_ptr[kNaMaxDevices] = {};
hipError_t g_na_err[kNaMaxDevices] = {};

// Device address of `scratch` on `device` (current), once per device.
hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNaMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (g_na_ptr[device]) return hipSuccess;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_na_err[device] = e;
  if (e == hipSuccess) g_na_ptr[device] = static_cast<unsigned char *>(ptr);
  return e;
}

unsigned char *noalloc_scratch(int device) {
  if (device < 0 || device >= kNaMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_na_mutex);
  return g_na_ptr[device];
}
}  // namespace

struct cv2_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // configuration
  bool box_set = false, lj_set = false, nl_set = false;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int bc[3] = {1, 1, 1};
  double ghost_cut = 0.0;
  double nl_cutoff = 0.0, nl_coef = 1.0;
  unsigned nl_flags = 0;
  int precision = CV2_PREC_FP64;
  LJDev lj{};
  LJDev *lj_dev = nullptr;
  // state
  DevState s{};
  BoxDev b{};
  int64_t n = 0, ng = 0;
  int64_t pos_cap = 0;  // capacity of the owned+ghost position arrays
  bool host_halo = false;
  bool forces_valid = false;
  bool energy_ok = false;  // pe and virial are those of the current positions
  bool need_build = true;  // configuration changed since the last forced build
  bool atoms = false;
  int64_t steps = 0;
  // MPI halo packing scratch (cv2_halo_pack / cv2_halo_unpack)
  int *halo_code = nullptr, *halo_src = nullptr, *halo_tiles = nullptr, *halo_toff = nullptr;
  int *halo_dn = nullptr;
  long long *halo_base = nullptr;
  double *halo_buf = nullptr;
  int64_t halo_n = 0, halo_buf_cap = 0;
  // owned-atom exchange (cv2_owned_pack / cv2_owned_unpack)
  int64_t own_cap = 0;       // owned atoms the per-atom arrays hold
  bool ids_sorted = false;   // owned order is ascending id (merge arrivals by id)
  int *type_by_id = nullptr; // cv2_set_type_table
  int64_t ntype_ids = 0;
  int *mig_moved = nullptr;  // device flag: some atom was shifted back into the domain
  unsigned long long *vstate = nullptr;  // CV2_NL_VERIFY: state of the build check (k_verify.hip)
  void *mig_aux = nullptr;   // received ids, sources and packet tables
  size_t mig_aux_bytes = 0;
  double virial = 0.0, pe = 0.0;
  int *h_flags = nullptr;
  double *h_red = nullptr;
  long long *h_fx = nullptr;
  long long *h_halo = nullptr;  // halo packing: the 28 results of the device scan
  long long *halo_res = nullptr;  // cv2_halo_pack: the device scan's 28 results
  int partial_blocks = 0;
  bool fuse = true;  // integrator fused into the FP32 task kernel where it runs
  bool step_open = false;  // cv2_step_begin ran, cv2_step_end has not
  UpdateArgs open_args{};
  int min_mpr = 0;   // task slots per cell row asked for after a row ran out of them
  int task_grid_cap = 0;  // CV2_TASK_GRID: cap of the force launch's grid (tests: tail launch)
  // HIP-event timing of kernel groups (cv2_set_timing): every group, or
  // (period > 1) the force launch of every period-th force call only
  bool timing = false;
  int timing_period = 1;
  int64_t force_seq = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Span {
    hipEvent_t a, b;
    int kind;  // 0 force, 1 build, 2 integrate
  };
  std::vector<Span> spans;
  double t_ms[3] = {0, 0, 0};
  int64_t force_calls = 0;
};

namespace {

hipEvent_t next_event(cv2_ctx *c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

// Open a timed span of the given kind (0 force, 1 build, 2 integrate) on the
// context stream (no-op when timing is off or the span is not sampled).
int span_begin(cv2_ctx *c, int kind) {
  if (!c->timing) return -1;
  if (c->timing_period > 1 && (kind != 0 || c->force_seq++ % c->timing_period != 0)) return -1;
  hipEvent_t a = next_event(c), b = next_event(c);
  if (!a || !b) return -1;
  (void)hipEventRecord(a, c->stream);
  c->spans.push_back({a, b, 0});
  return (int)c->spans.size() - 1;
}

void span_end(cv2_ctx *c, int k, int kind) {
  if (k < 0) return;
  c->spans[k].kind = kind;
  (void)hipEventRecord(c->spans[k].b, c->stream);
  if (kind == 0) c->force_calls++;
}

// Fold finished spans into the totals (stream must be synchronized).
void spans_collect(cv2_ctx *c) {
  for (const auto &s : c->spans) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) c->t_ms[s.kind] += ms;
  }
  c->spans.clear();
  c->ev_used = 0;
}

}  // namespace

namespace {

int fail(cv2_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(ctx, CV2_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <class T>
hipError_t dalloc(T **p, size_t count) {
  if (*p) {
    hipError_t e = hipFree(*p);
    *p = nullptr;
    if (e != hipSuccess) return e;
  }
  if (count == 0) count = 1;
  return hipMalloc((void **)p, sizeof(T) * count);
}

void free_state(DevState &s) {
  void *ptrs[] = {s.pos, s.pos_alt, s.pos_ref, s.q, s.vel, s.vel_alt, s.acc, s.acc_alt, s.tag,
                  s.tag_alt, s.id, s.id_alt, s.cell, s.slot, s.perm, s.order, s.box_dev, s.wstat, s.cell_count, s.cell_start, s.cell_local,
                  s.gcell_count, s.gcell_start, s.gcell_items, s.tile_sum, s.nbr, s.nnbr, s.flags,
                  s.partials, s.fx_acc, s.reduce_out, s.tpiece, s.tinfo, s.ctask, s.rowtasks,
                  s.ctinfo, s.ctpiece, s.q_alt, s.gtag, s.gslot, s.gtag_alt, s.gid_alt,
                  s.gpos_alt, s.mol};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  s = DevState{};
}

// Per-block statistics of the list build: one entry per block of 4 cells
// (wave-per-cell build) or of 256 owned atoms (per-atom build).
int ensure_wstat(cv2_ctx *ctx) {
  DevState &s = ctx->s;
  const int64_t need = std::max<int64_t>(((int64_t)s.ncell + 3) / 4, (ctx->n + 255) / 256) + 8;
  if (need > s.wstat_cap) {
    HIPCHK(dalloc(&s.wstat, need));
    s.wstat_cap = need;
  }
  return CV2_OK;
}

// Cell grid and box constants; requires box, neighbour settings.
int setup_box(cv2_ctx *ctx) {
  BoxDev &b = ctx->b;
  b = BoxDev{};
  const double ce = ctx->nl_cutoff * ctx->nl_coef;  // neighborlist.cpp:317
  b.ce_sq = ce * ce;
  b.thresh_sq = 0.25 * (ce - ctx->nl_cutoff) * (ce - ctx->nl_cutoff);  // :318
  b.ghost_cut = ctx->ghost_cut;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = ctx->lo[d];
    b.hi[d] = ctx->hi[d];
    b.L[d] = ctx->hi[d] - ctx->lo[d];
    b.halfL[d] = 0.5 * b.L[d];
    b.bc[d] = ctx->host_halo ? 0 : ctx->bc[d];
    b.wrap[d] = ctx->bc[d];
    if (ctx->bc[d] && !ctx->host_halo && b.L[d] <= 2.0 * ce)
      return fail(ctx, CV2_ERR_BAD_ARG, "periodic box must be wider than twice the list cutoff");
  }
  // Fully periodic boxes with at least 5 cells of width >= ce per axis use
  // the wave-per-cell build with the coarse filter (k_nlist_cells), and so
  // do host-halo boxes whose grid over the box and its ghost margin has at
  // least 3 such cells per axis (open grid: the ghosts are sorted into it);
  // every other box uses the per-atom build over cells at least ce/2 wide.
  bool fast = true;
  for (int d = 0; d < 3; ++d) {
    if (ctx->host_halo)
      fast = fast && std::floor((b.L[d] + 2.0 * ctx->ghost_cut) / (ce * (1.0 + 1e-12))) >= 3.0;
    else
      fast = fast && b.bc[d] == 1 && std::floor(b.L[d] / (ce * (1.0 + 1e-12))) >= 5.0;
  }
  int64_t ncell = 1;
  double wmax = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double margin = b.bc[d] ? 0.0 : ctx->ghost_cut;
    b.glo[d] = b.lo[d] - margin;
    const double ext = b.L[d] + 2.0 * margin;
    const double wmin = fast ? ce : 0.5 * ce;
    int nc = (int)std::floor(ext / (wmin * (1.0 + 1e-12)));
    nc = std::max(1, std::min(nc, 1024));
    b.nc[d] = nc;
    b.cinv[d] = nc / ext;
    wmax = std::max(wmax, ext / nc);
    ncell *= nc;
  }
  // Fixed-point units: the force kernel forms separations of list pairs
  // (below ce + skin < 2 ce); the coarse filter forms them between a home
  // cell's atoms and those of the adjacent cells (below 2 w).
  set_qscale(b, fast ? 2.0 * std::max(wmax, ce) : 2.0 * ce);
  {
    const bool ok = !fast || coarse_filter_setup(b, ce, wmax);
    if (!ok) return fail(ctx, CV2_ERR_BAD_ARG, "box too large for the coarse list filter");
    b.coarse = fast ? 1 : 0;
    b.open = fast && ctx->host_halo ? 1 : 0;
    // host-halo ghosts are explicit: no image needs the existence rule
    b.ghost_ok = ctx->ghost_cut >= ce || ctx->host_halo ? 1 : 0;
  }
  if (ncell > (1 << 26)) return fail(ctx, CV2_ERR_BAD_ARG, "cell grid too large");
  {
    int rc = ensure_wstat(ctx);
    if (rc) return rc;
  }
  if (!ctx->s.box_dev) HIPCHK(dalloc(&ctx->s.box_dev, 1));
  HIPCHK(hipMemcpyAsync(ctx->s.box_dev, &b, sizeof(BoxDev), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (ncell != ctx->s.ncell) {
    ctx->s.ncell = (int)ncell;
    for (int **p : {&ctx->s.cell_count, &ctx->s.cell_start, &ctx->s.cell_local, &ctx->s.gcell_count,
                    &ctx->s.gcell_start, &ctx->s.tile_sum}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    HIPCHK(dalloc(&ctx->s.cell_count, ncell + 1));
    // builds leave the counts cleared for the next one (k_rank_move)
    HIPCHK(hipMemsetAsync(ctx->s.cell_count, 0, sizeof(int) * (ncell + 1), ctx->stream));
    HIPCHK(dalloc(&ctx->s.cell_start, ncell + 1));
    HIPCHK(dalloc(&ctx->s.cell_local, ncell + 1));
    HIPCHK(dalloc(&ctx->s.gcell_count, ncell + 1));
    HIPCHK(dalloc(&ctx->s.gcell_start, ncell + 1));
    HIPCHK(dalloc(&ctx->s.tile_sum, ncell / 1024 + 2));
    HIPCHK(hipMemsetAsync(ctx->s.gcell_start, 0, sizeof(int) * (ncell + 1), ctx->stream));
  }
  return CV2_OK;
}

bool ref_rebuilds(const cv2_ctx *ctx) { return (ctx->nl_flags & CV2_NL_REFERENCE_REBUILDS) != 0; }

int ensure_ready(cv2_ctx *ctx) {
  if (!ctx->box_set || !ctx->lj_set || !ctx->nl_set || !ctx->atoms)
    return fail(ctx, CV2_ERR_STATE, "box, LJ parameters, neighbour settings and atoms must be set first");
  return CV2_OK;
}

// Launch grid of the force task kernel: one work-group per task slot, a
// multiple of 8 (one share per XCD), so that every build's tasks fit
// whatever their count (the work-groups beyond it return at once, which
// costs less than a second launch); at most cap (> 0) when set, the tail
// launch then taking the rest.
int task_grid_for(int64_t nslot, int cap) {
  int64_t g = std::max<int64_t>(nslot, 8);
  if (cap > 0) g = std::min<int64_t>(g, cap);
  return (int)((g + 7) / 8 * 8);
}

int check_flags(cv2_ctx *ctx) {
  HIPCHK(hipMemcpyAsync(ctx->h_flags, ctx->s.flags, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int *f = ctx->h_flags;
  DevState &s = ctx->s;
  if (f[F_TASKOVF]) {
    // some cell row wanted more task slots: its last task took the rest of
    // the row (correct, slower); widen the rows at the next build
    ctx->min_mpr = std::max(ctx->min_mpr, s.tg.mpr + 2);
    ctx->need_build = true;
    int zero = 0;
    HIPCHK(hipMemcpyAsync(s.flags + F_TASKOVF, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  if (f[F_OVERFLOW])  // the callers grow the lists and go on (grow_lists)
    return fail(ctx, CV2_ERR_CAPACITY,
                "neighbour list capacity " + std::to_string(ctx->s.cap) + " exceeded (longest list " +
                    std::to_string(f[F_MAXNBR]) + ")");
  if (f[F_NAN] || f[F_FXOVF]) {
    // fatal at a step: the device halted there.  The flags are cleared so
    // that the context can take a new state (cv2_upload_atoms) and go on.
    const std::string where = " (step " + std::to_string(f[F_STEP]) + " of the run)";
    const int code = f[F_NAN] ? CV2_ERR_NAN : CV2_ERR_FIXED_OVERFLOW;
    const std::string msg = f[F_NAN] ? "NaN acceleration, position or velocity" + where
                                     : "fixed-point force range exceeded (r < sigma/2)" + where;
    const int zero[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(s.flags + F_NAN, zero, sizeof(int) * 2, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(s.flags + F_HALT, zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return fail(ctx, code, msg);
  }
  if (f[F_STAGE])
    return fail(ctx, CV2_ERR_CAPACITY, "a force task's neighbourhood exceeds 65536 atoms (16-bit slots)");
  return CV2_OK;
}

// After a build halted on the list capacity (F_OVERFLOW): lists of
// 1.5 times the longest one (at least 1.5 times the old capacity), and the
// halt cleared; the build request stays set, so the next build runs again.
// Returns CV2_OK when the lists have grown.
int grow_lists(cv2_ctx *ctx) {
  DevState &s = ctx->s;
  const int64_t longest = ctx->h_flags[F_MAXNBR];
  int64_t cap = std::max<int64_t>(longest + longest / 2, s.cap + s.cap / 2) + 16;
  cap = (cap + 15) / 16 * 16;
  if (cap > (1 << 16))
    return fail(ctx, CV2_ERR_CAPACITY, "a neighbour list of " + std::to_string(longest) +
                                           " entries (more than 65536 slots would be needed)");
  HIPCHK(dalloc(&s.nbr, (size_t)cap * s.stride));
  s.cap = (int)cap;
  const int zero = 0;
  HIPCHK(hipMemcpyAsync(s.flags + F_OVERFLOW, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(s.flags + F_HALT, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->err.clear();
  return CV2_OK;
}

// List form of the next build: task slots (16-bit) for the FP32 kernel on
// wave-build boxes, 32-bit entries otherwise.  Called when the host forces
// a build, so the task geometry only changes with the list it belongs to.
int prepare_list_form(cv2_ctx *ctx) {
  DevState &s = ctx->s;
  // the task sentinel is more than sqrt(2) list cutoffs from its home atoms
  double rc2max = 0.0;
  for (int k = 0; k < ctx->lj.nt * ctx->lj.nt; ++k) rc2max = std::max(rc2max, ctx->lj.rc2[k]);
  // (ignore_intra_molecule runs on the gather kernels, 32-bit entries)
  s.slots = ctx->precision != CV2_PREC_FP64 && ctx->b.coarse && s.n > 0 &&
            rc2max < 1.9 * ctx->b.ce_sq && !s.mol && !ctx->vstate;
  if (!s.slots) return CV2_OK;
  const TaskGeom old = s.tg;
  s.tg = task_geom(s.n, ctx->b, ctx->min_mpr);
  if (s.tg.nslot > s.task_cap) {
    HIPCHK(dalloc(&s.tpiece, (size_t)s.tg.nslot * kTaskPieces));
    HIPCHK(dalloc(&s.tinfo, (size_t)s.tg.nslot));
    HIPCHK(dalloc(&s.ctinfo, (size_t)s.tg.nslot));
    HIPCHK(dalloc(&s.ctpiece, (size_t)s.tg.nslot * kTaskPieces));
    s.task_cap = s.tg.nslot;
  }
  const int rows = s.tg.ny * s.tg.nz;
  if (rows != old.ny * old.nz || !s.rowtasks) HIPCHK(dalloc(&s.rowtasks, rows));
  if (s.ncell > s.ctask_cap) {
    HIPCHK(dalloc(&s.ctask, s.ncell));
    s.ctask_cap = s.ncell;
  }
  // until a build has counted the tasks: about one per 200 atoms, plus one
  // per row for the row ends
  s.task_grid = task_grid_for(s.tg.nslot, ctx->task_grid_cap);
  const int nb = s.tg.nslot + 80;  // per-block energy partials (main grid <= nslot + 7, tail 64)
  if (nb > ctx->partial_blocks) {
    HIPCHK(dalloc(&s.partials, (size_t)nb * 8));
    ctx->partial_blocks = nb;
  }
  return CV2_OK;
}

int fetch_energy(cv2_ctx *ctx) {
  if (ctx->precision == CV2_PREC_FP64) {
    HIPCHK(hipMemcpyAsync(ctx->h_red, ctx->s.reduce_out, sizeof(double) * 2, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->virial = ctx->h_red[0];
    ctx->pe = ctx->h_red[1];
  } else {
    HIPCHK(hipMemcpyAsync(ctx->h_fx, ctx->s.fx_acc, sizeof(long long) * 2, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->virial = (double)ctx->h_fx[0] / kEnergyScale;
    ctx->pe = (double)ctx->h_fx[1] / kEnergyScale;
  }
  return CV2_OK;
}

// A list build as queued on the stream, followed (CV2_NL_VERIFY) by the
// independent check of the pair set it produced, before anything moves.
void build_checked(cv2_ctx *ctx, bool keep_acc, int step = 0) {
  launch_build(ctx->s, ctx->b, ctx->stream, ctx->host_halo, keep_acc, step);
  if (ctx->vstate) launch_verify(ctx->s, ctx->b, ctx->vstate, ctx->stream);
}

int queue_build(cv2_ctx *ctx, bool force, bool keep_acc, int step = 0) {
  force = force || ctx->need_build;
  ctx->need_build = false;
  if (force) {
    const int rc = prepare_list_form(ctx);
    if (rc) return rc;
  }
  if (force || (ctx->nl_flags & CV2_NL_REBUILD_EVERY_STEP))
    launch_set_flag(ctx->s.flags, F_REBUILD, 1, ctx->stream);
  build_checked(ctx, keep_acc, step);
  return CV2_OK;
}

int alloc_positions(cv2_ctx *ctx, int64_t ntot) {
  if (ntot <= ctx->pos_cap) return CV2_OK;
  // room for ghosts: 1/8 more than asked, so that a halo whose size
  // fluctuates from step to step does not reallocate on every new maximum
  const int64_t cap = ntot > ctx->n ? ntot + (ntot - ctx->n) / 8 + 4096 : ntot;
  DevState &s = ctx->s;
  // keep the owned part when growing for ghosts
  std::vector<double4> keep;
  std::vector<int> keep_id;
  if (ctx->pos_cap > 0 && ctx->n > 0) {
    keep.resize(ctx->n);
    keep_id.resize(ctx->n);
    HIPCHK(hipMemcpyAsync(keep.data(), s.pos, sizeof(double4) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(keep_id.data(), s.id, sizeof(int) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  // growing: the previous arrays are released (the owned positions and ids
  // were saved above; the rest is rewritten before use)
  for (void **p : {(void **)&s.pos, (void **)&s.pos_ref, (void **)&s.q, (void **)&s.q_alt,
                   (void **)&s.id, (void **)&s.cell, (void **)&s.slot, (void **)&s.gcell_items,
                   (void **)&s.gtag, (void **)&s.gslot, (void **)&s.gtag_alt, (void **)&s.gid_alt,
                   (void **)&s.gpos_alt}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  HIPCHK(dalloc(&s.pos, cap));
  HIPCHK(dalloc(&s.pos_ref, cap));
  HIPCHK(dalloc(&s.q, cap));
  HIPCHK(dalloc(&s.q_alt, cap));
  HIPCHK(dalloc(&s.id, cap));
  HIPCHK(dalloc(&s.cell, cap));
  HIPCHK(dalloc(&s.slot, cap));
  HIPCHK(dalloc(&s.gcell_items, cap));
  // ghost order of open wave builds (ghost counts are below cap)
  HIPCHK(dalloc(&s.gtag, cap));
  HIPCHK(dalloc(&s.gslot, cap));
  HIPCHK(dalloc(&s.gtag_alt, cap));
  HIPCHK(dalloc(&s.gid_alt, cap));
  HIPCHK(dalloc(&s.gpos_alt, cap));
  if (!keep.empty()) {
    HIPCHK(hipMemcpyAsync(s.pos, keep.data(), sizeof(double4) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(s.id, keep_id.data(), sizeof(int) * ctx->n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  ctx->pos_cap = cap;
  return CV2_OK;
}

// Reallocate *p for count elements keeping its first keep elements.
template <class T>
hipError_t grow_keep(T **p, size_t keep, size_t count, hipStream_t st) {
  T *q = nullptr;
  hipError_t e = hipMalloc((void **)&q, sizeof(T) * std::max<size_t>(count, 1));
  if (e != hipSuccess) return e;
  if (*p && keep > 0) {
    e = hipMemcpyAsync(q, *p, sizeof(T) * keep, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      (void)hipFree(q);
      return e;
    }
  }
  if (*p) (void)hipFree(*p);
  *p = q;
  return hipSuccess;
}

// Per-owned-atom arrays for at least n owned atoms (an owned exchange that
// brings atoms in): grown with 1/8 slack, keeping the state of the current
// owned atoms (positions and ids, velocities and accelerations, SoA with
// stride ctx->n), the list stride following the capacity.
int ensure_owned_cap(cv2_ctx *ctx, int64_t n) {
  if (n + ctx->ng > ctx->pos_cap) {  // positions and ids: the owned part is kept
    const int rc = alloc_positions(ctx, n + ctx->ng);
    if (rc) return rc;
  }
  if (n <= ctx->own_cap) return CV2_OK;
  DevState &s = ctx->s;
  const int64_t cap = n + n / 8 + 1024, n0 = ctx->n;
  HIPCHK(grow_keep(&s.vel, 3 * n0, 3 * cap, ctx->stream));
  HIPCHK(grow_keep(&s.acc, 3 * n0, 3 * cap, ctx->stream));
  HIPCHK(grow_keep(&s.tag, n0, cap, ctx->stream));
  HIPCHK(dalloc(&s.pos_alt, cap));
  HIPCHK(dalloc(&s.vel_alt, 3 * cap));
  HIPCHK(dalloc(&s.acc_alt, 3 * cap));
  HIPCHK(dalloc(&s.tag_alt, cap));
  HIPCHK(dalloc(&s.id_alt, cap));
  HIPCHK(dalloc(&s.perm, cap));
  HIPCHK(dalloc(&s.order, cap));
  HIPCHK(dalloc(&s.nnbr, cap));
  s.stride = (int)((cap + 63) / 64 * 64);
  HIPCHK(dalloc(&s.nbr, (size_t)s.cap * s.stride));
  const int nb = (int)((cap + 255) / 256) + 8;
  if (nb > ctx->partial_blocks) {
    HIPCHK(dalloc(&s.partials, (size_t)nb * 8));
    ctx->partial_blocks = nb;
  }
  ctx->own_cap = cap;
  return CV2_OK;
}

// The owned set after an exchange: n atoms in owned order in the state
// arrays.  The ghosts of the old halo go (the next ghost exchange brings the
// new one) and the next force pass rebuilds the list.
int adopt_owned(cv2_ctx *ctx, int64_t n) {
  DevState &s = ctx->s;
  launch_mig_adopt(s, (int)n, ctx->stream);
  ctx->n = n;
  s.n = (int)n;
  ctx->ng = 0;
  s.ng = 0;
  int rc = ensure_wstat(ctx);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->need_build = true;
  if (!ctx->step_open) ctx->forces_valid = ctx->energy_ok = false;
  return CV2_OK;
}

}  // namespace

namespace cv2 {
// Pair tables of the LJ force (lj.cpp:475-484), its energy and the FP32
// path's fixed-point overflow bound, for ntypes x ntypes row-major inputs.
bool build_lj_table(LJDev &L, int ntypes, const double *eps, const double *sigma,
                    const double *cutoff_list, double cutoff, const double *mass, std::string &err) {
  L = LJDev{};
  L.nt = ntypes;
  for (int t = 0; t < ntypes; ++t) {
    if (!(mass[t] > 0.0)) {
      err = "masses must be positive";
      return false;
    }
    L.mass[t] = mass[t];
    L.mass_inv[t] = 1.0 / mass[t];
  }
  for (int k = 0; k < ntypes * ntypes; ++k) {
    double rc = cutoff;
    if (cutoff_list) {
      if (cutoff_list[k] < 0.0) {
        err = "Cutoff list values must be non-negative.";
        return false;
      }
      rc = cutoff_list[k];
    }
    const double c = rc * rc;
    const double rcinv = 1.0 / c;
    const double s2 = sigma[k] * sigma[k];
    const double rho_c2 = s2 * rcinv, rho_c6 = rho_c2 * rho_c2 * rho_c2, rho_c12 = rho_c6 * rho_c6;
    L.eps4[k] = 4.0 * eps[k];
    L.sig2[k] = s2;
    L.rc2[k] = eps[k] == 0.0 ? -1.0 : c;  // eps == 0 pairs are skipped (lj.cpp:470-472)
    L.fshift[k] = 12.0 * rho_c12 * rcinv - 6.0 * rho_c6 * rcinv;
    L.ecut[k] = (6.0 * rho_c12 - 3.0 * rho_c6) * rcinv;
    L.ecut2[k] = 7.0 * rho_c12 - 4.0 * rho_c6;
    L.eps4f[k] = (float)L.eps4[k];
    L.e6f[k] = (float)(6.0 * L.eps4[k] * kForceScale);
    L.e12f[k] = (float)(12.0 * L.eps4[k] * kForceScale);
    L.fshf[k] = (float)(L.eps4[k] * L.fshift[k] * kForceScale);
    // FP32 path: largest r^2 at which |F(r)| = |forceCoef| r reaches 0.9 * 2^19
    // (bisection on the repulsive branch, where |F| falls with r)
    L.r2ovf[k] = 0.0;
    if (eps[k] != 0.0 && sigma[k] > 0.0) {
      auto force = [&](double r) {
        const double r2 = r * r, r2i = 1.0 / r2, s6 = std::pow(s2 * r2i, 3);
        return std::fabs(L.eps4[k] * (r2i * (s6 * (-12.0 * s6 + 6.0)) + L.fshift[k])) * r;
      };
      double a = 1e-3 * sigma[k], z = sigma[k];
      for (int it = 0; it < 200; ++it) {
        const double m = 0.5 * (a + z);
        (force(m) > 0.9 * 524288.0 ? a : z) = m;
      }
      L.r2ovf[k] = z * z;
    }
  }
  return true;
}

bool coarse_filter_setup(BoxDev &b, double ce, double wmax) {
  // Rounding to the fixed-point unit u moves a distance by less than
  // sqrt(sum u_d^2); the float arithmetic on coordinates relative to a home
  // atom (|x| < 2 w) adds at most e_f = 8 * 2^-24 * (2 w + ce) per axis.
  double u2 = 0.0;
  for (int d = 0; d < 3; ++d) u2 += b.qinv[d] * b.qinv[d];
  const double ef = 8.0 * 5.9604644775390625e-08 * (2.0 * wmax + ce) * std::sqrt(3.0);
  const double e = 1.01 * std::sqrt(u2) + ef + 1e-6 * ce;
  b.cthr_in = std::nextafter((float)((ce - e) * (ce - e) * (1.0 - 4e-6)), 0.0f);
  b.cthr_out = std::nextafter((float)((ce + e) * (ce + e) * (1.0 + 4e-6)), 1e30f);
  bool ok = ce > 2.0 * e;
  // periodic axes: int32 differences are minimum-image separations
  for (int d = 0; d < 3; ++d) ok = ok && (!b.bc[d] || 0.5 * b.L[d] > ce + 2.0 * e);
  return ok;
}

void set_qscale(BoxDev &b, double span) {
  // periodic axes: the largest k with span < L / 2^(k+1) (0.1% margin), or
  // 0 (int32 differences are then minimum-image separations)
  double ufine = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!b.bc[d]) continue;
    int k = 0;
    while (k < 30 && std::ldexp(span * 1.001, k + 2) < b.L[d]) ++k;
    b.qinv[d] = std::ldexp(b.L[d], -(32 + k));
    ufine = ufine == 0.0 ? b.qinv[d] : std::min(ufine, b.qinv[d]);
  }
  // open axes need only span < 2^31 u: take the finest periodic unit, which
  // satisfies it, or span / 2^30 when no axis is periodic
  if (ufine == 0.0) ufine = std::ldexp(span, -30);
  for (int d = 0; d < 3; ++d)
    if (!b.bc[d]) b.qinv[d] = ufine;
  for (int d = 0; d < 3; ++d) {
    b.qscale[d] = 1.0 / b.qinv[d];
    b.cunit[d] = (float)b.qinv[d];
  }
  b.qiso = b.qinv[0] == b.qinv[1] && b.qinv[1] == b.qinv[2];
}

namespace {

// MPI halo packing helpers (cv2_halo_pack / cv2_halo_unpack).
HaloGeomDev halo_geom(const cv2_halo_geom *g) {
  HaloGeomDev h{};
  for (int a = 0; a < 3; ++a) {
    h.gi[a] = g->grid_index[a];
    h.np[a] = g->nprocs[a];
    h.bc[a] = g->bc[a];
    h.lo[a] = g->lo_local[a];
    h.hi[a] = g->hi_local[a];
    h.size[a] = g->size_global[a];
  }
  h.cut = g->ghost_cutoff;
  h.with_vel = g->with_velocity ? 1 : 0;
  return h;
}

// Device scratch of the halo packing for n owned atoms.
int halo_scratch(cv2_ctx *ctx, int64_t n) {
  if (n <= ctx->halo_n && ctx->halo_code) return CV2_OK;
  // (dalloc releases the smaller buffers)
  const int64_t nt = halo_tiles((int)n);
  HIPCHK(dalloc(&ctx->halo_code, n));
  HIPCHK(dalloc(&ctx->halo_src, n));
  HIPCHK(dalloc(&ctx->halo_tiles, nt * 27));
  HIPCHK(dalloc(&ctx->halo_toff, nt * 27));
  if (!ctx->halo_dn) HIPCHK(dalloc(&ctx->halo_dn, 27));
  if (!ctx->halo_base) HIPCHK(dalloc(&ctx->halo_base, 3 * 64));
  if (!ctx->halo_res) HIPCHK(dalloc(&ctx->halo_res, 28));
  ctx->halo_n = n;
  return CV2_OK;
}

int halo_buffer(cv2_ctx *ctx, int64_t doubles) {
  if (doubles <= ctx->halo_buf_cap) return CV2_OK;
  if (ctx->halo_buf) (void)hipFree(ctx->halo_buf);
  ctx->halo_buf = nullptr;
  HIPCHK(dalloc(&ctx->halo_buf, doubles));
  ctx->halo_buf_cap = doubles;
  return CV2_OK;
}

}  // namespace

}  // namespace cv2

extern "C" {

int cv2_create(cv2_ctx **out, int device) {
  if (!out) return CV2_ERR_BAD_ARG;
  *out = nullptr;
  cv2_ctx *ctx = new cv2_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&ctx->lj_dev, sizeof(LJDev));
  if (e == hipSuccess) e = hipMalloc((void **)&ctx->s.flags, sizeof(int) * F_COUNT);
  if (e == hipSuccess) e = hipMemset(ctx->s.flags, 0, sizeof(int) * F_COUNT);
  if (e == hipSuccess) e = hipMalloc((void **)&ctx->s.reduce_out, sizeof(double) * 16);
  if (e == hipSuccess) e = hipMalloc((void **)&ctx->s.fx_acc, sizeof(long long) * 4);
  if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->h_flags, sizeof(int) * F_COUNT, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->h_red, sizeof(double) * 16, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->h_fx, sizeof(long long) * 4, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->h_halo, sizeof(long long) * 32, hipHostMallocDefault);
  // the no-allocation halo pack's scratch on this device (its error, if
  // any, is reported by that entry point)
  if (e == hipSuccess) (void)noalloc_scratch_init(device);
  // LDS staging capacity of the FP32 task kernel; CV2_STAGE_LIMIT lowers it
  // (tests use it to exercise the unstaged path)
  ctx->s.stage_limit = kStage;
  if (const char *lim = std::getenv("CV2_STAGE_LIMIT"))
    ctx->s.stage_limit = std::max(0, std::min(kStage, std::atoi(lim)));
  // CV2_FUSE=0 keeps the integrator in its own kernel (tests compare both)
  if (const char *fz = std::getenv("CV2_FUSE")) ctx->fuse = std::atoi(fz) != 0;
  // CV2_TASK_GRID caps the force launch's grid, so that the tail launch
  // takes the remaining tasks (tests compare both)
  if (const char *tg = std::getenv("CV2_TASK_GRID")) ctx->task_grid_cap = std::max(0, std::atoi(tg));
  if (e != hipSuccess) {
    fprintf(stderr, "cv2_create: %s\n", hipGetErrorString(e));
    cv2_destroy(ctx);
    return CV2_ERR_HIP;
  }
  *out = ctx;
  return CV2_OK;
}

int cv2_destroy(cv2_ctx *ctx) {
  if (!ctx) return CV2_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_state(ctx->s);
  for (void *p : {(void *)ctx->halo_code, (void *)ctx->halo_src, (void *)ctx->halo_tiles,
                  (void *)ctx->halo_toff, (void *)ctx->halo_dn, (void *)ctx->halo_base,
                  (void *)ctx->halo_buf, (void *)ctx->halo_res})
    if (p) (void)hipFree(p);
  if (ctx->type_by_id) (void)hipFree(ctx->type_by_id);
  if (ctx->vstate) (void)hipFree(ctx->vstate);
  if (ctx->mig_moved) (void)hipFree(ctx->mig_moved);
  if (ctx->mig_aux) (void)hipFree(ctx->mig_aux);
  if (ctx->lj_dev) (void)hipFree(ctx->lj_dev);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
  if (ctx->h_red) (void)hipHostFree(ctx->h_red);
  if (ctx->h_fx) (void)hipHostFree(ctx->h_fx);
  if (ctx->h_halo) (void)hipHostFree(ctx->h_halo);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CV2_OK;
}

const char *cv2_last_error(const cv2_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cv2_set_box(cv2_ctx *ctx, const double lo[3], const double hi[3], const int bc[3],
                double ghost_cutoff) {
  if (!ctx || !lo || !hi || !bc) return fail(ctx, CV2_ERR_BAD_ARG, "null argument");
  for (int d = 0; d < 3; ++d) {
    if (!(hi[d] > lo[d])) return fail(ctx, CV2_ERR_BAD_ARG, "upper must exceed lower on every axis");
    if (bc[d] != 0 && bc[d] != 1) return fail(ctx, CV2_ERR_BAD_ARG, "boundary condition must be 0 or 1");
    if (bc[d] == 1 && hi[d] - lo[d] <= 2.0 * ghost_cutoff)
      return fail(ctx, CV2_ERR_BAD_ARG, "periodic box must be wider than twice the ghost cutoff");
    ctx->lo[d] = lo[d];
    ctx->hi[d] = hi[d];
    ctx->bc[d] = bc[d];
  }
  if (ghost_cutoff < 0) return fail(ctx, CV2_ERR_BAD_ARG, "ghost cutoff must be non-negative");
  ctx->ghost_cut = ghost_cutoff;
  ctx->box_set = true;
  ctx->need_build = true;
  ctx->forces_valid = ctx->energy_ok = false;
  if (ctx->nl_set) return setup_box(ctx);
  return CV2_OK;
}

int cv2_set_lj(cv2_ctx *ctx, int ntypes, const double *eps, const double *sigma,
               const double *cutoff_list, double cutoff, const double *mass) {
  if (!ctx || !eps || !sigma || !mass) return fail(ctx, CV2_ERR_BAD_ARG, "null argument");
  if (ntypes < 1 || ntypes > kMaxTypes) return fail(ctx, CV2_ERR_BAD_ARG, "ntypes must be 1..8");
  if (cutoff < 0.0) return fail(ctx, CV2_ERR_BAD_ARG, "Force field cutoff must be non-negative.");
  {
    std::string err;
    const double jt = ctx->lj.jump_tol;  // set independently (cv2_set_jump_tol)
    if (!build_lj_table(ctx->lj, ntypes, eps, sigma, cutoff_list, cutoff, mass, err))
      return fail(ctx, CV2_ERR_BAD_ARG, err);
    ctx->lj.jump_tol = jt;
  }
  ctx->lj_set = true;
  ctx->need_build = true;
  ctx->forces_valid = ctx->energy_ok = false;
  hipError_t e = hipMemcpyAsync(ctx->lj_dev, &ctx->lj, sizeof(LJDev), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, CV2_ERR_HIP, hipGetErrorString(e));
  return CV2_OK;
}

int cv2_set_jump_tol(cv2_ctx *ctx, double jump_tol) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  ctx->lj.jump_tol = jump_tol > 0.0 ? jump_tol : 0.0;
  ctx->forces_valid = ctx->energy_ok = false;
  if (!ctx->lj_set) return CV2_OK;  // uploaded with the LJ table
  hipError_t e = hipMemcpyAsync(ctx->lj_dev, &ctx->lj, sizeof(LJDev), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, CV2_ERR_HIP, hipGetErrorString(e));
  return CV2_OK;
}

int cv2_set_molecules(cv2_ctx *ctx, int64_t n, const int32_t *molecule_index,
                      int ignore_intra_molecule) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "upload owned atoms before their molecule indices");
  DevState &s = ctx->s;
  (void)hipSetDevice(ctx->device);
  if (!ignore_intra_molecule) {
    if (s.mol) (void)hipFree(s.mol);
    s.mol = nullptr;
  } else {
    if (n != ctx->n) return fail(ctx, CV2_ERR_BAD_ARG, "one molecule index per owned atom");
    if (ctx->host_halo)
      return fail(ctx, CV2_ERR_STATE,
                  "ignore_intra_molecule on a host-halo share: the halo packets carry no molecule "
                  "indices (exchange_ghost_mpi_shared_atoms.cpp:405-470)");
    std::vector<int> m((size_t)n, -1);  // atom_data.cpp:625: -1 outside molecules
    if (molecule_index)
      for (int64_t i = 0; i < n; ++i) m[i] = molecule_index[i];
    if (!s.mol) HIPCHK(dalloc(&s.mol, std::max<int64_t>(n, 1)));
    HIPCHK(hipMemcpyAsync(s.mol, m.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  // the list form follows (32-bit entries while the flag is on)
  ctx->need_build = true;
  ctx->forces_valid = ctx->energy_ok = false;
  return CV2_OK;
}

int cv2_set_neighbor(cv2_ctx *ctx, double cutoff, double extra_coef, unsigned flags) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  if (cutoff <= 0.0) return fail(ctx, CV2_ERR_BAD_ARG, "cutoff have to non-negative.");
  if (extra_coef < 1.0) return fail(ctx, CV2_ERR_BAD_ARG, "cutoff_extra_coef must be larger than 1.0");
  if ((flags & CV2_NL_VERIFY) && ctx->host_halo)
    return fail(ctx, CV2_ERR_STATE, "CV2_NL_VERIFY checks periodic and open boxes, not host-supplied halos");
  if (flags & CV2_NL_VERIFY) {
    if (!ctx->vstate) {
      (void)hipSetDevice(ctx->device);
      HIPCHK(dalloc(&ctx->vstate, (size_t)verify_state_words()));
      HIPCHK(hipMemsetAsync(ctx->vstate, 0, sizeof(unsigned long long) * verify_state_words(), ctx->stream));
      launch_verify_arm(ctx->s.flags, ctx->vstate, ctx->stream);
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
  } else if (ctx->vstate) {
    (void)hipFree(ctx->vstate);
    ctx->vstate = nullptr;
  }
  ctx->nl_cutoff = cutoff;
  ctx->nl_coef = extra_coef;
  ctx->nl_flags = flags;
  ctx->nl_set = true;
  ctx->need_build = true;
  ctx->forces_valid = ctx->energy_ok = false;
  if (ctx->box_set) return setup_box(ctx);
  return CV2_OK;
}

int cv2_set_precision(cv2_ctx *ctx, int precision) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  if (precision != CV2_PREC_FP64 && precision != CV2_PREC_FP32_FIXED)
    return fail(ctx, CV2_ERR_BAD_ARG, "unknown precision");
  ctx->precision = precision;
  ctx->s.write_q = precision != CV2_PREC_FP64;
  ctx->forces_valid = ctx->energy_ok = false;
  ctx->need_build = true;  // the list form follows the precision
  return CV2_OK;
}

int cv2_upload_atoms(cv2_ctx *ctx, int64_t n, const double *pos, const double *vel,
                     const int32_t *type, const int64_t *id) {
  if (!ctx || n < 0 || (n > 0 && !pos)) return fail(ctx, CV2_ERR_BAD_ARG, "bad atom arrays");
  if (!ctx->box_set || !ctx->nl_set || !ctx->lj_set)
    return fail(ctx, CV2_ERR_STATE, "set box, LJ and neighbour parameters before uploading atoms");
  if (n + ctx->ng > (int64_t)kIndexMask) return fail(ctx, CV2_ERR_BAD_ARG, "too many atoms (2^26 max)");
  (void)hipSetDevice(ctx->device);
  DevState &s = ctx->s;
  const int ni = (int)n;
  // host staging in device layout
  std::vector<double4> p4(n);
  std::vector<double> v3(3 * n, 0.0), zero(3 * n, 0.0);
  std::vector<int> tg(n), idv(n);
  for (int64_t i = 0; i < n; ++i) {
    const int t = type ? type[i] : 0;
    if (t < 0 || t >= ctx->lj.nt) return fail(ctx, CV2_ERR_BAD_ARG, "atom type out of range");
    p4[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], (double)t);
    if (vel)
      for (int d = 0; d < 3; ++d) v3[d * n + i] = vel[3 * i + d];
    tg[i] = (int)i;
    idv[i] = id ? (int)id[i] : (int)i;
  }
  ctx->pos_cap = 0;
  ctx->n = n;
  s.n = ni;
  if (s.mol) (void)hipFree(s.mol);  // molecule indices belong to the previous atoms
  s.mol = nullptr;
  {
    int rc = alloc_positions(ctx, n + ctx->ng);
    if (rc) return rc;
  }
  HIPCHK(dalloc(&s.pos_alt, n));
  HIPCHK(dalloc(&s.vel, 3 * n));
  HIPCHK(dalloc(&s.vel_alt, 3 * n));
  HIPCHK(dalloc(&s.acc, 3 * n));
  HIPCHK(dalloc(&s.acc_alt, 3 * n));
  HIPCHK(dalloc(&s.tag, n));
  HIPCHK(dalloc(&s.tag_alt, n));
  HIPCHK(dalloc(&s.id_alt, n));
  HIPCHK(dalloc(&s.perm, n));
  HIPCHK(dalloc(&s.order, n));
  {
    int rc = ensure_wstat(ctx);
    if (rc) return rc;
  }
  HIPCHK(dalloc(&s.nnbr, n));
  // list capacity from the mean density: 1.6x the expected count + 64
  const double vol = (ctx->hi[0] - ctx->lo[0]) * (ctx->hi[1] - ctx->lo[1]) * (ctx->hi[2] - ctx->lo[2]);
  const double ce = ctx->nl_cutoff * ctx->nl_coef;
  const double expect = 4.18879020478639 * ce * ce * ce * (double)n / vol;
  int cap = (int)std::ceil(1.6 * expect + 64.0);
  // CV2_LIST_CAP lowers it (tests: lists that outgrow it during a run)
  if (const char *lc = std::getenv("CV2_LIST_CAP")) cap = std::max(16, std::min(cap, std::atoi(lc)));
  cap = (cap + 15) / 16 * 16;
  s.cap = cap;
  s.stride = (int)((n + 63) / 64 * 64);  // atoms per 4-entry group row
  HIPCHK(dalloc(&s.nbr, (size_t)s.cap * s.stride));
  const int nb = (int)((n + 255) / 256) + 8;
  if (nb > ctx->partial_blocks) {
    HIPCHK(dalloc(&s.partials, (size_t)nb * 8));
    ctx->partial_blocks = nb;
  }
  HIPCHK(hipMemcpyAsync(s.pos, p4.data(), sizeof(double4) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(s.vel, v3.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(s.acc, zero.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(s.tag, tg.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(s.id, idv.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(s.flags, 0, sizeof(int) * F_COUNT, ctx->stream));
  if (ctx->vstate)  // the build count starts again: so do the checks
    HIPCHK(hipMemsetAsync(ctx->vstate, 0, sizeof(unsigned long long) * verify_state_words(), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->atoms = true;
  ctx->need_build = true;
  ctx->forces_valid = ctx->energy_ok = false;
  ctx->step_open = false;
  ctx->steps = 0;
  ctx->own_cap = n;
  ctx->ids_sorted = true;
  for (int64_t i = 1; i < n; ++i)
    if (idv[i] <= idv[i - 1]) ctx->ids_sorted = false;
  return CV2_OK;
}

int cv2_set_ghosts(cv2_ctx *ctx, int64_t ng, const double *pos, const int32_t *type,
                   const int64_t *id) {
  if (!ctx || ng < 0 || (ng > 0 && !pos)) return fail(ctx, CV2_ERR_BAD_ARG, "bad ghost arrays");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "upload owned atoms before ghosts");
  if (ctx->s.mol)
    return fail(ctx, CV2_ERR_STATE, "host-supplied ghosts carry no molecule indices (ignore_intra_molecule is on)");
  if (ctx->vstate) return fail(ctx, CV2_ERR_STATE, "CV2_NL_VERIFY does not check host-supplied halos");
  (void)hipSetDevice(ctx->device);
  const bool count_changed = ng != ctx->ng || !ctx->host_halo;
  if (!ctx->host_halo) {
    ctx->host_halo = true;
    int rc = setup_box(ctx);
    if (rc) return rc;
  }
  {
    int rc = alloc_positions(ctx, ctx->n + ng);
    if (rc) return rc;
  }
  std::vector<double4> p4(ng);
  std::vector<int> idv(ng);
  for (int64_t g = 0; g < ng; ++g) {
    const int t = type ? type[g] : 0;
    if (t < 0 || t >= ctx->lj.nt) return fail(ctx, CV2_ERR_BAD_ARG, "ghost type out of range");
    p4[g] = make_double4(pos[3 * g], pos[3 * g + 1], pos[3 * g + 2], (double)t);
    idv[g] = id ? (int)id[g] : (int)(ctx->n + g);
  }
  DevState &s = ctx->s;
  // staged, then placed: ghost g at its slot of the last build's cell order
  // when the count is unchanged (k_ghost_place)
  HIPCHK(hipMemcpyAsync(s.gpos_alt, p4.data(), sizeof(double4) * ng, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(s.gid_alt, idv.data(), sizeof(int) * ng, hipMemcpyHostToDevice, ctx->stream));
  ctx->ng = ng;
  s.ng = (int)ng;
  launch_ghost_place(s, s.gpos_alt, s.gid_alt, (int)ng, ctx->b.open && !count_changed, ctx->stream);
  if (count_changed) ctx->need_build = true;  // neighborlist.cpp:118-121
  // ghosts compared by index (:139-144) unless the build is forced anyway;
  // the fixed-point copy is written by the build or the next force pass
  if (!count_changed) launch_check_disp(s, ctx->b, ref_rebuilds(ctx), ctx->stream);
  HIPCHK(hipStreamSynchronize(ctx->stream));  // p4/idv are stack buffers
  return CV2_OK;
}


int cv2_halo_pack(cv2_ctx *ctx, const cv2_halo_geom *g, int64_t counts[27], int64_t cap,
                  double *out, int64_t *need) {
  if (!ctx || !g || !counts || !need || cap < 0 || (cap > 0 && !out))
    return fail(ctx, CV2_ERR_BAD_ARG, "bad halo pack arguments");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "upload atoms before packing a halo");
  for (int a = 0; a < 3; ++a)
    if (g->nprocs[a] < 1 || g->grid_index[a] < 0 || g->grid_index[a] >= g->nprocs[a] ||
        !(g->hi_local[a] > g->lo_local[a]))
      return fail(ctx, CV2_ERR_BAD_ARG, "bad process grid or local box");
  (void)hipSetDevice(ctx->device);
  const int n = (int)ctx->n;
  int rc = halo_scratch(ctx, std::max(n, 1));
  if (rc) return rc;
  const HaloGeomDev h = halo_geom(g);
  const int per_atom = g->with_velocity ? 8 : 5;
  // tile offsets, per-direction totals and packet bases scanned on the
  // device; the host reads the 28 totals
  launch_halo_classify(ctx->s, h, ctx->halo_code, ctx->halo_src, ctx->halo_tiles, ctx->stream);
  launch_halo_scan(halo_tiles(n), per_atom, ctx->halo_tiles, ctx->halo_toff, ctx->halo_dn,
                   ctx->halo_base, ctx->halo_res, ctx->stream);
  HIPCHK(hipMemcpyAsync(ctx->h_halo, ctx->halo_res, sizeof(long long) * 28, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int d = 0; d < 27; ++d) counts[d] = n > 0 ? ctx->h_halo[d] : 0;
  const int64_t at = n > 0 ? ctx->h_halo[27] : 0;
  *need = at;
  if (at == 0 || at > cap) return CV2_OK;
  rc = halo_buffer(ctx, at);
  if (rc) return rc;
  launch_halo_write(ctx->s, h, ctx->halo_code, ctx->halo_src, ctx->halo_toff, ctx->halo_base,
                    ctx->halo_dn, ctx->halo_buf, ctx->stream);
  HIPCHK(hipMemcpyAsync(out, ctx->halo_buf, sizeof(double) * at, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CV2_OK;
}

int cv2_halo_pack_noalloc(cv2_ctx *ctx, const cv2_halo_geom *g, int64_t counts[27], int64_t cap,
                          double *out, int64_t *need) {
  if (!ctx || !g || !counts || !need || cap < 0 || (cap > 0 && !out))
    return fail(ctx, CV2_ERR_BAD_ARG, "bad halo pack arguments");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "upload atoms before packing a halo");
  for (int a = 0; a < 3; ++a)
    if (g->nprocs[a] < 1 || g->grid_index[a] < 0 || g->grid_index[a] >= g->nprocs[a] ||
        !(g->hi_local[a] > g->lo_local[a]))
      return fail(ctx, CV2_ERR_BAD_ARG, "bad process grid or local box");
  if (ctx->n > (int64_t)kNaMaxAtoms)
    return fail(ctx, CV2_ERR_BAD_ARG,
                "cv2_halo_pack_noalloc: " + std::to_string(ctx->n) +
                    " owned atoms exceed its scratch limit of " + std::to_string(kNaMaxAtoms));
  (void)hipSetDevice(ctx->device);
  unsigned char *sc = noalloc_scratch(ctx->device);
  if (!sc)
    return fail(ctx, CV2_ERR_HIP,
                std::string("cv2_halo_pack_noalloc: no scratch address on this device: ") +
                    hipGetErrorString(ctx->device >= 0 && ctx->device < kNaMaxDevices
                                          ? g_na_err[ctx->device] : hipErrorInvalidDevice));
  int *code = reinterpret_cast<int *>(sc + kNaCode), *src = reinterpret_cast<int *>(sc + kNaSrc);
  int *tcnt = reinterpret_cast<int *>(sc + kNaTcnt), *toff = reinterpret_cast<int *>(sc + kNaToff);
  int *dir_n = reinterpret_cast<int *>(sc + kNaDirN);
  long long *dir_base = reinterpret_cast<long long *>(sc + kNaBase);
  long long *res = reinterpret_cast<long long *>(sc + kNaRes);
  double *buf = reinterpret_cast<double *>(sc + kNaBuf);
  const HaloGeomDev h = halo_geom(g);
  const int n = (int)ctx->n;
  const int per_atom = g->with_velocity ? 8 : 5;
  // every scratch word read below is written first by this call's kernels
  launch_halo_classify(ctx->s, h, code, src, tcnt, ctx->stream);
  launch_halo_scan(halo_tiles(n), per_atom, tcnt, toff, dir_n, dir_base, res, ctx->stream);
  HIPCHK(hipMemcpyAsync(ctx->h_halo, res, sizeof(long long) * 28, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int d = 0; d < 27; ++d) counts[d] = n > 0 ? ctx->h_halo[d] : 0;
  const int64_t at = n > 0 ? ctx->h_halo[27] : 0;
  *need = at;
  if (at == 0 || at > cap) return CV2_OK;
  if (at > (int64_t)CV2_NOALLOC_MAX_DOUBLES)
    return fail(ctx, CV2_ERR_BAD_ARG,
                "cv2_halo_pack_noalloc: " + std::to_string(at) + " packet doubles exceed its scratch limit of " +
                    std::to_string((int64_t)CV2_NOALLOC_MAX_DOUBLES));
  launch_halo_write(ctx->s, h, code, src, toff, dir_base, dir_n, buf, ctx->stream);
  HIPCHK(hipMemcpyAsync(out, buf, sizeof(double) * at, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CV2_OK;
}

// cv2_halo_unpack / cv2_halo_unpack_at: packet p at packets + offsets[p]
// (offsets null: the packets back to back).
static int halo_unpack_impl(cv2_ctx *ctx, int32_t npackets, const int64_t *counts, const int64_t *offsets,
                     const double *packets, int32_t with_velocity) {
  if (!ctx || npackets < 0 || npackets > 64 || (npackets > 0 && !counts))
    return fail(ctx, CV2_ERR_BAD_ARG, "bad halo unpack arguments (at most 64 packets)");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "upload owned atoms before ghosts");
  if (ctx->s.mol)
    return fail(ctx, CV2_ERR_STATE, "halo packets carry no molecule indices (ignore_intra_molecule is on)");
  if (ctx->vstate) return fail(ctx, CV2_ERR_STATE, "CV2_NL_VERIFY does not check host-supplied halos");
  (void)hipSetDevice(ctx->device);
  const int64_t total = with_velocity ? 8 : 5;
  long long first[64], pbase[64], pn[64];
  int64_t ng = 0, at = 0;  // at: doubles of the buffer the packets lie in
  for (int p = 0; p < npackets; ++p) {
    if (counts[p] < 0) return fail(ctx, CV2_ERR_BAD_ARG, "negative packet count");
    const int64_t o = offsets ? offsets[p] : at;
    if (o < 0) return fail(ctx, CV2_ERR_BAD_ARG, "negative packet offset");
    if (counts[p] > 0 && !packets) return fail(ctx, CV2_ERR_BAD_ARG, "missing packets");
    first[p] = ng;
    pbase[p] = o;
    pn[p] = counts[p];
    // types are validated here, as cv2_set_ghosts does
    for (int64_t c = 0; c < counts[p]; ++c) {
      const double t = packets[o + counts[p] + c];
      if (!(t >= 0.0 && t < ctx->lj.nt && t == (double)(int)t))
        return fail(ctx, CV2_ERR_BAD_ARG, "ghost type out of range");
    }
    ng += counts[p];
    at = offsets ? std::max(at, o + total * counts[p]) : at + total * counts[p];
  }
  if (ng > 0 && !packets) return fail(ctx, CV2_ERR_BAD_ARG, "missing packets");
  const bool count_changed = ng != ctx->ng || !ctx->host_halo;
  if (!ctx->host_halo) {
    ctx->host_halo = true;
    int rc = setup_box(ctx);
    if (rc) return rc;
  }
  int rc = alloc_positions(ctx, ctx->n + ng);
  if (rc) return rc;
  rc = halo_scratch(ctx, std::max<int64_t>(ctx->n, 1));
  if (rc) return rc;
  rc = halo_buffer(ctx, std::max<int64_t>(at, 1));
  if (rc) return rc;
  DevState &s = ctx->s;
  if (at > 0 && ng > 0)
    HIPCHK(hipMemcpyAsync(ctx->halo_buf, packets, sizeof(double) * at, hipMemcpyHostToDevice,
                          ctx->stream));
  if (npackets > 0) {
    long long tab[3 * 64];
    for (int p = 0; p < npackets; ++p) {
      tab[p] = first[p];
      tab[64 + p] = pbase[p];
      tab[128 + p] = pn[p];
    }
    HIPCHK(hipMemcpyAsync(ctx->halo_base, tab, sizeof(tab), hipMemcpyHostToDevice, ctx->stream));
  }
  ctx->ng = ng;
  s.ng = (int)ng;
  launch_halo_unpack(s, ctx->halo_buf, npackets, ctx->halo_base, ctx->halo_base + 64,
                     ctx->halo_base + 128, (int)ng, ctx->b.open && !count_changed, ctx->stream);
  if (count_changed) ctx->need_build = true;  // neighborlist.cpp:118-121
  // ghosts compared by index (:139-144) unless the build is forced anyway;
  // the fixed-point copy is written by the build or the next force pass
  if (!count_changed) launch_check_disp(s, ctx->b, ref_rebuilds(ctx), ctx->stream);
  HIPCHK(hipStreamSynchronize(ctx->stream));  // tab is a stack buffer
  return CV2_OK;
}

int cv2_halo_unpack(cv2_ctx *ctx, int32_t npackets, const int64_t *counts, const double *packets,
                    int32_t with_velocity) {
  return halo_unpack_impl(ctx, npackets, counts, nullptr, packets, with_velocity);
}

int cv2_halo_unpack_at(cv2_ctx *ctx, int32_t npackets, const int64_t *counts,
                       const int64_t *offsets, const double *packets, int32_t with_velocity) {
  if (npackets > 0 && !offsets) return fail(ctx, CV2_ERR_BAD_ARG, "missing packet offsets");
  return halo_unpack_impl(ctx, npackets, counts, offsets, packets, with_velocity);
}

namespace {
// Bit d: direction d leads back to this domain (one process on every axis
// the direction moves along), as Domain::all[i][j][k] == me (domain.cpp:70-94).
unsigned self_directions(const cv2_halo_geom *g) {
  unsigned m = 0;
  for (int d = 0; d < 27; ++d) {
    if (d == 13) continue;
    const int c[3] = {d / 9, (d / 3) % 3, d % 3};
    bool self = true;
    for (int a = 0; a < 3; ++a)
      if (c[a] != 1 && g->nprocs[a] != 1) self = false;
    if (self) m |= 1u << d;
  }
  return m;
}
}  // namespace

int cv2_set_type_table(cv2_ctx *ctx, int64_t nid, const int32_t *type_by_id) {
  if (!ctx || nid < 0 || (nid > 0 && !type_by_id)) return fail(ctx, CV2_ERR_BAD_ARG, "bad type table");
  if (!ctx->lj_set) return fail(ctx, CV2_ERR_STATE, "set the LJ types before the type table");
  for (int64_t i = 0; i < nid; ++i)
    if (type_by_id[i] < 0 || type_by_id[i] >= ctx->lj.nt)
      return fail(ctx, CV2_ERR_BAD_ARG, "atom type out of range");
  (void)hipSetDevice(ctx->device);
  std::vector<int> t(type_by_id, type_by_id + nid);
  HIPCHK(dalloc(&ctx->type_by_id, std::max<int64_t>(nid, 1)));
  if (nid > 0)
    HIPCHK(hipMemcpyAsync(ctx->type_by_id, t.data(), sizeof(int) * nid, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->ntype_ids = nid;
  return CV2_OK;
}

int cv2_owned_pack(cv2_ctx *ctx, const cv2_halo_geom *g, int64_t counts[27], int64_t cap,
                   double *out, int64_t *need) {
  if (!ctx || !g || !counts || !need || cap < 0 || (cap > 0 && !out))
    return fail(ctx, CV2_ERR_BAD_ARG, "bad owned pack arguments");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "upload atoms before exchanging them");
  if (ctx->s.mol)
    return fail(ctx, CV2_ERR_STATE,
                "atoms in molecules move with their molecule's centre (exchange_owned_mpi_shared_atoms.cpp:"
                "208-300); the device exchange moves atoms outside molecules");
  for (int a = 0; a < 3; ++a)
    if (g->nprocs[a] < 1 || g->grid_index[a] < 0 || g->grid_index[a] >= g->nprocs[a] ||
        !(g->hi_local[a] > g->lo_local[a]))
      return fail(ctx, CV2_ERR_BAD_ARG, "bad process grid or local box");
  (void)hipSetDevice(ctx->device);
  const int n = (int)ctx->n;
  int rc = halo_scratch(ctx, std::max(n, 1));
  if (rc) return rc;
  if (!ctx->mig_moved) HIPCHK(dalloc(&ctx->mig_moved, 1));
  const HaloGeomDev h = halo_geom(g);
  const unsigned selfmask = self_directions(g);
  const int per_atom = mig_per_atom();
  HIPCHK(hipMemsetAsync(ctx->mig_moved, 0, sizeof(int), ctx->stream));
  launch_mig_classify(ctx->s, h, selfmask, ctx->halo_code, ctx->halo_src, ctx->halo_tiles,
                      ctx->mig_moved, ctx->stream);
  launch_halo_scan(halo_tiles(n), per_atom, ctx->halo_tiles, ctx->halo_toff, ctx->halo_dn,
                   ctx->halo_base, ctx->halo_res, ctx->stream, true);
  int moved = 0;
  HIPCHK(hipMemcpyAsync(ctx->h_halo, ctx->halo_res, sizeof(long long) * 28, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(&moved, ctx->mig_moved, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int64_t stay = n > 0 ? ctx->h_halo[13] : 0;
  for (int d = 0; d < 27; ++d) counts[d] = (n > 0 && d != 13) ? ctx->h_halo[d] : 0;
  const int64_t at = n > 0 ? ctx->h_halo[27] : 0;
  *need = at;
  if (at > cap) return CV2_OK;  // sizing: nothing changes until the packets fit
  if (at == 0) {
    if (moved) {  // only shifts back into the domain: in place, the list is rebuilt
      launch_mig_shift(ctx->s, h, selfmask, ctx->stream);
      HIPCHK(hipStreamSynchronize(ctx->stream));
      ctx->need_build = true;
    }
    return CV2_OK;
  }
  rc = halo_buffer(ctx, at);
  if (rc) return rc;
  launch_mig_write(ctx->s, h, selfmask, ctx->halo_code, ctx->halo_src, ctx->halo_toff,
                   ctx->halo_base, ctx->halo_dn, ctx->halo_buf, ctx->stream);
  HIPCHK(hipMemcpyAsync(out, ctx->halo_buf, sizeof(double) * at, hipMemcpyDeviceToHost, ctx->stream));
  return adopt_owned(ctx, stay);
}

int cv2_owned_unpack(cv2_ctx *ctx, int32_t npackets, const int64_t *counts, const double *packets) {
  if (!ctx || npackets < 0 || npackets > 64 || (npackets > 0 && !counts))
    return fail(ctx, CV2_ERR_BAD_ARG, "bad owned unpack arguments (at most 64 packets)");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "upload atoms before exchanging them");
  const int per_atom = mig_per_atom();
  long long pbase[64], pn[64];
  int64_t nr = 0, at = 0;
  for (int p = 0; p < npackets; ++p) {
    if (counts[p] < 0) return fail(ctx, CV2_ERR_BAD_ARG, "negative packet count");
    pbase[p] = at;
    pn[p] = counts[p];
    nr += counts[p];
    at += per_atom * counts[p];
  }
  if (nr == 0) return CV2_OK;
  if (!packets) return fail(ctx, CV2_ERR_BAD_ARG, "missing packets");
  if (!ctx->type_by_id)
    return fail(ctx, CV2_ERR_STATE, "cv2_set_type_table first: received atoms take their type by id");
  if (ctx->s.mol) return fail(ctx, CV2_ERR_STATE, "the device exchange moves atoms outside molecules");
  if (ctx->n + nr + ctx->ng > (int64_t)kIndexMask) return fail(ctx, CV2_ERR_BAD_ARG, "too many atoms (2^26 max)");
  // received atoms by id (a stable sort: equal ids keep packet order)
  std::vector<int> rid((size_t)nr);
  std::vector<int2> src((size_t)nr);
  {
    std::vector<std::pair<int, int2>> e((size_t)nr);
    int64_t k = 0;
    for (int p = 0; p < npackets; ++p)
      for (int64_t c = 0; c < counts[p]; ++c, ++k) {
        const double v = packets[pbase[p] + c];
        if (!(v >= 0.0 && v < (double)ctx->ntype_ids && v == (double)(int64_t)v))
          return fail(ctx, CV2_ERR_BAD_ARG, "received atom id outside the type table");
        if (packets[pbase[p] + 10 * counts[p] + c] > -1.0)
          return fail(ctx, CV2_ERR_STATE, "the device exchange moves atoms outside molecules");
        e[(size_t)k] = {(int)v, make_int2(p, (int)c)};
      }
    std::stable_sort(e.begin(), e.end(),
                     [](const std::pair<int, int2> &a, const std::pair<int, int2> &b) { return a.first < b.first; });
    for (size_t i = 0; i < e.size(); ++i) {
      rid[i] = e[i].first;
      src[i] = e[i].second;
    }
  }
  (void)hipSetDevice(ctx->device);
  const int64_t ns = ctx->n, n = ns + nr;
  launch_mig_to_owned(ctx->s, ctx->stream);  // the merge reads the owned atoms in owned order
  int rc = ensure_owned_cap(ctx, n);
  if (rc) return rc;
  rc = halo_buffer(ctx, at);
  if (rc) return rc;
  // ids, sources and the packet tables in one device buffer
  const size_t b_rid = 0, b_src = (sizeof(int) * nr + 15) / 16 * 16,
               b_pb = b_src + (sizeof(int2) * nr + 15) / 16 * 16, b_pn = b_pb + 8 * 64,
               bytes = b_pn + 8 * 64;
  if (bytes > ctx->mig_aux_bytes) {
    if (ctx->mig_aux) (void)hipFree(ctx->mig_aux);
    ctx->mig_aux = nullptr;
    HIPCHK(hipMalloc(&ctx->mig_aux, bytes));
    ctx->mig_aux_bytes = bytes;
  }
  unsigned char *aux = static_cast<unsigned char *>(ctx->mig_aux);
  HIPCHK(hipMemcpyAsync(ctx->halo_buf, packets, sizeof(double) * at, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(aux + b_rid, rid.data(), sizeof(int) * nr, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(aux + b_src, src.data(), sizeof(int2) * nr, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(aux + b_pb, pbase, 8 * npackets, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(aux + b_pn, pn, 8 * npackets, hipMemcpyHostToDevice, ctx->stream));
  launch_mig_merge(ctx->s, (int)ns, (int)nr, ctx->ids_sorted, reinterpret_cast<const int *>(aux + b_rid),
                   reinterpret_cast<const int2 *>(aux + b_src), ctx->halo_buf,
                   reinterpret_cast<const long long *>(aux + b_pb),
                   reinterpret_cast<const long long *>(aux + b_pn), ctx->type_by_id, ctx->stream);
  return adopt_owned(ctx, n);
}

int cv2_get_ids(cv2_ctx *ctx, int64_t *ids) {
  if (!ctx || (ctx->n > 0 && !ids)) return fail(ctx, CV2_ERR_BAD_ARG, "null argument");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "no atoms");
  (void)hipSetDevice(ctx->device);
  const int64_t n = ctx->n;
  std::vector<int> id((size_t)n), tg((size_t)n);
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(id.data(), ctx->s.id, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(tg.data(), ctx->s.tag, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int64_t k = 0; k < n; ++k) ids[tg[k]] = id[k];
  return CV2_OK;
}

int cv2_host_alloc(int64_t bytes, void **ptr) {
  if (!ptr || bytes < 0) return CV2_ERR_BAD_ARG;
  *ptr = nullptr;
  if (bytes == 0) return CV2_OK;
  return hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault) == hipSuccess ? CV2_OK : CV2_ERR_HIP;
}

int cv2_host_free(void *ptr) {
  if (!ptr) return CV2_OK;
  return hipHostFree(ptr) == hipSuccess ? CV2_OK : CV2_ERR_HIP;
}

int cv2_build_neighbors(cv2_ctx *ctx, int force, int *rebuilt) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  int rc = ensure_ready(ctx);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  int before = 0;
  if (rebuilt) {
    HIPCHK(hipMemcpy(&before, ctx->s.flags + F_REBUILDS, sizeof(int), hipMemcpyDeviceToHost));
  }
  if (!force) launch_check_disp(ctx->s, ctx->b, ref_rebuilds(ctx), ctx->stream);
  rc = queue_build(ctx, force != 0, true);
  if (rc) return rc;
  ctx->forces_valid = ctx->energy_ok = false;
  rc = check_flags(ctx);
  // a list cut at its capacity: grow the lists and build again
  for (int tries = 0; rc == CV2_ERR_CAPACITY && ctx->h_flags[F_OVERFLOW] && tries < 8; ++tries) {
    rc = grow_lists(ctx);
    if (rc) return rc;
    build_checked(ctx, true);
    rc = check_flags(ctx);
  }
  if (rebuilt) *rebuilt = ctx->h_flags[F_REBUILDS] != before;
  return rc;
}

namespace {
// One force pass at the current positions (after the build queued before
// it); a build that halted on the list capacity is run again with grown
// lists first.
int forces_pass(cv2_ctx *ctx, bool energy) {
  for (int tries = 0;; ++tries) {
    if (ctx->precision != CV2_PREC_FP64) launch_pack_q(ctx->s, ctx->b, ctx->stream);
    const int t = span_begin(ctx, 0);
    launch_forces(ctx->s, ctx->b, ctx->lj_dev, ctx->lj, ctx->precision, energy, ctx->stream);
    span_end(ctx, t, 0);
    int rc = check_flags(ctx);
    if (rc != CV2_ERR_CAPACITY || !ctx->h_flags[F_OVERFLOW] || tries == 8) return rc;
    rc = grow_lists(ctx);
    if (rc) return rc;
    build_checked(ctx, false);
  }
}
}  // namespace

int cv2_compute_forces(cv2_ctx *ctx) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  int rc = ensure_ready(ctx);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  if (ctx->need_build) {
    rc = queue_build(ctx, true, false);
    if (rc) return rc;
  }
  rc = forces_pass(ctx, true);
  if (rc) return rc;
  rc = fetch_energy(ctx);
  if (rc) return rc;
  ctx->forces_valid = ctx->energy_ok = true;
  return CV2_OK;
}

namespace {
// Potential energy and virial at the current positions, for cv2_thermo,
// when the last force pass did not sum them: one energy pass whose
// accelerations go to scratch, so that the state keeps the bits the step
// loop computed (on the FP32 path the energy kernel tests the cutoff in
// FP64 and the step kernel on r^-2 in FP32).
int energy_pass(cv2_ctx *ctx) {
  if (!ctx->forces_valid || ctx->need_build) return cv2_compute_forces(ctx);
  if (ctx->precision != CV2_PREC_FP64) launch_pack_q(ctx->s, ctx->b, ctx->stream);
  DevState tmp = ctx->s;
  tmp.acc = ctx->s.acc_alt;
  launch_forces(tmp, ctx->b, ctx->lj_dev, ctx->lj, ctx->precision, true, ctx->stream);
  int rc = check_flags(ctx);
  if (rc) return rc;
  rc = fetch_energy(ctx);
  if (rc) return rc;
  ctx->energy_ok = true;
  return CV2_OK;
}
}  // namespace

int cv2_compute_accelerations(cv2_ctx *ctx) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  int rc = ensure_ready(ctx);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  if (ctx->need_build) {
    rc = queue_build(ctx, true, false);
    if (rc) return rc;
  }
  rc = forces_pass(ctx, false);
  if (rc) return rc;
  ctx->forces_valid = true;
  return CV2_OK;
}

// Shared step loop.  On entry the forces of the current positions are
// known; each iteration closes the previous step and opens the next one in
// one integrator pass, runs the (device-conditional) list build, and
// computes forces (no energy: cv2_thermo sums it when asked).  Where the
// FP32 task kernel runs, the integrator pass of the next iteration is fused
// into its epilogue (launch_forces_fused), so that iteration starts with its
// build.
int run_steps(cv2_ctx *ctx, int64_t nsteps, UpdateArgs u) {
  u.writeq = ctx->precision != CV2_PREC_FP64;
  u.ref_rebuild = ref_rebuilds(ctx);
  const uint64_t s0 = (uint64_t)ctx->steps;
  int64_t k0 = 0;
  bool opened = false;  // this iteration's integrator pass already ran (fused)
  launch_set_flag(ctx->s.flags, F_STEP, 0, ctx->stream);  // steps are counted from this run's first
  for (int attempt = 0;; ++attempt) {
    bool closed = false;  // the last step's finishing half ran fused into its force pass
    for (int64_t k = k0; k < nsteps; ++k) {
      if (!opened) {
        u.finish = k > 0;
        u.start = true;
        u.step = s0 + (uint64_t)k - 1;
        const int t = span_begin(ctx, 2);
        launch_update(ctx->s, ctx->b, u, ctx->stream);
        span_end(ctx, t, 2);
      }
      int t = span_begin(ctx, 1);
      {
        const int rc = queue_build(ctx, false, false, (int)k);
        if (rc) return rc;
      }
      span_end(ctx, t, 1);
      t = span_begin(ctx, 0);
      opened = false;
      if (ctx->fuse && ctx->precision != CV2_PREC_FP64) {
        UpdateArgs f = u;
        f.finish = true;
        f.start = k + 1 < nsteps;   // the last step closes only
        f.step = s0 + (uint64_t)k;  // what iteration k + 1 would close and open
        if (launch_forces_fused(ctx->s, ctx->b, ctx->lj_dev, ctx->lj, f, ctx->stream)) {
          if (f.start) {
            std::swap(ctx->s.q, ctx->s.q_alt);
            opened = true;
          } else {
            closed = true;
          }
        }
      }
      if (!opened && !closed)
        launch_forces(ctx->s, ctx->b, ctx->lj_dev, ctx->lj, ctx->precision, false, ctx->stream);
      span_end(ctx, t, 0);
      if (ctx->spans.size() > 4000) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        spans_collect(ctx);
      }
    }
    if (nsteps > 0 && !closed) {
      u.finish = true;
      u.start = false;
      u.step = s0 + (uint64_t)nsteps - 1;
      const int t = span_begin(ctx, 2);
      launch_update(ctx->s, ctx->b, u, ctx->stream);
      span_end(ctx, t, 2);
    }
    int rc = check_flags(ctx);
    if (rc == CV2_ERR_CAPACITY && ctx->h_flags[F_OVERFLOW] && attempt < 8) {
      // The build of step F_STEP cut a list at its capacity and halted the
      // run there: that step's integrator pass had run, its forces had not.
      // Grow the lists and go on from that step (its build runs again; it
      // writes the fixed-point coordinates the force kernel reads).
      k0 = ctx->h_flags[F_STEP];
      opened = true;
      rc = grow_lists(ctx);
      if (rc) return rc;
      continue;
    }
    if (rc) return rc;
    break;
  }
  ctx->steps += nsteps;
  // the potential energy and virial of the new positions are summed when
  // cv2_thermo asks for them (energy_pass)
  if (nsteps > 0) ctx->energy_ok = false;
  return CV2_OK;
}

int cv2_run_vv(cv2_ctx *ctx, int64_t nsteps, double dt) {
  if (!ctx || nsteps < 0 || dt < 0) return fail(ctx, CV2_ERR_BAD_ARG, "bad step arguments");
  if (ctx->step_open) return fail(ctx, CV2_ERR_STATE, "a step is open: cv2_step_end first");
  int rc = ensure_ready(ctx);
  if (rc) return rc;
  if (!ctx->forces_valid) {
    rc = cv2_compute_forces(ctx);
    if (rc) return rc;
  }
  UpdateArgs u{};
  u.langevin = false;
  u.dt = dt;
  return run_steps(ctx, nsteps, u);
}

int cv2_run_langevin(cv2_ctx *ctx, int64_t nsteps, double dt, double gamma, double kbt,
                     uint64_t seed) {
  if (!ctx || nsteps < 0 || dt < 0) return fail(ctx, CV2_ERR_BAD_ARG, "bad step arguments");
  if (kbt < 0.0) return fail(ctx, CV2_ERR_BAD_ARG, "kbt have to non-negative.");
  if (ctx->step_open) return fail(ctx, CV2_ERR_STATE, "a step is open: cv2_step_end first");
  int rc = ensure_ready(ctx);
  if (rc) return rc;
  if (!ctx->forces_valid) {
    rc = cv2_compute_forces(ctx);
    if (rc) return rc;
  }
  UpdateArgs u{};
  u.langevin = true;
  u.dt = dt;
  // md_simulator.cpp:398-405
  u.la = (2.0 - gamma * dt) / (2.0 + gamma * dt);
  u.lb = std::sqrt(kbt * gamma * 0.5 * dt);
  u.lc = 2.0 * dt / (2.0 + gamma * dt);
  u.seed = seed;
  return run_steps(ctx, nsteps, u);
}

int cv2_step_begin(cv2_ctx *ctx, double dt, int langevin, double gamma, double kbt, uint64_t seed) {
  if (!ctx || dt < 0) return fail(ctx, CV2_ERR_BAD_ARG, "bad step arguments");
  if (langevin && kbt < 0.0) return fail(ctx, CV2_ERR_BAD_ARG, "kbt have to non-negative.");
  if (ctx->step_open) return fail(ctx, CV2_ERR_STATE, "a step is open: cv2_step_end first");
  int rc = ensure_ready(ctx);
  if (rc) return rc;
  if (!ctx->forces_valid) {
    rc = cv2_compute_forces(ctx);
    if (rc) return rc;
  }
  (void)hipSetDevice(ctx->device);
  UpdateArgs u{};
  u.langevin = langevin != 0;
  u.dt = dt;
  if (u.langevin) {  // md_simulator.cpp:398-405
    u.la = (2.0 - gamma * dt) / (2.0 + gamma * dt);
    u.lb = std::sqrt(kbt * gamma * 0.5 * dt);
    u.lc = 2.0 * dt / (2.0 + gamma * dt);
  }
  u.seed = seed;
  u.writeq = ctx->precision != CV2_PREC_FP64;
  u.ref_rebuild = ref_rebuilds(ctx);
  // the host packs the halo from the positions this half leaves: owned
  // atoms are back in the box by then
  u.wrap_owned = ctx->host_halo;
  u.finish = false;
  u.start = true;
  u.step = (uint64_t)ctx->steps - 1;  // the start half draws the noise of step ctx->steps
  launch_set_flag(ctx->s.flags, F_STEP, 0, ctx->stream);
  launch_update(ctx->s, ctx->b, u, ctx->stream);
  ctx->open_args = u;
  ctx->step_open = true;
  ctx->forces_valid = ctx->energy_ok = false;
  // no host round trip: a NaN raised here halts the device and is reported
  // by cv2_step_end
  return CV2_OK;
}

int cv2_step_end(cv2_ctx *ctx, int energy) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  if (!ctx->step_open) return fail(ctx, CV2_ERR_STATE, "no open step: cv2_step_begin first");
  (void)hipSetDevice(ctx->device);
  int rc = queue_build(ctx, false, false);
  if (rc) return rc;
  UpdateArgs u = ctx->open_args;
  u.finish = true;
  u.start = false;
  u.step = (uint64_t)ctx->steps;
  // forces and the finishing half with one host round trip: a build cut at
  // the list capacity halts both (F_HALT); the lists then grow and the
  // build, the forces and the half run again
  for (int tries = 0;; ++tries) {
    if (ctx->precision != CV2_PREC_FP64) launch_pack_q(ctx->s, ctx->b, ctx->stream);
    const int t = span_begin(ctx, 0);
    launch_forces(ctx->s, ctx->b, ctx->lj_dev, ctx->lj, ctx->precision, energy != 0, ctx->stream);
    span_end(ctx, t, 0);
    launch_update(ctx->s, ctx->b, u, ctx->stream);
    rc = check_flags(ctx);
    if (rc != CV2_ERR_CAPACITY || !ctx->h_flags[F_OVERFLOW] || tries == 8) break;
    rc = grow_lists(ctx);
    if (rc) return rc;
    build_checked(ctx, false);
  }
  if (rc) return rc;
  ctx->step_open = false;
  ctx->steps += 1;
  if (energy) {
    rc = fetch_energy(ctx);
    if (rc) return rc;
  }
  ctx->forces_valid = true;
  ctx->energy_ok = energy != 0;
  return CV2_OK;
}

int cv2_thermo(cv2_ctx *ctx, int n_r_df, double *ke, double *pe, double *virial) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  int rc = ensure_ready(ctx);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  if (ke) {
    launch_thermo(ctx->s, ctx->lj_dev, ctx->stream);
    HIPCHK(hipMemcpyAsync(ctx->h_red, ctx->s.reduce_out, sizeof(double) * 8, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double *r = ctx->h_red + 2;  // sum m v^2, sum m v (3), sum m
    double e = r[0];
    if (n_r_df >= 1 && n_r_df <= 3 && r[4] > 0.0)  // kinetic_energy.cpp:56-70
      e -= (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]) / r[4];
    *ke = 0.5 * e;
  }
  // (while a step is open: those of the last energy pass)
  if ((pe || virial) && !ctx->step_open && !(ctx->energy_ok && ctx->forces_valid)) {
    rc = energy_pass(ctx);
    if (rc) return rc;
  }
  if (pe) *pe = ctx->pe;
  if (virial) *virial = ctx->virial;
  return CV2_OK;
}

int cv2_scale_velocities(cv2_ctx *ctx, double lambda) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "no atoms");
  if (!std::isfinite(lambda)) return fail(ctx, CV2_ERR_BAD_ARG, "velocity scale is not finite");
  (void)hipSetDevice(ctx->device);
  launch_scale_vel(ctx->s, lambda, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CV2_OK;
}

int cv2_download(cv2_ctx *ctx, double *pos, double *vel, double *acc) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "no atoms");
  (void)hipSetDevice(ctx->device);
  const int64_t n = ctx->n;
  std::vector<double4> p4(n);
  std::vector<double> v(3 * n), a(3 * n);
  std::vector<int> tg(n);
  DevState &s = ctx->s;
  HIPCHK(hipMemcpyAsync(tg.data(), s.tag, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  if (pos) HIPCHK(hipMemcpyAsync(p4.data(), s.pos, sizeof(double4) * n, hipMemcpyDeviceToHost, ctx->stream));
  if (vel) HIPCHK(hipMemcpyAsync(v.data(), s.vel, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  if (acc) HIPCHK(hipMemcpyAsync(a.data(), s.acc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // The device wraps positions into the box at list builds; between builds
  // an atom may lie up to half the skin beyond a periodic face.  The copy
  // gets the reference's single wrap (Domain::fix_position,
  // domain.cpp:443-493), which a build would apply to the same value.
  const BoxDev &b = ctx->b;
  auto wrap = [&](double c, int d) {
    if (b.wrap[d] != 1) return c;
    if (c < b.lo[d]) return c + b.L[d];
    if (c > b.hi[d]) return c - b.L[d];
    return c;
  };
  for (int64_t k = 0; k < n; ++k) {
    const int64_t o = tg[k];
    if (pos) {
      pos[3 * o] = wrap(p4[k].x, 0);
      pos[3 * o + 1] = wrap(p4[k].y, 1);
      pos[3 * o + 2] = wrap(p4[k].z, 2);
    }
    for (int d = 0; d < 3; ++d) {
      if (vel) vel[3 * o + d] = v[d * n + k];
      if (acc) acc[3 * o + d] = a[d * n + k];
    }
  }
  return CV2_OK;
}

int cv2_get_pairs(cv2_ctx *ctx, int64_t cap, int64_t *npairs, int64_t *pairs) {
  if (!ctx || !npairs || (cap > 0 && !pairs)) return fail(ctx, CV2_ERR_BAD_ARG, "bad pair buffer");
  if (!ctx->atoms) return fail(ctx, CV2_ERR_STATE, "no atoms");
  (void)hipSetDevice(ctx->device);
  const int64_t n = ctx->n;
  DevState &s = ctx->s;
  std::vector<int> cnt(n), tg(n);
  std::vector<uint32_t> nb((size_t)s.cap * s.stride);
  HIPCHK(hipMemcpyAsync(cnt.data(), s.nnbr, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(tg.data(), s.tag, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(nb.data(), s.nbr, sizeof(uint32_t) * nb.size(), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // task-slot form: rebuild the 32-bit entries (atom | image) from the task
  // piece tables; the image of j follows from its cell relative to i's cell
  // (periodic grid, at least 5 cells per axis)
  std::vector<int> cst, ctk, gtg;
  std::vector<int4> tpc;
  if (ctx->b.open && s.ng > 0) {  // ghosts in cell order: their host indices
    gtg.resize(s.ng);
    HIPCHK(hipMemcpyAsync(gtg.data(), s.gtag, sizeof(int) * s.ng, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  if (s.slots) {
    cst.resize(s.ncell + 1);
    ctk.resize(s.ncell);
    tpc.resize((size_t)s.tg.nslot * kTaskPieces);
    HIPCHK(hipMemcpyAsync(cst.data(), s.cell_start, sizeof(int) * cst.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctk.data(), s.ctask, sizeof(int) * ctk.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(tpc.data(), s.tpiece, sizeof(int4) * tpc.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  const TaskGeom &g = s.tg;
  auto cell_of = [&](int64_t a) {
    return (int)(std::upper_bound(cst.begin(), cst.end(), (int)a) - cst.begin()) - 1;
  };
  auto entry = [&](int64_t i, int m) -> uint32_t {
    if (!s.slots) return nb[((size_t)(m >> 2) * s.stride + i) * 4 + (m & 3)];
    const uint16_t *nb16 = reinterpret_cast<const uint16_t *>(nb.data());
    const int sl = nb16[((size_t)(m >> 3) * s.stride + i) * 8 + (m & 7)];
    const int ci = cell_of(i);
    const int cx = ci % g.nx, cy = (ci / g.nx) % g.ny, cz = ci / (g.nx * g.ny);
    const int4 *tp = &tpc[(size_t)ctk[ci] * kTaskPieces];
    int p = 0;
    for (int k2 = 1; k2 < kTaskPieces; ++k2) p = tp[k2].y <= sl ? k2 : p;
    const int j = tp[p].x + (sl - tp[p].y);
    if (g.open) return (uint32_t)j | (kNoImage << 26);  // no images on open grids
    const int cj = cell_of(j);
    const int d[3] = {cj % g.nx - cx, (cj / g.nx) % g.ny - cy, cj / (g.nx * g.ny) - cz};
    uint32_t img = 0;
    for (int a = 0; a < 3; ++a) img |= (uint32_t)((d[a] > 1 ? 1 : (d[a] < -1 ? -1 : 0)) + 1) << (26 + 2 * a);
    return (uint32_t)j | img;
  };
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int m = 0; m < cnt[i]; ++m) {
      const uint32_t e = entry(i, m);
      const int64_t j = e & kIndexMask;
      const uint32_t img = e >> 26;
      const int code = (int)(img & 3u) + 3 * (int)((img >> 2) & 3u) + 9 * (int)(img >> 4);
      int64_t tj;
      if (j >= n) {
        tj = gtg.empty() ? j : n + gtg[j - n];  // host-halo ghost: n + host index
      } else {
        tj = tg[j];
        if (img == kNoImage && tj < tg[i]) continue;  // half list: owned j > i
      }
      if (k < cap) {
        pairs[3 * k] = tg[i];
        pairs[3 * k + 1] = tj;
        pairs[3 * k + 2] = code;
      }
      ++k;
    }
  }
  *npairs = k;
  return CV2_OK;
}

int cv2_get_stats(cv2_ctx *ctx, cv2_stats *out) {
  if (!ctx || !out) return CV2_ERR_BAD_ARG;
  (void)hipSetDevice(ctx->device);
  HIPCHK(hipMemcpyAsync(ctx->h_flags, ctx->s.flags, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *out = cv2_stats{};
  out->natoms = ctx->n;
  out->nghost = ctx->ng;
  out->steps = ctx->steps;
  out->rebuilds = ctx->h_flags[F_REBUILDS];
  out->list_entries = (uint32_t)ctx->h_flags[F_ENTRIES];
  out->max_neighbors = ctx->h_flags[F_MAXNBR];
  out->capacity = ctx->s.cap;
  for (int d = 0; d < 3; ++d) out->ncell[d] = ctx->b.nc[d];
  out->precision = ctx->precision;
  out->build_kind = ctx->b.coarse ? CV2_BUILD_CELL_WAVE : CV2_BUILD_PER_ATOM;
  out->force_tasks = ctx->s.slots ? ctx->h_flags[F_NTASK] : 0;
  if (ctx->vstate) {
    unsigned long long v[8];
    HIPCHK(hipMemcpy(v, ctx->vstate, sizeof(v), hipMemcpyDeviceToHost));
    out->verified_builds = (int64_t)v[6];
    out->verify_mismatches = (int64_t)v[7];
    out->verify_sum[0] = v[2];
    out->verify_pairs[0] = (int64_t)v[3];
    out->verify_sum[1] = v[4];
    out->verify_pairs[1] = (int64_t)v[5];
  }
  return CV2_OK;
}

int cv2_get_build_filter(cv2_ctx *ctx, cv2_build_filter *out) {
  if (!ctx || !out) return CV2_ERR_BAD_ARG;
  if (!ctx->box_set || !ctx->nl_set) return fail(ctx, CV2_ERR_STATE, "set the box and neighbour settings first");
  const BoxDev &b = ctx->b;
  *out = cv2_build_filter{};
  for (int d = 0; d < 3; ++d) {
    out->qunit[d] = b.qinv[d];
    out->lo[d] = b.lo[d];
    out->ncell[d] = b.nc[d];
    out->cell_width[d] = b.cinv[d] > 0.0 ? 1.0 / b.cinv[d] : 0.0;
  }
  out->ce = std::sqrt(b.ce_sq);
  out->cthr_in = b.cthr_in;
  out->cthr_out = b.cthr_out;
  out->coarse = b.coarse;
  return CV2_OK;
}

int cv2_synchronize(cv2_ctx *ctx) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CV2_OK;
}

int cv2_set_timing(cv2_ctx *ctx, int enable) {
  if (!ctx) return CV2_ERR_BAD_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  spans_collect(ctx);
  if (enable < 0) return fail(ctx, CV2_ERR_BAD_ARG, "timing period must be >= 0");
  ctx->timing = enable != 0;
  if (enable) {
    ctx->timing_period = enable;
    ctx->force_seq = 0;
    ctx->t_ms[0] = ctx->t_ms[1] = ctx->t_ms[2] = 0.0;
    ctx->force_calls = 0;
  }
  return CV2_OK;
}

int cv2_get_timing(cv2_ctx *ctx, cv2_timing *out) {
  if (!ctx || !out) return CV2_ERR_BAD_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  spans_collect(ctx);
  out->force_ms = ctx->t_ms[0];
  out->force_calls = ctx->force_calls;
  out->build_ms = ctx->t_ms[1];
  out->integrate_ms = ctx->t_ms[2];
  return CV2_OK;
}

}  // extern "C"
