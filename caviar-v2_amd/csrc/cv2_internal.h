// cv2_internal.h -- shared device/host declarations of the MD-step library.
//
// HBM layout (all arrays in cell-sorted order of the owned atoms, rebuilt at
// every neighbour-list build; host-halo ghosts follow the owned atoms in
// host order):
//   pos   double4[n_owned + n_ghost]  x, y, z, type (as double)     32 B/atom
//   vel   double[3][n_owned]          SoA                            24 B/atom
//   acc   double[3][n_owned]          SoA                            24 B/atom
//   q     int4[n_owned + n_ghost]     fixed-point x, y, z, type      16 B/atom
//                                     (pack_q)
//   tag   int32[n_owned]              upload index (inverse of the sort)
//   id    int32[n_owned + n_ghost]    user id (virial/energy pair rule, noise key)
//   pos_ref double4[n_owned+n_ghost]  positions at the last build
//   nbr   uint32[cap/4][stride][4]    grouped ELL Verlet list (entry k of atom i
//                                     at ((k/4)*stride + i)*4 + k%4): entry
//                                     j | (vx+1)<<26 | (vy+1)<<28 | (vz+1)<<30,
//                                     the image of j being pos_j - v*L
//   nnbr  int32[n_owned]              list length per atom
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace cv2 {

constexpr int kMaxTypes = 8;
constexpr uint32_t kIndexMask = (1u << 26) - 1u;
constexpr uint32_t kNoImage = 0x15u;  // (1,1,1) in 2-bit fields: no shift on any axis

// Flags word layout in device memory.
enum Flag : int {
  F_REBUILD = 0,   // a rebuild is pending (set by the displacement check)
  F_OVERFLOW = 1,  // some list exceeded its capacity
  F_MAXNBR = 2,    // longest list of the last build
  F_NAN = 3,       // NaN acceleration, position or velocity seen
  F_FXOVF = 4,     // fixed-point overflow seen
  F_REBUILDS = 5,  // number of builds done on the device
  F_ENTRIES = 6,   // list entries of the last build (low 32 bits)
  F_STAGE = 7,     // a force task's neighbourhood exceeds 16-bit slots
  F_STATBLK = 8,   // blocks of k_build_stats done (reset by the last one)
  F_BINBLK = 9,    // blocks of k_wrap_bin done (reset by the last one)
  F_NTASK = 10,    // force tasks of the current build (compact task map length)
  F_TASKOVF = 11,  // some cell row needed more task slots than it has
  F_HALT = 12,     // a fatal condition (NaN, fixed-point overflow, list capacity) was
                   // raised: every later kernel of the run returns at once, so the
                   // state stays at the step that raised it
  F_STEP = 13,     // index (within the run) of the step whose kernels run now;
                   // frozen by F_HALT
  F_COUNT = 14
};

// A list build runs when one is requested and the run has not halted.
__device__ inline bool build_pending(const int *flags) {
  return flags[F_REBUILD] != 0 && flags[F_HALT] == 0;
}

// Raise a fatal flag and halt the run (Md_simulator::integrate_* ends the
// program at the step of a NaN acceleration, md_simulator.cpp:646-652; here
// the device stops advancing at that step and the host reports it).
__device__ inline void raise_fatal(int *flags, int which) {
  atomicOr(&flags[which], 1);
  atomicOr(&flags[F_HALT], 1);
}

struct BoxDev {
  double lo[3], hi[3], L[3], halfL[3];
  double glo[3];   // cell-grid origin
  double cinv[3];  // cells per unit length
  double ghost_cut;
  double ce_sq;     // (cutoff * extra_coef)^2, the list predicate
  double thresh_sq; // squared half-skin displacement that triggers a rebuild
  double qscale[3]; // fixed-point units per length (set_qscale)
  double qinv[3];   // length per fixed-point unit
  int bc[3];        // 1 = periodic (images generated on the device)
  int wrap[3];      // 1 = owned atoms are wrapped into [lo, hi) at list builds: periodic
                    // axes, also of a host-halo share that is its own neighbour
  int nc[3];
  // Coarse filter of the wave-per-cell list build (fully periodic boxes,
  // coarse = 1), on the 32-bit fixed-point coordinates q: a candidate whose
  // float squared distance from them is below cthr_in is within the list
  // cutoff for certain, one at or above cthr_out is outside it for certain;
  // the band between is decided by the FP64 predicate.
  float cunit[3];    // length per q unit (qinv as float)
  float cthr_in, cthr_out;
  int coarse;
  int ghost_ok;  // ghost cutoff >= list cutoff: every image within it exists
  int qiso;      // one fixed-point unit on all three axes
  // open = 1: the wave build over a host-halo box (coarse = 1, no periodic
  // axis): the cell grid spans the box and its ghost margin, the ghosts are
  // sorted into it at each build (ghost section of the arrays in cell order)
  // and candidates are the owned and the ghost atoms of adjacent cells
  int open;
};

struct LJDev {
  int nt;
  double eps4[kMaxTypes * kMaxTypes];   // 4 * eps
  double sig2[kMaxTypes * kMaxTypes];   // sigma^2
  double rc2[kMaxTypes * kMaxTypes];    // cutoff^2 (per pair with cutoff_list)
  double fshift[kMaxTypes * kMaxTypes]; // 12 rho_c^12 / rc^2 - 6 rho_c^6 / rc^2
  double ecut[kMaxTypes * kMaxTypes];   // U(r) = 4e[(s6^2 - s6) + eA r^2 - eB]: eA
  double ecut2[kMaxTypes * kMaxTypes];  //                                     eB
  double mass_inv[kMaxTypes];
  double mass[kMaxTypes];
  // FP32 constants of the mixed-precision kernel, forces scaled by 2^32:
  // forceCoef * 2^32 = r2i * s6 * (e6 - e12 * s6) + fsh
  float eps4f[kMaxTypes * kMaxTypes];
  float e6f[kMaxTypes * kMaxTypes];   // 6 * 4 eps * 2^32
  float e12f[kMaxTypes * kMaxTypes];  // 12 * 4 eps * 2^32
  float fshf[kMaxTypes * kMaxTypes];  // 4 eps * fshift * 2^32
  double r2ovf[kMaxTypes * kMaxTypes];  // r^2 below which |F| may exceed 2^19
  // jump_fix (lj.cpp:459-468): a listed pair closer than sqrt(jump_tol) takes
  // the force of r^2 = jump_tol along its own direction; 0 = off.  Served by
  // the species-table kernels only (MULTI), which runs that select.
  double jump_tol;
};

// Fixed-point scales of the FP32 path: forces (and accelerations) in units of
// 2^-32, energies and virials in units of 2^-24.
constexpr double kForceScale = 4294967296.0;
constexpr double kEnergyScale = 16777216.0;

// Fixed-point coordinates (FP32 pair kernel and coarse list filter).  Axis d
// is (x - lo) / u_d rounded to an integer and kept modulo 2^32.  On a
// periodic axis u_d = L_d / 2^(32 + k_d), so a box length is a multiple of
// 2^32 units and int32 differences are separations modulo L_d / 2^k_d: they
// are the minimum-image separations for every pair closer than
// L_d / 2^(k_d + 1), and k_d is the largest shift that keeps the widest
// separation a kernel forms below that (set_qscale).  The unit is then
// between span / 2^31 and span / 2^30 whatever the box size (2.6e-9 to
// 5.2e-9 for a 2.8 list cutoff), and a separation needs one integer
// subtraction and one conversion.  w = type.
__device__ inline int4 pack_q(double x, double y, double z, int type, const double lo[3],
                              const double scale[3]) {
  const double c[3] = {x, y, z};
  int q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    q[d] = (int)(unsigned)(unsigned long long)llrint((c[d] - lo[d]) * scale[d]);
  return make_int4(q[0], q[1], q[2], type);
}

// ---- Philox4x32-10 ----
struct U4 {
  uint32_t x, y, z, w;
};

__device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Three standard normals for (seed, id, step).
// Three standard normals for (seed, id, step): Box-Muller on the four
// Philox words.  F32 evaluates the transform in FP32 with the hardware log,
// square root and sine/cosine (the FP32+fixed-point form), otherwise in FP64;
// the uniforms are the same 32-bit words either way.
template <bool F32>
__device__ inline void gauss3(uint64_t seed, int id, uint64_t step, double &g0, double &g1,
                              double &g2) {
  const U4 r = philox4x32_10(U4{(uint32_t)step, (uint32_t)(step >> 32), (uint32_t)id, 0x4c414e47u},
                             (uint32_t)seed, (uint32_t)(seed >> 32));
  if (F32) {
    constexpr float k2m32f = 2.3283064365386963e-10f;
    const float u0 = ((float)r.x + 0.5f) * k2m32f, u1 = (float)r.y * k2m32f;  // u0 in (0, 1]
    const float u2 = ((float)r.z + 0.5f) * k2m32f, u3 = (float)r.w * k2m32f;
    const float ra = __builtin_sqrtf(-2.0f * __logf(u0)), rb = __builtin_sqrtf(-2.0f * __logf(u2));
    // v_sin_f32 / v_cos_f32 take the angle in turns: sin(2 pi u)
    g0 = (double)(ra * __builtin_amdgcn_cosf(u1));
    g1 = (double)(ra * __builtin_amdgcn_sinf(u1));
    g2 = (double)(rb * __builtin_amdgcn_cosf(u3));
  } else {
    constexpr double k2m32 = 2.3283064365386963e-10;
    const double u0 = ((double)r.x + 0.5) * k2m32, u1 = (double)r.y * k2m32;
    const double u2 = ((double)r.z + 0.5) * k2m32, u3 = (double)r.w * k2m32;
    const double ra = sqrt(-2.0 * log(u0)), rb = sqrt(-2.0 * log(u2));
    double sa, ca, sb, cb;
    sincospi(2.0 * u1, &sa, &ca);
    sincospi(2.0 * u3, &sb, &cb);
    g0 = ra * ca;
    g1 = ra * sa;
    g2 = rb * cb;
    (void)sb;
  }
}

struct StepParams {
  double dt;
  double la, lb, lc;  // Langevin a, b, c (md_simulator.cpp:398-405)
  uint64_t seed;
  uint64_t step;      // step index s of the finishing half; s+1 starts
  double thresh_sq;
  double lo[3], qscale[3];
  // CV2_NL_REFERENCE_REBUILDS: also rebuild when an atom leaves the box
  // (outside [wlo, whi] on a periodic axis) or crosses the ghost band of a
  // face (below blo / above bhi, i.e. within the ghost cutoff of it) since
  // the last build; open axes carry infinite bounds
  int ref_rebuild;
  double wlo[3], whi[3], blo[3], bhi[3];
  // the start half of cv2_step_begin on a host-halo box wraps owned atoms
  // that left the box back in, on the axes with wL[d] > 0 (the owned
  // exchange that precedes the ghost exchange, md_simulator.cpp:547-557)
  int wrap;
  double hi[3], wL[3];
};

// The reference's other rebuild triggers for one atom at x, built at r:
// Atom_data::exchange_owned_single_md_domain reports a wrap
// (exchange_owned_single_md_domain.cpp:48-56, domain.cpp:443-493) and
// Neighborlist::update_is_needed sees a changed ghost set
// (neighborlist.cpp:117-120,139-144) whenever an atom enters or leaves the
// ghost band of a face.
__device__ inline bool ref_trigger(const double4 &x, const double4 &r, const StepParams &p) {
  const double c[3] = {x.x, x.y, x.z}, o[3] = {r.x, r.y, r.z};
  bool t = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    t = t || c[d] < p.wlo[d] || c[d] > p.whi[d];
    t = t || (c[d] < p.blo[d]) != (o[d] < p.blo[d]) || (c[d] > p.bhi[d]) != (o[d] > p.bhi[d]);
  }
  return t;
}


// The integrator for one atom (k_update, and the epilogue of the fused FP32
// force kernel), on loaded values: FINISH applies the velocity update that
// closes step p.step (after its forces a), START the updates that open step
// p.step + 1 (velocities and positions).  F32NOISE draws the Langevin noise
// with the FP32 transform (the FP32+fixed-point form).
template <bool LANGEVIN, bool FINISH, bool START, bool F32NOISE>
__device__ inline void integrate_values(double4 &x, double v[3], int id, const double a[3],
                                        const StepParams &p) {
  // explicit FMAs: every kernel that inlines this rounds the same way
  const double hdt = 0.5 * p.dt;
  if (FINISH) {
    if (LANGEVIN) {
      double e[3];
      gauss3<F32NOISE>(p.seed, id, p.step, e[0], e[1], e[2]);
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = fma(a[d], hdt, fma(e[d], p.lb, v[d] * p.la));  // md_simulator.cpp:860
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = fma(a[d], hdt, v[d]);
    }
  }
  if (START) {
    if (LANGEVIN) {
      double e[3];
      gauss3<F32NOISE>(p.seed, id, p.step + 1, e[0], e[1], e[2]);
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = v[d] + fma(e[d], p.lb, a[d] * hdt);  // :778
      x.x = fma(v[0], p.lc, x.x);                                          // :835
      x.y = fma(v[1], p.lc, x.y);
      x.z = fma(v[2], p.lc, x.z);
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = fma(a[d], hdt, v[d]);
      x.x = fma(v[0], p.dt, x.x);
      x.y = fma(v[1], p.dt, x.y);
      x.z = fma(v[2], p.dt, x.z);
    }
  }
}

// Stores of one integrated atom: velocities; with START the positions,
// fixed-point coordinates (WRITEQ) and the rebuild check of
// Neighborlist::update_is_needed (neighborlist.cpp:100-148) against the
// build position r, raising F_REBUILD.
// EXTRA = false leaves out the owned wrap and the reference rebuild
// triggers (callers that never set them: fewer live kernel parameters).
template <bool START, bool WRITEQ, bool EXTRA = true>
__device__ inline void integrate_store(int i, int n, const double4 &x, const double v[3],
                                       const double4 &r, double4 *__restrict__ pos,
                                       double *__restrict__ vel, int4 *__restrict__ q,
                                       const StepParams &p, int *__restrict__ flags) {
  if (isnan(x.x) | isnan(x.y) | isnan(x.z) | isnan(v[0]) | isnan(v[1]) | isnan(v[2]))
    raise_fatal(flags, F_NAN);
  if (START) {
    double4 y = x;
    if (EXTRA && p.wrap) {  // same rule as the wrap at list builds (k_wrap_bin)
      double *c[3] = {&y.x, &y.y, &y.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (!(p.wL[d] > 0.0)) continue;
        if (*c[d] < p.lo[d])
          *c[d] += p.wL[d];
        else if (*c[d] > p.hi[d])
          *c[d] -= p.wL[d];
      }
    }
    pos[i] = y;
    if (WRITEQ) q[i] = pack_q(y.x, y.y, y.z, (int)y.w, p.lo, p.qscale);
    // a wrapped atom has moved by a box length: the list is rebuilt
    const double dx = y.x - r.x, dy = y.y - r.y, dz = y.z - r.z;
    const bool trig =
        dx * dx + dy * dy + dz * dz > p.thresh_sq || (EXTRA && p.ref_rebuild && ref_trigger(y, r, p));
    // one atomic per wavefront (many atoms trigger together at a melt or a
    // wrap): the lowest triggering lane
    const unsigned long long m = __ballot(trig);
    if (trig && __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u)) == 0)
      atomicOr(&flags[F_REBUILD], 1);
  }
  vel[i] = v[0];
  vel[n + i] = v[1];
  vel[2 * n + i] = v[2];
}

// Chooses the fixed-point units of b (bc and L set): span is the largest
// separation along one axis that any kernel forms from q.  Sets qscale,
// qinv, cunit and qiso.
void set_qscale(BoxDev &b, double span);

// Coarse-filter thresholds of the wave-per-cell build (cthr_in, cthr_out)
// for list cutoff ce and cell width at most wmax, from the fixed-point unit
// (set_qscale first); false when the filter cannot decide (box too large).
bool coarse_filter_setup(BoxDev &b, double ce, double wmax);

// Force tasks of the FP32 path on wave-build boxes.  A task is a run of
// consecutive cells [x0, x1) of one x-row of the cell grid (its atoms are
// contiguous in the sorted order).  Its neighbourhood -- the x-cells
// x0-1 .. x1 (periodic) of the 9 rows at dy, dz in {-1, 0, 1} -- is 1 or 2
// contiguous atom ranges ("pieces") per row, staged back to back in LDS by
// the force kernel; the list of the FP32 path stores, for each neighbour, its
// 16-bit position ("slot") in that staging area.
//
// The rows are cut into tasks at every build from the actual cell counts
// (k_scatter): greedily, each task takes cells while it holds at most
// kTaskBlock home atoms and its neighbourhood fits the LDS stage, so a
// lattice whose cells hold uneven counts gets tasks of even size.  Row r
// owns task slots [r * mpr, (r + 1) * mpr); a row that would need more takes
// the rest of the row into its last slot (and raises F_TASKOVF, so the host
// widens mpr at the next forced build).  The used slots of all rows, in row
// order, are copied to compact tables (ctinfo, ctpiece) that the force
// kernel walks.
constexpr int kTaskPieces = 18;  // 9 rows x 2 pieces; piece 2r + p of row r
struct TaskGeom {
  int nx, ny, nz;  // cell grid
  int mpr;         // task slots per x-row
  int nslot;       // ny * nz * mpr
  int open;        // BoxDev::open: rows do not wrap; piece 2r + 1 holds ghosts
};

// x-cells [a0, a1) of piece p of a task over [x0, x1): the whole row once
// when the neighbourhood covers it, else the direct range and, at a row
// end, the wrapped cell.
__host__ __device__ inline void task_piece_cells(int nx, int x0, int x1, int p, int &a0, int &a1) {
  if (x1 - x0 + 2 >= nx) {
    a0 = 0;
    a1 = p == 0 ? nx : 0;
  } else if (x0 == 0) {
    a0 = p == 0 ? 0 : nx - 1;
    a1 = p == 0 ? x1 + 1 : nx;
  } else if (x1 == nx) {
    a0 = p == 0 ? 0 : x0 - 1;
    a1 = p == 0 ? 1 : nx;
  } else {
    a0 = p == 0 ? x0 - 1 : 0;
    a1 = p == 0 ? x1 + 1 : 0;
  }
}

// Open grids: x-cells [a0, a1) of both pieces (owned, ghost) of a task
// over [x0, x1): x0 - 1 .. x1, clipped to the grid.
__host__ __device__ inline void task_piece_cells_open(int nx, int x0, int x1, int &a0, int &a1) {
  a0 = x0 > 0 ? x0 - 1 : 0;
  a1 = x1 < nx ? x1 + 1 : nx;
}

// First cell of row r (dy = r % 3 - 1, dz = r / 3 - 1, periodic) around
// row (cy, cz); on open grids -1 for a row outside the grid.
__host__ __device__ inline int task_row_base(const TaskGeom &g, int cy, int cz, int r) {
  int y = cy + r % 3 - 1, z = cz + r / 3 - 1;
  if (g.open) {
    if (y < 0 || y >= g.ny || z < 0 || z >= g.nz) return -1;
  } else {
    y += y < 0 ? g.ny : (y >= g.ny ? -g.ny : 0);
    z += z < 0 ? g.nz : (z >= g.nz ? -g.nz : 0);
  }
  return g.nx * (y + g.ny * z);
}

struct DevState {
  int n;        // owned atoms
  int ng;       // host-halo ghosts
  int stride;   // ELL stride (>= n, multiple of 64)
  int cap;      // ELL slots per atom
  double4 *pos, *pos_alt, *pos_ref;
  int4 *q;
  int4 *q_alt;                  // second fixed-point buffer (fused FP32 step)
  double *vel, *vel_alt;  // 3 * n each (SoA: x block, y block, z block)
  double *acc, *acc_alt;
  int *tag, *tag_alt, *id, *id_alt;
  // Lj::ignore_intra_molecule (lj.cpp:446): molecule index of each owned
  // atom by upload index (read through tag); null when the flag is off
  int *mol;
  BoxDev *box_dev;              // device copy of the box constants
  int2 *wstat;                  // per-block build statistics (longest list, entries)
  int64_t wstat_cap;
  int *cell, *slot, *perm, *order;  // per owned atom during a build
  int *cell_count, *cell_start; // ncell (+1)
  int *cell_local;               // tile-local cell starts during a build (ncell)
  int *tile_sum;                // per scan tile (ncell / 1024 + 1)
  int *gcell_count, *gcell_start, *gcell_items; // host-halo ghost cell list
  // open wave builds sort the ghosts: gtag[r] = host index of device ghost
  // r, gslot its inverse (where an upload of an unchanged ghost count
  // writes host ghost g); gpos_alt / gid_alt / gtag_alt: the new order
  int *gtag, *gslot, *gtag_alt, *gid_alt;
  double4 *gpos_alt;
  uint32_t *nbr;                // u32 entries, or u16 slots when `slots`
  int *nnbr;
  bool slots;                   // FP32 path on a wave-build box: task-slot list
  TaskGeom tg;
  int4 *tpiece;                 // per task slot: kTaskPieces x (first atom, slot base, a0, a1)
  int4 *tinfo;                  // per task slot: first home atom, end, staged atoms S, x0 | x1 << 16
  int *ctask;                   // per cell: its task slot
  int ctask_cap;
  int *rowtasks;                // per x-row: tasks used
  int4 *ctinfo, *ctpiece;       // the used slots' tinfo / tpiece in row order (count: flags[F_NTASK])
  int64_t task_cap;             // task slots allocated
  int task_grid;                // work-groups of a force launch (tasks beyond loop)
  int stage_limit;              // staged atoms per block held in LDS
  int *flags;
  double *partials;            // per-block reduction partials
  long long *fx_acc;           // fixed-point reductions (energy, virial)
  double *reduce_out;          // final reductions
  int ncell;
  bool write_q;                // keep q current (FP32 pair kernel in use)
};

// Ensemble of independent periodic systems sharing one set of arrays: the
// atoms of a system are contiguous, and so are its cells (global cell index
// = cell_off[s] + local index).  A system is rebuilt when its flag is set.
struct EnsDev {
  int nsys;
  const int *sys;       // per atom: system
  const int *cell_sys;  // per cell: system
  const int *cell_off;  // per system: first global cell
  const int *atom_off;  // per system: first atom (nsys + 1 entries)
  const BoxDev *box;    // per system: box, cell grid, fixed-point scales
  int *sflag;           // per system: rebuild requested
  int *srebuilds;       // per system: builds done
  // the flagged systems of a build (k_ens_begin): act[0] = their number m,
  // act[1] = their atoms, act[2 .. 2+m] = atom prefix, act[3+nsys ..] their
  // first atoms; the per-atom build kernels walk only these atoms.  From
  // act[ens_act_cells(nsys)]: the same for the flagged systems on the wave
  // build and their cells (count, total cells, cell prefix, first cells).
  int *act;
};
// first index of the cell part of EnsDev::act (4 nsys + 8 ints in all)
__host__ __device__ inline int ens_act_cells(int nsys) { return 3 + 2 * nsys; }

// One ensemble atom's velocity-Verlet pass (k_ens_update, and the epilogue of
// the fused ensemble force kernel): FINISH closes step s with a (v += a dt/2),
// START opens step s+1 (v += a dt/2, x += v dt, fixed-point coordinates into
// q_out with the atom's own system unit, displacement test raising that
// system's flag and F_REBUILD).  Both forms run this one function, so they
// give the same bits.
template <bool FINISH, bool START>
__device__ inline void ens_step_atom(int i, int n, const EnsDev &e, double4 *__restrict__ pos,
                                     const double4 *__restrict__ pos_ref, double *__restrict__ vel,
                                     const double a[3], int4 *__restrict__ q_out, double dt,
                                     double thresh_sq, int *__restrict__ flags) {
  double v[3] = {vel[i], vel[n + i], vel[2 * n + i]};
  StepParams p{};
  p.dt = dt;
  double4 x = make_double4(0.0, 0.0, 0.0, 0.0);
  if (START) x = pos[i];
  integrate_values<false, FINISH, START, true>(x, v, 0, a, p);
  if (START) {
    const int sy = e.sys[i];
    const BoxDev &b = e.box[sy];
    pos[i] = x;
    q_out[i] = pack_q(x.x, x.y, x.z, (int)x.w, b.lo, b.qscale);
    const double4 r = pos_ref[i];
    const double dx = x.x - r.x, dy = x.y - r.y, dz = x.z - r.z;
    if (dx * dx + dy * dy + dz * dz > thresh_sq) {
      atomicOr(&e.sflag[sy], 1);
      atomicOr(&flags[F_REBUILD], 1);
    }
  }
  vel[i] = v[0];
  vel[n + i] = v[1];
  vel[2 * n + i] = v[2];
}

// Host: LJ pair tables (false + message on invalid input).
bool build_lj_table(LJDev &L, int ntypes, const double *eps, const double *sigma,
                    const double *cutoff_list, double cutoff, const double *mass, std::string &err);

// ---- launchers (defined in the .hip files) ----
void launch_pack_q(const DevState &s, const BoxDev &b, hipStream_t st);
// keep_acc: carry the accelerations into the new order (false when a force
// pass follows the build before anything reads them)
// step: index of the step within the run (kept in flags[F_STEP] while the
// run has not halted)
void launch_build(DevState &s, const BoxDev &b, hipStream_t st, bool host_halo, bool keep_acc,
                  int step = 0);
void launch_forces(DevState &s, const BoxDev &b, const LJDev *lj_dev, const LJDev &lj,
                   int precision, bool energy, hipStream_t st);
// Task slots for n atoms on b's cell grid: mpr slots per x-row, at least
// min_mpr; staged-atom capacity of the force kernel's LDS and home atoms per
// task (one thread each).
TaskGeom task_geom(int n, const BoxDev &b, int min_mpr);
constexpr int kStage = 2560;
constexpr int kTaskBlock = 256;
// One pass of the integrator: finish = close step `step` (velocity update
// after its forces), start = open step `step + 1` (velocity and position
// update, fixed-point coordinates when writeq, rebuild check).
struct UpdateArgs {
  bool langevin, finish, start, writeq;
  bool ref_rebuild;  // CV2_NL_REFERENCE_REBUILDS
  bool wrap_owned;   // StepParams::wrap
  double dt, la, lb, lc;
  uint64_t seed, step;
};
void launch_update(DevState &s, const BoxDev &b, const UpdateArgs &u, hipStream_t st);
StepParams step_params(const BoxDev &b, const UpdateArgs &u);
// FP32 forces on a wave-build box (slot lists) with the integrator fused
// into the epilogue: forces of the current positions, then u's finish +
// start for every owned atom (accelerations are not stored); the
// fixed-point coordinates of the new positions go to s.q_alt, and the host
// swaps s.q and s.q_alt after the launch.  u.start = false (the last step
// of a run): the finishing velocity half only, the accelerations stored,
// positions and q unchanged (no swap).  Returns false (nothing launched)
// where the fused form does not apply.
bool launch_forces_fused(DevState &s, const BoxDev &b, const LJDev *lj_dev, const LJDev &lj,
                         const UpdateArgs &u, hipStream_t st);
// Displacement check (and, ref_rebuild, the wrap / ghost-band triggers)
// of the current positions against the build positions.
void launch_check_disp(DevState &s, const BoxDev &b, bool ref_rebuild, hipStream_t st);
void launch_thermo(DevState &s, const LJDev *lj_dev, hipStream_t st);
void launch_scale_vel(DevState &s, double lambda, hipStream_t st);
void launch_set_flag(int *flags, int which, int value, hipStream_t st);

// Ensemble: list build of the flagged systems (policy_all: a request from
// any system rebuilds all), FP32+fixed-point forces (eatom: per-atom
// fixed-point virial and energy when energy is set), velocity Verlet with
// the per-system rebuild test, per-system (ke, pe, virial).
void launch_build_ens(DevState &s, const EnsDev &e, int ncell, int block, bool policy_all,
                      bool per_atom, hipStream_t st, const int4 *q_from = nullptr);
// fuse_dt > 0: the integrator pass of the next step runs in the force
// kernel's epilogue (new coordinates into s.q_alt)
void launch_forces_ens(DevState &s, const EnsDev &e, const LJDev *lj_dev, int block, bool iso, bool energy,
                       long long *eatom, hipStream_t st, double fuse_dt = 0.0, double thresh_sq = 0.0);
void launch_update_ens(DevState &s, const EnsDev &e, bool finish, bool start, double dt,
                       double thresh_sq, hipStream_t st);
void launch_thermo_ens(const DevState &s, const EnsDev &e, const LJDev *lj_dev,
                       const long long *eatom, double *out, hipStream_t st);
void launch_fill_int(int *p, int n, int value, hipStream_t st);

// MPI halo packets (k_halo.hip): the domain's place in the process grid.
struct HaloGeomDev {
  int gi[3], np[3], bc[3];
  double lo[3], hi[3], size[3];
  double cut;
  int with_vel;
};
int halo_tiles(int n);
void launch_halo_classify(const DevState &s, const HaloGeomDev &g, int *code_by_tag,
                          int *src_by_tag, int *tile_counts, hipStream_t st);
// Device-side scan of the tile counts: tile offsets, per-direction totals
// (dir_n, res[0..26]), packet bases (dir_base, doubles) and their total
// (res[27]); per_atom = doubles per atom (5, or 8 with velocities).
// skip13: direction 13 (the owned exchange's staying atoms) holds no packet.
void launch_halo_scan(int ntiles, int per_atom, const int *tile_counts, int *tile_off, int *dir_n,
                      long long *dir_base, long long *res, hipStream_t st, bool skip13 = false);
void launch_halo_write(const DevState &s, const HaloGeomDev &g, const int *code_by_tag,
                       const int *src_by_tag, const int *tile_off, const long long *dir_base,
                       const int *dir_n, double *out, hipStream_t st);
void launch_halo_unpack(DevState &s, const double *in, int npk, const long long *pk_first,
                        const long long *pk_base, const long long *pk_n, int ng, bool remap,
                        hipStream_t st);
// cv2_set_ghosts: ghosts staged on the device, placed as k_halo_unpack does
void launch_ghost_place(DevState &s, const double4 *gpos_in, const int *gid_in, int ng, bool remap,
                        hipStream_t st);

// Independent check of each list build (k_verify.hip, CV2_NL_VERIFY): vs
// holds verify_state_words() words (zeroed at allocation).
int verify_state_words();
void launch_verify(const DevState &s, const BoxDev &b, unsigned long long *vs, hipStream_t st);
void launch_verify_arm(const int *flags, unsigned long long *vs, hipStream_t st);

// Owned-atom exchange (k_migrate.hip).  selfmask bit d: direction d leads
// back to this domain.  classify: direction codes by upload index (13 =
// staying), per-tile counts, *moved |= 1 when an atom is shifted back into
// the domain instead of leaving; write: the leavers' packets (mig_per_atom()
// doubles each) and the staying atoms, compacted and shifted, into the alt
// arrays with stride dir_n[13]; adopt: the alt arrays into the state arrays
// for n atoms in owned order; shift: the in-place shifts alone (nothing
// leaves); merge: owned (ns, stride ns) and received atoms (nr, sorted by
// id) into the alt arrays with stride ns + nr.
int mig_per_atom();
void launch_mig_classify(DevState &s, const HaloGeomDev &g, unsigned selfmask, int *code_by_tag,
                         int *src_by_tag, int *tile_counts, int *moved, hipStream_t st);
void launch_mig_write(DevState &s, const HaloGeomDev &g, unsigned selfmask, const int *code_by_tag,
                      const int *src_by_tag, const int *tile_off, const long long *dir_base,
                      const int *dir_n, double *out, hipStream_t st);
void launch_mig_shift(DevState &s, const HaloGeomDev &g, unsigned selfmask, hipStream_t st);
void launch_mig_adopt(DevState &s, int n, hipStream_t st);
// the state arrays of the s.n owned atoms put in owned order (tag = index)
void launch_mig_to_owned(DevState &s, hipStream_t st);
void launch_mig_merge(DevState &s, int ns, int nr, bool merge, const int *rid, const int2 *src,
                      const double *in, const long long *pbase, const long long *pn,
                      const int *type_by_id, hipStream_t st);

}  // namespace cv2
