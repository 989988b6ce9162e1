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
  F_NAN = 3,       // NaN acceleration seen
  F_FXOVF = 4,     // fixed-point overflow seen
  F_REBUILDS = 5,  // number of builds done on the device
  F_ENTRIES = 6,   // list entries of the last build (low 32 bits)
  F_STAGE = 7,     // a force task's neighbourhood exceeds 16-bit slots
  F_STATBLK = 8,   // blocks of k_build_stats done (reset by the last one)
  F_COUNT = 10
};

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

// Chooses the fixed-point units of b (bc and L set): span is the largest
// separation along one axis that any kernel forms from q.  Sets qscale,
// qinv, cunit and qiso.
void set_qscale(BoxDev &b, double span);

// Force tasks of the FP32 path on wave-build boxes.  A task is a run of
// consecutive cells [x0, x1) of one x-row of the cell grid (its atoms are
// contiguous in the sorted order).  Its neighbourhood -- the x-cells
// x0-1 .. x1 (periodic) of the 9 rows at dy, dz in {-1, 0, 1} -- is 1 or 2
// contiguous atom ranges ("pieces") per row, staged back to back in LDS by
// the force kernel; the list of the FP32 path stores, for each neighbour, its
// 16-bit position ("slot") in that staging area.
constexpr int kTaskPieces = 18;  // 9 rows x 2 pieces; piece 2r + p of row r
struct TaskGeom {
  int nx, ny, nz;  // cell grid
  int tpr;         // tasks per x-row
  int ntask;       // ny * nz * tpr
};

__host__ __device__ inline int task_x0(const TaskGeom &g, int t) {
  return t * g.nx / g.tpr;  // nx, tpr <= 1024
}

// Task (within its row) of x-cell cx.
__host__ __device__ inline int task_of_x(const TaskGeom &g, int cx) {
  int t = cx * g.tpr / g.nx;
  while (t + 1 < g.tpr && task_x0(g, t + 1) <= cx) ++t;
  return t;
}

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

// First cell of row r (dy = r % 3 - 1, dz = r / 3 - 1, periodic) around
// row (cy, cz).
__host__ __device__ inline int task_row_base(const TaskGeom &g, int cy, int cz, int r) {
  int y = cy + r % 3 - 1, z = cz + r / 3 - 1;
  y += y < 0 ? g.ny : (y >= g.ny ? -g.ny : 0);
  z += z < 0 ? g.nz : (z >= g.nz ? -g.nz : 0);
  return g.nx * (y + g.ny * z);
}

struct DevState {
  int n;        // owned atoms
  int ng;       // host-halo ghosts
  int stride;   // ELL stride (>= n, multiple of 64)
  int cap;      // ELL slots per atom
  double4 *pos, *pos_alt, *pos_ref;
  int4 *q;
  double *vel, *vel_alt;  // 3 * n each (SoA: x block, y block, z block)
  double *acc, *acc_alt;
  int *tag, *tag_alt, *id, *id_alt;
  BoxDev *box_dev;              // device copy of the box constants
  int2 *wstat;                  // per-block build statistics (longest list, entries)
  int64_t wstat_cap;
  int *cell, *slot, *perm, *order;  // per owned atom during a build
  int *cell_count, *cell_start; // ncell (+1)
  int *tile_sum;                // per scan tile (ncell / 1024 + 1)
  int *gcell_count, *gcell_start, *gcell_items; // host-halo ghost cell list
  uint32_t *nbr;                // u32 entries, or u16 slots when `slots`
  int *nnbr;
  bool slots;                   // FP32 path on a wave-build box: task-slot list
  TaskGeom tg;
  int4 *tpiece;                 // per task: kTaskPieces x (first atom, slot base, x0, x1)
  int *tstage;                  // per task: staged atoms
  int64_t task_cap;
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
};

// Host: LJ pair tables (false + message on invalid input).
bool build_lj_table(LJDev &L, int ntypes, const double *eps, const double *sigma,
                    const double *cutoff_list, double cutoff, const double *mass, std::string &err);

// ---- launchers (defined in the .hip files) ----
void launch_pack_q(const DevState &s, const BoxDev &b, hipStream_t st);
// keep_acc: carry the accelerations into the new order (false when a force
// pass follows the build before anything reads them)
void launch_build(DevState &s, const BoxDev &b, hipStream_t st, bool host_halo, bool keep_acc);
void launch_forces(DevState &s, const BoxDev &b, const LJDev *lj_dev, const LJDev &lj,
                   int precision, bool energy, hipStream_t st);
// Task geometry for n atoms on b's cell grid (same result for the build and
// the force kernel); staged-atom capacity of the force kernel's LDS.
TaskGeom task_geom(int n, const BoxDev &b);
constexpr int kStage = 2560;
// One pass of the integrator: finish = close step `step` (velocity update
// after its forces), start = open step `step + 1` (velocity and position
// update, fixed-point coordinates when writeq, rebuild check).
struct UpdateArgs {
  bool langevin, finish, start, writeq;
  double dt, la, lb, lc;
  uint64_t seed, step;
};
void launch_update(DevState &s, const BoxDev &b, const UpdateArgs &u, hipStream_t st);
void launch_check_disp(DevState &s, const BoxDev &b, hipStream_t st);
void launch_thermo(DevState &s, const LJDev *lj_dev, hipStream_t st);
void launch_set_flag(int *flags, int which, int value, hipStream_t st);

// Ensemble: list build of the flagged systems (policy_all: a request from
// any system rebuilds all), FP32+fixed-point forces (eatom: per-atom
// fixed-point virial and energy when energy is set), velocity Verlet with
// the per-system rebuild test, per-system (ke, pe, virial).
void launch_build_ens(DevState &s, const EnsDev &e, int ncell, int block, bool policy_all,
                      hipStream_t st);
void launch_forces_ens(DevState &s, const EnsDev &e, const LJDev *lj_dev, int block, bool iso, bool energy,
                       long long *eatom, hipStream_t st);
void launch_update_ens(DevState &s, const EnsDev &e, bool finish, bool start, double dt,
                       double thresh_sq, hipStream_t st);
void launch_thermo_ens(const DevState &s, const EnsDev &e, const LJDev *lj_dev,
                       const long long *eatom, double *out, hipStream_t st);
void launch_fill_int(int *p, int n, int value, hipStream_t st);

}  // namespace cv2
