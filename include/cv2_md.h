/*
 * cv2_md.h -- C ABI of the MI355X-native CAVIAR2 molecular-dynamics step.
 *
 * One opaque context holds one periodic box (or one domain's share plus a
 * host-supplied ghost halo) resident in HBM: cell-sorted atom state,
 * cell list, Verlet list and thermodynamic accumulators.  Every call returns
 * CV2_OK (0) or an error code; cv2_last_error() gives the message.  Host
 * arrays passed in are copied, never retained.  Calls on one context are
 * ordered on its own HIP stream and are not thread-safe; separate contexts
 * are independent.
 *
 * The reference exposes this path only as C++ classes (no C ABI); each entry
 * point names the reference interface it stands in for.
 */
#ifndef CV2_MD_H
#define CV2_MD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define CV2_OK 0
#define CV2_ERR_HIP 1            /* HIP runtime failure */
#define CV2_ERR_BAD_ARG 2        /* invalid argument (reference: std::runtime_error / error_all) */
#define CV2_ERR_NAN 3            /* NaN acceleration (md_simulator.cpp:646-652) */
#define CV2_ERR_FIXED_OVERFLOW 4 /* fixed-point force/energy out of range */
#define CV2_ERR_CAPACITY 5       /* neighbour-list capacity exceeded */
#define CV2_ERR_STATE 6          /* call out of order (e.g. run before upload) */

/* force precision forms */
#define CV2_PREC_FP64 0          /* FP64 pair arithmetic, FP64 accumulation */
#define CV2_PREC_FP32_FIXED 1    /* FP32 pair arithmetic, 64-bit fixed-point accumulation */

/* neighbour-list flags (cv2_set_neighbor) */
#define CV2_NL_REBUILD_EVERY_STEP 1u /* Neighborlist::set_rebuild_test (neighborlist.cpp:107-110) */
/* Rebuild also where the reference's single-domain step does besides the
 * half-skin test: when an owned atom leaves the box through a periodic face
 * (exchange_owned_single_md_domain.cpp:48-56 -> md_simulator.cpp:425-431)
 * and when the set of single-domain ghosts changes because an atom entered
 * or left the ghost band of a face (count or order change,
 * neighborlist.cpp:117-120,139-144).  Off by default: on large boxes some
 * atom crosses a face on almost every step, so the reference rebuilds on
 * almost every step; the pair set within the cutoff does not depend on it. */
#define CV2_NL_REFERENCE_REBUILDS 2u
/* Check every list build: right after each build the device recomputes the
 * pair set from scratch -- every ordered pair of owned atoms at the minimum
 * periodic image, kept where the reference's single-domain ghost of that
 * image exists, the FP64 predicate of neighborlist.cpp:261-266, no cells,
 * coarse filter or fixed point -- and compares a 64-bit checksum and the
 * count of its pairs with those of the built list (cv2_stats
 * verified_builds, verify_mismatches).  O(N^2) per build: for tests.  The
 * list is kept in its 32-bit form while it is on; host-supplied halos are
 * refused. */
#define CV2_NL_VERIFY 4u

typedef struct cv2_ctx cv2_ctx;

typedef struct {
  int64_t natoms, nghost;   /* owned atoms; host-halo ghosts */
  int64_t steps;            /* integration steps taken */
  int64_t rebuilds;         /* neighbour-list builds */
  int64_t list_entries;     /* entries in the current list (full list) */
  int32_t max_neighbors;    /* longest list of the current build */
  int32_t capacity;         /* list slots per atom */
  int32_t ncell[3];         /* cell grid */
  int32_t precision;
  int32_t build_kind;       /* CV2_BUILD_CELL_WAVE or CV2_BUILD_PER_ATOM */
  int32_t force_tasks;      /* force tasks of the current build (FP32 form, wave-build boxes), else 0 */
  int64_t verified_builds;  /* CV2_NL_VERIFY: builds checked */
  int64_t verify_mismatches; /* CV2_NL_VERIFY: checked builds whose pair set differed */
  /* CV2_NL_VERIFY, the last checked build: [0] the built list, [1] the
   * recomputation -- sum over ordered pairs (i, j, image) of
   * splitmix64((id_i << 38) ^ (id_j << 6) ^ image), image = (vx + 1) |
   * (vy + 1) << 2 | (vz + 1) << 4 for j's position x_j - v L, and the count
   * of pairs */
  uint64_t verify_sum[2];
  int64_t verify_pairs[2];
} cv2_stats;

/* list build in use (cv2_stats.build_kind): one wavefront per cell with a
 * coarse distance filter (fully periodic boxes with >= 5 cells of width
 * >= cutoff*coef per axis), or one thread per atom otherwise */
#define CV2_BUILD_PER_ATOM 0
#define CV2_BUILD_CELL_WAVE 1

/* Context lifetime.  Stands in for Caviar2 (caviar2.hpp:39-66), which owns
 * Atom_data, Domain, Neighborlist and Md_simulator by value. */
int cv2_create(cv2_ctx **ctx, int device);
int cv2_destroy(cv2_ctx *ctx);
const char *cv2_last_error(const cv2_ctx *ctx);

/* Box and ghost cutoff.  Domain::set_lower_global/set_upper_global/
 * set_boundary_condition + generate (domain.hpp:38-42, domain.cpp:295-328)
 * and Atom_data::set_ghost_cutoff (atom_data.cpp:302-305).  Periodic axes
 * need hi-lo > 2*ghost_cutoff. */
int cv2_set_box(cv2_ctx *ctx, const double lo[3], const double hi[3], const int bc[3],
                double ghost_cutoff);

/* Lennard-Jones parameters.  force_field::Lj::set_cutoff / set_epsilon /
 * set_sigma / set_cutoff_list (lj.hpp:67-79, lj.cpp:74-140) and
 * Atom_data::add_type_mass (atom_data.cpp:368-380).  eps, sigma and
 * cutoff_list (nullable) are ntypes*ntypes row-major [t_i][t_j]; mass has
 * ntypes entries. */
int cv2_set_lj(cv2_ctx *ctx, int ntypes, const double *eps, const double *sigma,
               const double *cutoff_list, double cutoff, const double *mass);

/* jump_fix.  force_field::Lj::set_jump_tol (lj.hpp; applied at
 * lj.cpp:459-468): a listed pair closer than sqrt(jump_tol) takes the force,
 * virial and energy of r^2 = jump_tol along its own direction, which caps the
 * force of overlapping atoms (a relaxation from random positions).
 * jump_tol <= 0 turns it off.  Kept across cv2_set_lj. */
int cv2_set_jump_tol(cv2_ctx *ctx, double jump_tol);

/* Verlet-list parameters.  Neighborlist::set_cutoff /
 * set_cutoff_extra_coef / set_rebuild_test (neighborlist.hpp:46-54,
 * neighborlist.cpp:202-234).  The list holds pairs closer than
 * cutoff*extra_coef; it is rebuilt when an atom has moved more than half
 * the skin since the last build (neighborlist.cpp:100-148, :318). */
int cv2_set_neighbor(cv2_ctx *ctx, double cutoff, double extra_coef, unsigned flags);

/* CV2_PREC_FP64 or CV2_PREC_FP32_FIXED. */
int cv2_set_precision(cv2_ctx *ctx, int precision);

/* Owned atoms.  Atom_data::add_atom(id, type, pos, vel) (atom_data.cpp:603-637)
 * for n atoms at once; pos and vel are AoS xyz (Vector3d<double>); id may be
 * NULL (id = index, as add_unique_atom_molecule.cpp:31 assigns). */
int cv2_upload_atoms(cv2_ctx *ctx, int64_t n, const double *pos, const double *vel,
                     const int32_t *type, const int64_t *id);

/* Molecule indices and force_field::Lj::ignore_intra_molecule_flag
 * (lj.hpp; applied at lj.cpp:446).  molecule_index holds n = owned-atom
 * entries in upload order (Atom_struct::molecule_index, -1 outside
 * molecules, atom_data.cpp:625; NULL = all -1).  With the flag on, a listed
 * pair whose two molecule indices are equal adds no force, virial or
 * energy; as at lj.cpp:446 this includes two atoms that are both outside
 * molecules (-1 == -1).  Periodic images carry their owner's index (the
 * single-domain ghost copy, exchange_ghost_single_md_domain.cpp:54-60).
 * The pair set is unchanged.  The flag runs on the species-table gather
 * kernels (32-bit lists); a host-halo share refuses it, because its
 * ghosts' indices do not travel in the halo packets.  Cleared by
 * cv2_upload_atoms; flag 0 turns it off. */
int cv2_set_molecules(cv2_ctx *ctx, int64_t n, const int32_t *molecule_index,
                      int ignore_intra_molecule);

/* Host-supplied ghost halo for one domain's share (the images that
 * exchange_ghost_mpi_shared_atoms.cpp:274-311 would deliver).  Switches the
 * context to host-halo mode: the box is then treated as open and only these
 * ghosts provide neighbours across its faces; owned atoms on the axes set
 * periodic in cv2_set_box are still wrapped into the box at list builds (a
 * share that is its own neighbour, as in a 1x1x1 process grid).  Call again
 * each step with the refreshed halo (between cv2_step_begin and
 * cv2_step_end); a changed count forces a rebuild (neighborlist.cpp:118-121). */
int cv2_set_ghosts(cv2_ctx *ctx, int64_t ng, const double *pos, const int32_t *type,
                   const int64_t *id);

/* MPI halo packets: the device side of Atom_data::exchange_ghost_mpi_shared_atoms
 * (exchange_ghost_mpi_shared_atoms.cpp), so that a domain's share exchanges its
 * halo in the reference's own packet format without staging the owned atoms
 * on the host.  The domain's place in the process grid, as Domain holds it
 * (grid_index_*, nprocs_*, lower_local / upper_local, size_global,
 * boundary_condition) plus the ghost cutoff and make_ghost_velocity. */
typedef struct {
  int32_t grid_index[3];
  int32_t nprocs[3];
  double lo_local[3], hi_local[3];
  double size_global[3];
  int32_t bc[3];
  double ghost_cutoff;
  int32_t with_velocity; /* 8 doubles per atom (with velocities), else 5 */
} cv2_halo_geom;

/* The packing half (:113-325): every owned atom within ghost_cutoff of a face
 * is sent to each of the up to seven directions (x, y, z, xy, xz, yz, xyz) its
 * face flags span; direction d = i*9 + j*3 + k with i = x_val + 1, etc.
 * counts[27] receives the atoms per direction (counts[13] = 0).  The packets
 * are written to out concatenated in direction order, each
 * [N ids][N types][3N positions][3N velocities] as doubles, atoms in upload
 * order, positions shifted by size_global where they cross a periodic grid
 * edge.  *need receives the doubles required; out is written only when cap
 * suffices (call with cap 0 to size it). */
int cv2_halo_pack(cv2_ctx *ctx, const cv2_halo_geom *g, int64_t counts[27], int64_t cap,
                  double *out, int64_t *need);

/* cv2_halo_pack without device memory allocation: the same counts and
 * packets (bit for bit), from the same kernels, with every device buffer the
 * packing needs -- per-atom direction codes, per-tile counts and offsets,
 * per-direction totals and bases, the packet staging area -- taken from one
 * static per-device scratch array of the library (about 25 MiB, held by each
 * GPU the library runs on until the process ends) instead of context
 * buffers allocated on first use; the tile offsets are scanned on the device.
 * No counterpart in the reference: this is for callers that cannot call
 * hipMalloc, such as a per-step exchange loop that must not allocate or code
 * around HIP graph capture.  Limits: at most CV2_NOALLOC_MAX_ATOMS owned
 * atoms and CV2_NOALLOC_MAX_DOUBLES packet doubles (CV2_ERR_BAD_ARG beyond).
 * Every call on one GPU uses that GPU's one array: calls that overlap on one
 * GPU (two contexts on different host threads) race; no state carries over
 * from one call to the next. */
#define CV2_NOALLOC_MAX_ATOMS 1048576
#define CV2_NOALLOC_MAX_DOUBLES 2097152
int cv2_halo_pack_noalloc(cv2_ctx *ctx, const cv2_halo_geom *g, int64_t counts[27], int64_t cap,
                          double *out, int64_t *need);

/* The unpacking half (:405-470): the ghosts become the atoms of npackets
 * received packets, appended in the order given (the reference appends the
 * received packets in direction order, then the domain's own periodic
 * images), packet p holding counts[p] atoms in the layout above.  Equivalent
 * to cv2_set_ghosts with the packets' positions, types and ids; ghost
 * velocities are not used by the force path. */
int cv2_halo_unpack(cv2_ctx *ctx, int32_t npackets, const int64_t *counts, const double *packets,
                    int32_t with_velocity);

/* cv2_halo_unpack for packets that lie anywhere in one host buffer, packet p
 * at packets + offsets[p] (doubles), appended in the order given: the
 * receive buffers of an MPI exchange posted at offsets of one allocation
 * (the reference keeps one recv_data buffer per direction,
 * exchange_ghost_mpi_shared_atoms.cpp:405-470), forwarded without a host
 * copy.  The buffer is copied to the device once, up to the end of the
 * farthest packet. */
int cv2_halo_unpack_at(cv2_ctx *ctx, int32_t npackets, const int64_t *counts,
                       const int64_t *offsets, const double *packets, int32_t with_velocity);

/* Owned-atom exchange: the device side of
 * Atom_data::exchange_owned_mpi_shared_atoms (exchange_owned_mpi_shared_atoms.cpp:44-602),
 * called where the reference calls it, between the drift and the ghost
 * exchange of a step (Md_simulator::boundary_condition, md_simulator.cpp:547-557),
 * with the geometry of cv2_halo_pack (ghost_cutoff and with_velocity unused).
 *
 * cv2_owned_pack: every owned atom strictly outside [lo_local, hi_local]
 * gets the direction d = i*9 + j*3 + k of its three components (-1 below,
 * +1 above, zeroed at a non-periodic grid edge, :136-190).  Where direction
 * d leads back to this domain (one process on every axis it moves along)
 * the atom stays and is shifted by the local box length (:571-583);
 * otherwise it leaves this domain in direction d's packet:
 * [N ids][3N positions][3N velocities][3N accelerations][N molecule
 * indices][N bond counts] as doubles (:330-369, none of the optional
 * blocks), atoms in owned order, positions shifted by size_global across a
 * periodic grid edge (:402-431).  counts[27] receives the atoms per
 * direction (counts[13] = 0) and *need the doubles of the packets,
 * concatenated in direction order.  The state changes only when the
 * packets fit in cap (out written): the leavers go, the others keep their
 * owned order, the host-supplied ghosts are dropped (the ghost exchange
 * that follows brings the new halo) and the list is rebuilt at the next
 * force pass.  A call with cap 0 sizes the packets and changes nothing.
 * Atoms in molecules, which the reference moves with their molecule's
 * centre (:208-300), are refused (cv2_set_molecules). */
int cv2_owned_pack(cv2_ctx *ctx, const cv2_halo_geom *g, int64_t counts[27], int64_t cap,
                   double *out, int64_t *need);

/* The receiving half (:534-569): the atoms of npackets received packets
 * (packet p with counts[p] atoms, back to back in the layout above) join
 * the owned atoms with their positions, velocities and accelerations; each
 * takes its type from the table of cv2_set_type_table by its id, as every
 * rank of the reference holds every atom's type.  The owned order stays
 * ascending by id when it was (atoms uploaded in ascending id order, the
 * reference's index order when ids are the indices,
 * add_unique_atom_molecule.cpp:31); otherwise the arrivals follow the
 * owned atoms in ascending id order.  The list is rebuilt at the next force
 * pass. */
int cv2_owned_unpack(cv2_ctx *ctx, int32_t npackets, const int64_t *counts, const double *packets);

/* Type of every atom id (nid entries): Atom_struct::type of the reference's
 * shared atoms, which every rank holds; read by cv2_owned_unpack. */
int cv2_set_type_table(cv2_ctx *ctx, int64_t nid, const int32_t *type_by_id);

/* Ids of the owned atoms in owned order (the order of cv2_download), which
 * owned exchanges change; ids[n]. */
int cv2_get_ids(cv2_ctx *ctx, int64_t *ids);

/* Page-locked host memory for packet buffers (hipHostMalloc): device copies
 * from and to it run at the link's full rate.  No reference counterpart
 * (the reference's buffers are std::vector).  *ptr = NULL for bytes = 0. */
int cv2_host_alloc(int64_t bytes, void **ptr);
int cv2_host_free(void *ptr);

/* Neighbour-list build.  Neighborlist::build(bool) (neighborlist.cpp:77-98):
 * wraps positions into the box (exchange_owned_single_md_domain.cpp:29-57),
 * re-sorts atoms by cell and rebuilds the list when force != 0 or when
 * update_is_needed() holds.  *rebuilt (nullable) reports whether it did. */
int cv2_build_neighbors(cv2_ctx *ctx, int force, int *rebuilt);

/* Accelerations, virial and potential energy at the current positions.
 * Atom_data::reset_owned_acceleration + force_field::Lj::calculate_acceleration
 * (atom_data.cpp:424-432, lj.cpp:393-517). */
int cv2_compute_forces(cv2_ctx *ctx);

/* Accelerations only, through the same force kernel the step loop runs
 * (for CV2_PREC_FP32_FIXED on wave-build boxes: the branchless pair
 * arithmetic with the cutoff tested on r^-2 in FP32); cv2_thermo sums the
 * potential energy and virial when asked.  Same reference counterpart as
 * cv2_compute_forces (lj.cpp:393-517 without the virial sum of
 * :487-490,516). */
int cv2_compute_accelerations(cv2_ctx *ctx);

/* nsteps of velocity Verlet.  Md_simulator::integrate_velocity_verlet
 * (md_simulator.cpp:617-729) including re_calculate_acc (:422-437).  The
 * state must be consistent (forces computed) on entry; it is on exit.  No
 * force pass of the run sums the potential energy: cv2_thermo does, at the
 * positions the run ends on, when asked. */
int cv2_run_vv(cv2_ctx *ctx, int64_t nsteps, double dt);

/* nsteps of the Langevin integrator of Md_simulator::
 * integrate_velocity_verlet_langevin (md_simulator.cpp:731-869) with
 * a, b, c of md_simulator.cpp:398-405.  The Gaussian noise is Philox4x32-10
 * keyed by (seed, atom id, step) instead of the reference's mt19937 streams. */
int cv2_run_langevin(cv2_ctx *ctx, int64_t nsteps, double dt, double gamma, double kbt,
                     uint64_t seed);

/* One step in two halves, for callers that exchange the halo of a domain's
 * share within the step, as Md_simulator does in boundary_condition between
 * the drift and the forces (md_simulator.cpp:422-437; the MPI exchange of
 * exchange_ghost_mpi_shared_atoms.cpp): cv2_step_begin applies the first
 * half of integrate_velocity_verlet (langevin = 0, md_simulator.cpp:654) or
 * integrate_velocity_verlet_langevin (langevin = 1, :778,835; a, b, c of
 * :398-405), velocities and positions, and the rebuild test; the caller may
 * then replace the ghosts (cv2_set_ghosts, cv2_halo_unpack) or pack its own
 * halo (cv2_halo_pack); cv2_step_end builds the list when needed, computes
 * the forces at the new positions (re_calculate_acc, :422-437) and applies
 * the second velocity half (:678 / :860).  Together they give the bits of
 * cv2_run_vv / cv2_run_langevin with nsteps = 1.  cv2_run_* refuse to run
 * while a step is open.  energy != 0: the force pass also sums the potential
 * energy and virial that cv2_thermo reports; energy = 0: the force pass of
 * the steps of a run, and cv2_thermo sums them when asked.  A NaN raised by
 * cv2_step_begin's half is reported by cv2_step_end. */
int cv2_step_begin(cv2_ctx *ctx, double dt, int langevin, double gamma, double kbt, uint64_t seed);
int cv2_step_end(cv2_ctx *ctx, int energy);

/* Thermodynamics of the current state: kinetic energy (kinetic_energy.cpp:
 * 26-70, n_r_df as set), potential energy and virial (Atom_data::
 * virialForce, lj.cpp:487-490,516) of the current positions: when the last
 * force pass did not sum them (a run, cv2_compute_accelerations,
 * cv2_step_end with energy = 0), an energy pass does, without changing the
 * state's accelerations; while a step is open, those of the last energy
 * pass.  Any pointer may be NULL. */
int cv2_thermo(cv2_ctx *ctx, int n_r_df, double *ke, double *pe, double *virial);

/* Multiply every owned velocity by lambda: the velocity rescale of the NVE
 * constraint (constraint::Nve::apply_thermostat, nve.cpp:186-189), whose
 * lambda (nve.cpp:147-184) the host forms from cv2_thermo's kinetic energy.
 * Acts on the state between runs (velocities of the last completed step). */
int cv2_scale_velocities(cv2_ctx *ctx, double lambda);

/* Copy state back in upload order (AoS xyz; any pointer may be NULL).
 * Positions are wrapped into [lo, hi] on periodic axes with the single
 * image shift of Domain::fix_position (domain.cpp:443-493). */
int cv2_download(cv2_ctx *ctx, double *pos, double *vel, double *acc);

/* The current Verlet list as pairs in the reference's half-list convention
 * (Neighborlist::neighlist, neighborlist.hpp:102: owned j > i, plus every
 * ghost image in range).  Each pair is three int64: upload index of i, upload
 * index of j (or of the ghost's owner / host ghost index + n in host-halo
 * mode), image code (vx+1)+3(vy+1)+9(vz+1) with image = pos_j - v*L.
 * *npairs receives the total; at most cap pairs are written. */
int cv2_get_pairs(cv2_ctx *ctx, int64_t cap, int64_t *npairs, int64_t *pairs);

int cv2_get_stats(cv2_ctx *ctx, cv2_stats *out);

/* The fixed-point grid and the coarse distance filter of the wave-per-cell
 * list build, as the library set them up for the current box (for checks of
 * the filter's soundness against the FP64 predicate).  A coordinate is
 * round((x - lo) / qunit) kept modulo 2^32 (a separation is the int32
 * difference of two, converted to float and multiplied by (float)qunit);
 * a candidate whose float squared distance is below cthr_in is inside the
 * list cutoff ce for certain, at or above cthr_out outside it for certain,
 * and in between the FP64 predicate decides.  coarse = 0: the per-atom
 * build (no coarse filter) is in use.  No counterpart in the reference,
 * whose list build has no coarse filter. */
typedef struct {
  double qunit[3];
  double lo[3];
  double ce;
  double cell_width[3];
  float cthr_in, cthr_out;
  int32_t coarse;
  int32_t ncell[3];
} cv2_build_filter;
int cv2_get_build_filter(cv2_ctx *ctx, cv2_build_filter *out);

/* Kernel timing with HIP events recorded on the context's stream around each
 * kernel group (no reference counterpart; the reference only logs the wall
 * time of Md_simulator::run, md_simulator.cpp:325-353). */
typedef struct {
  double force_ms;      /* summed duration of the pair-force kernel launches */
  int64_t force_calls;
  double build_ms;      /* rebuild-check + build sequence (also when it exits early) */
  double integrate_ms;  /* integrator kernels */
} cv2_timing;

/* enable != 0 starts (and zeroes) timing; 0 stops it.  enable = 1 times
 * every kernel group; enable = k > 1 times only the force launch of every
 * k-th force call (build_ms and integrate_ms stay 0), so that the events
 * barely disturb the run they measure. */
int cv2_set_timing(cv2_ctx *ctx, int enable);
int cv2_get_timing(cv2_ctx *ctx, cv2_timing *out);

/* Wait for all work queued on the context's stream. */
int cv2_synchronize(cv2_ctx *ctx);

/* ---- Batched ensemble ----------------------------------------------------
 * Many independent periodic single-species LJ systems of uneven size,
 * advanced together: one set of device arrays, each step one captured HIP
 * graph (update, list build of the systems that need one, FP32+fixed-point
 * forces).  The reference has no counterpart: it runs one system per
 * Caviar2 object; an ensemble stands in for nsys independent
 * Md_simulator::run calls (md_simulator.cpp:317-356) with velocity Verlet,
 * the same force law, list rule and rebuild test as the single-box context. */
typedef struct cv2_ens cv2_ens;

typedef struct {
  int32_t force_block;   /* threads per block of the force kernel: 64, 128 or 256 */
  int32_t build_block;   /* threads per block of the list kernel: 64, 128 or 256 */
  int32_t rebuild_all;   /* 0: rebuild only the systems that ask; 1: rebuild all when any asks */
  double ms_per_step;    /* cv2_ens_tune: measured time per step of the chosen setting */
  double cell_scale;     /* cell width in list cutoffs, 1 to 2: each system's cells are at least
                            this wide where that leaves two per axis, else one list cutoff */
} cv2_ens_tuning;

typedef struct {
  int64_t natoms;        /* all systems */
  int32_t nsys;
  int32_t capacity;      /* list slots per atom */
  int64_t steps;         /* steps taken */
  int64_t build_steps;   /* steps on which any system was rebuilt */
  int64_t system_builds; /* sum over systems of their builds */
  int64_t list_entries;  /* entries of the current lists */
  int32_t max_neighbors; /* longest list built so far */
  int32_t graph_nodes;   /* nodes of the captured step graph (0 before the first run) */
} cv2_ens_stats;

int cv2_ens_create(cv2_ens **e, int device);
int cv2_ens_destroy(cv2_ens *e);
const char *cv2_ens_last_error(const cv2_ens *e);
/* One species: eps, sigma, LJ cutoff, mass (lj.cpp:475-484 force law). */
int cv2_ens_set_lj(cv2_ens *e, double eps, double sigma, double cutoff, double mass);
/* List cutoff and skin coefficient, as cv2_set_neighbor. */
int cv2_ens_set_neighbor(cv2_ens *e, double cutoff, double extra_coef);
/* nsys periodic boxes [lo, hi) (3 doubles each), natoms[s] atoms each;
 * pos and vel are AoS xyz of all systems in system order.  Each box must be
 * wider than twice the list cutoff. */
int cv2_ens_upload(cv2_ens *e, int32_t nsys, const int64_t *natoms, const double *lo,
                   const double *hi, const double *pos, const double *vel);
/* Set the launch shapes and rebuild policy (defaults 256, 128, 0). */
int cv2_ens_set_tuning(cv2_ens *e, const cv2_ens_tuning *t);
/* Autotuner: times trial_steps velocity-Verlet steps (after a warm-up) for
 * every force block x build block x rebuild policy, from the same uploaded
 * state each time, keeps the fastest and restores the state. */
int cv2_ens_tune(cv2_ens *e, double dt, int32_t trial_steps, cv2_ens_tuning *best);
/* nsteps of velocity Verlet for every system. */
int cv2_ens_run_vv(cv2_ens *e, int64_t nsteps, double dt);
/* Rebuild every system's list now and recompute forces (as
 * Neighborlist::build(true), neighborlist.cpp:77-98, per system). */
int cv2_ens_build_neighbors(cv2_ens *e);
/* Per-system kinetic energy (raw), potential energy and virial of the
 * current state (arrays of nsys; any may be NULL). */
int cv2_ens_thermo(cv2_ens *e, double *ke, double *pe, double *virial);
/* State in upload order (AoS xyz of all systems; any pointer may be NULL). */
int cv2_ens_download(cv2_ens *e, double *pos, double *vel, double *acc);
/* Pairs of system sys, as cv2_get_pairs (indices local to the system). */
int cv2_ens_get_pairs(cv2_ens *e, int32_t sys, int64_t cap, int64_t *npairs, int64_t *pairs);
int cv2_ens_get_stats(cv2_ens *e, cv2_ens_stats *out);
/* enable != 0 (default): steps 2..n replay the captured step graph;
 * 0: the same launches are issued one by one (tests compare the two). */
int cv2_ens_use_graph(cv2_ens *e, int enable);

#ifdef __cplusplus
}
#endif
#endif
