/*
 * cv2_oracle.h -- CPU restatement of the CAVIAR2 molecular-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle for the HIP
 * kernels: plain C99, FP64, one thread, written to reproduce the reference's
 * floating-point operation order so that its results are bit-identical to the
 * compiled reference (checked against committed fixtures).  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and
 * only as a checker or as the timed CPU baseline; the product path never
 * calls it.
 *
 * Every function names the reference file:line it restates.  Arrays of
 * 3-vectors are AoS (x0,y0,z0,x1,...) like the reference's Vector3d<double>
 * vectors.  Ghost image codes are (vx+1) + 3*(vy+1) + 9*(vz+1), where the
 * image of owner p is p - v*L (13 = no shift).
 */
#ifndef CV2_ORACLE_H
#define CV2_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double lo[3], hi[3]; /* global box [lower_global, upper_global] */
  int bc[3];           /* 1 = periodic, 0 = open */
} orc_box;

typedef struct {
  int ntypes;
  const double *eps;         /* ntypes*ntypes, row-major [ti][tj] */
  const double *sigma;       /* ntypes*ntypes */
  const double *cutoff_list; /* ntypes*ntypes or NULL (cutoff_list inactive) */
  double cutoff;             /* Lj::cutoff */
  const double *mass_inv;    /* ntypes */
  double jump_tol;           /* <= 0: jump_fix disabled */
  int ignore_intra_molecule; /* Lj::ignore_intra_molecule (lj.cpp:446) */
  const int32_t *mol;        /* owned molecule indices (-1 outside molecules); used with the flag */
  const int32_t *gmol;       /* ghost molecule indices (copies of the owners', */
                             /* exchange_ghost_single_md_domain.cpp:54-60) */
} orc_lj_params;

/* domain.cpp:443-493 via exchange_owned_single_md_domain.cpp:29-57.
 * Returns 1 if any coordinate was wrapped.  msd (AoS int32, may be NULL)
 * accumulates the image crossings. */
int orc_wrap(int64_t n, double *pos, const orc_box *b, int32_t *msd);

/* exchange_ghost_single_md_domain.cpp:24-216.  Emits ghosts in the
 * reference order; returns the total count (writes at most cap). */
int64_t orc_ghosts(int64_t n, const double *pos, const orc_box *b, double ghost_cutoff,
                   int64_t cap, double *gpos, int64_t *owner, int32_t *code);

/* Halo packets of the MPI ghost exchange, exchange_ghost_mpi_shared_atoms.cpp:
 * direction of each owned atom in index order (:113-203: below lower_local +
 * g gives -1, above upper_local - g gives +1, zeroed at a non-periodic grid
 * edge; an atom goes to each of the x, y, z, xy, xz, yz, xyz directions its
 * nonzero components span), packet layout (:222-241: N ids, N types, 3N
 * positions, 3N velocities when with_vel) and the periodic shift of
 * positions sent across the global box (:266-305).  Direction d = i*9 + j*3
 * + k with i = x_val + 1 etc.; counts[27] receives the atoms per direction
 * and out the packets concatenated in d order.  Returns the doubles needed
 * (writes at most cap). */
int64_t orc_halo_pack(int64_t n, const double *pos, const double *vel, const int64_t *id,
                      const int32_t *type, const int32_t grid_index[3], const int32_t nprocs[3],
                      const double lo_local[3], const double hi_local[3],
                      const double size_global[3], const int32_t bc[3], double ghost_cutoff,
                      int32_t with_vel, int64_t counts[27], int64_t cap, double *out);

/* The owned-atom exchange of exchange_owned_mpi_shared_atoms.cpp for one
 * domain's owned atoms in index order, all outside molecules: an atom
 * strictly below lo_local on an axis gets -1 there, strictly above hi_local
 * +1, zeroed at a non-periodic grid edge (:131-190); direction d = i*9 +
 * j*3 + k.  Where d leads back to the domain (one process on every axis d
 * moves along: me == all[i][j][k]) the atom stays and its position is
 * shifted by the local box length (:571-583); otherwise it leaves in packet
 * d: [N ids][3N positions][3N velocities][3N accelerations][N molecule
 * indices = -1][N bond counts = 0] (:308-369), positions shifted by
 * size_global across a periodic grid edge (:402-431), packets concatenated
 * in d order.  counts[27] (counts[13] = 0); stay[i] (nullable) = 1 for
 * staying atoms and pos_after (nullable, 3n) their positions after the
 * shift (the leavers' unshifted).  Returns the doubles needed (writes out,
 * stay and pos_after only when cap suffices). */
int64_t orc_owned_pack(int64_t n, const double *pos, const double *vel, const double *acc,
                       const int64_t *id, const int32_t grid_index[3], const int32_t nprocs[3],
                       const double lo_local[3], const double hi_local[3],
                       const double size_global[3], const int32_t bc[3], int64_t counts[27],
                       int64_t cap, double *out, int32_t *stay, double *pos_after);

/* neighborlist.cpp:247-297 (O(N^2) default).  CSR output: offs[n+1],
 * nbr[...] with ghost g stored as n+g.  Returns the total number of entries
 * (writes at most cap). */
int64_t orc_nlist_brute(int64_t n, const double *pos, int64_t ng, const double *gpos,
                        double cutoff, double extra_coef, int64_t cap, int64_t *offs,
                        uint32_t *nbr);

/* neighborlist.cpp:303-455,496-672 (cell-list path; unsorted order). */
int64_t orc_nlist_cell(int64_t n, const double *pos, int64_t ng, const double *gpos,
                       const orc_box *b, double cutoff, double extra_coef, int64_t cap,
                       int64_t *offs, uint32_t *nbr);

/* lj.cpp:393-517 with potential energy added (U consistent with the
 * reference force law, U(rc)=0).  acc, virial and pe are accumulated (+=). */
void orc_lj(int64_t n, const double *pos, const int32_t *type, const int64_t *id,
            int64_t ng, const double *gpos, const int32_t *gtype, const int64_t *gid,
            const int64_t *offs, const uint32_t *nbr, const orc_lj_params *p, double *acc,
            double *virial, double *pe);

/* Pair potential of the reference force law (lj.cpp:475-484 integrated). */
double orc_lj_pair_energy(double r2, double eps, double sigma, double rc2);

/* kinetic_energy.cpp:26-55 (n_r_df 0) and :56-70 with
 * owned_velocity_cm() :409-453 (n_r_df 1..3). */
double orc_kinetic_energy(int64_t n, const double *vel, const int32_t *type,
                          const double *mass, int n_r_df);

/* ---- whole-step driver: md_simulator.cpp:445-611,617-869 ---- */
typedef struct orc_sim orc_sim;

typedef struct {
  orc_box box;
  double ghost_cutoff;
  double nl_cutoff, nl_coef;
  int use_cell_list;         /* set_make_verlet_list_from_cell_list */
  int rebuild_every_step;    /* Neighborlist::rebuild_test */
  int ntypes;
  const double *eps, *sigma, *cutoff_list, *mass;
  double lj_cutoff;
  int integrator;            /* 0 = velocity Verlet, 1 = Langevin (Kroger) */
  double dt;
  double langevin_gamma, langevin_kbt;
  double k_b;
  int n_r_df;
} orc_sim_params;

orc_sim *orc_sim_create(const orc_sim_params *p, int64_t n, const double *pos,
                        const double *vel, const int32_t *type, const int64_t *id);
void orc_sim_destroy(orc_sim *s);
/* Atom_struct::molecule_index of every atom (NULL: all -1, atom_data.cpp:625)
 * and Lj::ignore_intra_molecule_flag (lj.cpp:93-96, applied at :446). */
void orc_sim_set_molecules(orc_sim *s, const int32_t *mol, int ignore);
/* Md_simulator::setup(), md_simulator.cpp:565-611. */
void orc_sim_setup(orc_sim *s);
/* One step; eta = 3n Gaussian numbers (AoS) for Langevin, ignored for VV.
 * Returns 1 if the neighbour list was rebuilt during the step. */
int orc_sim_step(orc_sim *s, const double *eta);
void orc_sim_get(const orc_sim *s, double *pos, double *vel, double *acc);
int64_t orc_sim_nghost(const orc_sim *s);
int64_t orc_sim_npairs(const orc_sim *s);
void orc_sim_thermo(const orc_sim *s, double *ke, double *temperature, double *virial,
                    double *pe);
/* Nve::apply_thermostat, nve.cpp:147-184: the factor that takes the kinetic
 * energy ke_now of dof degrees of freedom to the target set by exactly one of
 * energy_per_dof, energy_tot, kbt or kb*temperature (the others <= 0).
 * With ke_now = 0 the reference warns that the step is ignored (:147-150)
 * and then divides by the zero; 1 here, the step ignored as the warning says. */
double orc_nve_lambda(double ke_now, int dof, double energy_per_dof, double energy_tot,
                      double kbt, double kb, double temperature);
/* nve.cpp:186-189: every velocity times lambda (kinetic energy refreshed). */
void orc_sim_scale_velocities(orc_sim *s, double lambda);

#ifdef __cplusplus
}
#endif
#endif
