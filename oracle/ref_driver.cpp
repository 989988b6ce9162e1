// ref_driver.cpp -- C-ABI harness around the compiled CAVIAR2 reference.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/
// together with the reference sources (compiled where they lie, never copied).
// It drives the reference's own classes exactly as a user program does
// (tests/t001_lj.cpp:23-99) and exposes setup / step / getters so that the
// fixture generator and the CPU tests can compare the oracle restatement with
// the reference on identical inputs.
//
// The reference's default-constructed members leave fields uninitialised
// (Domain::nprocs_*/grid_index_*/me, Atom_data::n_r_df,
// stochastic_force_present, make_ghost_velocity, record_owned_*,
// Md_simulator::initialized/use_time, the Langevin seeds); this harness sets
// every one of them before use.
#include <cstdint>
#include <cstring>
#include <vector>

#include "caviar2/caviar2.hpp"
#include "caviar2/atom_data.hpp"
#include "caviar2/domain.hpp"
#include "caviar2/force/lj.hpp"
#include "caviar2/md_simulator.hpp"
#include "caviar2/neighborlist.hpp"

using caviar2::Vector3d;

namespace {

// Read-only access to Neighborlist::pos_old (protected) through a
// pointer-to-member of a derived class.
struct NlSpy : caviar2::Neighborlist {
  static const std::vector<Vector3d<double>> &pos_old_of(const caviar2::Neighborlist &nl) {
    return nl.*(&NlSpy::pos_old);
  }
};

// Read access to Lj's private tables (epsilon, sigma, cutoff_list) through
// pointers to members named in explicit template instantiations, where
// access checking does not apply ([temp.spec.general]): the harness reads
// what Lj::verify_settings formed without changing the reference.
using Table = std::vector<std::vector<double>>;
template <typename Tag, typename Tag::type M>
struct MemberOf {
  friend typename Tag::type member(Tag) { return M; }
};
struct LjEps { using type = Table caviar2::force_field::Lj::*; friend type member(LjEps); };
struct LjSig { using type = Table caviar2::force_field::Lj::*; friend type member(LjSig); };
struct LjCut { using type = Table caviar2::force_field::Lj::*; friend type member(LjCut); };
template struct MemberOf<LjEps, &caviar2::force_field::Lj::epsilon>;
template struct MemberOf<LjSig, &caviar2::force_field::Lj::sigma>;
template struct MemberOf<LjCut, &caviar2::force_field::Lj::cutoff_list>;

struct RefSim {
  caviar2::Caviar2 *cv = nullptr;
  caviar2::force_field::Lj *lj = nullptr;
  int integrator = 0;
};

}  // namespace

extern "C" {

// integrator: 0 = Velocity_verlet, 1 = Velocity_verlet_langevin
void *ref_sim_create(const double lo[3], const double hi[3], const int bc[3], double ghost_cutoff,
                     double nl_cutoff, double nl_coef, int use_cell_list, int rebuild_test,
                     int ntypes, const double *eps, const double *sigma,
                     const double *cutoff_list, const double *mass, double lj_cutoff,
                     int integrator, double dt, double gamma, double kbt, double k_b,
                     int n_r_df, int64_t n, const double *pos, const double *vel,
                     const int32_t *type) {
  RefSim *s = new RefSim();
  s->cv = new caviar2::Caviar2();
  caviar2::Caviar2 &c = *s->cv;
  s->integrator = integrator;

  c.domain.nprocs_x = c.domain.nprocs_y = c.domain.nprocs_z = 1;
  c.domain.grid_index_x = c.domain.grid_index_y = c.domain.grid_index_z = 0;
  c.domain.me = 0;
  c.domain.nprocs = 1;
  c.domain.set_lower_global(Vector3d<double>{lo[0], lo[1], lo[2]});
  c.domain.set_upper_global(Vector3d<double>{hi[0], hi[1], hi[2]});
  c.domain.set_boundary_condition(Vector3d<int>{bc[0], bc[1], bc[2]});
  c.domain.generate();

  c.atom_data.set_record_owned_position_old(false);
  c.atom_data.set_record_owned_velocity_old(false);
  c.atom_data.set_record_owned_acceleration_old(false);
  c.atom_data.make_ghost_velocity = false;
  c.atom_data.stochastic_force_present = (integrator == 1);
  c.atom_data.set_n_r_df(n_r_df);
  c.atom_data.set_ghost_cutoff(ghost_cutoff);
  for (int64_t i = 0; i < n; ++i) {
    c.atom_data.add_atom((size_t)i, (size_t)type[i],
                         Vector3d<double>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]},
                         Vector3d<double>{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]});
  }
  for (int t = 0; t < ntypes; ++t) c.atom_data.add_type_mass(t, mass[t]);
  c.atom_data.set_k_b(k_b);
  c.atom_data.set_temperature_process(true);

  c.neighborlist.set_cutoff(nl_cutoff);
  c.neighborlist.set_cutoff_extra_coef(nl_coef);
  c.neighborlist.set_dt(dt);
  if (use_cell_list) c.neighborlist.set_make_verlet_list_from_cell_list(true);
  if (rebuild_test) c.neighborlist.set_rebuild_test(true);

  s->lj = new caviar2::force_field::Lj(&c);
  s->lj->set_cutoff(lj_cutoff);
  for (int a = 0; a < ntypes; ++a)
    for (int b = 0; b < ntypes; ++b) {
      s->lj->set_sigma(a, b, sigma[a * ntypes + b]);
      s->lj->set_epsilon(a, b, eps[a * ntypes + b]);
    }
  if (cutoff_list) {
    std::vector<std::vector<double>> cl(ntypes, std::vector<double>(ntypes));
    for (int a = 0; a < ntypes; ++a)
      for (int b = 0; b < ntypes; ++b) cl[a][b] = cutoff_list[a * ntypes + b];
    s->lj->set_cutoff_list(cl);
  }
  c.add_force_field(s->lj);

  caviar2::Md_simulator &md = c.md_simulator;
  md.initialized = false;
  md.use_time = false;
  md.current_step = 0;
  md.time = 0.0;
  md.set_integrator_type(integrator == 1 ? caviar2::Integrator_t::Velocity_verlet_langevin
                                         : caviar2::Integrator_t::Velocity_verlet);
  md.set_initial_step(0);
  md.set_final_step(1);
  md.set_dt(dt);
  md.langevin_param.friction = gamma;
  md.langevin_param.kbt = kbt;
  md.langevin_param.kb = 0.0;
  md.langevin_param.temperature = 0.0;
  md.langevin_param.rnd_tools_x.seed(1);
  md.langevin_param.rnd_tools_y.seed(2);
  md.langevin_param.rnd_tools_z.seed(3);
  return s;
}

// Atom_struct::molecule_index of every atom (add_atom sets -1,
// atom_data.cpp:625) and Lj::ignore_intra_molecule_flag (lj.cpp:93-96).
// The molecules themselves (molecule_struct_owned, read by
// set_atoms_mpi_rank, atom_data.cpp:817-838) list their atoms' ids.
void ref_sim_set_molecules(void *h, const int32_t *mol, int ignore) {
  RefSim *s = static_cast<RefSim *>(h);
  caviar2::Atom_data &ad = s->cv->atom_data;
  auto &mi = ad.atom_struct_owned.molecule_index;
  int nmol = 0;
  for (size_t i = 0; i < mi.size(); ++i) {
    mi[i] = mol[i];
    if (mol[i] + 1 > nmol) nmol = mol[i] + 1;
  }
  ad.molecule_struct_owned.resize((size_t)nmol);
  for (size_t i = 0; i < mi.size(); ++i)
    if (mol[i] > -1) ad.molecule_struct_owned[(size_t)mol[i]].atom_list.push_back((int)ad.atom_struct_owned.id[i]);
  s->lj->ignore_intra_molecule_flag(ignore != 0);
}

void ref_sim_destroy(void *h) {
  RefSim *s = static_cast<RefSim *>(h);
  delete s->lj;
  delete s->cv;
  delete s;
}

// Md_simulator::step() setup branch (md_simulator.cpp:445-467) without the
// first integration step.
void ref_sim_setup(void *h) {
  RefSim *s = static_cast<RefSim *>(h);
  caviar2::Md_simulator &md = s->cv->md_simulator;
  md.initialize();
  s->cv->atom_data.reset_virial();
  md.setup();
}

// One Md_simulator::step(); returns 1 if the Verlet list was rebuilt.
int ref_sim_step(void *h) {
  RefSim *s = static_cast<RefSim *>(h);
  s->cv->md_simulator.step();
  s->cv->md_simulator.current_step++;
  const auto &po = NlSpy::pos_old_of(s->cv->neighborlist);
  const auto &p = s->cv->atom_data.atom_struct_owned.position;
  if (po.size() != p.size()) return 1;
  return std::memcmp(po.data(), p.data(), p.size() * sizeof(Vector3d<double>)) == 0 ? 1 : 0;
}

int64_t ref_sim_natoms(void *h) {
  return (int64_t) static_cast<RefSim *>(h)->cv->atom_data.atom_struct_owned.position.size();
}

static void copy_vec(const std::vector<Vector3d<double>> &v, double *out) {
  if (out) std::memcpy(out, v.data(), v.size() * sizeof(Vector3d<double>));
}

void ref_sim_get(void *h, double *pos, double *vel, double *acc) {
  const auto &a = static_cast<RefSim *>(h)->cv->atom_data.atom_struct_owned;
  copy_vec(a.position, pos);
  copy_vec(a.velocity, vel);
  copy_vec(a.acceleration, acc);
}

void ref_sim_thermo(void *h, double *ke, double *temperature, double *virial) {
  caviar2::Atom_data &ad = static_cast<RefSim *>(h)->cv->atom_data;
  if (ke) *ke = ad.kinetic_energy();
  if (temperature) *temperature = ad.temperature();
  if (virial) *virial = ad.virialForce;
}

// Langevin noise of the last step (md_simulator.cpp:748-757), AoS 3n.
void ref_sim_eta(void *h, double *eta) {
  const auto &L = static_cast<RefSim *>(h)->cv->md_simulator.langevin_param;
  for (size_t i = 0; i < L.eta_x.size(); ++i) {
    eta[3 * i] = L.eta_x[i];
    eta[3 * i + 1] = L.eta_y[i];
    eta[3 * i + 2] = L.eta_z[i];
  }
}

int64_t ref_sim_nghost(void *h) {
  return (int64_t) static_cast<RefSim *>(h)->cv->atom_data.atom_struct_ghost.position.size();
}

void ref_sim_ghosts(void *h, double *gpos, int64_t *gid) {
  const auto &g = static_cast<RefSim *>(h)->cv->atom_data.atom_struct_ghost;
  copy_vec(g.position, gpos);
  if (gid)
    for (size_t i = 0; i < g.id.size(); ++i) gid[i] = (int64_t)g.id[i];
}

// Neighbour list as CSR; returns the number of entries (writes at most cap).
int64_t ref_sim_nlist(void *h, int64_t cap, int64_t *offs, uint32_t *nbr) {
  const auto &nl = static_cast<RefSim *>(h)->cv->neighborlist.neighlist;
  int64_t k = 0;
  for (size_t i = 0; i < nl.size(); ++i) {
    if (offs) offs[i] = k;
    for (unsigned j : nl[i]) {
      if (k < cap) nbr[k] = j;
      ++k;
    }
  }
  if (offs) offs[nl.size()] = k;
  return k;
}

// Atom_data::add_random_velocity (atom_data.cpp:713-734) on a fresh
// 1-type system: the libstdc++ stream the drop-in host API must reproduce.
void ref_random_velocity(int64_t n, unsigned seed, double amplitude, double *vel) {
  caviar2::Caviar2 c;
  c.domain.nprocs_x = c.domain.nprocs_y = c.domain.nprocs_z = 1;
  c.domain.grid_index_x = c.domain.grid_index_y = c.domain.grid_index_z = 0;
  c.domain.me = 0;
  c.domain.nprocs = 1;
  c.domain.set_lower_global(Vector3d<double>{0, 0, 0});
  c.domain.set_upper_global(Vector3d<double>{1, 1, 1});
  c.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  c.domain.generate();
  c.atom_data.set_record_owned_position_old(false);
  c.atom_data.set_record_owned_velocity_old(false);
  c.atom_data.set_record_owned_acceleration_old(false);
  for (int64_t i = 0; i < n; ++i)
    c.atom_data.add_atom((size_t)i, 0, Vector3d<double>{0.5, 0.5, 0.5}, Vector3d<double>{0, 0, 0});
  c.atom_data.add_type_mass(0, 1.0);
  c.atom_data.add_random_velocity(seed, amplitude);
  copy_vec(c.atom_data.atom_struct_owned.velocity, vel);
}

// Lj::verify_settings (lj.cpp:180-391) as a user program would call it:
// owned atoms of the given types (the tables cover them), then the setters
// in the order epsilon_atom / sigma_atom (per-atom input), set_epsilon /
// set_sigma (a, b, value triples), enable_wca, set_cutoff_list, set_lambda_s,
// set_lambda_e (ragged rows: lengths, values concatenated; none when the
// row count is 0), verify_settings.  The tables it leaves are returned as
// dense k x k (missing entries 0).  Both kinds of input together end the
// process in the reference (error_all): callers do not pass both.
static std::vector<std::vector<double>> ragged(int rows, const int32_t *len, const double *v) {
  std::vector<std::vector<double>> m((size_t)rows);
  for (int r = 0, at = 0; r < rows; ++r)
    for (int k = 0; k < len[r]; ++k) m[(size_t)r].push_back(v[at++]);
  return m;
}

void ref_lj_tables(int64_t n, const int32_t *type, int neatom, const double *eatom, int nsatom,
                   const double *satom, int ne, const int32_t *eab, const double *ev, int ns,
                   const int32_t *sab, const double *sv, int wca, int ncl, const int32_t *cl_len,
                   const double *cl_v, int nls, const int32_t *ls_len, const double *ls_v, int nle,
                   const int32_t *le_len, const double *le_v, int k, double *eps_out,
                   double *sig_out, double *cl_out) {
  caviar2::Caviar2 c;
  c.domain.nprocs_x = c.domain.nprocs_y = c.domain.nprocs_z = 1;
  c.domain.grid_index_x = c.domain.grid_index_y = c.domain.grid_index_z = 0;
  c.domain.me = 0;
  c.domain.nprocs = 1;
  c.domain.set_lower_global(Vector3d<double>{0, 0, 0});
  c.domain.set_upper_global(Vector3d<double>{10, 10, 10});
  c.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  c.domain.generate();
  c.atom_data.set_record_owned_position_old(false);
  c.atom_data.set_record_owned_velocity_old(false);
  c.atom_data.set_record_owned_acceleration_old(false);
  for (int64_t i = 0; i < n; ++i)
    c.atom_data.add_atom((size_t)i, (size_t)type[i], Vector3d<double>{0.5, 0.5, 0.5},
                         Vector3d<double>{0, 0, 0});
  caviar2::force_field::Lj lj(&c);
  for (int t = 0; t < neatom; ++t) lj.set_epsilon_atom(t, eatom[t]);
  for (int t = 0; t < nsatom; ++t) lj.set_sigma_atom(t, satom[t]);
  for (int q = 0; q < ne; ++q) lj.set_epsilon(eab[2 * q], eab[2 * q + 1], ev[q]);
  for (int q = 0; q < ns; ++q) lj.set_sigma(sab[2 * q], sab[2 * q + 1], sv[q]);
  if (wca) lj.enable_wca(true);
  if (ncl > 0) lj.set_cutoff_list(ragged(ncl, cl_len, cl_v));
  if (nls > 0) lj.set_lambda_s(ragged(nls, ls_len, ls_v));
  if (nle > 0) lj.set_lambda_e(ragged(nle, le_len, le_v));
  lj.verify_settings();
  auto dense = [k](const std::vector<std::vector<double>> &m, double *out) {
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b)
        out[a * k + b] = (a < (int)m.size() && b < (int)m[(size_t)a].size()) ? m[(size_t)a][(size_t)b] : 0.0;
  };
  dense(lj.*member(LjEps{}), eps_out);
  dense(lj.*member(LjSig{}), sig_out);
  dense(lj.*member(LjCut{}), cl_out);
}

}  // extern "C"
