// caviar2_api.cpp -- host side of the caviar2 drop-in API over the C ABI.
//
// The state objects keep the reference's interface; the MD step itself runs
// on the device through include/cv2_md.h.  Md_simulator::run() creates one
// device session (box, LJ tables, list settings, atoms), advances it in
// chunks between writer steps and copies the state back only when a writer
// or the caller needs it.
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "caviar2/caviar2.hpp"
#include "caviar2/constraint/nve.hpp"
#include "caviar2/tools/atom.hpp"
#include "caviar2/writer/atom_writer.hpp"
#include "cv2_md.h"

namespace caviar2 {

// ------------------------------------------------------------------ Log
void Log::info(const std::string &s, int level) {
  if (level <= verbosity_) std::cout << "[INF] " << s << std::endl;
}
void Log::warning(const std::string &s) { std::cout << "[WRN] " << s << std::endl; }
void Log::error_all(const std::string &s) {
  std::cout << "[ERR] " << s << std::endl;
  std::exit(1);
}
void Log::error_all(const char *file, int line, const char *func, const std::string &s) {
  std::cout << "[ERR] " << file << ":" << line << " " << func << ": " << s << std::endl;
  std::exit(1);
}

// ------------------------------------------------------------------ Domain
void Domain::generate() {
  // one process: the local box is the global box (domain.cpp:309-328)
  lower_local = lower_global + (upper_global - lower_global) * 0.0;
  upper_local = upper_global;
  update_after_domain_change();
}

void Domain::update_after_domain_change() {
  size_local = upper_local - lower_local;
  size_global = upper_global - lower_global;
  half_edge = 0.5 * (upper_global - lower_global);
}

// ------------------------------------------------------------------ Caviar2
Caviar2::Caviar2() {
  atom_data.set_caviar(this);
  domain.set_caviar(this);
  md_simulator.set_caviar(this);
  neighborlist.set_caviar(this);
  log.set_caviar(this);
}

Caviar2::~Caviar2() = default;  // plugins are not owned (caviar2.cpp:30-38)

void Caviar2::add_force_field(Force_field *x) {
  log.info("Force_field added to the list");
  forces.push_back(x);
}

// ------------------------------------------------------------------ device session
struct DeviceSession {
  cv2_ctx *ctx = nullptr;
  Caviar2 *cv = nullptr;
  bool device_current = false;  // device holds the latest state
  ~DeviceSession() {
    if (ctx) cv2_destroy(ctx);
  }
  void check(int rc, const char *what) {
    if (rc) cv->log.error_all(std::string(what) + ": " + cv2_last_error(ctx));
  }
};

namespace {

force_field::Lj *find_lj(Caviar2 *cv) {
  force_field::Lj *lj = nullptr;
  for (auto *f : cv->forces) {
    auto *l = dynamic_cast<force_field::Lj *>(f);
    if (!l) cv->log.error_all("only force_field::Lj runs on the device path");
    if (lj) cv->log.error_all("one force_field::Lj per system on the device path");
    lj = l;
  }
  if (!lj) cv->log.error_all("no force field registered");
  return lj;
}

void configure(DeviceSession &s, Precision precision, int device) {
  Caviar2 *cv = s.cv;
  if (!s.ctx) s.check(cv2_create(&s.ctx, device), "cv2_create");
  const Domain &d = cv->domain;
  const double lo[3] = {d.lower_global.x, d.lower_global.y, d.lower_global.z};
  const double hi[3] = {d.upper_global.x, d.upper_global.y, d.upper_global.z};
  const int bc[3] = {d.boundary_condition.x, d.boundary_condition.y, d.boundary_condition.z};
  s.check(cv2_set_box(s.ctx, lo, hi, bc, cv->atom_data.ghost_cutoff), "cv2_set_box");
  Atom_data &ad = cv->atom_data;
  const int nt = (int)ad.atom_type_params.mass.size();
  if (nt < 1) cv->log.error_all("no atom type masses (Atom_data::add_type_mass)");
  force_field::Lj *lj = find_lj(cv);
  std::vector<double> eps, sig, cl;
  bool has_cl = false;
  lj->tables(nt, eps, sig, cl, has_cl);
  s.check(cv2_set_lj(s.ctx, nt, eps.data(), sig.data(), has_cl ? cl.data() : nullptr, lj->cutoff,
                     ad.atom_type_params.mass.data()),
          "cv2_set_lj");
  s.check(cv2_set_jump_tol(s.ctx, lj->device_jump_tol()), "cv2_set_jump_tol");
  Neighborlist &nl = cv->neighborlist;
  s.check(cv2_set_neighbor(s.ctx, nl.cutoff, nl.cutoff_extra_coef,
                           (nl.rebuild_test ? CV2_NL_REBUILD_EVERY_STEP : 0u) |
                               (nl.reference_rebuilds ? CV2_NL_REFERENCE_REBUILDS : 0u) |
                               (nl.verify_builds ? CV2_NL_VERIFY : 0u)),
          "cv2_set_neighbor");
  s.check(cv2_set_precision(s.ctx, precision == Precision::fp64 ? CV2_PREC_FP64 : CV2_PREC_FP32_FIXED),
          "cv2_set_precision");
}

void upload(DeviceSession &s) {
  Atom_data &ad = s.cv->atom_data;
  const auto &o = ad.atom_struct_owned;
  const size_t n = o.position.size();
  std::vector<int32_t> type(n);
  std::vector<int64_t> id(n);
  for (size_t i = 0; i < n; ++i) {
    type[i] = (int32_t)o.type[i];
    id[i] = (int64_t)o.id[i];
  }
  s.check(cv2_upload_atoms(s.ctx, (int64_t)n, &o.position[0].x, &o.velocity[0].x, type.data(), id.data()),
          "cv2_upload_atoms");
  // lj.cpp:446 compares the owned molecule indices (ghosts: their owners')
  if (find_lj(s.cv)->ignores_intra_molecule()) {
    std::vector<int32_t> mol(n, -1);
    for (size_t i = 0; i < n && i < o.molecule_index.size(); ++i) mol[i] = (int32_t)o.molecule_index[i];
    s.check(cv2_set_molecules(s.ctx, (int64_t)n, mol.data(), 1), "cv2_set_molecules");
  }
  s.device_current = true;
}

void download(DeviceSession &s) {
  Atom_data &ad = s.cv->atom_data;
  auto &o = ad.atom_struct_owned;
  s.check(cv2_download(s.ctx, &o.position[0].x, &o.velocity[0].x, &o.acceleration[0].x), "cv2_download");
  double ke, pe, vir;
  s.check(cv2_thermo(s.ctx, 0, nullptr, &pe, &vir), "cv2_thermo");
  (void)ke;
  ad.virialForce = vir;
  ad.potential_energy = pe;
}

double wall_seconds() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

}  // namespace

// ------------------------------------------------------------------ Atom_data
Atom_data::Atom_data(Caviar2 *c) { set_caviar(c); }

void Atom_data::set_caviar(Caviar2 *c) {
  caviar_ = c;
  domain = &c->domain;
}

void Atom_data::add_type_mass(int t, double m) {
  if ((int)atom_type_params.mass.size() < t + 1) {
    atom_type_params.mass.resize(t + 1);
    atom_type_params.mass_inv.resize(t + 1);
  }
  atom_type_params.mass[t] = m;
  atom_type_params.mass_inv[t] = m == 0.0 ? 0.0 : 1.0 / m;
}

void Atom_data::add_type_charge(int t, double c) {
  if ((int)atom_type_params.charge.size() < t + 1) atom_type_params.charge.resize(t + 1);
  atom_type_params.charge[t] = c;
}

void Atom_data::add_type_radius(int t, double r) {
  if ((int)atom_type_params.radius.size() < t + 1) atom_type_params.radius.resize(t + 1);
  atom_type_params.radius[t] = r;
}

void Atom_data::set_position(int i, const Vector3d<double> &v) {
  if (i < 0 || i >= (int)atom_struct_owned.position.size()) return;
  atom_struct_owned.position[i] = v;
}
void Atom_data::set_velocity(int i, const Vector3d<double> &v) {
  if (i < 0 || i >= (int)atom_struct_owned.velocity.size()) return;
  atom_struct_owned.velocity[i] = v;
}
void Atom_data::set_acceleration(int i, const Vector3d<double> &v) {
  if (i < 0 || i >= (int)atom_struct_owned.acceleration.size()) return;
  atom_struct_owned.acceleration[i] = v;
}

bool Atom_data::position_inside_global_domain(const Vector3d<double> &p) const {
  const Domain &d = *domain;
  return !(p.x < d.lower_global.x || p.x >= d.upper_global.x || p.y < d.lower_global.y ||
           p.y >= d.upper_global.y || p.z < d.lower_global.z || p.z >= d.upper_global.z);
}

bool Atom_data::add_atom(tools::Atom &a) {
  const size_t id = (size_t)get_num_of_atoms_global();
  const Vector3d<double> p = a.pos_tot();
  if (!position_inside_global_domain(p)) return false;
  return add_atom(id, a.type, p, a.vel_tot());
}

bool Atom_data::add_atom(size_t id, size_t type, const Vector3d<double> &pos, const Vector3d<double> &vel) {
  atom_struct_owned.id.push_back(id);
  atom_struct_owned.type.push_back(type);
  atom_struct_owned.position.push_back(pos);
  atom_struct_owned.velocity.push_back(vel);
  atom_struct_owned.acceleration.emplace_back(0.0, 0.0, 0.0);
  atom_struct_owned.mpi_rank.push_back(0);
  atom_struct_owned.molecule_index.push_back(-1);
  return true;
}

Vector3d<double> Atom_data::owned_velocity_cm() {
  double v[3] = {0.0, 0.0, 0.0}, m_sum = 0.0;
  const auto &o = atom_struct_owned;
  for (size_t i = 0; i < o.velocity.size(); ++i) {
    const double m = atom_type_params.mass[o.type[i]];
    m_sum += m;
    v[0] += o.velocity[i].x * m;
    v[1] += o.velocity[i].y * m;
    v[2] += o.velocity[i].z * m;
  }
  return Vector3d<double>{v[0], v[1], v[2]} / m_sum;
}

void Atom_data::add_random_velocity(unsigned int seed, double amplitude) {
  std::mt19937 mt(seed);
  std::uniform_real_distribution<double> dist(-amplitude, amplitude);
  for (auto &v : atom_struct_owned.velocity) {
    v.x = dist(mt);
    v.y = dist(mt);
    v.z = dist(mt);
  }
  const Vector3d<double> vcm = owned_velocity_cm();
  for (size_t i = 0; i < atom_struct_owned.velocity.size(); ++i)
    if (atom_type_params.mass[atom_struct_owned.type[i]] != 0.0) atom_struct_owned.velocity[i] -= vcm;
}

int Atom_data::get_n_r_df() {
  if (n_r_df == -1) {
    if (stochastic_force_present) {
      n_r_df = 0;
    } else {
      const auto bc = domain->boundary_condition;
      const int s = bc.x + bc.y + bc.z;
      n_r_df = s == 0 ? 6 : (s == 3 ? 3 : 6 - s);
    }
  }
  return n_r_df;
}

double Atom_data::kinetic_energy(int t) {
  const auto &o = atom_struct_owned;
  double e = 0.0;
  const int nr = get_n_r_df();
  const Vector3d<double> vcm = (nr >= 1 && nr <= 3) ? owned_velocity_cm() : Vector3d<double>{};
  for (size_t i = 0; i < o.velocity.size(); ++i) {
    if (t > -1 && t != (int)o.type[i]) continue;
    const Vector3d<double> v = o.velocity[i] - vcm;
    e += atom_type_params.mass[o.type[i]] * (v * v);
  }
  return 0.5 * e;
}

int Atom_data::degree_of_freedoms() { return 3 * (int)atom_struct_owned.position.size() - get_n_r_df(); }

void Atom_data::finalize_temperature() {
  if (!temperature_process) return;
  if (k_b < 0) caviar_->log.error_all("k_b is not set.");
  temperature_ = 2.0 * kinetic_energy() / (k_b * degree_of_freedoms());
}

void Atom_data::finalize_pressure() {
  if (!pressure_process) return;
  const double vol = domain->volume_local();
  const double n = (double)atom_struct_owned.position.size();
  pressure_ = n * k_b * temperature_ / vol + (virialConstraint + virialForce) / (3.0 * vol);
}

void Atom_data::reset_owned_acceleration() {
  for (auto &a : atom_struct_owned.acceleration) a = Vector3d<double>{0.0, 0.0, 0.0};
}

// ------------------------------------------------------------------ Neighborlist
void Neighborlist::set_cutoff(double v) {
  if (v <= 0.0) caviar_->log.error_all("cutoff have to non-negative.");
  cutoff = v;
}
void Neighborlist::set_cutoff_extra_coef(double v) {
  if (v < 1.0) caviar_->log.error_all("cutoff_extra_coef must be larger than 1.0");
  cutoff_extra_coef = v;
}
void Neighborlist::set_dt(double v) {
  if (v < 0.0) caviar_->log.error_all("dt have to non-negative.");
  dt = v;
}

void Neighborlist::init() {
  if (cutoff <= 0.0) caviar_->log.error_all("cutoff have to non-negative.");
  cutoff_extra = cutoff * cutoff_extra_coef;
  threshold_distance_sq = 0.25 * (cutoff_extra - cutoff) * (cutoff_extra - cutoff);
}

bool Neighborlist::build(bool update_neighborlist) {
  init();
  DeviceSession *s = caviar_->md_simulator.device_session();
  int rebuilt = 0;
  s->check(cv2_build_neighbors(s->ctx, update_neighborlist ? 1 : 0, &rebuilt), "cv2_build_neighbors");
  return rebuilt != 0;
}

void Neighborlist::update_host_list() {
  DeviceSession *s = caviar_->md_simulator.device_session();
  int64_t np = 0;
  s->check(cv2_get_pairs(s->ctx, 0, &np, nullptr), "cv2_get_pairs");
  std::vector<int64_t> pairs(3 * (size_t)np);
  s->check(cv2_get_pairs(s->ctx, np, &np, pairs.data()), "cv2_get_pairs");
  Atom_data &ad = caviar_->atom_data;
  caviar_->md_simulator.sync_host_mirror();
  // ghost images in the order exchange_ghost_single_md_domain.cpp:67-215 makes them
  const Domain &d = caviar_->domain;
  const auto &o = ad.atom_struct_owned;
  auto &g = ad.atom_struct_ghost;
  g = atom_data::Atom_struct{};
  std::map<std::pair<int64_t, int>, unsigned> ghost_index;
  const double gc = ad.ghost_cutoff;
  const double lo[3] = {d.lower_global.x, d.lower_global.y, d.lower_global.z};
  const double hi[3] = {d.upper_global.x, d.upper_global.y, d.upper_global.z};
  const int bc[3] = {d.boundary_condition.x, d.boundary_condition.y, d.boundary_condition.z};
  const double w[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const size_t n = o.position.size();
  for (size_t i = 0; i < n; ++i) {
    const double p[3] = {o.position[i].x, o.position[i].y, o.position[i].z};
    int v[3];
    for (int k = 0; k < 3; ++k) {
      v[k] = p[k] < lo[k] + gc ? -1 : (p[k] > hi[k] - gc ? 1 : 0);
      v[k] *= bc[k];
    }
    const int combos[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (const auto &cmb : combos) {
      bool ok = true;
      int sv[3];
      for (int k = 0; k < 3; ++k) {
        sv[k] = cmb[k] ? v[k] : 0;
        if (cmb[k] && v[k] == 0) ok = false;
      }
      if (!ok) continue;
      const int code = (sv[0] + 1) + 3 * (sv[1] + 1) + 9 * (sv[2] + 1);
      ghost_index[{(int64_t)i, code}] = (unsigned)g.position.size();
      g.position.emplace_back(p[0] - sv[0] * w[0], p[1] - sv[1] * w[1], p[2] - sv[2] * w[2]);
      g.id.push_back(o.id[i]);
      g.type.push_back(o.type[i]);
    }
  }
  neighlist.assign(n, {});
  if (all_atom_test) {
    // Neighborlist::all_atom_test_function (neighborlist.cpp:170-200): every
    // owned j > i and every ghost, whatever the distance.  The device keeps
    // its distance list: pairs beyond the cutoff add nothing to the forces
    // (lj.cpp:456), so both give the same accelerations.
    const unsigned ng = (unsigned)g.position.size();
    for (size_t i = 0; i < n; ++i) {
      for (size_t j = i + 1; j < n; ++j) neighlist[i].push_back((unsigned)j);
      for (unsigned q = 0; q < ng; ++q) neighlist[i].push_back((unsigned)(n + q));
    }
    return;
  }
  for (int64_t k = 0; k < np; ++k) {
    const int64_t i = pairs[3 * k], j = pairs[3 * k + 1];
    const int code = (int)pairs[3 * k + 2];
    if (code == 13) {
      neighlist[i].push_back((unsigned)j);
    } else {
      auto it = ghost_index.find({j, code});
      if (it != ghost_index.end()) neighlist[i].push_back((unsigned)(n + it->second));
    }
  }
  for (auto &l : neighlist) std::sort(l.begin(), l.end());
}

// ------------------------------------------------------------------ Force_field / Lj
void Force_field::set_caviar(Caviar2 *c) {
  caviar_ = c;
  atom_data = &c->atom_data;
  domain = &c->domain;
  neighborlist = &c->neighborlist;
}

double Force_field::energy() {
  caviar_->log.error_all("Force_field::energy: not implemented for this force field");
}

namespace force_field {

Lj::Lj(Caviar2 *c) : Force_field(c) {}

void Lj::set_cutoff(double v) {
  if (v < 0.0) throw std::runtime_error("Force field cutoff must be non-negative.");
  cutoff = v;
}
void Lj::set_jump_tol(double v) {
  jump_fix = true;
  jump_tol = v;
}
void Lj::ignore_intra_molecule_flag(bool f) { ignore_intra_molecule = f; }
void Lj::enable_wca(bool f) {
  wca = f;
  if (f) cutoff_list_activated = true;  // the cutoffs are formed by verify_settings
}
void Lj::set_cutoff_list(const std::vector<std::vector<double>> &values) {
  for (auto &v : values)
    for (double x : v)
      if (x < 0.0) throw std::runtime_error("Cutoff list values must be non-negative.");
  cutoff_list = values;
  cutoff_list_activated = true;
}
void Lj::set_epsilon(const int t1, const int t2, const double value) {
  if ((int)epsilon.size() < t1 + 1) epsilon.resize(t1 + 1);
  if ((int)epsilon[t1].size() < t2 + 1) epsilon[t1].resize(t2 + 1, 0.0);
  epsilon[t1][t2] = value;
  input_by_array = true;
}
void Lj::set_sigma(const int t1, const int t2, const double value) {
  if ((int)sigma.size() < t1 + 1) sigma.resize(t1 + 1);
  if ((int)sigma[t1].size() < t2 + 1) sigma[t1].resize(t2 + 1, 0.0);
  sigma[t1][t2] = value;
  input_by_array = true;
}
void Lj::set_lambda_s(const std::vector<std::vector<double>> &v) {
  for (auto &r : v)
    for (double x : r)
      if (x < 0.0) throw std::runtime_error("Lambda_s values must be non-negative.");
  lambda_s = v;
  lambda_s_is_set = true;
}
void Lj::set_lambda_e(const std::vector<std::vector<double>> &v) {
  for (auto &r : v)
    for (double x : r)
      if (x < 0.0) throw std::runtime_error("Lambda_e values must be non-negative.");
  lambda_e = v;
  lambda_e_is_set = true;
}
void Lj::set_epsilon_atom(const int t, const double v) {
  if ((int)epsilon_atom.size() < t + 1) epsilon_atom.resize(t + 1, 0.0);
  epsilon_atom[t] = v;
  input_by_atom = true;
}
void Lj::set_sigma_atom(const int t, const double v) {
  if ((int)sigma_atom.size() < t + 1) sigma_atom.resize(t + 1, 0.0);
  sigma_atom[t] = v;
  input_by_atom = true;
}

void Lj::verify_settings() {
  if (input_by_atom && input_by_array) caviar_->log.error_all(FC_FILE_LINE_FUNC, "(input_by_atom && input_by_array)");
  // every table covers the types of the owned atoms
  size_t ntypes = 0;
  for (auto t : atom_data->atom_struct_owned.type) ntypes = std::max(ntypes, (size_t)t + 1);
  auto square = [](std::vector<std::vector<double>> &m, size_t k, double fill) {
    if (m.size() < k) m.resize(k);
    for (auto &r : m)
      if (r.size() < k) r.resize(k, fill);
  };
  size_t k = 0;
  if (input_by_atom) {  // lj.cpp:204-258
    if (epsilon_atom.size() != sigma_atom.size()) caviar_->log.warning("lj:: (epsilon_atom_size != sigma_atom_size)");
    k = std::max(epsilon_atom.size(), sigma_atom.size());
    if (k < ntypes) {
      caviar_->log.warning("lj:: (atom_max_size < atom_data_type_max + 1)");
      k = ntypes;
    }
    epsilon_atom.resize(k, 0.0);
    sigma_atom.resize(k, 0.0);
    epsilon.assign(k, std::vector<double>(k, 0.0));
    sigma.assign(k, std::vector<double>(k, 0.0));
    for (size_t i = 0; i < k; ++i)
      for (size_t j = 0; j < k; ++j) {
        // the diagonal takes the atom values as given (no mixing round-off)
        sigma[i][j] = i == j ? sigma_atom[i] : 0.5 * (sigma_atom[i] + sigma_atom[j]);
        epsilon[i][j] = i == j ? epsilon_atom[i] : std::sqrt(epsilon_atom[i] * epsilon_atom[j]);
      }
  } else if (input_by_array) {  // lj.cpp:260-309
    if (epsilon.size() != sigma.size()) caviar_->log.warning("lj:: (epsilon_size != sigma_size)");
    k = std::max(epsilon.size(), sigma.size());
    if (k < ntypes) {
      caviar_->log.warning("lj:: (es_max_size < atom_data_type_max + 1)");
      k = ntypes;
    }
    square(epsilon, k, 0.0);
    square(sigma, k, 0.0);
  }
  if (wca) {  // lj.cpp:312-328: cutoffs at the potential's minimum
    const double coef = std::pow(2.0, 1.0 / 6.0);
    cutoff_list.assign(k, std::vector<double>(k, 0.0));
    for (size_t i = 0; i < k; ++i)
      for (size_t j = 0; j < k; ++j) {
        cutoff_list[i][j] = coef * sigma[i][j];
        caviar_->log.info("lj wca cutoff " + std::to_string(i) + "," + std::to_string(j) + " : " +
                          std::to_string(cutoff_list[i][j]));
      }
  } else if (cutoff_list_activated) {  // lj.cpp:330-344
    if (cutoff_list.size() < k) caviar_->log.warning("lj:: (cutoff_list.size() < atom_max_size)");
    square(cutoff_list, k, 0.0);
  }
  // lj.cpp:347-383: scalings of sigma and epsilon, missing entries 1
  auto scale = [&](std::vector<std::vector<double>> &lam, std::vector<std::vector<double>> &m) {
    square(lam, ntypes, 1.0);
    square(m, lam.size(), 0.0);
    for (size_t i = 0; i < lam.size(); ++i)
      for (size_t j = 0; j < lam[i].size(); ++j) {
        if (m[i].size() <= j) m[i].resize(j + 1, 0.0);
        m[i][j] *= lam[i][j];
      }
  };
  if (lambda_s_is_set) scale(lambda_s, sigma);
  if (lambda_e_is_set) scale(lambda_e, epsilon);
}

void Lj::tables(int nt, std::vector<double> &eps, std::vector<double> &sig, std::vector<double> &cl,
                bool &has_cl) const {
  auto at = [nt](const std::vector<std::vector<double>> &m, int a, int b) {
    (void)nt;
    return (a < (int)m.size() && b < (int)m[a].size()) ? m[a][b] : 0.0;
  };
  eps.assign((size_t)nt * nt, 0.0);
  sig.assign((size_t)nt * nt, 0.0);
  cl.assign((size_t)nt * nt, 0.0);
  has_cl = cutoff_list_activated;
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      eps[a * nt + b] = at(epsilon, a, b);
      sig[a * nt + b] = at(sigma, a, b);
      cl[a * nt + b] = at(cutoff_list, a, b);
    }
}

// lj.cpp:393-517 on the device: adds the accelerations of the current
// positions into the host mirror and the virial into virialForce.
void Lj::calculate_acceleration() {
  DeviceSession *s = caviar_->md_simulator.device_session();
  s->check(cv2_compute_forces(s->ctx), "cv2_compute_forces");
  auto &o = atom_data->atom_struct_owned;
  std::vector<Vector3d<double>> acc(o.position.size());
  s->check(cv2_download(s->ctx, nullptr, nullptr, &acc[0].x), "cv2_download");
  for (size_t i = 0; i < acc.size(); ++i) o.acceleration[i] += acc[i];
  double pe, vir;
  s->check(cv2_thermo(s->ctx, 0, nullptr, &pe, &vir), "cv2_thermo");
  atom_data->virialForce += vir;
  atom_data->potential_energy = pe;
}

double Lj::energy() { return atom_data->potential_energy; }

}  // namespace force_field

// ------------------------------------------------------------------ Md_simulator
Md_simulator::~Md_simulator() { delete session_; }

void Md_simulator::set_caviar(Caviar2 *c) {
  caviar_ = c;
  atom_data = &c->atom_data;
}

void Md_simulator::langevin_kbt(double v) {
  langevin_param.kbt = v;
  if (v < 0.0) caviar_->log.error_all("kbt have to non-negative.");
}

void Md_simulator::set_integrator_type(Integrator_t t) {
  integrator_type = t;
  if (t == Integrator_t::Unknown) caviar_->log.error_all("invalid or non-implemented integrator");
  // Leap_frog (md_simulator.cpp:871-980) is the kick-drift-kick form
  // v += a dt/2; x += v dt; forces; v += a dt/2 -- the form the device's
  // velocity Verlet runs -- so both take cv2_run_vv.
  if (t != Integrator_t::Velocity_verlet && t != Integrator_t::Leap_frog &&
      t != Integrator_t::Velocity_verlet_langevin)
    caviar_->log.error_all("only Velocity_verlet, Leap_frog and Velocity_verlet_langevin run on the device path");
}

void Md_simulator::verify_settings() {
  if (use_time && use_step) caviar_->log.error_all("simulator md: verify_settings: cannot use both 'time' and 'step'.");
  if (!(use_time || use_step))
    caviar_->log.error_all("simulator md: verify_settings: please assign 'time' or 'step' variables.");
  if (dt < 0.0) caviar_->log.error_all("expected 'dt' input");
  for (auto *c : caviar_->constraints) {
    if (!c->device_capable())
      caviar_->log.error_all("only velocity-rescale constraints (constraint::Nve) run on the device path");
    // The reference's Langevin integrator applies the thermostat between the
    // first half-kick and the drift (md_simulator.cpp:766-775); the device
    // step has no such stop, so the pairing is refused rather than run with
    // the rescale at another point of the step.
    if (integrator_type == Integrator_t::Velocity_verlet_langevin)
      caviar_->log.error_all("constraint::Nve with Velocity_verlet_langevin is not supported on the device path");
    c->verify_settings();
  }
}

DeviceSession *Md_simulator::device_session() {
  if (!session_) {
    session_ = new DeviceSession();
    session_->cv = caviar_;
    configure(*session_, precision, device);
    upload(*session_);
  }
  return session_;
}

void Md_simulator::sync_host_mirror() {
  if (session_ && session_->device_current) download(*session_);
}

void Md_simulator::setup() {
  if (use_time) {
    initial_step = steps_of_time(initial_time, dt);
    final_step = steps_of_time(final_time, dt);
  }
  if (caviar_->forces.empty()) caviar_->log.warning("Md_simulator::setup: force_field.size() = 0");
  if (integrator_type == Integrator_t::Velocity_verlet_langevin) {
    // md_simulator.cpp:398-405
    auto &L = langevin_param;
    atom_data->stochastic_force_present = true;
    L.a = (2.0 - L.friction * dt) / (2.0 + L.friction * dt);
    if (L.kbt > 0)
      L.b = std::sqrt(L.kbt * L.friction * 0.5 * dt);
    else if (L.kb > 0 && L.temperature > 0)
      L.b = std::sqrt(L.kb * L.temperature * L.friction * 0.5 * dt);
    else
      caviar_->log.error_all("expected ('temperature'  & 'kb') or 'kbt' input");
    L.c = 2.0 * dt / (2.0 + L.friction * dt);
  }
  caviar_->neighborlist.init();
  if (session_) {  // a session created before run() may hold stale settings
    delete session_;
    session_ = nullptr;
  }
  DeviceSession *s = device_session();
  s->check(cv2_compute_forces(s->ctx), "cv2_compute_forces");
  initialized = true;
  time = dt * (double)initial_step;
  write_outputs(0, 0.0);
}

void Md_simulator::cleanup() {}

void Md_simulator::advance(int64_t nsteps) {
  if (nsteps <= 0) return;
  DeviceSession *s = device_session();
  int rc;
  if (integrator_type == Integrator_t::Velocity_verlet_langevin) {
    const double kbt = langevin_param.kbt > 0 ? langevin_param.kbt : langevin_param.kb * langevin_param.temperature;
    rc = cv2_run_langevin(s->ctx, nsteps, dt, langevin_param.friction, kbt, langevin_seed);
  } else {
    rc = cv2_run_vv(s->ctx, nsteps, dt);
  }
  if (rc == CV2_ERR_NAN)
    caviar_->log.error_all(FC_FILE_LINE_FUNC, "an atom has NaN acceleration");
  s->check(rc, "device step");
  s->device_current = true;
  if (caviar_->neighborlist.verify_builds) {
    cv2_stats st;
    s->check(cv2_get_stats(s->ctx, &st), "cv2_get_stats");
    if (st.verify_mismatches != 0)
      caviar_->log.error_all(FC_FILE_LINE_FUNC, std::to_string(st.verify_mismatches) + " of " +
                                                    std::to_string(st.verified_builds) +
                                                    " neighbour list builds differ from the all-pairs check");
  }
}

void Md_simulator::write_outputs(int64_t step_i, double t) {
  bool any = false;
  for (auto *w : caviar_->writers) any = any || w->next_output_step(step_i) == step_i;
  if (!any) return;
  sync_host_mirror();
  atom_data->finalize_temperature();
  atom_data->finalize_pressure();
  for (auto *w : caviar_->writers) w->write(step_i, t);
}

// The constraints acting at step_i, after its velocity update and before its
// writers (md_simulator.cpp:683-689,725-728).
void Md_simulator::apply_constraints(int64_t step_i) {
  for (auto *c : caviar_->constraints) {
    if (c->next_device_step(step_i) != step_i) continue;
    DeviceSession *s = device_session();
    double ke = 0.0;
    s->check(cv2_thermo(s->ctx, atom_data->get_n_r_df(), &ke, nullptr, nullptr), "cv2_thermo");
    const double lam = c->device_velocity_scale(step_i, ke, atom_data->degree_of_freedoms());
    if (lam != 1.0) s->check(cv2_scale_velocities(s->ctx, lam), "cv2_scale_velocities");
  }
}

void Md_simulator::step() {
  if (!initialized) setup();
  advance(1);
  apply_constraints(current_step);
  write_outputs(current_step, time);
  time += dt;
}

bool Md_simulator::run() {
  caviar_->log.info("MD started.");
  verify_settings();
  const double wall0 = wall_seconds();
  if (!initialized) setup();
  current_step = initial_step;
  while (current_step < final_step) {
    // run up to (and including) the next step at which a writer reports or
    // a constraint acts
    int64_t stop = final_step - 1;
    for (auto *w : caviar_->writers) {
      const int64_t e = w->next_output_step(current_step);
      if (e >= current_step && e < stop) stop = e;
    }
    for (auto *c : caviar_->constraints) {
      const int64_t e = c->next_device_step(current_step);
      if (e >= current_step && e < stop) stop = e;
    }
    const int64_t k = stop - current_step + 1;
    advance(k);
    apply_constraints(stop);
    time = dt * (double)stop;
    write_outputs(stop, time);
    current_step = stop + 1;
    time = dt * (double)current_step;
  }
  sync_host_mirror();
  atom_data->finalize_temperature();
  const double wall1 = wall_seconds();
  cleanup();
  caviar_->log.info("MD finished.");
  caviar_->log.info("md_simulation run call:Wall Time = " + std::to_string(wall1 - wall0) + " sec. , ");
  return true;
}

// ------------------------------------------------------------------ constraint::Nve
namespace constraint {

Nve::Nve(Caviar2 *caviar) : Constraint(caviar), atom_data(&caviar->atom_data) {
  constraint_type = Constraint_t::Nve;
}

void Nve::verify_settings() {  // nve.cpp:99-130
  const bool kbt_set = kbt > 0.0, temperature_set = temperature > 0.0, kb_set = kb > 0.0;
  const bool kb_temperature_set = temperature_set && kb_set;
  if (kb_temperature_set != (temperature_set || kb_set))
    caviar_->log.error_all(FC_FILE_LINE_FUNC, "kb and temperature should be set together.");
  const bool kbt_kb_temperature_set = kb_temperature_set || kbt_set;
  const bool energy_set = energy_per_dof > 0.0 || energy_tot > 0.0;
  if (!(energy_set || kbt_kb_temperature_set))
    caviar_->log.error_all(FC_FILE_LINE_FUNC, "energy or temperature is not set.");
  if (energy_set && kbt_kb_temperature_set)
    caviar_->log.error_all(FC_FILE_LINE_FUNC, "Only one of 'energy' and 'temperature' can be set.");
  if (energy_per_dof > 0.0 && energy_tot > 0.0)
    caviar_->log.error_all(FC_FILE_LINE_FUNC, "Only one of 'energy_per_dof' and 'energy_tot' can be set.");
  if (kbt_set && kb_temperature_set)
    caviar_->log.error_all(FC_FILE_LINE_FUNC, "Only one of 'kbt' and 'temperatue kb' can be set.");
  if (step <= 0) caviar_->log.error_all(FC_FILE_LINE_FUNC, "step have to non-negative.");
}

// nve.cpp:147-184.  With a zero kinetic energy the reference warns that the
// step is ignored; lambda is then 1 (the reference divides by the zero).
double Nve::lambda(double ke_now, int dof, int64_t timestep) const {
  if (ke_now == 0.0) {
    caviar_->log.warning("energy = 0. NVE thermostat step is ignored at " + std::to_string(timestep));
    return 1.0;
  }
  if (energy_per_dof > 0.0) return std::sqrt(energy_per_dof * dof / ke_now);
  if (energy_tot > 0.0) return std::sqrt(energy_tot / ke_now);
  if (kbt > 0.0) return std::sqrt(dof * kbt * 0.5 / ke_now);
  if (temperature > 0.0) return std::sqrt(dof * kb * temperature * 0.5 / ke_now);
  return 1.0;
}

void Nve::apply_thermostat(int64_t timestep, bool &recalculate_temperature) {
  if (timestep % step != 0) return;
  recalculate_temperature = true;
  const double lam = lambda(atom_data->kinetic_energy(), atom_data->degree_of_freedoms(), timestep);
  for (auto &v : atom_data->atom_struct_owned.velocity) v *= lam;
}

int64_t Nve::next_device_step(int64_t from) const {
  if (step <= 0) return -1;
  const int64_t r = from % step;
  return r == 0 ? from : from + (step - r);
}

double Nve::device_velocity_scale(int64_t timestep, double ke_now, int dof) {
  return lambda(ke_now, dof, timestep);
}

}  // namespace constraint

// ------------------------------------------------------------------ Atom_writer
namespace writer {

namespace {
int64_t next_multiple(int64_t from, int step) {
  if (step <= 0) return -1;
  const int64_t r = from % step;
  return r == 0 ? from : from + (step - r);
}
}  // namespace

Atom_writer::Atom_writer(Caviar2 *c) : Writer(c), atom_data(&c->atom_data) {}

Atom_writer::~Atom_writer() = default;

int64_t Atom_writer::next_output_step(int64_t from) const {
  int64_t best = -1;
  const std::pair<bool, int> outs[5] = {{output_xyz, xyz_step}, {output_energy, energy_step},
                                        {output_temperature, temperature_step},
                                        {output_pressure, pressure_step}, {output_pe, pe_step}};
  for (const auto &o : outs) {
    if (!o.first) continue;
    const int64_t e = next_multiple(from, o.second);
    if (e >= 0 && (best < 0 || e < best)) best = e;
  }
  return best;
}

void Atom_writer::write(int64_t i, double t) {
  if (output_energy && i % energy_step == 0) {
    if (!ofs_energy.is_open()) ofs_energy.open(file_name_energy + ".txt");
    ofs_energy << i << " " << t << " " << atom_data->kinetic_energy() << std::endl;
  }
  if (output_temperature && i % temperature_step == 0) {
    if (!ofs_temperature.is_open()) ofs_temperature.open(file_name_temperature + ".txt");
    ofs_temperature << i << " " << t << " " << atom_data->temperature() << std::endl;
  }
  if (output_pressure && i % pressure_step == 0) {
    if (!ofs_pressure.is_open()) ofs_pressure.open(file_name_pressure + ".txt");
    ofs_pressure << i << " " << t << " " << atom_data->pressure() << std::endl;
  }
  if (output_pe && i % pe_step == 0) {
    if (!ofs_pe.is_open()) ofs_pe.open(file_name_pe + ".txt");
    ofs_pe << i << " " << t << " " << atom_data->potential_energy << std::endl;
  }
  if (output_xyz && i % xyz_step == 0) {
    if (!ofs_xyz.is_open()) ofs_xyz.open(file_name_xyz + ".xyz");
    const auto &o = atom_data->atom_struct_owned;
    ofs_xyz << o.position.size() << "\nAtom\n";
    for (size_t k = 0; k < o.position.size(); ++k) {
      ofs_xyz << o.type[k];
      if (xyz_output_id) ofs_xyz << " " << o.id[k];
      ofs_xyz << " " << o.position[k].x << " " << o.position[k].y << " " << o.position[k].z;
      if (xyz_output_velocity) ofs_xyz << " " << o.velocity[k].x << " " << o.velocity[k].y << " " << o.velocity[k].z;
      if (xyz_output_acceleration)
        ofs_xyz << " " << o.acceleration[k].x << " " << o.acceleration[k].y << " " << o.acceleration[k].z;
      ofs_xyz << "\n";
    }
    ofs_xyz << std::flush;
  }
}

}  // namespace writer

}  // namespace caviar2
