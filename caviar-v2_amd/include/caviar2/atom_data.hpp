// caviar2/atom_data.hpp -- particle state (reference inc/caviar2/atom_data.hpp,
// atom_data/utility/atom_struct.hpp:28-100, atom_type_params.hpp:26-53).
//
// The vectors below are the host mirror of the device state.  While an
// Md_simulator run is in progress the device copy is authoritative; the
// mirror is refreshed at writer steps and when the run ends.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

#include "caviar2/vector3d.hpp"

namespace caviar2 {

class Caviar2;
class Domain;
namespace tools {
class Atom;
}

namespace atom_data {
struct Atom_struct {
  std::vector<Vector3d<double>> position, velocity, acceleration;
  std::vector<size_t> id, type;
  std::vector<int> mpi_rank;
  std::vector<int> molecule_index;
};
struct Atom_type_params {
  std::vector<double> mass, mass_inv, charge, radius;
};
}  // namespace atom_data

class Atom_data {
 public:
  Atom_data() = default;
  explicit Atom_data(Caviar2 *c);
  void set_caviar(Caviar2 *c);

  // settings (atom_data.cpp:302-345)
  void set_ghost_cutoff(double v) { ghost_cutoff = v; }
  void set_neighborlist_cutoff(double v) { neighborlist_cutoff = v; }
  void set_cutoff_extra(double v) { cutoff_extra = v; }
  void set_k_b(double v) { k_b = v; }
  void set_msd_process(int) {}
  void set_debug_virial(int) {}
  void set_pressure_process(int f) { pressure_process = f != 0; }
  void set_temperature_process(int f) { temperature_process = f != 0; }
  void set_n_r_df(int v) { n_r_df = v; }
  void add_type_mass(int type, double m);
  void add_type_charge(int type, double c);
  void add_type_radius(int type, double r);

  void set_position(int i, const Vector3d<double> &v);
  void set_velocity(int i, const Vector3d<double> &v);
  void set_acceleration(int i, const Vector3d<double> &v);

  // add_unique_atom_molecule.cpp:29-38 (id = current count; atoms outside
  // [lower, upper) are dropped) and atom_data.cpp:603-637
  bool add_atom(tools::Atom &a);
  bool add_atom(size_t id, size_t type, const Vector3d<double> &pos, const Vector3d<double> &vel);
  // atom_data.cpp:713-734: mt19937(seed) + uniform(-a, a) per component,
  // then the centre-of-mass velocity removed
  void add_random_velocity(unsigned int seed, double amplitude);

  long get_num_of_atoms_local() const { return (long)atom_struct_owned.id.size(); }
  long get_num_of_atoms_global() const { return (long)atom_struct_owned.id.size(); }
  bool position_inside_global_domain(const Vector3d<double> &p) const;

  // thermodynamics of the host mirror (kinetic_energy.cpp:26-198,206-303)
  Vector3d<double> owned_velocity_cm();
  double kinetic_energy(int t = -1);
  int degree_of_freedoms();
  int get_n_r_df();
  void finalize_temperature();
  void finalize_pressure();
  double temperature() const { return temperature_; }
  double pressure() const { return pressure_; }
  bool get_pressure_process() const { return pressure_process; }
  bool get_temperature_process() const { return temperature_process; }
  void reset_virial() { virialForce = virialConstraint = virialExternalForce = 0.0; }
  void reset_owned_acceleration();

  atom_data::Atom_struct atom_struct_owned, atom_struct_ghost;
  atom_data::Atom_type_params atom_type_params;
  double ghost_cutoff = 0.0, neighborlist_cutoff = 0.0, cutoff_extra = 0.0;
  double k_b = -1.0;
  int n_r_df = -1;
  bool stochastic_force_present = false;
  double virialForce = 0.0, virialConstraint = 0.0, virialExternalForce = 0.0;
  double potential_energy = 0.0;  // of the last force pass (not in the reference)

 private:
  Caviar2 *caviar_ = nullptr;
  Domain *domain = nullptr;
  bool pressure_process = false, temperature_process = true;
  double temperature_ = 0.0, pressure_ = 0.0;
};

}  // namespace caviar2
