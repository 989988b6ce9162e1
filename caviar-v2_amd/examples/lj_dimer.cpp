// LJ dimer in a large periodic box, velocity Verlet, written with the caviar2
// API calls a reference user program makes (Caviar2, tools::Atom, Domain,
// Atom_data, Neighborlist, force_field::Lj, writer::Atom_writer,
// Md_simulator).  Prints the final separation and energies.
#include <cstdio>

#include "caviar2/caviar2.hpp"
#include "caviar2/tools/all.hpp"
#include "caviar2/writer/atom_writer.hpp"

using namespace caviar2;

int main() {
  Caviar2 sim;
  const double rc = 5.0, dt = 0.001;

  tools::Atom left(&sim), right(&sim);
  left.set_type(0);
  left.set_position(Vector3d<double>{-0.6, 0.0, 0.0});
  right.set_type(0);
  right.set_position(Vector3d<double>{0.6, 0.0, 0.0});

  sim.domain.set_lower_global(Vector3d<double>{-20, -20, -20});
  sim.domain.set_upper_global(Vector3d<double>{20, 20, 20});
  sim.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  sim.domain.generate();

  sim.atom_data.set_ghost_cutoff(rc * 1.1);
  sim.atom_data.add_atom(left);
  sim.atom_data.add_atom(right);
  sim.atom_data.add_type_mass(0, 1.0);
  sim.atom_data.add_type_charge(0, 0.0);
  sim.atom_data.set_k_b(1.0);
  sim.atom_data.set_temperature_process(true);

  sim.neighborlist.set_cutoff(rc);
  sim.neighborlist.set_cutoff_extra_coef(1.1);
  sim.neighborlist.set_dt(dt);

  force_field::Lj lj(&sim);
  lj.set_cutoff(rc);
  lj.set_sigma(0, 0, 1.0);
  lj.set_epsilon(0, 0, 1.0);
  sim.add_force_field(&lj);

  writer::Atom_writer out(&sim);
  out.set_energy_step(100);
  out.set_file_name_energy("dimer_energy");
  sim.add_writer(&out);

  sim.md_simulator.set_integrator_type(Integrator_t::Velocity_verlet);
  sim.md_simulator.set_initial_step(0);
  sim.md_simulator.set_final_step(2000);
  sim.md_simulator.set_dt(dt);
  sim.md_simulator.run();

  const auto &p = sim.atom_data.atom_struct_owned.position;
  const double ke = sim.atom_data.kinetic_energy();
  std::printf("separation %.12f ke %.12e pe %.12e\n", p[1].x - p[0].x, ke, lj.energy());
  return 0;
}
