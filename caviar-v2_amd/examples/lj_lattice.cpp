// 8^3 simple-cubic lattice (spacing 5) of LJ atoms with uniform random
// velocities from Atom_data::add_random_velocity(1, 0.5), velocity Verlet,
// energy and temperature written every 500 steps.  Usage: lj_lattice [steps]
#include <cstdio>
#include <cstdlib>

#include "caviar2/caviar2.hpp"
#include "caviar2/tools/all.hpp"
#include "caviar2/writer/atom_writer.hpp"

using namespace caviar2;

int main(int argc, char **argv) {
  const long steps = argc > 1 ? std::atol(argv[1]) : 5000;
  Caviar2 sim;
  const double rc = 5.0, dt = 0.001;
  sim.domain.set_lower_global(Vector3d<double>{-51, -50, -50});
  sim.domain.set_upper_global(Vector3d<double>{51, 50, 50});
  sim.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  sim.domain.generate();
  sim.atom_data.set_ghost_cutoff(rc);
  for (int i = -20; i < 20; i += 5)
    for (int j = -20; j < 20; j += 5)
      for (int k = -20; k < 20; k += 5) {
        tools::Atom a(&sim);
        a.set_type(0);
        a.set_position(Vector3d<double>{(double)i, (double)j, (double)k});
        sim.atom_data.add_atom(a);
      }
  sim.atom_data.add_type_mass(0, 1.0);
  sim.atom_data.set_k_b(1.0);
  sim.atom_data.set_temperature_process(true);
  sim.atom_data.add_random_velocity(1, 0.5);
  sim.neighborlist.set_cutoff(rc);
  sim.neighborlist.set_cutoff_extra_coef(1.0523);
  sim.neighborlist.set_dt(dt);
  force_field::Lj lj(&sim);
  lj.set_cutoff(rc);
  lj.set_sigma(0, 0, 1.0);
  lj.set_epsilon(0, 0, 1.0);
  sim.add_force_field(&lj);
  writer::Atom_writer out(&sim);
  out.set_energy_step(500);
  out.set_temperature_step(500);
  out.set_file_name_energy("lattice_energy");
  out.set_file_name_temperature("lattice_temperature");
  sim.add_writer(&out);
  sim.md_simulator.set_integrator_type(Integrator_t::Velocity_verlet);
  sim.md_simulator.set_initial_step(0);
  sim.md_simulator.set_final_step((size_t)steps);
  sim.md_simulator.set_dt(dt);
  sim.md_simulator.run();
  std::printf("natoms %zu ke %.12e T %.12e\n", sim.atom_data.atom_struct_owned.position.size(),
              sim.atom_data.kinetic_energy(), sim.atom_data.temperature());
  return 0;
}
