// 864-atom LJ fluid (6^3 FCC cells, rho* = 0.8442) started from uniform
// random velocities, velocity Verlet on the device with the velocity-rescale
// constraint constraint::Nve (kbt = 0.9 every 50 steps), temperature written
// every 50 steps.  Usage: lj_nve [steps]
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "caviar2/caviar2.hpp"
#include "caviar2/constraint/nve.hpp"
#include "caviar2/tools/all.hpp"
#include "caviar2/writer/atom_writer.hpp"

using namespace caviar2;

int main(int argc, char **argv) {
  const long steps = argc > 1 ? std::atol(argv[1]) : 500;
  const int nc = 6;
  const double a = std::cbrt(4.0 / 0.8442), rc = 2.5, dt = 0.005;
  Caviar2 sim;
  sim.domain.set_lower_global(Vector3d<double>{0, 0, 0});
  sim.domain.set_upper_global(Vector3d<double>{nc * a, nc * a, nc * a});
  sim.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  sim.domain.generate();
  sim.atom_data.set_ghost_cutoff(rc * 1.12);
  const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j)
      for (int k = 0; k < nc; ++k)
        for (const auto &b : basis) {
          tools::Atom at(&sim);
          at.set_type(0);
          at.set_position(Vector3d<double>{(i + b[0]) * a, (j + b[1]) * a, (k + b[2]) * a});
          sim.atom_data.add_atom(at);
        }
  sim.atom_data.add_type_mass(0, 1.0);
  sim.atom_data.set_k_b(1.0);
  sim.atom_data.set_temperature_process(true);
  sim.atom_data.add_random_velocity(7, 1.5);
  sim.neighborlist.set_cutoff(rc);
  sim.neighborlist.set_cutoff_extra_coef(1.12);
  sim.neighborlist.set_dt(dt);
  force_field::Lj lj(&sim);
  lj.set_cutoff(rc);
  lj.set_sigma(0, 0, 1.0);
  lj.set_epsilon(0, 0, 1.0);
  sim.add_force_field(&lj);
  constraint::Nve nve(&sim);
  nve.kbt = 0.9;
  nve.step = 50;
  sim.add_constraint(&nve);
  writer::Atom_writer out(&sim);
  out.set_temperature_step(50);
  out.set_file_name_temperature("nve_temperature");
  sim.add_writer(&out);
  sim.md_simulator.set_integrator_type(Integrator_t::Velocity_verlet);
  sim.md_simulator.set_initial_step(0);
  sim.md_simulator.set_final_step((size_t)steps);
  sim.md_simulator.set_dt(dt);
  sim.md_simulator.run();
  std::printf("natoms %zu T %.15e\n", sim.atom_data.atom_struct_owned.position.size(),
              sim.atom_data.temperature());
  return 0;
}
