// Host side of constraint::Nve through the drop-in headers: builds a few
// atoms with given velocities, applies the thermostat at a step and prints
// the kinetic energy before and after and lambda (hex floats); with
// argument "bad" it calls verify_settings on conflicting settings, which
// ends the program through the reference's fatal-error path.
#include <cstdio>
#include <cstring>

#include "caviar2/caviar2.hpp"
#include "caviar2/constraint/nve.hpp"
#include "caviar2/tools/all.hpp"

using namespace caviar2;

int main(int argc, char **argv) {
  Caviar2 sim;
  sim.domain.set_lower_global(Vector3d<double>{-10, -10, -10});
  sim.domain.set_upper_global(Vector3d<double>{10, 10, 10});
  sim.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  sim.domain.generate();
  sim.atom_data.add_type_mass(0, 2.0);
  sim.atom_data.set_k_b(1.0);
  const double v[4][3] = {{1.0, -0.5, 0.25}, {-0.75, 0.5, 1.5}, {0.0, 2.0, -1.0}, {0.5, 0.5, 0.5}};
  for (int i = 0; i < 4; ++i) {
    tools::Atom a(&sim);
    a.set_type(0);
    a.set_position(Vector3d<double>{(double)i, 0.0, 0.0});
    a.set_velocity(Vector3d<double>{v[i][0], v[i][1], v[i][2]});
    sim.atom_data.add_atom(a);
  }
  constraint::Nve nve(&sim);
  if (argc > 1 && std::strcmp(argv[1], "bad") == 0) {
    nve.kbt = 1.0;
    nve.energy_tot = 3.0;
    nve.verify_settings();
    return 0;
  }
  nve.kbt = 1.3;
  nve.step = 10;
  nve.verify_settings();
  const double ke0 = sim.atom_data.kinetic_energy();
  bool recalc = false;
  nve.apply_thermostat(7, recalc);  // not a multiple of step: no change
  const double ke7 = sim.atom_data.kinetic_energy();
  nve.apply_thermostat(20, recalc);
  const double ke1 = sim.atom_data.kinetic_energy();
  std::printf("%a %a %a %d %d %lld\n", ke0, ke7, ke1, sim.atom_data.degree_of_freedoms(), (int)recalc,
              (long long)nve.next_device_step(21));
  return 0;
}
