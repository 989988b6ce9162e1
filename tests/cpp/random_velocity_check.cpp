// Prints the velocities Atom_data::add_random_velocity(seed, amplitude)
// gives n atoms (host-only part of the caviar2 API; no device use).
#include <cstdio>
#include <cstdlib>

#include "caviar2/caviar2.hpp"

using namespace caviar2;

int main(int argc, char **argv) {
  const int n = std::atoi(argv[1]);
  const unsigned seed = (unsigned)std::atoi(argv[2]);
  const double amp = std::atof(argv[3]);
  Caviar2 c;
  c.domain.set_lower_global(Vector3d<double>{0, 0, 0});
  c.domain.set_upper_global(Vector3d<double>{1, 1, 1});
  c.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  c.domain.generate();
  for (int i = 0; i < n; ++i)
    c.atom_data.add_atom((size_t)i, 0, Vector3d<double>{0.5, 0.5, 0.5}, Vector3d<double>{0, 0, 0});
  c.atom_data.add_type_mass(0, 1.0);
  c.atom_data.add_random_velocity(seed, amp);
  for (const auto &v : c.atom_data.atom_struct_owned.velocity) std::printf("%a %a %a\n", v.x, v.y, v.z);
  return 0;
}
