// caviar2/tools/atom.hpp -- single-atom builder used with
// Atom_data::add_atom (reference inc/caviar2/tools/atom.hpp:33-70).
#pragma once
#include <cstddef>

#include "caviar2/vector3d.hpp"

namespace caviar2 {
class Caviar2;
namespace tools {

class Atom {
 public:
  Atom() = default;
  explicit Atom(Caviar2 *c) : caviar_(c) {}
  void set_type(size_t t) { type = t; }
  void set_position(const Vector3d<double> &v) { position = v; }
  void set_velocity(const Vector3d<double> &v) { velocity = v; }
  Vector3d<double> pos_tot() const { return position; }
  Vector3d<double> vel_tot() const { return velocity; }
  void set_caviar(Caviar2 *c) { caviar_ = c; }

  Vector3d<double> position, velocity;
  size_t type = 0;

 private:
  Caviar2 *caviar_ = nullptr;
};

}  // namespace tools
}  // namespace caviar2
