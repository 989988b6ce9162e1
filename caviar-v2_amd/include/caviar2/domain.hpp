// caviar2/domain.hpp -- simulation box (reference inc/caviar2/domain.hpp).
// One process holds the whole box, so the local box equals the global one.
#pragma once
#include "caviar2/vector3d.hpp"

namespace caviar2 {

class Caviar2;

class Domain {
 public:
  Domain() = default;
  explicit Domain(Caviar2 *c) : caviar_(c) {}
  void set_caviar(Caviar2 *c) { caviar_ = c; }
  void set_lower_global(const Vector3d<double> &v) { lower_global = v; }
  void set_upper_global(const Vector3d<double> &v) { upper_global = v; }
  void set_boundary_condition(const Vector3d<int> &v) { boundary_condition = v; }
  // domain.cpp:295-328,495-502 with a 1x1x1 process grid
  void generate();
  void update_after_domain_change();
  double volume_local() const { return size_local.x * size_local.y * size_local.z; }
  double volume_global() const { return size_global.x * size_global.y * size_global.z; }

  Vector3d<int> boundary_condition{0, 0, 0};
  Vector3d<double> lower_global, upper_global;
  Vector3d<double> lower_local, upper_local;
  Vector3d<double> size_local, size_global, half_edge;
  int me = 0, nprocs = 1;

 private:
  Caviar2 *caviar_ = nullptr;
};

}  // namespace caviar2
