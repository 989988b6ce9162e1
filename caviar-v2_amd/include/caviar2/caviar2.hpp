// caviar2/caviar2.hpp -- composition root of the caviar2 API
// (reference inc/caviar2/caviar2.hpp:39-66): owns the state objects by
// value and holds force fields, constraints and writers by non-owning
// pointer.
#pragma once
#include <vector>

#include "caviar2/atom_data.hpp"
#include "caviar2/constraint.hpp"
#include "caviar2/domain.hpp"
#include "caviar2/force/lj.hpp"
#include "caviar2/force_field.hpp"
#include "caviar2/log.hpp"
#include "caviar2/md_simulator.hpp"
#include "caviar2/neighborlist.hpp"
#include "caviar2/vector3d.hpp"
#include "caviar2/writer.hpp"

namespace caviar2 {

class Caviar2 {
 public:
  Caviar2();
  ~Caviar2();
  Caviar2(const Caviar2 &) = delete;
  Caviar2 &operator=(const Caviar2 &) = delete;

  Atom_data atom_data;
  Domain domain;
  Md_simulator md_simulator;
  Neighborlist neighborlist;
  Log log;

  std::vector<Force_field *> forces;
  std::vector<Constraint *> constraints;
  std::vector<Writer *> writers;

  void add_force_field(Force_field *x);
  void add_constraint(Constraint *x) { constraints.push_back(x); }
  void add_writer(Writer *x) { writers.push_back(x); }
};

}  // namespace caviar2
