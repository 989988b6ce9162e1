// caviar2/force_field.hpp -- force-field plugin base
// (reference inc/caviar2/force_field.hpp:36-128).  Plugins add into
// atom_data->atom_struct_owned.acceleration and atom_data->virialForce.
#pragma once
#include "caviar2/vector3d.hpp"

namespace caviar2 {

class Caviar2;
class Atom_data;
class Domain;
class Neighborlist;

class Force_field {
 public:
  Force_field() = default;
  explicit Force_field(Caviar2 *c) { set_caviar(c); }
  virtual ~Force_field() = default;
  virtual void set_caviar(Caviar2 *c);
  virtual void calculate_acceleration() = 0;
  virtual double energy();
  virtual void verify_settings() {}

  double cutoff = 0.0;
  Atom_data *atom_data = nullptr;
  Domain *domain = nullptr;
  Neighborlist *neighborlist = nullptr;

 protected:
  Caviar2 *caviar_ = nullptr;
};

}  // namespace caviar2
