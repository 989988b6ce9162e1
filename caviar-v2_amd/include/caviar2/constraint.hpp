// caviar2/constraint.hpp -- constraint hook base (reference
// inc/caviar2/constraint.hpp).  Md_simulator runs on the device; a
// constraint takes part there when it acts only by rescaling velocities at
// given steps (device_capable): the run stops at those steps, forms the
// factor from the device's kinetic energy and applies it on the device.
// Other constraints stop a device run with an error.
#pragma once
#include <cstdint>

namespace caviar2 {

class Caviar2;

enum class Constraint_t {
  Atom_molarity,
  Atoms_molarity,
  Berendsen,
  Cm_motion,
  M_shake,
  Nose_hoover,
  Nve,
  Rattle,
  Shake,
  Unknown
};

class Constraint {
 public:
  Constraint() = default;
  explicit Constraint(Caviar2 *c) : caviar_(c) {}
  virtual ~Constraint() = default;
  virtual void verify_settings() {}
  virtual void apply(int64_t) {}
  virtual void apply_thermostat(int64_t, bool &) {}
  virtual void apply_barostat(int64_t, bool &) {}
  virtual void fix_position(int64_t) {}
  virtual void fix_velocity(int64_t, bool &) {}
  virtual void fix_acceleration(int64_t) {}
  virtual void apply_shake(int64_t) {}

  // Device path.  The first step >= from at which the constraint acts (-1:
  // never), and the velocity factor it applies at such a step, given the
  // kinetic energy then (Atom_data::kinetic_energy) and the degrees of
  // freedom (Atom_data::degree_of_freedoms).
  virtual bool device_capable() const { return false; }
  virtual int64_t next_device_step(int64_t) const { return -1; }
  virtual double device_velocity_scale(int64_t, double, int) { return 1.0; }

  Constraint_t constraint_type = Constraint_t::Unknown;

 protected:
  Caviar2 *caviar_ = nullptr;
};

}  // namespace caviar2
