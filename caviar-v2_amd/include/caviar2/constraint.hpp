// caviar2/constraint.hpp -- constraint hook base (reference
// inc/caviar2/constraint.hpp).  Hooks run on the host; a run with
// constraints registered steps the device one step at a time so that the
// hooks see every step.
#pragma once
#include <cstdint>

namespace caviar2 {

class Caviar2;

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

 protected:
  Caviar2 *caviar_ = nullptr;
};

}  // namespace caviar2
