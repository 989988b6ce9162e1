// caviar2/constraint/nve.hpp -- velocity-rescale constraint (reference
// inc/caviar2/constraint/nve.hpp, src/constraint/nve.cpp).
//
// Every `step` timesteps, after the velocity update, all owned velocities
// are multiplied by lambda so that the kinetic energy equals the target:
// energy_per_dof * dof, energy_tot, dof * kbt / 2 or dof * kb * temperature
// / 2, whichever is set.  On the device path (Md_simulator::run/step) the
// kinetic energy comes from the device and the rescale runs there.
#pragma once
#include <cstdint>

#include "caviar2/constraint.hpp"

namespace caviar2 {

class Atom_data;

namespace constraint {

class Nve : public Constraint {
 public:
  explicit Nve(Caviar2 *caviar);
  ~Nve() override = default;

  // nve.cpp:132-190 on the host state (Atom_data velocities)
  void apply_thermostat(int64_t timestep, bool &recalculate_temperature) override;
  // nve.cpp:99-130
  void verify_settings() override;

  bool device_capable() const override { return true; }
  int64_t next_device_step(int64_t from) const override;
  double device_velocity_scale(int64_t timestep, double ke_now, int dof) override;

  // lambda of nve.cpp:147-184 for a kinetic energy and degrees of freedom
  double lambda(double ke_now, int dof, int64_t timestep) const;

  double energy_per_dof = -1.0, energy_tot = -1.0;
  double temperature = -1.0, kb = -1.0, kbt = -1.0;
  int step = 1;  // apply after each 'step' timesteps
  Atom_data *atom_data = nullptr;
};

}  // namespace constraint
}  // namespace caviar2
