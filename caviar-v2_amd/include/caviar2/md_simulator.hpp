// caviar2/md_simulator.hpp -- MD driver (reference inc/caviar2/md_simulator.hpp).
//
// run() moves the system to the device once, advances it there in chunks
// between writer steps, and brings the host mirror back at writer steps and
// at the end.  Additions to the reference interface: set_precision,
// set_device and set_langevin_seed.
#pragma once
#include <cstddef>
#include <cstdint>

namespace caviar2 {

class Caviar2;
class Atom_data;
struct DeviceSession;

enum class Integrator_t { Verlet, Velocity_verlet, Leap_frog, Velocity_verlet_langevin, Unknown };

// Step index of a time, as Md_simulator::setup converts the initial and
// final time (md_simulator.cpp:569-572: (int)time / dt, the time truncated
// to an integer before the division, the quotient truncated on assignment;
// 2.5 with dt = 0.005 gives 400 steps).
inline int64_t steps_of_time(double t, double dt) { return (int64_t)((double)(int)t / dt); }

// Pair-force precision on the device.
enum class Precision { fp64, fp32_fixed };

class Md_simulator {
 public:
  Md_simulator() = default;
  explicit Md_simulator(Caviar2 *c) { set_caviar(c); }
  virtual ~Md_simulator();
  virtual void set_caviar(Caviar2 *c);

  virtual bool run();
  virtual void step();
  virtual void step(int64_t i) {
    current_step = i;
    step();
  }
  virtual void setup();
  virtual void cleanup();
  virtual void verify_settings();

  void set_initial_time(double t) { initial_time = t; use_time = true; }
  void set_final_time(double t) { final_time = t; use_time = true; }
  void set_initial_step(size_t i) { initial_step = (int64_t)i; use_step = true; }
  void set_final_step(size_t i) { final_step = (int64_t)i; use_step = true; }
  void set_final_add(size_t i) { final_step += (int64_t)i; }
  void langevin_temperature(double v) { langevin_param.temperature = v; }
  void langevin_friction(double v) { langevin_param.friction = v; }
  void langevin_kb(double v) { langevin_param.kb = v; }
  void langevin_kbt(double v);
  void set_integrator_type(Integrator_t t);
  void set_dt(double v) { dt = v; }
  void set_precision(Precision p) { precision = p; }
  void set_device(int d) { device = d; }
  void set_langevin_seed(uint64_t s) { langevin_seed = s; }

  // The device copy of this system (created, configured and uploaded on
  // first use) and the copy of its state back into Atom_data.
  DeviceSession *device_session();
  void sync_host_mirror();

  Atom_data *atom_data = nullptr;
  double time = 0.0, dt = -1.0;
  double initial_time = 0.0, final_time = 0.0;
  bool use_time = false, use_step = false;
  int64_t current_step = 0, initial_step = 0, final_step = 0;
  bool initialized = false;
  Integrator_t integrator_type = Integrator_t::Velocity_verlet;
  struct {
    double temperature = 0.0, friction = 0.0, kb = 0.0, kbt = 0.0;
    double a = 0.0, b = 0.0, c = 0.0;
  } langevin_param;
  Precision precision = Precision::fp64;
  int device = 0;
  uint64_t langevin_seed = 1;

 private:
  void advance(int64_t nsteps);
  void sync_host();
  void write_outputs(int64_t step, double t);
  void apply_constraints(int64_t step);
  Caviar2 *caviar_ = nullptr;
  DeviceSession *session_ = nullptr;
};

}  // namespace caviar2
