// caviar2/writer/atom_writer.hpp -- text dumps at a step cadence
// (reference inc/caviar2/writer/atom_writer.hpp): xyz frames
// (dump_xyz_serial.cpp:51-64), kinetic energy "step time value"
// (dump_energy.cpp:42-53), temperature and pressure in the same format, and
// (an addition; the reference defines no potential energy) the LJ potential
// energy of the last force pass, "step time value" as well.
#pragma once
#include <cstdint>
#include <fstream>
#include <string>

#include "caviar2/writer.hpp"

namespace caviar2 {
class Atom_data;
namespace writer {

class Atom_writer : public Writer {
 public:
  explicit Atom_writer(Caviar2 *c);
  ~Atom_writer() override;

  void write(int64_t step, double time) override;
  int64_t next_output_step(int64_t from) const override;

  void set_xyz_step(int s) { xyz_step = s; output_xyz = s > 0; }
  void set_energy_step(int s) { energy_step = s; output_energy = s > 0; }
  void set_temperature_step(int s) { temperature_step = s; output_temperature = s > 0; }
  void set_pressure_step(int s) { pressure_step = s; output_pressure = s > 0; }
  void set_potential_energy_step(int s) { pe_step = s; output_pe = s > 0; }
  void set_file_name_potential_energy(const std::string &f) { file_name_pe = f; }
  void set_xyz_ghost_step(int) {}
  void set_povray_step(int) {}
  void set_msd_step(int) {}
  void set_msd_initial_step(int) {}
  void set_volume_step(int) {}
  // MPI placement switches: one process writes everything here
  void set_xyz_mpi_rank0(bool) {}
  void set_xyz_mpi_per_process(bool) {}
  void set_xyz_ghost_mpi_rank0(bool) {}
  void set_xyz_ghost_mpi_per_process(bool) {}
  void set_energy_mpi_rank0(bool) {}
  void set_energy_mpi_per_process(bool) {}
  void set_temperature_mpi_rank0(bool) {}
  void set_temperature_mpi_per_process(bool) {}
  void set_pressure_mpi_rank0(bool) {}
  void set_pressure_mpi_per_process(bool) {}
  void set_povray_mpi_rank0(bool) {}
  void set_povray_mpi_per_process(bool) {}
  void set_msd_mpi_rank0(bool) {}
  void set_msd_mpi_per_process(bool) {}
  void set_volume_mpi_rank0(bool) {}
  void set_volume_mpi_per_process(bool) {}
  void set_file_name_xyz(const std::string &f) { file_name_xyz = f; }
  void set_file_name_energy(const std::string &f) { file_name_energy = f; }
  void set_file_name_temperature(const std::string &f) { file_name_temperature = f; }
  void set_file_name_pressure(const std::string &f) { file_name_pressure = f; }
  void set_xyz_output_id(int v) { xyz_output_id = v != 0; }
  void set_xyz_output_velocity(int v) { xyz_output_velocity = v != 0; }
  void set_xyz_output_acceleration(int v) { xyz_output_acceleration = v != 0; }

 private:
  Atom_data *atom_data = nullptr;
  bool output_xyz = false, output_energy = false, output_temperature = false, output_pressure = false;
  bool output_pe = false;
  int xyz_step = 0, energy_step = 0, temperature_step = 0, pressure_step = 0, pe_step = 0;
  bool xyz_output_id = false, xyz_output_velocity = false, xyz_output_acceleration = false;
  std::string file_name_xyz = "o_xyz", file_name_energy = "o_energy";
  std::string file_name_temperature = "o_temperature", file_name_pressure = "o_pressure";
  std::string file_name_pe = "o_potential_energy";
  std::ofstream ofs_xyz, ofs_energy, ofs_temperature, ofs_pressure, ofs_pe;
};

}  // namespace writer
}  // namespace caviar2
