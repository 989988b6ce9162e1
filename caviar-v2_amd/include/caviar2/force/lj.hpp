// caviar2/force/lj.hpp -- Lennard-Jones force field
// (reference inc/caviar2/force/lj.hpp:33-150).  calculate_acceleration()
// runs the device kernel through the C ABI; inside Md_simulator::run the
// whole step runs on the device and this object only supplies parameters.
#pragma once
#include <vector>

#include "caviar2/force_field.hpp"

namespace caviar2 {
namespace force_field {

class Lj : public Force_field {
 public:
  explicit Lj(Caviar2 *c);
  ~Lj() override = default;

  // lj.cpp:180-391: per-atom parameters mixed into the tables (sigma_ij =
  // (sigma_i + sigma_j) / 2, epsilon_ij = sqrt(epsilon_i epsilon_j)), tables
  // sized to the atom types present, WCA cutoffs 2^(1/6) sigma_ij, then the
  // lambda_s / lambda_e scalings.  The reference's Md_simulator does not
  // call it; a program that does gets the tables it forms on the device.
  void verify_settings() override;
  void calculate_acceleration() override;
  double energy() override;

  // lj.cpp:74-176
  void set_cutoff(double value);
  void set_jump_tol(double value);
  void enable_monitor_jump(bool flag = true) { monitor_jump = flag; }
  void ignore_intra_molecule_flag(bool flag = true);
  void enable_wca(bool flag = true);
  void enable_off_diagonal_vectors(bool flag = true) { make_off_diagonal_vectors = flag; }
  void set_cutoff_list(const std::vector<std::vector<double>> &values);
  void set_epsilon(const int type1, const int type2, const double value);
  void set_sigma(const int type1, const int type2, const double value);
  void set_lambda_s(const std::vector<std::vector<double>> &values);
  void set_lambda_e(const std::vector<std::vector<double>> &values);
  void set_epsilon_atom(const int type, const double value);
  void set_sigma_atom(const int type, const double value);

  // Dense ntypes x ntypes tables as the device uses them (missing entries 0).
  void tables(int ntypes, std::vector<double> &eps, std::vector<double> &sigma,
              std::vector<double> &cutoff_list, bool &has_cutoff_list) const;
  // The jump_fix clamp of r^2 the device applies (0 when jump_fix is off).
  double device_jump_tol() const { return jump_fix ? jump_tol : 0.0; }
  bool ignores_intra_molecule() const { return ignore_intra_molecule; }

 private:
  std::vector<std::vector<double>> epsilon, sigma, cutoff_list, lambda_s, lambda_e;
  std::vector<double> epsilon_atom, sigma_atom;
  bool cutoff_list_activated = false, wca = false, jump_fix = false, monitor_jump = false;
  bool ignore_intra_molecule = false, make_off_diagonal_vectors = false;
  // which inputs were given (lj.cpp:28-176); verify_settings forms the
  // tables from them
  bool input_by_array = false, input_by_atom = false, lambda_s_is_set = false,
       lambda_e_is_set = false;
  double jump_tol = 1e-6;
};

}  // namespace force_field
}  // namespace caviar2
