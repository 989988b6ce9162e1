// Drop-in checks: a user program of the caviar2 API (as the reference's own
// tests/t00x programs are), driven by tests/test_dropin.py.
//
//   dropin_checks steps <time> <dt>          print steps_of_time(time, dt) (no device)
//   dropin_checks run <input> <vv|lf> <steps> <cadence>
//       read the system, run Velocity_verlet or Leap_frog with every writer
//       (energy, temperature, pressure, potential energy, xyz with id,
//       velocity and acceleration) at the cadence, print the final state
//   dropin_checks molrun <input> <steps>
//       molecule indices (runs of 4 atoms, every fifth run outside
//       molecules) with Lj::ignore_intra_molecule_flag, run Velocity_verlet,
//       print the final state
//   dropin_checks neighlist <input> <all>
//       build the list (Neighborlist::build(true)), materialise neighlist
//       (all = 1: the all_atom_test form), print it and the ghosts
//   dropin_checks ljtables <input> <atom|array|both>
//       types i % 3, Lj parameters given per atom (with WCA and lambda
//       scalings) or as arrays, Lj::verify_settings, print the 3 x 3
//       epsilon, sigma and cutoff-list tables (no device)
//   dropin_checks ljrun <input> <steps>
//       types i % 2 (masses 1, 1.5), per-atom epsilon / sigma, WCA,
//       lambda_e, Lj::verify_settings, run Velocity_verlet, print the final
//       state
//
// Input: n L rc coef ghost_cutoff, then n lines "x y z vx vy vz" (box [0, L)^3,
// periodic, one species eps = sigma = mass = 1).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "caviar2/caviar2.hpp"
#include "caviar2/tools/all.hpp"
#include "caviar2/writer/atom_writer.hpp"

using namespace caviar2;

static void load(Caviar2 &sim, const char *path, double &rc, double &coef) {
  std::ifstream in(path);
  long n = 0;
  double L = 0.0, gc = 0.0;
  in >> n >> L >> rc >> coef >> gc;
  sim.domain.set_lower_global(Vector3d<double>{0.0, 0.0, 0.0});
  sim.domain.set_upper_global(Vector3d<double>{L, L, L});
  sim.domain.set_boundary_condition(Vector3d<int>{1, 1, 1});
  sim.domain.generate();
  sim.atom_data.set_ghost_cutoff(gc);
  for (long i = 0; i < n; ++i) {
    double x, y, z, vx, vy, vz;
    in >> x >> y >> z >> vx >> vy >> vz;
    sim.atom_data.add_atom((size_t)i, 0, Vector3d<double>{x, y, z}, Vector3d<double>{vx, vy, vz});
  }
  sim.atom_data.add_type_mass(0, 1.0);
  sim.atom_data.set_k_b(1.0);
  sim.atom_data.set_n_r_df(3);
  sim.atom_data.set_temperature_process(true);
  sim.atom_data.set_pressure_process(true);
  sim.neighborlist.set_cutoff(rc);
  sim.neighborlist.set_cutoff_extra_coef(coef);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  if (mode == "steps") {
    std::printf("%lld\n", (long long)steps_of_time(std::atof(argv[2]), std::atof(argv[3])));
    return 0;
  }
  Caviar2 sim;
  double rc = 0.0, coef = 1.0;
  load(sim, argv[2], rc, coef);
  force_field::Lj lj(&sim);
  lj.set_cutoff(rc);
  const bool own_params = mode == "ljtables" || mode == "ljrun";
  if (!own_params) {
    lj.set_sigma(0, 0, 1.0);
    lj.set_epsilon(0, 0, 1.0);
  }
  sim.add_force_field(&lj);
  auto print_state = [&sim]() {
    const auto &o = sim.atom_data.atom_struct_owned;
    for (size_t i = 0; i < o.position.size(); ++i)
      std::printf("%a %a %a %a %a %a\n", o.position[i].x, o.position[i].y, o.position[i].z,
                  o.velocity[i].x, o.velocity[i].y, o.velocity[i].z);
  };
  if (mode == "ljtables") {
    auto &ty = sim.atom_data.atom_struct_owned.type;
    for (size_t i = 0; i < ty.size(); ++i) ty[i] = i % 3;
    const std::string c = argv[3];
    if (c == "atom" || c == "both") {
      const double e[3] = {1.0, 0.5, 0.8}, sg[3] = {1.0, 0.88, 1.2};
      for (int t = 0; t < 3; ++t) {
        lj.set_epsilon_atom(t, e[t]);
        lj.set_sigma_atom(t, sg[t]);
      }
      lj.enable_wca(true);
      lj.set_lambda_s({{1.0, 0.9}, {0.9}});  // partial rows: the rest is 1
      lj.set_lambda_e({{1.0, 0.5, 1.0}, {0.5, 1.0, 1.0}, {1.0, 1.0, 0.7}});
    }
    if (c == "array" || c == "both") {
      lj.set_epsilon(0, 0, 1.0);
      lj.set_epsilon(0, 1, 0.5);
      lj.set_epsilon(1, 0, 0.5);
      lj.set_epsilon(1, 1, 0.8);
      lj.set_sigma(0, 0, 1.0);
      lj.set_sigma(1, 1, 0.9);
      lj.set_cutoff_list({{2.5}});
    }
    lj.verify_settings();
    std::vector<double> e, sg, cl;
    bool has_cl = false;
    lj.tables(3, e, sg, cl, has_cl);
    std::printf("has_cutoff_list %d\n", has_cl ? 1 : 0);
    for (const auto *tab : {&e, &sg, &cl}) {
      for (int k = 0; k < 9; ++k) std::printf("%a%c", (*tab)[k], k == 8 ? '\n' : ' ');
    }
    return 0;
  }
  if (mode == "ljrun") {
    auto &ty = sim.atom_data.atom_struct_owned.type;
    for (size_t i = 0; i < ty.size(); ++i) ty[i] = i % 2;
    sim.atom_data.add_type_mass(1, 1.5);
    lj.set_epsilon_atom(0, 1.0);
    lj.set_epsilon_atom(1, 0.6);
    lj.set_sigma_atom(0, 1.0);
    lj.set_sigma_atom(1, 0.9);
    lj.enable_wca(true);
    lj.set_lambda_e({{1.0, 0.8}, {0.8, 1.0}});
    lj.verify_settings();
    sim.md_simulator.set_integrator_type(Integrator_t::Velocity_verlet);
    sim.md_simulator.set_initial_step(0);
    sim.md_simulator.set_final_step((size_t)std::atol(argv[3]));
    sim.md_simulator.set_dt(0.005);
    sim.md_simulator.run();
    print_state();
    return 0;
  }
  if (mode == "run") {
    const bool lf = std::strcmp(argv[3], "lf") == 0;
    const long steps = std::atol(argv[4]);
    const int cad = std::atoi(argv[5]);
    // optional: every list build of the run checked on the device
    sim.neighborlist.set_verify_builds(argc > 6 && std::strcmp(argv[6], "verify") == 0);
    writer::Atom_writer out(&sim);
    out.set_energy_step(cad);
    out.set_temperature_step(cad);
    out.set_pressure_step(cad);
    out.set_potential_energy_step(cad);
    out.set_xyz_step(cad);
    out.set_xyz_output_id(1);
    out.set_xyz_output_velocity(1);
    out.set_xyz_output_acceleration(1);
    out.set_file_name_energy("w_energy");
    out.set_file_name_temperature("w_temperature");
    out.set_file_name_pressure("w_pressure");
    out.set_file_name_potential_energy("w_pe");
    out.set_file_name_xyz("w_xyz");
    sim.add_writer(&out);
    sim.md_simulator.set_integrator_type(lf ? Integrator_t::Leap_frog : Integrator_t::Velocity_verlet);
    sim.md_simulator.set_initial_step(0);
    sim.md_simulator.set_final_step((size_t)steps);
    sim.md_simulator.set_dt(0.005);
    sim.md_simulator.run();
    const auto &o = sim.atom_data.atom_struct_owned;
    for (size_t i = 0; i < o.position.size(); ++i)
      std::printf("%a %a %a %a %a %a\n", o.position[i].x, o.position[i].y, o.position[i].z,
                  o.velocity[i].x, o.velocity[i].y, o.velocity[i].z);
    return 0;
  }
  if (mode == "molrun") {
    auto &mi = sim.atom_data.atom_struct_owned.molecule_index;
    for (size_t i = 0; i < mi.size(); ++i) mi[i] = (i / 4) % 5 == 4 ? -1 : (int)(i / 4);
    lj.ignore_intra_molecule_flag(true);
    sim.md_simulator.set_integrator_type(Integrator_t::Velocity_verlet);
    sim.md_simulator.set_initial_step(0);
    sim.md_simulator.set_final_step((size_t)std::atol(argv[3]));
    sim.md_simulator.set_dt(0.005);
    sim.md_simulator.run();
    const auto &o = sim.atom_data.atom_struct_owned;
    for (size_t i = 0; i < o.position.size(); ++i)
      std::printf("%a %a %a %a %a %a\n", o.position[i].x, o.position[i].y, o.position[i].z,
                  o.velocity[i].x, o.velocity[i].y, o.velocity[i].z);
    return 0;
  }
  if (mode == "neighlist") {
    sim.neighborlist.set_all_atom_test(std::atoi(argv[3]) != 0);
    sim.neighborlist.build(true);
    sim.neighborlist.update_host_list();
    const auto &nl = sim.neighborlist.neighlist;
    const auto &g = sim.atom_data.atom_struct_ghost;
    std::printf("%zu %zu\n", nl.size(), g.position.size());
    for (size_t q = 0; q < g.position.size(); ++q)
      std::printf("%zu %a %a %a\n", g.id[q], g.position[q].x, g.position[q].y, g.position[q].z);
    for (const auto &l : nl) {
      std::printf("%zu", l.size());
      for (unsigned j : l) std::printf(" %u", j);
      std::printf("\n");
    }
    return 0;
  }
  return 2;
}
