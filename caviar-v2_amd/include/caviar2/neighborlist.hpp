// caviar2/neighborlist.hpp -- Verlet list settings and build
// (reference inc/caviar2/neighborlist.hpp:36-199).
//
// The list itself lives on the device (full list, cell-sorted order).
// neighlist, the reference's half list of owned j > i plus ghost images
// (ghost g stored as N + g, ghosts in atom_struct_ghost order), is filled on
// the host only by update_host_list(); with set_all_atom_test(true) it is
// the reference's all-pairs debug list (neighborlist.cpp:170-200), while the
// device keeps its distance list, which gives the same forces.
#pragma once
#include <vector>

#include "caviar2/vector3d.hpp"

namespace caviar2 {

class Caviar2;

class Neighborlist {
 public:
  Neighborlist() = default;
  explicit Neighborlist(Caviar2 *c) : caviar_(c) {}
  virtual ~Neighborlist() = default;
  virtual void set_caviar(Caviar2 *c) { caviar_ = c; }

  void set_cutoff(double v);
  void set_cutoff_extra_coef(double v);
  void set_make_verlet_list_from_cell_list(bool v) {
    make_verlet_list_from_cell_list = v;
    build_cell_list = v;
  }
  void set_build_cell_list(bool v) { build_cell_list = v; }
  void set_all_atom_test(bool v) { all_atom_test = v; }
  void set_rebuild_test(bool v) { rebuild_test = v; }
  // Addition (default false): also rebuild on the reference's other
  // triggers, an atom's wrap and a changed ghost set, so that the rebuild
  // steps are the reference's (the device's images need neither)
  void set_reference_rebuilds(bool v) { reference_rebuilds = v; }
  // Addition (default false): check every list build on the device against
  // an all-pairs recomputation of the pair set (O(N^2) per build, for
  // tests); a build that differs stops the run with an error
  void set_verify_builds(bool v) { verify_builds = v; }
  void set_dt(double v);

  // neighborlist.cpp:303-390 (cutoff_extra and the rebuild threshold)
  virtual void init();
  // neighborlist.cpp:77-98: rebuild when asked or when update_is_needed()
  virtual bool build(bool update_neighborlist);
  // Materialise neighlist (and the ghost images it indexes) on the host.
  void update_host_list();

  std::vector<std::vector<unsigned int>> neighlist;
  bool make_verlet_list_from_cell_list = false;
  bool build_cell_list = false;
  bool all_atom_test = false;
  bool rebuild_test = false;
  bool reference_rebuilds = false;
  bool verify_builds = false;
  double cutoff = 0.0;
  double cutoff_extra = 0.0;
  double cutoff_extra_coef = 1.12347;
  double threshold_distance_sq = 0.0;
  double dt = 0.0;

 protected:
  Caviar2 *caviar_ = nullptr;
};

}  // namespace caviar2
