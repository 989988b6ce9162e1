"""Generate the golden fixtures from the compiled reference.

Runs only where the reference tree exists (it drives
``oracle/_ref/libcaviar2_ref.so``, built by ``make -C oracle ref``).  Each
fixture is an ``.npz`` of inputs and the reference's outputs; the manifest
``golden.json`` records a SHA-256 per file.  Regenerate with

    make -C oracle oracle ref && python tests/golden/make_golden.py [fixture ...]

(naming fixtures regenerates only those).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "caviar-v2_amd"))

import oracle as O  # noqa: E402
from cv2md.systems import ka_mixture, lj_fluid  # noqa: E402

SYS_KEYS = ("lo", "hi", "bc", "pos", "vel", "type", "eps", "sigma", "cutoff_list", "mass",
            "lj_cutoff", "nl_cutoff", "nl_coef", "ghost_cutoff", "dt", "integrator", "gamma",
            "kbt", "k_b", "n_r_df")


def sys_arrays(s):
    out = {}
    for k in SYS_KEYS:
        v = s.get(k)
        if v is None:
            continue
        out["in_" + k] = np.asarray(v)
    return out


def setup_fixture(name, s, use_cell_list=False):
    s = dict(s, use_cell_list=int(use_cell_list))
    with O.RefSim(s) as r:
        r.setup()
        pos, vel, acc = r.state()
        th = r.thermo()
        gpos, gid = r.ghosts()
        offs, nbr = r.nlist()
    d = sys_arrays(s)
    d.update(in_use_cell_list=np.asarray(int(use_cell_list)), pos=pos, acc=acc, virial=th["virial"],
             ke=th["ke"], temperature=th["temperature"], gpos=gpos, gid=gid, offs=offs, nbr=nbr)
    return name, d


def traj_fixture(name, s, nsteps, keep=(1, 10, 100), use_cell_list=False, with_eta=False, mol=None):
    s = dict(s, use_cell_list=int(use_cell_list))
    d = sys_arrays(s)
    d["in_use_cell_list"] = np.asarray(int(use_cell_list))
    if mol is not None:  # Lj::ignore_intra_molecule on, with these molecule indices
        d["in_mol"] = np.asarray(mol, np.int32)
    rebuilt, kes, etas = [], [], []
    with O.RefSim(s) as r:
        if mol is not None:
            r.set_molecules(mol, True)
        r.setup()
        for k in range(1, nsteps + 1):
            rebuilt.append(r.step())
            if with_eta:
                etas.append(r.eta())
            kes.append(r.thermo()["ke"])
            if k in keep:
                p, v, a = r.state()
                d[f"pos_{k}"], d[f"vel_{k}"], d[f"acc_{k}"] = p, v, a
    d.update(nsteps=np.asarray(nsteps), rebuilt=np.asarray(rebuilt, np.int8), ke=np.asarray(kes),
             keep=np.asarray(keep))
    if with_eta:
        d["eta"] = np.asarray(etas)
    return name, d


def dimer():
    # tests/t001_lj.cpp:30-71 geometry, with the LJ cutoff equal to the list cutoff.
    return dict(lo=np.array([-51.0, -50.0, -50.0]), hi=np.array([51.0, 50.0, 50.0]), bc=(1, 1, 1),
                pos=np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), vel=np.zeros((2, 3)),
                type=np.zeros(2, np.int32), eps=np.ones((1, 1)), sigma=np.ones((1, 1)),
                cutoff_list=None, mass=np.ones(1), lj_cutoff=5.0, nl_cutoff=5.0, nl_coef=1.0523,
                ghost_cutoff=5.0, dt=0.001, integrator=0, k_b=1.0, n_r_df=3)


def grid512():
    # tests/t002_lj.cpp:46-64: 8^3 atoms, spacing 5, velocities uniform(+-0.5) seed 1.
    g = np.arange(-20, 20, 5, dtype=np.float64)
    pos = np.array([[i, j, k] for i in g for j in g for k in g])
    vel = O.ref_random_velocity(len(pos), 1, 0.5)
    s = dimer()
    s.update(pos=pos, vel=vel, type=np.zeros(len(pos), np.int32))
    return s


def molecules(n):
    """Molecule indices: runs of 4 consecutive atoms share one index, every
    fifth run stays outside molecules (-1)."""
    m = np.arange(n) // 4
    return np.where(m % 5 == 4, -1, m).astype(np.int32)


# Lj::verify_settings inputs of tests/cpp/dropin_checks.cpp's ljtables mode
# (12 owned atoms of types i % 3)
LJ_ATOM = dict(eps_atom=(1.0, 0.5, 0.8), sigma_atom=(1.0, 0.88, 1.2), wca=True,
               lambda_s=[[1.0, 0.9], [0.9]],
               lambda_e=[[1.0, 0.5, 1.0], [0.5, 1.0, 1.0], [1.0, 1.0, 0.7]])
LJ_ARRAY = dict(eps_set=[(0, 0, 1.0), (0, 1, 0.5), (1, 0, 0.5), (1, 1, 0.8)],
                sigma_set=[(0, 0, 1.0), (1, 1, 0.9)], cutoff_list=[[2.5]])


def lj_tables():
    """The tables the reference's Lj::verify_settings forms: per-atom input
    with WCA and ragged lambda rows ("atom"), arrays with a short cutoff list
    ("array")."""
    types = np.arange(12) % 3
    d = {"types": types.astype(np.int32)}
    for case, kw in (("atom", LJ_ATOM), ("array", LJ_ARRAY)):
        e, sg, cl = O.ref_lj_tables(types, 3, **kw)
        d[case + "_eps"], d[case + "_sigma"], d[case + "_cutoff_list"] = e, sg, cl
    return "lj_tables", d


def main(only=None):
    O.ref_lib()
    if only:  # regenerate the named fixtures, keep the others and their hashes
        gen = {"mol_traj": lambda: traj_fixture("mol_traj", lj_fluid(5, temperature=1.0, seed=1006), 20,
                                                keep=(1, 10, 20), use_cell_list=True,
                                                mol=molecules(500)),
               "lj_tables": lj_tables}
        mpath = os.path.join(HERE, "golden.json")
        meta = json.load(open(mpath))
        for name in only:
            name_, d = gen[name]()
            path = os.path.join(HERE, name_ + ".npz")
            np.savez_compressed(path, **d)
            meta["files"][name_ + ".npz"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
            print(f"{name_}: {os.path.getsize(path) / 1024:.0f} KiB")
        with open(mpath, "w") as f:
            json.dump(meta, f, indent=1)
        return
    fixtures = []
    c1 = lj_fluid(10, rho=0.8442, temperature=1.44, seed=1001)
    fixtures.append(setup_fixture("c1_setup_brute", c1, use_cell_list=False))
    fixtures.append(setup_fixture("c1_setup_cell", c1, use_cell_list=True))
    fixtures.append(traj_fixture("c1_traj_brute", c1, 100))
    fixtures.append(traj_fixture("c1_traj_cell", c1, 20, keep=(1, 10, 20), use_cell_list=True))
    fixtures.append(traj_fixture("dimer_traj", dimer(), 200, keep=(1, 10, 100, 200)))
    fixtures.append(setup_fixture("ka_setup_cell", ka_mixture(500, seed=1005), use_cell_list=True))
    fixtures.append(traj_fixture("langevin_traj", lj_fluid(4, temperature=1.0, seed=1003,
                                                           integrator=1, gamma=1.0),
                                 20, keep=(1, 10, 20), with_eta=True))
    fixtures.append(traj_fixture("grid512_traj", grid512(), 50, keep=(1, 50)))
    fixtures.append(traj_fixture("mol_traj", lj_fluid(5, temperature=1.0, seed=1006), 20, keep=(1, 10, 20),
                                 use_cell_list=True, mol=molecules(500)))
    fixtures.append(("random_velocity", dict(n=np.asarray(512), seed=np.asarray(1),
                                             amplitude=np.asarray(0.5),
                                             vel=O.ref_random_velocity(512, 1, 0.5))))
    fixtures.append(lj_tables())
    manifest = {}
    for name, d in fixtures:
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **d)
        manifest[name + ".npz"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        print(f"{name}: {os.path.getsize(path) / 1024:.0f} KiB")
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "reference": "caviar-md/caviar-v2 compiled by oracle/Makefile (g++ -O2 -ffp-contract=off)",
                   "files": manifest}, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
