"""Long-run fixtures from the compiled reference: NVE energy drift and
Langevin temperature statistics (4000-atom LJ fluid).

Runs only where the reference tree exists (it drives
``oracle/_ref/libcaviar2_ref.so``).  The reference has no potential energy;
E = KE + PE uses the reference's kinetic energy and the pair energy of the
oracle (the potential consistent with the reference force law) evaluated on
the reference's positions.  Regenerate with

    make -C oracle oracle ref && python tests/golden/make_long_fixtures.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "caviar-v2_amd"))

import oracle as O  # noqa: E402
from cv2md.systems import lj_fluid  # noqa: E402


def potential_energy(s, pos):
    box = O.make_box(s["lo"], s["hi"], s.get("bc", (1, 1, 1)))
    p, _ = O.wrap(pos, box)
    gpos, owner, _code = O.ghosts(p, box, s["ghost_cutoff"])
    offs, nbr = O.nlist_cell(p, gpos, box, s["nl_cutoff"], s["nl_coef"])
    n = len(p)
    ids = np.arange(n)
    typ = np.asarray(s["type"])
    _acc, _vir, pe = O.lj(p, typ, ids, gpos, typ[owner], ids[owner], offs, nbr, s["eps"], s["sigma"],
                          s["lj_cutoff"], s["mass"])
    return pe


def nve(nsteps=10000, every=100):
    s = lj_fluid(10, temperature=1.44, seed=1001)
    s = dict(s, use_cell_list=1, integrator=0, n_r_df=3)
    steps, ke, pe = [], [], []
    with O.RefSim(s) as r:
        r.setup()
        for k in range(nsteps + 1):
            if k % every == 0:
                pos, _vel, _acc = r.state()
                steps.append(k)
                ke.append(r.thermo()["ke"])
                pe.append(potential_energy(s, pos))
            if k < nsteps:
                r.step()
    e = np.asarray(ke) + np.asarray(pe)
    return dict(in_pos=s["pos"], in_vel=s["vel"], in_lo=np.asarray(s["lo"]), in_hi=np.asarray(s["hi"]),
                in_dt=np.asarray(s["dt"]), in_seed=np.asarray(1001), in_temperature=np.asarray(1.44),
                steps=np.asarray(steps), ke=np.asarray(ke), pe=np.asarray(pe), etot=e)


def langevin(nsteps=10000, equil=2000, every=10):
    s = lj_fluid(10, temperature=1.44, seed=1002, integrator=1, gamma=1.0)
    s = dict(s, use_cell_list=1, n_r_df=0)
    temps = []
    with O.RefSim(s) as r:
        r.setup()
        for k in range(equil + nsteps):
            r.step()
            if k >= equil and (k - equil) % every == 0:
                temps.append(r.thermo()["temperature"])
    t = np.asarray(temps)
    return dict(in_pos=s["pos"], in_vel=s["vel"], in_lo=np.asarray(s["lo"]), in_hi=np.asarray(s["hi"]),
                in_dt=np.asarray(s["dt"]), in_gamma=np.asarray(1.0), in_kbt=np.asarray(s["kbt"]),
                in_seed=np.asarray(1002), equil=np.asarray(equil), every=np.asarray(every),
                temperature=t, t_mean=np.asarray(t.mean()), t_var=np.asarray(t.var()))


def main():
    if not O.ref_available():
        sys.exit("the compiled reference (oracle/_ref) is required")
    man_path = os.path.join(HERE, "golden.json")
    man = json.load(open(man_path))
    for name, fn in (("c1_nve_energy", nve), ("c1_langevin_temperature", langevin)):
        t0 = time.time()
        d = fn()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **d)
        man["files"][name + ".npz"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        print(f"{name}: {time.time() - t0:.0f} s", flush=True)
    json.dump(man, open(man_path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
