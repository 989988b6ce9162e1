"""Long runs on the GPU against the compiled reference's own long runs
(fixtures from tests/golden/make_long_fixtures.py, 4000-atom LJ fluid).

NVE: the total energy E = KE + PE over 10^4 velocity-Verlet steps must drift
no more than the reference's does.  Chaotic trajectories differ after a few
hundred steps, so only the statistics of E(t) (sampled every 100 steps) are
compared, with allowances for their sampling noise: the standard deviation
of E(t) per atom within 1.5x the reference's, max |E(t) - E(0)| per atom
within 2x, and the least-squares slope per atom within 1.5x plus 5e-9 per
step (about five standard errors of the reference's slope estimate).

Langevin: the kinetic temperature over 10^4 steps (after 2000 of
equilibration, sampled every 10 steps) must match the reference's mean and
variance within the statistical error of the two series (block averages of
the reference series set the error; tolerance 5 standard errors).
"""
import os

import numpy as np
import pytest

from cv2md import device as D
from cv2md.systems import lj_fluid

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name + ".npz"))


def _drift(steps, e, n):
    de = np.abs(e - e[0]).max() / n
    slope = np.polyfit(steps, e, 1)[0] / n
    return de, abs(slope), e.std() / n


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_nve_drift_not_worse_than_reference(prec):
    f = _load("c1_nve_energy")
    s = lj_fluid(10, temperature=1.44, seed=1001)
    assert np.array_equal(s["pos"], f["in_pos"]) and np.array_equal(s["vel"], f["in_vel"])
    n = len(s["pos"])
    steps = f["steps"]
    every = int(steps[1] - steps[0])
    print(f"[nve] prec={prec} steps={steps[-1]}", flush=True)
    es = []
    with D.Context.from_sysdef(s, precision=prec) as c:
        c.compute_forces()
        for k in range(len(steps)):
            if k:
                c.run_vv(every, s["dt"])
            t = c.thermo(n_r_df=3)
            es.append(t["ke"] + t["pe"])
    e = np.asarray(es)
    ref_de, ref_slope, ref_sd = _drift(steps, f["etot"], n)
    de, slope, sd = _drift(steps, e, n)
    print(f"E0 gpu {e[0]:.8f} ref {f['etot'][0]:.8f}; max|dE|/N gpu {de:.3e} ref {ref_de:.3e}; "
          f"sd/N gpu {sd:.3e} ref {ref_sd:.3e}; slope/N gpu {slope:.3e} ref {ref_slope:.3e}")
    # same initial state: FP64 energies differ only in summation order; the
    # FP32 pair energies carry FP32 rounding (measured 6e-7 per atom here,
    # 1.7e-7 of |E|/N)
    assert abs(e[0] - f["etot"][0]) / n < (1e-9 if prec == D.CV2_PREC_FP64 else 2e-6)
    assert sd <= 1.5 * ref_sd
    assert de <= 2.0 * ref_de
    assert slope <= 1.5 * ref_slope + 5e-9


def _block_se(x, nblocks=20):
    b = np.array_split(x, nblocks)
    m = np.array([v.mean() for v in b])
    return m.std(ddof=1) / np.sqrt(nblocks)


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_langevin_temperature_matches_reference(prec):
    f = _load("c1_langevin_temperature")
    s = lj_fluid(10, temperature=1.44, seed=1002, integrator=1, gamma=1.0)
    assert np.array_equal(s["pos"], f["in_pos"])
    n = len(s["pos"])
    equil, every = int(f["equil"]), int(f["every"])
    nsamp = len(f["temperature"])
    print(f"[langevin] prec={prec} samples={nsamp}", flush=True)
    temps = []
    with D.Context.from_sysdef(s, precision=prec) as c:
        c.compute_forces()
        c.run_langevin(equil, s["dt"], s["gamma"], s["kbt"], 11)
        for _ in range(nsamp):
            c.run_langevin(every, s["dt"], s["gamma"], s["kbt"], 11)
            temps.append(2.0 * c.thermo(n_r_df=0)["ke"] / (3.0 * n))
    t = np.asarray(temps)
    rt = f["temperature"]
    se = np.hypot(_block_se(t), _block_se(rt))
    var_se = np.hypot(_block_se((t - t.mean()) ** 2), _block_se((rt - rt.mean()) ** 2))
    print(f"<T> gpu {t.mean():.5f} ref {rt.mean():.5f} (se {se:.5f}); "
          f"Var T gpu {t.var():.3e} ref {rt.var():.3e} (se {var_se:.2e}); target {s['kbt']}")
    assert abs(t.mean() - rt.mean()) < 5 * se
    assert abs(t.var() - rt.var()) < 5 * var_se
