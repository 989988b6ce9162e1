"""cv2_thermo's potential energy and virial after force passes that did not
sum them (the steps of a run, cv2_compute_accelerations, cv2_step_end with
energy = 0): an energy pass at the current positions sums them on demand.

- After a Langevin run the values equal the oracle's at the positions the
  run ended on (lj.cpp:393-517; 1e-10 relative in FP64, 1e-5 in the FP32 +
  fixed-point form, the bounds of the force-pass parity tests).
- The energy pass leaves the state's accelerations bit for bit as the step
  loop left them, and the run continues with the same bits as a run that
  never asked (thermo between two runs changes nothing).
- While a step is open the last summed values are reported unchanged.
- A run split into several runs gives the bits of one run (each run closes
  its last step and the next opens the following one).
"""
import numpy as np
import pytest

from cv2md import device as D
from cv2md.systems import lj_fluid
from helpers import oracle_forces

pytestmark = pytest.mark.gpu

PRECS = {"fp64": (D.CV2_PREC_FP64, 1e-10), "fp32": (D.CV2_PREC_FP32_FIXED, 1e-5)}


@pytest.mark.parametrize("prec", sorted(PRECS))
def test_energy_on_demand_after_run(prec):
    p_id, tol = PRECS[prec]
    s = lj_fluid(10, temperature=1.2, seed=91, integrator=1, gamma=1.0)
    finals = []
    for ask in (True, False):
        with D.Context.from_sysdef(s, precision=p_id) as c:
            c.compute_forces()
            c.run_langevin(30, s["dt"], 1.0, 1.2, 3)
            pos, vel, acc0 = c.download()
            if ask:
                th = c.thermo()
                assert np.array_equal(c.download()[2], acc0)  # accelerations untouched
                assert c.thermo() == th  # summed once, then kept
            c.run_langevin(30, s["dt"], 1.0, 1.2, 4)
            finals.append(c.download())
    for a, b in zip(*finals):
        assert np.array_equal(a, b)
    _, _, vir, pe = oracle_forces(pos, s)
    n = len(pos)
    print(f"[thermo] {prec}: pe {th['pe']:.12e} vs {pe:.12e}, virial {th['virial']:.12e} vs {vir:.12e}",
          flush=True)
    assert abs(th["pe"] - pe) < tol * max(abs(pe), n)
    assert abs(th["virial"] - vir) < tol * max(abs(vir), n)


def test_open_step_reports_last_values():
    s = lj_fluid(8, temperature=1.2, seed=92)
    with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP64) as c:
        c.compute_forces()
        th0 = c.thermo()
        c.step_begin(s["dt"])
        th1 = c.thermo()
        assert th1["pe"] == th0["pe"] and th1["virial"] == th0["virial"]
        c.step_end(energy=False)
        th2 = c.thermo()  # summed at the step's new positions
        _, _, vir, pe = oracle_forces(c.download()[0], s)
        assert abs(th2["pe"] - pe) < 1e-10 * max(abs(pe), len(s["pos"]))
        assert th2["pe"] != th0["pe"]


@pytest.mark.parametrize("prec", sorted(PRECS))
def test_runs_in_pieces_give_the_same_bits(prec):
    # a run closes its last step (on the FP32 task path inside the fused
    # force kernel) and the next run opens the following one: 1 + 1 + 3 + 7
    # + 12 steps in five runs give the bits of one 24-step run, rebuilds
    # included
    p_id, _ = PRECS[prec]
    s = lj_fluid(10, temperature=1.2, seed=93, integrator=1, gamma=1.0)
    out = []
    for pieces in ((24,), (1, 1, 3, 7, 12)):
        with D.Context.from_sysdef(s, precision=p_id) as c:
            c.compute_forces()
            for k in pieces:
                c.run_langevin(k, s["dt"], 1.0, 1.2, 11)
            out.append((c.download(), c.stats()["rebuilds"]))
    assert out[0][1] == out[1][1] and out[0][1] > 1
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
