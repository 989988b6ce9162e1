"""Lj::ignore_intra_molecule on the device (lj.cpp:446), through the C ABI.

Pairs whose two molecule indices are equal (two atoms outside molecules
included, -1 == -1) add no force, virial or energy; the pair set is
unchanged.  Checked against the reference's own trajectory with the flag on
(tests/golden/mol_traj.npz, written by the compiled reference) and against
the oracle at dynamic positions.
"""
import numpy as np
import pytest

from conftest import sysdef_from
from cv2md import device as D
from helpers import force_error, oracle_forces, oracle_pairs, sort_rows

pytestmark = pytest.mark.gpu

# FP64: summation order only; FP32 + fixed point: the bar of the LJ liquid
FP64_TOL = 1e-12
FP32_TOL = 1e-6


def _ctx(s, prec):
    c = D.Context.from_sysdef(s, precision=prec)
    c.set_molecules(s["mol"], True)
    return c


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_forces_match_reference_step(golden, prec):
    """Accelerations at the reference's positions after step 10 equal its own
    accelerations there (FP64 within summation order)."""
    d = golden("mol_traj")
    s = sysdef_from(d)
    s["pos"] = d["pos_10"]
    print(f"[mol] forces prec={prec} n={len(s['pos'])}", flush=True)
    with _ctx(s, prec) as c:
        c.compute_forces()
        acc = c.download()[2]
        pairs = sort_rows(c.pairs())
    emax, _ = force_error(acc, d["acc_10"])
    assert emax < (FP64_TOL if prec == D.CV2_PREC_FP64 else FP32_TOL), emax
    assert np.array_equal(pairs, oracle_pairs(s["pos"], s)), "the flag must not change the pair set"


def test_trajectory_matches_reference(golden):
    """20 velocity-Verlet steps with the flag on, from the reference's start:
    positions and velocities within 1e-10 of the reference's (FP64)."""
    d = golden("mol_traj")
    s = sysdef_from(d)
    print("[mol] trajectory", flush=True)
    with _ctx(s, D.CV2_PREC_FP64) as c:
        c.compute_forces()
        c.run_vv(20, s["dt"])
        p, v, a = c.download()
    L = s["hi"] - s["lo"]
    dp = p - d["pos_20"]
    dp -= L * np.round(dp / L)
    assert np.abs(dp).max() < 1e-10
    assert np.abs(v - d["vel_20"]).max() < 1e-10
    assert force_error(a, d["acc_20"])[0] < 1e-10


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_energy_and_virial_skip_molecule_pairs(golden, prec):
    """Virial and potential energy leave out the same pairs as the forces."""
    d = golden("mol_traj")
    s = sysdef_from(d)
    s["pos"] = d["pos_20"]
    _, a_ref, vir_ref, pe_ref = oracle_forces(s["pos"], s)
    with _ctx(s, prec) as c:
        c.compute_forces()
        th = c.thermo()
    tol = 1e-10 if prec == D.CV2_PREC_FP64 else 1e-5
    assert abs(th["virial"] - vir_ref) <= tol * max(1.0, abs(vir_ref))
    assert abs(th["pe"] - pe_ref) <= tol * max(1.0, abs(pe_ref))


def test_flag_off_and_refusals(golden):
    """Flag off restores the plain forces; a host-supplied halo is refused
    while the flag is on (its ghosts carry no molecule indices)."""
    d = golden("mol_traj")
    s = sysdef_from(d)
    s["pos"] = d["pos_10"]
    plain = dict(s)
    plain.pop("mol")
    _, a_plain, _, _ = oracle_forces(s["pos"], plain)
    with _ctx(s, D.CV2_PREC_FP64) as c:
        c.set_molecules(None, False)
        c.compute_forces()
        assert force_error(c.download()[2], a_plain)[0] < FP64_TOL
        c.set_molecules(s["mol"], True)
        with pytest.raises(D.Cv2Error) as e:
            c.set_ghosts(np.zeros((1, 3)))
        assert e.value.code == 6  # CV2_ERR_STATE
        with pytest.raises(D.Cv2Error) as e:
            c.set_molecules(s["mol"][:-1], True)
        assert e.value.code == 2  # CV2_ERR_BAD_ARG
