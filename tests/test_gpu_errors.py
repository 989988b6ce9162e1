"""Error paths, the list capacity and the rebuild trigger, through the C ABI.

- A NaN position or velocity ends a force pass or run with CV2_ERR_NAN at
  the step it appeared (md_simulator.cpp:646-652 stops the program there);
  overlapping atoms in the FP32 form give CV2_ERR_FIXED_OVERFLOW; bad
  arguments give CV2_ERR_BAD_ARG with the reference's messages where it has
  them, calls out of order CV2_ERR_STATE.  These are device flags, not
  faults: every kernel still exits normally.
- A list that outgrows its capacity (a dense cluster at the first build, or
  lists that grow during a run) is regrown and the build redone; in a run the
  device halts at that step and the host resumes from it, so the trajectory
  is the one a large enough capacity gives, bit for bit.
- The rebuild test of Neighborlist::update_is_needed (neighborlist.cpp:
  100-148): a displacement just above half the skin rebuilds, just below
  does not; the rebuild-every-step mode (:107-110); and the reference's
  wrap / ghost-band triggers (CV2_NL_REFERENCE_REBUILDS) give the oracle's
  rebuild steps.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cv2md import device as D
from cv2md.systems import lj_fluid
from helpers import force_error, oracle_forces, oracle_pairs, sort_rows

pytestmark = pytest.mark.gpu


def code_of(fn):
    try:
        fn()
    except D.Cv2Error as e:
        return e.code, str(e)
    return 0, ""


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_nan_position_is_reported(prec):
    s = lj_fluid(10)
    s["pos"] = s["pos"].copy()
    s["pos"][17, 1] = np.nan
    with D.Context.from_sysdef(s, precision=prec) as c:
        code, msg = code_of(c.compute_forces)
    print(f"[nan] prec={prec}: {msg}", flush=True)
    assert code == 3 and "NaN" in msg


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_nan_velocity_halts_the_run_at_its_step(prec):
    s = lj_fluid(10)
    s["vel"] = s["vel"].copy()
    s["vel"][5, 0] = np.nan
    with D.Context.from_sysdef(s, precision=prec) as c:
        c.compute_forces()
        code, msg = code_of(lambda: c.run_vv(50, 0.005))
    print(f"[nan-run] prec={prec}: {msg}", flush=True)
    assert code == 3 and "step 0 of the run" in msg


def test_overlapping_atoms_fixed_point_overflow():
    s = lj_fluid(10)
    s["pos"] = s["pos"].copy()
    s["pos"][1] = s["pos"][0] + np.array([0.3, 0.0, 0.0])
    with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP32_FIXED) as c:
        code, msg = code_of(c.compute_forces)
    assert code == 4 and "fixed-point" in msg


def test_bad_arguments_and_state():
    s = lj_fluid(6)
    L = D.lib()
    with D.Context() as c:
        lo = np.zeros(3)
        hi = np.array([5.0, -1.0, 5.0])
        bc = np.ones(3, np.int32)
        assert L.cv2_set_box(c.h, D._p(lo, D._d), D._p(hi, D._d), D._p(bc, D._i32), 1.0) == 2
        assert b"upper must exceed lower" in L.cv2_last_error(c.h)
        assert code_of(lambda: c.set_neighbor(2.5, 0.9))[0] == 2  # cutoff_extra_coef < 1
        assert "cutoff_extra_coef must be larger than 1.0" in code_of(lambda: c.set_neighbor(2.5, 0.9))[1]
        assert code_of(lambda: c.set_lj(np.ones((9, 9)), np.ones((9, 9)), 2.5, np.ones(9)))[0] == 2
        assert code_of(lambda: c.set_lj(1.0, 1.0, -1.0, 1.0))[1].endswith("Force field cutoff must be non-negative.")
        assert code_of(c.compute_forces)[0] == 6  # nothing uploaded yet
    with D.Context.from_sysdef(s) as c:
        t = np.zeros(len(s["pos"]), np.int32)
        t[3] = 4  # one species only
        assert code_of(lambda: c.upload(s["pos"], s["vel"], t))[0] == 2
    with D.Context.from_sysdef(s) as c:
        assert L.cv2_run_vv(c.h, -1, 0.005) == 2
        assert code_of(lambda: c.run_langevin(10, 0.005, 1.0, -1.0, 1))[1].endswith("kbt have to non-negative.")


def dense_cluster():
    """A 4000-atom liquid with 300 of its atoms packed into a cube of side
    2.6 (about 17 atoms per sigma^3): their lists hold every other atom of
    the cluster and exceed the capacity chosen from the mean density."""
    s = lj_fluid(10, seed=71)
    rng = np.random.default_rng(71)
    pos = s["pos"].copy()
    pos[:300] = 4.0 + rng.uniform(0.0, 2.6, (300, 3))
    return dict(s, pos=pos)


def test_dense_cluster_lists_regrow():
    s = dense_cluster()
    with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP64) as c:
        cap0 = c.stats()["capacity"]
        c.compute_forces()
        st = c.stats()
        got = sort_rows(c.pairs())
        acc = c.download()[2]
    print(f"[regrow] capacity {cap0} -> {st['capacity']}, longest list {st['max_neighbors']}", flush=True)
    assert st["max_neighbors"] > cap0 and st["capacity"] >= st["max_neighbors"]
    assert np.array_equal(got, oracle_pairs(s["pos"], s))
    _, a_ref, _, _ = oracle_forces(s["pos"], s)
    assert force_error(acc, a_ref)[0] < 1e-12


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_lists_outgrowing_capacity_in_a_run_resume(prec, monkeypatch):
    """Capacity 80 fits the FCC start (78 neighbours each) but not the melted
    liquid: some build inside the run halts, the host grows the lists and
    resumes at that step.  The result equals a run with the default
    capacity bit for bit, with the same number of builds."""
    s = lj_fluid(10, temperature=1.44, seed=72)
    out = {}
    for cap in (None, "80"):
        if cap:
            monkeypatch.setenv("CV2_LIST_CAP", cap)
        with D.Context.from_sysdef(s, precision=prec) as c:
            c.compute_forces()
            cap0 = c.stats()["capacity"]
            c.run_vv(300, 0.005)
            st = c.stats()
            out[cap] = (c.download(), c.thermo(), st["rebuilds"], cap0, st["capacity"], st["max_neighbors"])
        monkeypatch.delenv("CV2_LIST_CAP", raising=False)
    (pa, va, aa), tha, ra, _, _, _ = out[None]
    (pb, vb, ab), thb, rb, c0, c1, mx = out["80"]
    print(f"[resume] prec={prec}: capacity {c0} -> {c1} (longest {mx}), builds {ra} / {rb}", flush=True)
    assert c0 == 80 and c1 > 80 and mx > 80
    assert ra == rb
    assert np.array_equal(pa, pb) and np.array_equal(va, vb) and np.array_equal(aa, ab)
    assert tha == thb


def _one_mover(delta_factor):
    """FCC lattice at rest (forces vanish by symmetry), atom 0 moving along x
    so that one velocity-Verlet step displaces it by delta_factor times half
    the skin (the rebuild threshold, neighborlist.cpp:318)."""
    s = lj_fluid(10)
    ce = s["nl_cutoff"] * s["nl_coef"]
    thresh_sq = 0.25 * (ce - s["nl_cutoff"]) ** 2
    v = np.zeros_like(s["pos"])
    v[0, 0] = delta_factor * np.sqrt(thresh_sq) / s["dt"]
    return dict(s, vel=v), thresh_sq


@pytest.mark.parametrize("factor", [1.0 - 1e-6, 1.0 + 1e-6])
def test_rebuild_trigger_at_half_skin(factor):
    s, thresh_sq = _one_mover(factor)
    with D.Context.from_sysdef(s) as c:
        c.compute_forces()
        r0 = c.stats()["rebuilds"]
        c.run_vv(1, s["dt"])
        p = c.download()[0]
        r1 = c.stats()["rebuilds"]
    d = p[0] - s["pos"][0]
    L = s["hi"] - s["lo"]
    d -= L * np.round(d / L)  # the atom starts on a face: its copy may come back wrapped
    d2 = float((d ** 2).sum())
    expect = d2 > thresh_sq  # update_is_needed: disp * disp > threshold_distance_sq
    print(f"[trigger] factor {factor}: d^2/thresh {d2 / thresh_sq:.9f}, rebuilt {r1 - r0}", flush=True)
    assert (r1 - r0 == 1) == expect
    assert expect == (factor > 1.0)


def test_rebuild_every_step_mode():
    s = lj_fluid(8)
    with D.Context.from_sysdef(s, rebuild_every_step=True) as c:
        c.compute_forces()
        r0 = c.stats()["rebuilds"]
        c.run_vv(25, s["dt"])
        assert c.stats()["rebuilds"] - r0 == 25


def test_reference_rebuild_steps_match_oracle():
    """With CV2_NL_REFERENCE_REBUILDS the device rebuilds at the steps the
    reference does (half skin, a wrap through a periodic face, a changed
    single-domain ghost set) on a 4000-atom liquid, step by step."""
    s = lj_fluid(10, temperature=1.44, seed=73)
    rng = np.random.default_rng(73)
    s["pos"] = s["pos"] + rng.normal(0.0, 0.05, s["pos"].shape)
    with O.OracleSim(dict(s, use_cell_list=1)) as o, D.Context.from_sysdef(s, reference_rebuilds=True) as c:
        o.setup()
        c.compute_forces()
        dev, ref = [], []
        for _ in range(40):
            r0 = c.stats()["rebuilds"]
            c.run_vv(1, s["dt"])
            dev.append(c.stats()["rebuilds"] - r0)
            ref.append(int(o.step()))
    print(f"[ref-rebuilds] device {dev}\n               oracle {ref}", flush=True)
    assert dev == ref and sum(ref) > 0
