"""Parity of the HIP path against the oracle, through the C ABI.

Pair sets: bit-exact, as (i, j, image) keys in the reference's half-list
convention (owned j > i plus every ghost image in range).
Forces: FP64 path within 1e-12 (relative to max(|a_i|, rms|a|, 1)) --
the only difference from the oracle is summation order; FP32+fixed-point
path within the tolerance stated in FP32_TOL.
"""
import numpy as np
import pytest

import oracle as O
from cv2md import device as D
from cv2md.systems import ka_mixture, lj_fluid
from helpers import force_error, oracle_forces, oracle_pairs, sort_rows

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
# FP32 pair arithmetic, any kernel, small systems: per-atom |da| / max(|a|, rms|a|)
FP32_TOL_MAX = 2e-5
FP32_TOL_RMS = 5e-6


def perturbed(s, amp, seed):
    s = dict(s)
    rng = np.random.default_rng(seed)
    s["pos"] = s["pos"] + rng.normal(0.0, amp, s["pos"].shape)
    return s


def t001_like():
    # ghost cutoff (2.8) below the list cutoff (2.5 * 1.2 = 3.0): images whose
    # owner is between 2.8 and 3.0 from a face are absent, as in the reference
    s = perturbed(lj_fluid(6, seed=5), 0.08, 5)
    s["nl_coef"] = 1.2
    s["ghost_cutoff"] = 2.8
    return s


def open_z():
    s = perturbed(lj_fluid(6, seed=6), 0.08, 6)
    s["bc"] = (1, 1, 0)
    return s


CASES = {
    "c1_fcc": lambda: lj_fluid(10),
    "c1_perturbed": lambda: perturbed(lj_fluid(10), 0.1, 1),
    "ka_500": lambda: ka_mixture(500, seed=3),
    "ka_4000": lambda: ka_mixture(4000, seed=4),
    "small_box_2cells": lambda: perturbed(lj_fluid(4), 0.05, 2),
    "ghost_below_list_cutoff": t001_like,
    "open_z": open_z,
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    return request.param, CASES[request.param]()


def test_pair_set_bit_exact(case):
    name, s = case
    print(f"[pairs] {name} n={len(s['pos'])}", flush=True)
    with D.Context.from_sysdef(s) as c:
        c.build(True)
        got = sort_rows(c.pairs())
        st = c.stats()
    exp = oracle_pairs(s["pos"], s)
    assert st["max_neighbors"] <= st["capacity"]
    assert len(got) == len(exp)
    assert np.array_equal(got, exp)


def wave_ghost_below_cutoff():
    # cell-wave build (5 cells of width >= 3.0 per axis) with images whose
    # ghost may not exist: those candidates always take the FP64 path
    s = perturbed(lj_fluid(10, seed=7), 0.08, 7)
    s["nl_coef"] = 1.2
    s["ghost_cutoff"] = 2.8
    return s


def wave_dense_cells():
    # about 65 atoms per cell and 1750 candidates per cell: several home-atom
    # chunks and several candidate passes (unstaged stores) per wavefront
    rng = np.random.default_rng(11)
    s = lj_fluid(4)
    L = 5.2 * 2.8
    n = int(2.6 * L ** 3)
    s["pos"] = rng.uniform(0.0, L, (n, 3))
    s["vel"] = np.zeros((n, 3))
    s["type"] = np.zeros(n, np.int64)
    s["lo"], s["hi"] = (0.0, 0.0, 0.0), (L, L, L)
    return s


WAVE_CASES = {
    "c1_fcc": lambda: lj_fluid(10),
    "ka_4000": lambda: ka_mixture(4000, seed=4),
    "wave_ghost_below_cutoff": wave_ghost_below_cutoff,
    "wave_dense_cells": wave_dense_cells,
}


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
@pytest.mark.parametrize("name", sorted(WAVE_CASES))
def test_pair_set_cell_wave_build(name, prec):
    """Systems that take the wavefront-per-cell build: bit-exact pair sets,
    from the 32-bit list (FP64) and from the 16-bit task-slot list (FP32)."""
    s = WAVE_CASES[name]()
    print(f"[pairs/cell-wave] {name} prec={prec} n={len(s['pos'])}", flush=True)
    with D.Context.from_sysdef(s, precision=prec) as c:
        c.build(True)
        st = c.stats()
        got = sort_rows(c.pairs())
    assert st["build_kind"] == "cell_wave"
    exp = oracle_pairs(s["pos"], s)
    assert st["max_neighbors"] <= st["capacity"]
    assert len(got) == len(exp)
    assert np.array_equal(got, exp)


def test_per_atom_build_in_use_for_small_boxes():
    with D.Context.from_sysdef(ka_mixture(500, seed=3)) as c:
        c.build(True)
        assert c.stats()["build_kind"] == "per_atom"


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_forces_virial_energy(case, prec):
    name, s = case
    print(f"[forces] {name} prec={prec}", flush=True)
    with D.Context.from_sysdef(s, precision=prec) as c:
        c.compute_forces()
        pos, _, acc = c.download()
        th = c.thermo()
    p, a_ref, vir, pe = oracle_forces(s["pos"], s)
    assert np.array_equal(pos, p)
    scale = max(float(np.sqrt((np.linalg.norm(a_ref, axis=1) ** 2).mean())), 1.0)
    d = np.linalg.norm(acc - a_ref, axis=1) / np.maximum(np.linalg.norm(a_ref, axis=1), scale)
    emax, erms = float(d.max()), float(np.sqrt((d ** 2).mean()))
    print(f"  force error max {emax:.3e} rms {erms:.3e}; virial {th['virial']:.9e} vs {vir:.9e}; "
          f"pe {th['pe']:.9e} vs {pe:.9e}")
    esc = max(abs(pe), float(len(p)))
    vsc = max(abs(vir), float(len(p)))
    if prec == D.CV2_PREC_FP64:
        assert emax < FP64_TOL
        # global sums over ~1e5 pairs in a different order: 1e-10 relative
        assert abs(th["virial"] - vir) < 1e-10 * vsc
        assert abs(th["pe"] - pe) < 1e-10 * esc
    else:
        assert emax < FP32_TOL_MAX and erms < FP32_TOL_RMS
        assert abs(th["virial"] - vir) < 1e-5 * vsc
        assert abs(th["pe"] - pe) < 1e-5 * esc


@pytest.mark.parametrize("name,mid", [("c1_fcc", "1000"), ("ka_4000", "1440")])
def test_fp32_unstaged_tasks_identical(name, mid, monkeypatch):
    """Force tasks whose neighbourhood exceeds the LDS staging limit read
    their neighbours from global memory instead: the same bits (limit 0: every
    task; mid: about the mean staged count, so some tasks of each kind)."""
    s = WAVE_CASES[name]()
    out = {}
    for lim in ("3072", "0", mid):
        monkeypatch.setenv("CV2_STAGE_LIMIT", lim)
        print(f"[stage-limit] {name} limit={lim}", flush=True)
        with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP32_FIXED) as c:
            c.compute_forces()
            th = c.thermo()
            out[lim] = (c.download()[2], th["virial"], th["pe"])
    monkeypatch.delenv("CV2_STAGE_LIMIT")
    for lim in ("0", mid):
        assert np.array_equal(out[lim][0], out["3072"][0])
        assert out[lim][1:] == out["3072"][1:]


@pytest.mark.parametrize("fuse", ["1", "0"])
def test_task_tail_launch_identical(fuse, monkeypatch):
    """Force tasks beyond the main launch's grid go to the tail launch
    (CV2_TASK_GRID caps the grid at 8 work-groups, one per XCD: all but 8 of
    the 4000-atom liquid's tasks run there); trajectories, energy and virial carry the
    same bits as with the full grid, across rebuilds."""
    s = perturbed(lj_fluid(10, temperature=1.44, seed=14), 0.05, 14)
    monkeypatch.setenv("CV2_FUSE", fuse)
    out = {}
    for cap in ("0", "8"):
        monkeypatch.setenv("CV2_TASK_GRID", cap)
        print(f"[tail] fuse={fuse} cap={cap}", flush=True)
        with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP32_FIXED) as c:
            c.compute_forces()
            c.run_langevin(60, 0.005, 1.0, 1.44, 5)
            c.compute_forces()
            th = c.thermo()
            st = c.stats()
            out[cap] = (c.download(), th, st["rebuilds"], st["force_tasks"])
    monkeypatch.delenv("CV2_TASK_GRID")
    monkeypatch.delenv("CV2_FUSE")
    (p1, v1, a1), th1, r1, nt1 = out["0"]
    (p0, v0, a0), th0, r0, nt0 = out["8"]
    assert nt1 == nt0 and nt1 > 8 and r1 == r0 and r1 > 1
    assert np.array_equal(p1, p0) and np.array_equal(v1, v0) and np.array_equal(a1, a0)
    assert th1 == th0


@pytest.mark.parametrize("langevin", [False, True])
def test_fused_integrator_identical(langevin, monkeypatch):
    """The FP32 task kernel with the integrator in its epilogue gives the
    same bits as separate force and integrator kernels, across rebuilds."""
    s = perturbed(lj_fluid(10, temperature=1.44, seed=12), 0.05, 12)
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("CV2_FUSE", fuse)
        print(f"[fused] langevin={langevin} fuse={fuse}", flush=True)
        with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP32_FIXED) as c:
            c.compute_forces()
            if langevin:
                c.run_langevin(120, 0.005, 1.0, 1.44, 31)
            else:
                c.run_vv(120, 0.005)
            th = c.thermo()
            out[fuse] = (c.download(), th, c.stats()["rebuilds"])
    monkeypatch.delenv("CV2_FUSE")
    (p1, v1, a1), th1, r1 = out["1"]
    (p0, v0, a0), th0, r0 = out["0"]
    assert r1 == r0 and r1 > 1
    assert np.array_equal(p1, p0) and np.array_equal(v1, v0) and np.array_equal(a1, a0)
    assert th1 == th0


def test_pairs_after_dynamics_bit_exact():
    """Pair set at a later rebuild, as a function of the device's positions."""
    s = perturbed(lj_fluid(8, temperature=1.44, seed=9), 0.05, 9)
    with D.Context.from_sysdef(s) as c:
        c.run_vv(200, 0.005)
        c.build(True)
        pos = c.download()[0]
        got = sort_rows(c.pairs())
        assert c.stats()["rebuilds"] >= 3
    s2 = dict(s, pos=pos)
    assert np.array_equal(got, oracle_pairs(pos, s2))


def test_short_trajectory_matches_oracle():
    s = perturbed(lj_fluid(10, temperature=1.44, seed=11), 0.05, 11)
    s["use_cell_list"] = 1
    with O.OracleSim(s) as o, D.Context.from_sysdef(s) as c:
        o.setup()
        for _ in range(10):
            o.step()
        c.run_vv(10, s["dt"])
        po, vo, _ = o.state()
        pg, vg, _ = c.download()
    # the reference wraps every step, the device at list builds: compare modulo L
    L = s["hi"] - s["lo"]
    d = pg - po
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-10
    assert np.abs(vg - vo).max() < 1e-8


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_deterministic(prec):
    s = perturbed(lj_fluid(8, seed=12), 0.05, 12)
    out = []
    for _ in range(2):
        with D.Context.from_sysdef(s, precision=prec) as c:
            c.run_vv(100, 0.005)
            out.append(c.download())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_energy_conservation_short():
    s = lj_fluid(8, temperature=1.44, seed=13)
    with D.Context.from_sysdef(s) as c:
        c.compute_forces()
        t0 = c.thermo()
        c.run_vv(2000, 0.005)
        t1 = c.thermo()
    e0 = t0["ke"] + t0["pe"]
    e1 = t1["ke"] + t1["pe"]
    print(f"E0 {e0:.6f} E1 {e1:.6f} per atom drift {(e1 - e0) / len(s['pos']):.3e}")
    assert abs(e1 - e0) / len(s["pos"]) < 2e-3


def test_host_halo_matches_periodic():
    """One domain's share with a host-supplied halo: the ghosts are the
    reference's periodic images, so forces must equal the periodic answer."""
    s = ka_mixture(4000, seed=21)
    box = O.make_box(s["lo"], s["hi"], s["bc"])
    p, _ = O.wrap(s["pos"], box)
    gpos, owner, _ = O.ghosts(p, box, s["ghost_cutoff"])
    t = np.asarray(s["type"], np.int32)
    with D.Context.from_sysdef(dict(s, pos=p)) as c:
        c.set_ghosts(gpos, t[owner], owner)
        c.compute_forces()
        acc = c.download()[2]
        th = c.thermo()
        pairs = sort_rows(c.pairs())
    offs, nbr = O.nlist_brute(p, gpos, s["nl_cutoff"], s["nl_coef"])
    n = len(p)
    exp = sort_rows(np.stack([np.repeat(np.arange(n), np.diff(offs)), nbr.astype(np.int64),
                              np.full(len(nbr), 13)], 1))
    assert np.array_equal(pairs, exp)
    _, a_ref, vir, pe = oracle_forces(s["pos"], s)
    assert np.abs(acc - a_ref).max() < 1e-12 * max(np.abs(a_ref).max(), 1.0)
    assert abs(th["virial"] - vir) < 1e-12 * abs(vir)


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
@pytest.mark.parametrize("mix", [False, True])
def test_jump_fix_forces(prec, mix):
    """jump_fix on overlapping random positions (a relaxation start): forces
    and virial against the oracle with the same clamp, one species (the
    dispatch then takes the species-table kernels) and the KA mixture."""
    rng = np.random.default_rng(41)
    if mix:
        s = ka_mixture(3000, seed=42)
    else:
        s = lj_fluid(8, seed=43)
    L = s["hi"] - s["lo"]
    s = dict(s, pos=rng.uniform(0.0, 1.0, s["pos"].shape) * L + s["lo"], jump_tol=0.64)
    _, a_ref, vir, _pe = oracle_forces(s["pos"], s)
    with D.Context.from_sysdef(s, precision=prec) as c:
        c.set_jump_tol(s["jump_tol"])
        c.compute_forces()
        acc = c.download()[2]
        th = c.thermo()
    emax, erms = force_error(acc, a_ref)
    print(f"[jump] mix={mix} prec={prec}: max {emax:.2e} rms {erms:.2e}, |a| max {np.abs(a_ref).max():.1f}")
    if prec == D.CV2_PREC_FP64:
        assert emax < FP64_TOL
        assert abs(th["virial"] - vir) < 1e-10 * abs(vir)
    else:
        assert emax < FP32_TOL_MAX and erms < FP32_TOL_RMS
        assert abs(th["virial"] - vir) < 1e-5 * abs(vir)


def test_nve_rescale_trajectory_matches_oracle():
    """Velocity Verlet with the velocity-rescale constraint (constraint::Nve,
    kbt = 1.1 every 5 steps) through the C ABI: the device's kinetic energy
    gives lambda (nve.cpp:165-172) and cv2_scale_velocities applies it, as
    Md_simulator does at the constraint's steps; the oracle runs the same
    steps with its restatement.  Every lambda and the final state agree, and
    the kinetic energy after each rescale is dof * kbt / 2."""
    s = perturbed(lj_fluid(10, temperature=1.44, seed=13), 0.05, 13)
    s["use_cell_list"] = 1
    n = len(s["pos"])
    dof = 3 * n - 3
    with O.OracleSim(s) as o, D.Context.from_sysdef(s) as c:
        o.setup()
        step = 0
        while step < 20:
            # steps up to the next multiple of 5, then the rescale
            k = 1 if step == 0 else 5 - (step - 1) % 5
            for _ in range(k):
                o.step()
            c.run_vv(k, s["dt"])
            step += k
            lo = O.nve_lambda(o.thermo()["ke"], dof, kbt=1.1)
            lg = O.nve_lambda(c.thermo()["ke"], dof, kbt=1.1)
            print(f"[nve] after step {step - 1}: lambda {lg:.15f} (oracle {lo:.15f})", flush=True)
            assert abs(lg - lo) < 1e-12
            o.scale_velocities(lo)
            c.scale_velocities(lg)
            assert c.thermo()["ke"] == pytest.approx(0.5 * dof * 1.1, rel=1e-13)
        po, vo, _ = o.state()
        pg, vg, _ = c.download()
    L = s["hi"] - s["lo"]
    d = pg - po
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-10
    assert np.abs(vg - vo).max() < 1e-8
