"""Pins for the oracle (the CPU restatement the HIP path is checked against).

1. Bit-exact against golden fixtures written by the compiled reference
   (tests/golden/make_golden.py).
2. Against the compiled reference directly on fresh random inputs, where
   oracle/_ref is built (this container).
3. Closed forms that a plausible slip (dropped term, wrong sign, index or
   operand order) would break: the dimer force, FCC shell counts, lattice
   symmetry, Newton's third law, and the potential energy as the integral of
   the force.
"""
import numpy as np
import pytest

import oracle as O
from conftest import sysdef_from
from cv2md.systems import fcc_lattice, ka_mixture, lj_fluid


def _box(s):
    return O.make_box(s["lo"], s["hi"], s["bc"])


@pytest.mark.parametrize("name", ["c1_setup_brute", "c1_setup_cell", "ka_setup_cell"])
def test_setup_matches_reference(golden, name):
    d = golden(name)
    s = sysdef_from(d)
    box = _box(s)
    pos, _ = O.wrap(s["pos"], box)
    assert np.array_equal(pos, d["pos"])
    gpos, owner, code = O.ghosts(pos, box, s["ghost_cutoff"])
    assert np.array_equal(gpos, d["gpos"])
    assert np.array_equal(owner, d["gid"])  # ids are insertion indices
    if s["use_cell_list"]:
        offs, nbr = O.nlist_cell(pos, gpos, box, s["nl_cutoff"], s["nl_coef"])
    else:
        offs, nbr = O.nlist_brute(pos, gpos, s["nl_cutoff"], s["nl_coef"])
    assert np.array_equal(offs, d["offs"]) and np.array_equal(nbr, d["nbr"])
    n = len(pos)
    ids = np.arange(n)
    acc, vir, pe = O.lj(pos, s["type"], ids, gpos, s["type"][owner], owner, offs, nbr,
                        s["eps"], s["sigma"], s["lj_cutoff"], s["mass"], s.get("cutoff_list"))
    assert np.array_equal(acc, d["acc"])
    assert vir == d["virial"]
    assert O.kinetic_energy(s["vel"], s["type"], s["mass"], s["n_r_df"]) == d["ke"]


@pytest.mark.parametrize("name", ["c1_traj_brute", "c1_traj_cell", "dimer_traj", "grid512_traj",
                                  "langevin_traj", "mol_traj"])
def test_trajectory_matches_reference(golden, name):
    d = golden(name)
    s = sysdef_from(d)
    keep = set(int(k) for k in d["keep"])
    with O.OracleSim(s) as o:
        if "mol" in s:  # Lj::ignore_intra_molecule on (mol_traj)
            o.set_molecules(s["mol"])
        o.setup()
        for k in range(1, int(d["nsteps"]) + 1):
            eta = d["eta"][k - 1] if "eta" in d else None
            assert o.step(eta) == bool(d["rebuilt"][k - 1]), f"rebuild flag differs at step {k}"
            if k in keep:
                p, v, a = o.state()
                assert np.array_equal(p, d[f"pos_{k}"]), f"pos differs at step {k}"
                assert np.array_equal(v, d[f"vel_{k}"]), f"vel differs at step {k}"
                assert np.array_equal(a, d[f"acc_{k}"]), f"acc differs at step {k}"
            if s["integrator"] == 0:
                assert o.thermo()["ke"] == d["ke"][k - 1]


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_systems_match_compiled_reference(seed):
    rng = np.random.default_rng(seed)
    s = ka_mixture(int(rng.integers(300, 700)), rho=float(rng.uniform(0.7, 1.2)), seed=seed)
    s["integrator"] = 0
    s["n_r_df"] = 3
    s["use_cell_list"] = seed % 2
    with O.RefSim(s) as r, O.OracleSim(s) as o:
        r.setup(); o.setup()
        for _ in range(15):
            assert r.step() == o.step()
        pr, vr, ar = r.state(); po, vo, ao = o.state()
        assert np.array_equal(pr, po) and np.array_equal(vr, vo) and np.array_equal(ar, ao)
        assert r.thermo()["virial"] == o.thermo()["virial"]


def test_dimer_closed_form():
    # F on i = forceCoef*(r_j - r_i), forceCoef = 4e(-12 s^12/r^14 + 6 s^6/r^8 + 12 s^12/rc^14 - 6 s^6/rc^8)
    for r in (0.95, 1.0, 2.0 ** (1 / 6), 1.5, 2.4):
        pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        offs = np.array([0, 1, 1]); nbr = np.array([1], np.uint32)
        acc, vir, pe = O.lj(pos, [0, 0], [0, 1], None, None, None, offs, nbr,
                            [[1.0]], [[1.0]], 2.5, [1.0])
        rc = 2.5
        fc = 4 * (-12 / r ** 14 + 6 / r ** 8 + 12 / rc ** 14 - 6 / rc ** 8)
        assert acc[0, 0] == pytest.approx(fc * r, rel=1e-14, abs=1e-14)
        assert acc[1, 0] == pytest.approx(-fc * r, rel=1e-14, abs=1e-14)
        assert vir == pytest.approx(-fc * r * r, rel=1e-14, abs=1e-14)
        u = 4 * (r ** -12 - r ** -6) + 4 * (6 * rc ** -12 - 3 * rc ** -6) * (r / rc) ** 2 \
            - 4 * (7 * rc ** -12 - 4 * rc ** -6)
        assert pe == pytest.approx(u, rel=1e-13, abs=1e-13)


def test_pair_energy_is_integral_of_force():
    # U(rc) = 0 and dU/dr = forceCoef * r (central finite differences)
    rc2 = 2.5 ** 2
    assert abs(O.pair_energy(rc2, 1.0, 1.0, rc2)) < 1e-15
    for eps, sig in ((1.0, 1.0), (1.5, 0.8), (0.5, 0.88)):
        for r in np.linspace(0.8 * sig, 2.4 * sig, 9):
            h = 1e-6
            du = (O.pair_energy((r + h) ** 2, eps, sig, rc2) - O.pair_energy((r - h) ** 2, eps, sig, rc2)) / (2 * h)
            fc = 4 * eps * (-12 * sig ** 12 / r ** 14 + 6 * sig ** 6 / r ** 8 + 12 * sig ** 12 / rc2 ** 7 - 6 * sig ** 6 / rc2 ** 4)
            assert du == pytest.approx(fc * r, rel=1e-6, abs=1e-7)


def test_fcc_shells_and_symmetry():
    pos, L = fcc_lattice(6, 0.8442)
    box = O.make_box([0, 0, 0], [L, L, L])
    gpos, owner, _ = O.ghosts(pos, box, 2.8000000000000003)
    for cut, coef, expect in ((2.5, 1.12, 78), (2.5, 1.0, 54)):
        offs, nbr = O.nlist_cell(pos, gpos, box, cut, coef)
        # half list + all ghosts: count pairs per atom from both sides
        n = len(pos)
        cnt = np.diff(offs).copy()
        own = nbr[nbr < n]
        np.add.at(cnt, own, 1)
        assert np.all(cnt == expect), (cut * coef, np.unique(cnt))
    offs, nbr = O.nlist_brute(pos, gpos, 2.5, 1.12)
    acc, vir, pe = O.lj(pos, np.zeros(len(pos)), np.arange(len(pos)), gpos, np.zeros(len(gpos)), owner,
                        offs, nbr, [[1.0]], [[1.0]], 2.5, [1.0])
    assert np.abs(acc).max() < 1e-11                     # lattice symmetry
    # PE per atom = sum over shells n_k U(r_k) / 2
    a = (4 / 0.8442) ** (1 / 3)
    shells = {1: 12, 2: 6, 3: 24, 4: 12}  # r^2 = k a^2 / 2 (the k=5 shell, r=2.656, lies beyond rc)
    u = sum(m * O.pair_energy(k * a * a / 2, 1.0, 1.0, 6.25) for k, m in shells.items())
    assert pe / len(pos) == pytest.approx(u / 2, rel=1e-12)


def test_newton_third_law_and_wrap():
    s = lj_fluid(5, seed=7)
    rng = np.random.default_rng(3)
    box = _box(s)
    p = s["pos"] + rng.normal(0, 0.05, s["pos"].shape)
    p[0, 0] = -1e-9          # just below lower: wraps by +L
    p[1, 1] = s["hi"][1]     # exactly upper: stays (strict >)
    pw, flag = O.wrap(p, box)
    assert flag and pw[0, 0] == -1e-9 + s["hi"][0] and pw[1, 1] == s["hi"][1]
    gpos, owner, code = O.ghosts(pw, box, 2.8)
    L = s["hi"] - s["lo"]
    shift = np.stack([(code % 3) - 1, (code // 3) % 3 - 1, code // 9 - 1], 1)
    assert np.array_equal(gpos, pw[owner] - shift * L)
    offs, nbr = O.nlist_brute(pw, gpos, 2.5, 1.12)
    acc, _, _ = O.lj(pw, np.zeros(len(pw)), np.arange(len(pw)), gpos, np.zeros(len(gpos)), owner,
                     offs, nbr, [[1.0]], [[1.0]], 2.5, [1.0])
    assert np.abs(acc.sum(0)).max() < 1e-9


def test_cell_list_equals_brute_force_as_sets():
    s = ka_mixture(800, rho=0.9, seed=11)
    box = _box(s)
    pos, _ = O.wrap(s["pos"], box)
    gpos, _, _ = O.ghosts(pos, box, s["ghost_cutoff"])
    a = O.nlist_brute(pos, gpos, s["nl_cutoff"], s["nl_coef"])
    b = O.nlist_cell(pos, gpos, box, s["nl_cutoff"], s["nl_coef"])
    n = len(pos)
    for i in range(n):
        assert sorted(a[1][a[0][i]:a[0][i + 1]]) == sorted(b[1][b[0][i]:b[0][i + 1]])


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference not built here")
def test_random_velocity_fixture(golden):
    d = golden("random_velocity")
    assert np.array_equal(O.ref_random_velocity(int(d["n"]), int(d["seed"]), float(d["amplitude"])), d["vel"])


def _packets(counts, out, with_vel=False):
    """Split the concatenated halo packets: {direction: (id, type, pos, vel)}."""
    total = 8 if with_vel else 5
    res, at = {}, 0
    for d in range(27):
        n = int(counts[d])
        if n:
            pk = out[at:at + total * n]
            res[d] = (pk[:n], pk[n:2 * n], pk[2 * n:5 * n].reshape(n, 3),
                      pk[5 * n:8 * n].reshape(n, 3) if with_vel else None)
        at += total * n
    return res


def test_halo_pack_one_domain_equals_single_domain_ghosts():
    """With a 1x1x1 process grid every direction is the domain itself, and
    the union of the 26 packets must be the single-domain ghost set of
    exchange_ghost_single_md_domain (pinned against the compiled reference):
    the same images, ids and positions bit for bit."""
    s = lj_fluid(5, seed=17)
    box = O.make_box(s["lo"], s["hi"])
    p, _ = O.wrap(s["pos"] + np.random.default_rng(3).normal(0, 0.05, s["pos"].shape), box)
    n = len(p)
    ids = np.arange(n) + 100
    L = s["hi"] - s["lo"]
    counts, out = O.halo_pack(p, None, ids, np.zeros(n, np.int32), (0, 0, 0), (1, 1, 1), s["lo"], s["hi"],
                              L, (1, 1, 1), 2.8)
    assert counts[13] == 0  # (1, 1, 1) is never a direction
    got = np.concatenate([np.column_stack([i, x]) for i, _t, x, _v in _packets(counts, out).values()])
    gpos, owner, _code = O.ghosts(p, box, 2.8)
    exp = np.column_stack([ids[owner], gpos])
    key = lambda a: a[np.lexsort(a.T[::-1])]
    assert len(got) == len(exp)
    assert np.array_equal(key(got), key(exp))


def test_halo_pack_known_answer_two_domains():
    """2x1x1 grid, box [0, 10)^3 split at x = 5, ghost cutoff 1.  Domain 0
    sends atoms below x = 1 to -x (shifted by +10: it wraps to domain 1's
    far face) and atoms above x = 4 to +x (unshifted), in index order, with
    the y/z directions and their combinations (periodic, one process per
    axis) alongside; velocities follow the positions."""
    pos = np.array([[0.5, 5.0, 5.0],    # -x
                    [4.5, 5.0, 5.0],    # +x
                    [2.5, 0.5, 5.0],    # -y
                    [0.2, 9.7, 5.0],    # -x, +y and -x+y
                    [2.5, 5.0, 5.0],    # interior
                    [4.9, 0.1, 9.9]])   # +x, -y, +z and all combinations
    vel = np.arange(18, dtype=float).reshape(6, 3)
    ids = np.array([7, 8, 9, 10, 11, 12])
    ty = np.array([0, 1, 0, 1, 0, 1], np.int32)
    counts, out = O.halo_pack(pos, vel, ids, ty, (0, 0, 0), (2, 1, 1), (0, 0, 0), (5, 10, 10),
                              (10, 10, 10), (1, 1, 1), 1.0, with_vel=True)
    pk = _packets(counts, out, True)
    D = lambda i, j, k: i * 9 + j * 3 + k
    exp_sets = {D(0, 1, 1): [7, 10], D(2, 1, 1): [8, 12], D(1, 0, 1): [9, 12], D(1, 2, 1): [10],
                D(1, 1, 2): [12], D(0, 2, 1): [10], D(2, 0, 1): [12], D(2, 1, 2): [12],
                D(1, 0, 2): [12], D(2, 0, 2): [12]}
    assert sorted(pk) == sorted(exp_sets)
    for d, want in exp_sets.items():
        assert list(pk[d][0].astype(int)) == want, d
    i0, t0, x0, v0 = pk[D(0, 1, 1)]
    assert np.array_equal(x0, [[10.5, 5.0, 5.0], [10.2, 9.7, 5.0]])
    assert np.array_equal(t0, [0, 1]) and np.array_equal(v0, vel[[0, 3]])
    i2, _t2, x2, _v2 = pk[D(2, 1, 1)]
    assert np.array_equal(x2, pos[[1, 5]])  # sent to +x inside the global box: no shift
    _i, _t, xy, _v = pk[D(1, 2, 1)]
    assert np.array_equal(xy, [[0.2, 9.7 - 10.0, 5.0]])  # +y wraps (one process along y)
    _i, _t, xyz, _v = pk[D(2, 0, 2)]
    assert np.allclose(xyz, [[4.9, 10.1, -0.1]]) and xyz[0, 1] == 0.1 + 10.0 and xyz[0, 2] == 9.9 - 10.0


def test_jump_fix_dimer_closed_form():
    """jump_fix (lj.cpp:459-468): a pair closer than sqrt(jump_tol) takes the
    force coefficient of r^2 = jump_tol, along its own separation; a pair
    beyond it is untouched."""
    eps, sig, rc, jt = 1.0, 1.0, 2.5, 0.81
    for r, clamped in ((0.7, True), (0.95, False)):
        pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        acc, vir, pe = O.lj(pos, [0, 0], [0, 1], None, None, None, [0, 1, 1], [1], [[eps]], [[sig]], rc,
                            [1.0], None, jt)
        r2 = jt if clamped else r * r
        c2 = rc * rc
        coef = 4 * eps * (-12 * (sig ** 2 / r2) ** 6 / r2 + 6 * (sig ** 2 / r2) ** 3 / r2
                          + 12 * (sig ** 2 / c2) ** 6 / c2 - 6 * (sig ** 2 / c2) ** 3 / c2)
        assert np.isclose(acc[0, 0], coef * r, rtol=1e-13) and np.isclose(acc[1, 0], -coef * r, rtol=1e-13)
        assert np.isclose(vir, -coef * r2, rtol=1e-13)


@pytest.mark.parametrize("mode", ["energy_per_dof", "energy_tot", "kbt", "kb_temperature"])
@pytest.mark.parametrize("n_r_df", [0, 3])
def test_nve_rescale_reaches_target(mode, n_r_df):
    """constraint::Nve (nve.cpp:147-189): after the rescale the kinetic energy
    is the target.  KE(lambda v) = lambda^2 KE(v) with or without the
    centre-of-mass term, so each mode's closed form pins lambda; a target
    formula with a dropped factor (dof, 1/2, kb) fails one of them."""
    s = lj_fluid(3, seed=3)
    rng = np.random.default_rng(2)
    s = dict(s, vel=rng.normal(0.3, 1.0, s["pos"].shape), n_r_df=n_r_df)
    n = len(s["pos"])
    dof = 3 * n - n_r_df
    arg = dict(energy_per_dof=dict(energy_per_dof=0.7), energy_tot=dict(energy_tot=450.0),
               kbt=dict(kbt=0.9), kb_temperature=dict(kb=2.0, temperature=0.45))[mode]
    target = dict(energy_per_dof=0.7 * dof, energy_tot=450.0, kbt=0.5 * dof * 0.9,
                  kb_temperature=0.5 * dof * 2.0 * 0.45)[mode]
    with O.OracleSim(s) as o:
        o.setup()
        ke0 = o.thermo()["ke"]
        assert ke0 > 0
        lam = O.nve_lambda(ke0, dof, **arg)
        o.scale_velocities(lam)
        ke1 = o.thermo()["ke"]
        v = o.state()[1]
    assert ke1 == pytest.approx(target, rel=1e-13)
    assert np.allclose(v, lam * s["vel"], rtol=0, atol=1e-15 * np.abs(s["vel"]).max())
    assert O.nve_lambda(0.0, dof, kbt=0.9) == 1.0  # the ignored step


def test_ignore_intra_molecule_drops_exactly_the_same_molecule_pairs():
    """lj.cpp:446: with the flag on, a pair adds nothing when its two
    molecule indices are equal (two atoms outside molecules, -1 and -1,
    included).  Pinned against the unflagged force over the list with those
    pairs removed (a different route to the same sum), and by the two
    extremes: one molecule per atom changes nothing, all atoms outside
    molecules leaves no force at all."""
    s = lj_fluid(4, temperature=1.0, seed=7)
    box = _box(s)
    pos, _ = O.wrap(s["pos"] + np.random.default_rng(7).normal(0, 0.05, s["pos"].shape), box)
    gpos, owner, _ = O.ghosts(pos, box, s["ghost_cutoff"])
    offs, nbr = O.nlist_cell(pos, gpos, box, s["nl_cutoff"], s["nl_coef"])
    n = len(pos)
    ids = np.arange(n)
    args = (pos, s["type"], ids, gpos, s["type"][owner], owner)
    prm = (s["eps"], s["sigma"], s["lj_cutoff"], s["mass"])
    mol = (np.arange(n) // 4).astype(np.int32)
    mol[(np.arange(n) // 4) % 5 == 4] = -1
    gmol = mol[owner]
    acc_m, vir_m, pe_m = O.lj(*args, offs, nbr, *prm, mol=mol, gmol=gmol)
    # the same list without the same-molecule pairs, flag off
    jm = np.where(nbr >= n, gmol[np.minimum(nbr.astype(np.int64) - n, len(gmol) - 1)], mol[np.minimum(nbr, n - 1)])
    im = np.repeat(mol, np.diff(offs))
    keep = jm != im
    offs2 = np.concatenate([[0], np.cumsum(np.add.reduceat(keep.astype(np.int64), offs[:-1]) * (np.diff(offs) > 0))])
    acc_f, vir_f, pe_f = O.lj(*args, offs2, nbr[keep], *prm)
    assert keep.sum() < len(nbr)
    assert np.array_equal(acc_m, acc_f) and vir_m == vir_f and pe_m == pe_f
    acc_0, vir_0, pe_0 = O.lj(*args, offs, nbr, *prm)
    acc_u, vir_u, pe_u = O.lj(*args, offs, nbr, *prm, mol=np.arange(n), gmol=owner)
    assert np.array_equal(acc_u, acc_0) and vir_u == vir_0 and pe_u == pe_0
    acc_n, vir_n, pe_n = O.lj(*args, offs, nbr, *prm, mol=np.full(n, -1), gmol=np.full(len(owner), -1))
    assert not acc_n.any() and vir_n == 0.0 and pe_n == 0.0
