"""The caviar2 drop-in C++ API (caviar-v2_amd/include/caviar2, libcaviar2_amd.so).

CPU: the reference's own user programs compile unchanged against these
headers (where the reference tree exists), and the host-side velocity
initialisation reproduces the reference's libstdc++ stream bit for bit.
GPU: the example programs run on the device and agree with the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

PKG = os.path.join(ROOT, "caviar-v2_amd")
INC = os.path.join(PKG, "include")
LIB = os.path.join(PKG, "lib")
EX = os.path.join(PKG, "build", "examples")
REF_TESTS = "/root/reference/tests"


def _compile(src, out):
    cmd = ["g++", "-O2", "-std=c++17", f"-I{INC}", src, f"-L{LIB}", "-lcaviar2_amd", "-lcv2md",
           f"-Wl,-rpath,{LIB}", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.skipif(not os.path.isdir(REF_TESTS), reason="reference tree not present")
@pytest.mark.parametrize("prog", ["t001_lj.cpp", "t002_lj.cpp"])
def test_reference_programs_compile_unchanged(prog, tmp_path):
    _compile(os.path.join(REF_TESTS, prog), str(tmp_path / prog.replace(".cpp", "")))


def test_random_velocity_matches_reference_stream(golden, tmp_path):
    d = golden("random_velocity")
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "random_velocity_check.cpp"), str(tmp_path / "rv"))
    out = subprocess.run([exe, str(int(d["n"])), str(int(d["seed"])), str(float(d["amplitude"]))],
                         capture_output=True, text=True, check=True).stdout.split()
    v = np.array([float.fromhex(x) for x in out]).reshape(-1, 3)
    assert np.array_equal(v, d["vel"])


def test_examples_built():
    for name in ("lj_dimer", "lj_lattice"):
        assert os.access(os.path.join(EX, name), os.X_OK), name


def _dimer_oracle(steps):
    s = dict(lo=np.full(3, -20.0), hi=np.full(3, 20.0), bc=(1, 1, 1),
             pos=np.array([[-0.6, 0.0, 0.0], [0.6, 0.0, 0.0]]), vel=np.zeros((2, 3)),
             type=np.zeros(2, np.int32), eps=np.ones((1, 1)), sigma=np.ones((1, 1)), cutoff_list=None,
             mass=np.ones(1), lj_cutoff=5.0, nl_cutoff=5.0, nl_coef=1.1, ghost_cutoff=5.5, dt=0.001,
             integrator=0, k_b=1.0, n_r_df=3)
    with O.OracleSim(s) as o:
        o.setup()
        for _ in range(steps):
            o.step()
        p, v, a = o.state()
    return p[1, 0] - p[0, 0]


@pytest.mark.gpu
def test_dimer_example_matches_oracle(tmp_path):
    r = subprocess.run([os.path.join(EX, "lj_dimer")], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("separation")][0].split()
    sep = float(line[1])
    assert abs(sep - _dimer_oracle(2000)) < 1e-9
    e = np.loadtxt(tmp_path / "dimer_energy.txt")
    # setup writes step 0, then every 100th step of 0..1999 (as md_simulator.cpp:607,726 do)
    assert e.shape == (21, 3) and list(e[:2, 0]) == [0, 0] and e[-1, 0] == 1900


@pytest.mark.gpu
def test_lattice_example_runs(tmp_path):
    r = subprocess.run([os.path.join(EX, "lj_lattice"), "3000"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    t = np.loadtxt(tmp_path / "lattice_temperature.txt")
    assert t.shape[1] == 3 and np.all(np.isfinite(t[:, 2]))
    assert [int(x) for x in t[:, 0]] == [0] + list(range(0, 3000, 500))


def test_nve_constraint_host(tmp_path):
    """constraint::Nve on the host state: at a multiple of `step` the kinetic
    energy becomes dof * kbt / 2 (nve.cpp:165-172,186-189), elsewhere it is
    untouched; conflicting settings end the program (nve.cpp:119-120)."""
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "nve_check.cpp"), str(tmp_path / "nve"))
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    ke0, ke7, ke1 = (float.fromhex(x) for x in out[:3])
    dof, recalc, nxt = int(out[3]), int(out[4]), int(out[5])
    vel = np.array([[1.0, -0.5, 0.25], [-0.75, 0.5, 1.5], [0.0, 2.0, -1.0], [0.5, 0.5, 0.5]])
    ref = O.kinetic_energy(vel, np.zeros(4, np.int32), np.array([2.0]), 3)
    assert dof == 3 * 4 - 3 and recalc == 1 and nxt == 30
    assert ke0 == pytest.approx(ref, rel=1e-15) and ke7 == ke0
    assert ke1 == pytest.approx(0.5 * dof * 1.3, rel=1e-14)
    r = subprocess.run([exe, "bad"], capture_output=True, text=True)
    assert r.returncode != 0 and "Only one of 'energy' and 'temperature' can be set." in r.stdout + r.stderr


@pytest.mark.gpu
def test_nve_example_rescales_on_device(tmp_path):
    """The velocity-rescale constraint in a device run (Md_simulator chunks at
    its steps): the temperature the writer reports at each multiple of 50 --
    after the rescale, as md_simulator.cpp:683-689,725-728 order them -- is
    the constraint's kbt / k_b."""
    r = subprocess.run([os.path.join(EX, "lj_nve"), "500"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    t = np.loadtxt(tmp_path / "nve_temperature.txt")
    assert [int(x) for x in t[:, 0]] == [0] + list(range(0, 500, 50))
    assert abs(t[0, 2] - 0.9) > 0.1  # the random start is far from the target
    assert np.allclose(t[1:, 2], 0.9, rtol=1e-12, atol=0)


# ---- writers, Leap_frog, host neighlist, all_atom_test, time -> steps ------

def test_time_to_step_matches_reference_expression(tmp_path):
    """Md_simulator::setup converts times with (int)time / dt
    (md_simulator.cpp:569-572): the time is truncated before the division."""
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "dropin_checks.cpp"), str(tmp_path / "dc"))
    for t, dt, want in ((2.5, 0.005, 400), (10.0, 0.005, 2000), (0.9, 0.001, 0), (7.99, 0.5, 14)):
        out = subprocess.run([exe, "steps", repr(t), repr(dt)], capture_output=True, text=True, check=True)
        assert int(out.stdout) == want == int(int(t) / dt)


def _small_fluid():
    """256 atoms (4^3 FCC cells, rho* = 0.8442) perturbed, T* = 1.44; rc 2.5,
    list coefficient 1.12, ghost cutoff 2.8."""
    from cv2md.systems import lj_fluid
    s = lj_fluid(4, temperature=1.44, seed=61)
    rng = np.random.default_rng(61)
    box = O.make_box(s["lo"], s["hi"])
    s["pos"], _ = O.wrap(s["pos"] + rng.normal(0.0, 0.05, s["pos"].shape), box)
    return s


def _write_input(path, s):
    L = float(s["hi"][0])
    with open(path, "w") as f:
        f.write(f"{len(s['pos'])} {L!r} {s['nl_cutoff']!r} {s['nl_coef']!r} {s['ghost_cutoff']!r}\n")
        for p, v in zip(s["pos"], s["vel"]):
            f.write(" ".join(repr(float(x)) for x in (*p, *v)) + "\n")


def _read_series(path):
    return np.loadtxt(path, ndmin=2)


@pytest.mark.gpu
@pytest.mark.parametrize("integ", ["vv", "lf"])
def test_writers_and_leapfrog_match_oracle(integ, tmp_path):
    """A drop-in run with every writer at a cadence of 10 over 50 steps:
    - the files have the reference's formats ("step time value" lines,
      dump_energy.cpp:42-53; xyz frames "N", "Atom", "type id x y z vx vy vz
      ax ay az", dump_xyz_serial.cpp:51-64), written at setup and after the
      steps 0, 10, ..., 40 (md_simulator.cpp:607,726);
    - kinetic energy, temperature, pressure (kinetic_energy.cpp:206-303),
      potential energy and positions agree with the oracle's run;
    - Leap_frog (kick-drift-kick, md_simulator.cpp:871-980) gives the
      velocity-Verlet trajectory;
    - the velocity-Verlet run has Neighborlist::set_verify_builds on: every
      list build is checked on the device and none differs (a difference
      would end the run with an error)."""
    s = _small_fluid()
    inp = tmp_path / "in.txt"
    _write_input(inp, s)
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "dropin_checks.cpp"), str(tmp_path / "dc"))
    r = subprocess.run([exe, "run", str(inp), integ, "50", "10"] + (["verify"] if integ == "vv" else []),
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the final state: lines of six hex floats (the log's info lines share stdout)
    rows = [l.split() for l in r.stdout.split("\n")]
    fin = np.array([[float.fromhex(x) for x in t] for t in rows
                    if len(t) == 6 and all(x.lstrip("-").startswith("0x") for x in t)])
    assert fin.shape == (len(s["pos"]), 6)
    n = len(s["pos"])
    V = float(np.prod(s["hi"] - s["lo"]))
    o_state = {}
    with O.OracleSim(dict(s, use_cell_list=1)) as o:
        o.setup()
        o_state[-1] = (o.thermo(), o.state()[0])
        for k in range(50):
            o.step()
            if k % 10 == 0 or k == 49:
                o_state[k] = (o.thermo(), o.state())
    steps = [-1, 0, 10, 20, 30, 40]  # -1: the setup line, written as step 0
    e = _read_series(tmp_path / "w_energy.txt")
    t = _read_series(tmp_path / "w_temperature.txt")
    pr = _read_series(tmp_path / "w_pressure.txt")
    pe = _read_series(tmp_path / "w_pe.txt")
    for arr in (e, t, pr, pe):
        assert arr.shape == (6, 3)
        assert list(arr[:, 0]) == [0, 0, 10, 20, 30, 40]
        assert np.allclose(arr[:, 1], [0.0, 0.0, 0.05, 0.1, 0.15, 0.2], rtol=0, atol=1e-12)
    # the values are printed with the stream's default 6 significant digits,
    # as the reference's writers print them
    for row, k in enumerate(steps):
        th = o_state[k][0]
        assert e[row, 2] == pytest.approx(th["ke"], rel=6e-6)
        assert t[row, 2] == pytest.approx(th["temperature"], rel=6e-6)
        p_ref = n * th["temperature"] / V + th["virial"] / (3.0 * V)
        assert pr[row, 2] == pytest.approx(p_ref, rel=6e-6)
        assert pe[row, 2] == pytest.approx(th["pe"], rel=6e-6)
    # xyz frames
    lines = open(tmp_path / "w_xyz.xyz").read().split("\n")
    frames = []
    at = 0
    while at < len(lines) and lines[at].strip():
        cnt = int(lines[at])
        assert lines[at + 1] == "Atom"
        rows = np.array([[float(x) for x in l.split()] for l in lines[at + 2:at + 2 + cnt]])
        assert rows.shape == (cnt, 11)
        frames.append(rows)
        at += 2 + cnt
    assert len(frames) == 6 and all(len(f) == n for f in frames)
    for f, k in zip(frames[1:], steps[1:]):
        p_o = o_state[k][1][0]
        assert np.array_equal(f[:, 0], np.zeros(n)) and np.array_equal(f[:, 1], np.arange(n))
        assert np.allclose(f[:, 2:5], p_o, rtol=6e-6, atol=1e-6)
    pf, vf = fin[:, :3], fin[:, 3:]
    po, vo, _ = o_state[49][1]
    assert np.abs(pf - po).max() < 1e-10 and np.abs(vf - vo).max() < 1e-9


@pytest.mark.gpu
def test_ignore_intra_molecule_dropin_matches_oracle(tmp_path):
    """A user program that sets Atom_struct::molecule_index and calls
    Lj::ignore_intra_molecule_flag (lj.cpp:446) runs 40 velocity-Verlet steps
    on the device; the final state equals the oracle's run with the same
    molecules."""
    s = _small_fluid()
    inp = tmp_path / "in.txt"
    _write_input(inp, s)
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "dropin_checks.cpp"), str(tmp_path / "dc"))
    r = subprocess.run([exe, "molrun", str(inp), "40"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.split() for l in r.stdout.split("\n")]
    fin = np.array([[float.fromhex(x) for x in t] for t in rows
                    if len(t) == 6 and all(x.lstrip("-").startswith("0x") for x in t)])
    n = len(s["pos"])
    assert fin.shape == (n, 6)
    mol = np.where((np.arange(n) // 4) % 5 == 4, -1, np.arange(n) // 4)
    with O.OracleSim(dict(s, use_cell_list=1)) as o:
        o.set_molecules(mol)
        o.setup()
        for _ in range(40):
            o.step()
        po, vo, _ = o.state()
    assert np.abs(fin[:, :3] - po).max() < 1e-10 and np.abs(fin[:, 3:] - vo).max() < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("all_pairs", [0, 1])
def test_host_neighlist_matches_oracle(all_pairs, tmp_path):
    """Neighborlist::update_host_list after a forced build: the ghosts are the
    reference's single-domain images in its order
    (exchange_ghost_single_md_domain.cpp:67-215), and neighlist equals the
    reference's Verlet list (neighborlist.cpp:247-297: owned j > i within the
    list cutoff, then ghosts as N + g), or with all_atom_test its all-pairs
    list (neighborlist.cpp:170-200)."""
    s = _small_fluid()
    inp = tmp_path / "in.txt"
    _write_input(inp, s)
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "dropin_checks.cpp"), str(tmp_path / "dc"))
    r = subprocess.run([exe, "neighlist", str(inp), str(all_pairs)], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = [l for l in r.stdout.split("\n") if not l.startswith("[")]  # without the log's lines
    n, ng = (int(x) for x in out[0].split())
    gid = np.array([int(l.split()[0]) for l in out[1:1 + ng]])
    gpos = np.array([[float.fromhex(x) for x in l.split()[1:]] for l in out[1:1 + ng]])
    lists = [[int(x) for x in l.split()[1:]] for l in out[1 + ng:1 + ng + n]]
    box = O.make_box(s["lo"], s["hi"])
    g_ref, owner, _ = O.ghosts(s["pos"], box, s["ghost_cutoff"])
    assert np.array_equal(gpos, g_ref) and np.array_equal(gid, owner)
    if all_pairs:
        for i, l in enumerate(lists):
            assert l == list(range(i + 1, n)) + list(range(n, n + ng))
        return
    offs, nbr = O.nlist_brute(s["pos"], g_ref, s["nl_cutoff"], s["nl_coef"])
    for i, l in enumerate(lists):
        assert l == sorted(int(x) for x in nbr[offs[i]:offs[i + 1]])


# ---- Lj::verify_settings: per-atom mixing, WCA, lambda scalings --------------

def _ljtables(exe, inp, case, tmp_path):
    r = subprocess.run([exe, "ljtables", str(inp), case], capture_output=True, text=True, cwd=tmp_path,
                       timeout=60)
    if r.returncode != 0:
        return r, None
    rows = [l.split() for l in r.stdout.split("\n") if l and not l.startswith("[")]
    assert rows[0] == ["has_cutoff_list", "1"]
    t = [np.array([float.fromhex(x) for x in row]).reshape(3, 3) for row in rows[1:4]]
    return r, t


def test_lj_verify_settings_tables_match_reference(golden, tmp_path):
    """Lj::verify_settings (lj.cpp:180-391) in a user program, no device:
    the epsilon, sigma and cutoff-list tables equal bit for bit those the
    compiled reference forms from the same inputs (tests/golden/lj_tables,
    made by make_golden.py through oracle/ref_driver.cpp), and follow the
    closed forms: Lorentz-Berthelot mixing of the per-atom values, WCA
    cutoffs 2^(1/6) sigma_ij before the lambda_s scaling, lambda_s / lambda_e
    products with missing entries 1, array tables zero-filled to the atom
    types present.  Per-atom and array input together are refused."""
    g = golden("lj_tables")
    s = _small_fluid()
    inp = tmp_path / "in.txt"
    _write_input(inp, s)
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "dropin_checks.cpp"), str(tmp_path / "dc"))
    for case in ("atom", "array"):
        r, (e, sg, cl) = _ljtables(exe, inp, case, tmp_path)
        assert np.array_equal(e, g[case + "_eps"]), case
        assert np.array_equal(sg, g[case + "_sigma"]), case
        assert np.array_equal(cl, g[case + "_cutoff_list"]), case
    # closed forms of the per-atom case
    ea, sa = np.array([1.0, 0.5, 0.8]), np.array([1.0, 0.88, 1.2])
    lam_s = np.array([[1.0, 0.9, 1.0], [0.9, 1.0, 1.0], [1.0, 1.0, 1.0]])
    lam_e = np.array([[1.0, 0.5, 1.0], [0.5, 1.0, 1.0], [1.0, 1.0, 0.7]])
    s_mix = 0.5 * (sa[:, None] + sa[None, :])
    e_mix = np.sqrt(ea[:, None] * ea[None, :])
    np.fill_diagonal(s_mix, sa)
    np.fill_diagonal(e_mix, ea)
    e, sg, cl = (g["atom_eps"], g["atom_sigma"], g["atom_cutoff_list"])
    assert np.allclose(sg, s_mix * lam_s, rtol=1e-15, atol=0)
    assert np.allclose(e, e_mix * lam_e, rtol=1e-15, atol=0)
    assert np.allclose(cl, 2.0 ** (1.0 / 6.0) * s_mix, rtol=1e-15, atol=0)
    # the array case: entries as set, the rest 0; the cutoff list [[2.5]] zero-filled
    assert np.array_equal(g["array_eps"], [[1.0, 0.5, 0.0], [0.5, 0.8, 0.0], [0.0, 0.0, 0.0]])
    assert np.array_equal(g["array_sigma"], [[1.0, 0.0, 0.0], [0.0, 0.9, 0.0], [0.0, 0.0, 0.0]])
    assert np.array_equal(g["array_cutoff_list"], [[2.5, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    r, _ = _ljtables(exe, inp, "both", tmp_path)
    assert r.returncode != 0 and "input_by_atom && input_by_array" in (r.stdout + r.stderr)


@pytest.mark.gpu
def test_lj_verify_settings_run_matches_oracle(tmp_path):
    """A user program with two species (masses 1, 1.5) given per atom
    (epsilon 1, 0.6; sigma 1, 0.9), WCA cutoffs and a lambda_e scaling, which
    calls Lj::verify_settings: 40 velocity-Verlet steps on the device follow
    the oracle's run with the tables of the closed forms (mixing, 2^(1/6)
    sigma_ij cutoffs, lambda_e products)."""
    s = _small_fluid()
    inp = tmp_path / "in.txt"
    _write_input(inp, s)
    exe = _compile(os.path.join(ROOT, "tests", "cpp", "dropin_checks.cpp"), str(tmp_path / "dc"))
    r = subprocess.run([exe, "ljrun", str(inp), "40"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l.split() for l in r.stdout.split("\n")]
    fin = np.array([[float.fromhex(x) for x in t] for t in rows
                    if len(t) == 6 and all(x.lstrip("-").startswith("0x") for x in t)])
    n = len(s["pos"])
    assert fin.shape == (n, 6)
    sa, ea = np.array([1.0, 0.9]), np.array([1.0, 0.6])
    sig = 0.5 * (sa[:, None] + sa[None, :])
    eps = np.sqrt(ea[:, None] * ea[None, :])
    np.fill_diagonal(sig, sa)
    np.fill_diagonal(eps, ea)
    eps = eps * np.array([[1.0, 0.8], [0.8, 1.0]])
    cl = 2.0 ** (1.0 / 6.0) * sig
    o_sys = dict(s, type=(np.arange(n) % 2).astype(np.int32), eps=eps, sigma=sig, cutoff_list=cl,
                 mass=np.array([1.0, 1.5]), use_cell_list=1)
    with O.OracleSim(o_sys) as o:
        o.setup()
        for _ in range(40):
            o.step()
        po, vo, _ = o.state()
    print(f"[ljrun] max |dx| {np.abs(fin[:, :3] - po).max():.2e}, |dv| {np.abs(fin[:, 3:] - vo).max():.2e}")
    assert np.abs(fin[:, :3] - po).max() < 1e-10 and np.abs(fin[:, 3:] - vo).max() < 1e-9
