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
