"""ctypes bindings for the oracle (test infrastructure only).

Two libraries are wrapped:

* ``liboracle.so`` -- the C restatement of the CAVIAR2 MD step
  (``oracle/cv2_oracle.c``), available everywhere the repository is built;
* ``_ref/libcaviar2_ref.so`` -- the reference itself, compiled from its own
  sources plus ``ref_driver.cpp`` where the reference tree exists.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, as a checker or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcaviar2_ref.so")

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u32 = C.POINTER(C.c_uint32)
_vp = C.c_void_p


class Box(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("bc", C.c_int * 3)]


class LjParams(C.Structure):
    _fields_ = [("ntypes", C.c_int), ("eps", _d), ("sigma", _d), ("cutoff_list", _d),
                ("cutoff", C.c_double), ("mass_inv", _d), ("jump_tol", C.c_double),
                ("ignore_intra_molecule", C.c_int), ("mol", _i32), ("gmol", _i32)]


class SimParams(C.Structure):
    _fields_ = [("box", Box), ("ghost_cutoff", C.c_double), ("nl_cutoff", C.c_double),
                ("nl_coef", C.c_double), ("use_cell_list", C.c_int),
                ("rebuild_every_step", C.c_int), ("ntypes", C.c_int), ("eps", _d),
                ("sigma", _d), ("cutoff_list", _d), ("mass", _d), ("lj_cutoff", C.c_double),
                ("integrator", C.c_int), ("dt", C.c_double), ("langevin_gamma", C.c_double),
                ("langevin_kbt", C.c_double), ("k_b", C.c_double), ("n_r_df", C.c_int)]


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def make_box(lo, hi, bc=(1, 1, 1)):
    b = Box()
    for d in range(3):
        b.lo[d], b.hi[d], b.bc[d] = float(lo[d]), float(hi[d]), int(bc[d])
    return b


_orc = None


def oracle_lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_wrap.argtypes = [C.c_int64, _d, C.POINTER(Box), _i32]
        L.orc_wrap.restype = C.c_int
        L.orc_ghosts.argtypes = [C.c_int64, _d, C.POINTER(Box), C.c_double, C.c_int64, _d, _i64, _i32]
        L.orc_ghosts.restype = C.c_int64
        L.orc_halo_pack.argtypes = [C.c_int64, _d, _d, _i64, _i32, _i32, _i32, _d, _d, _d, _i32,
                                    C.c_double, C.c_int32, _i64, C.c_int64, _d]
        L.orc_halo_pack.restype = C.c_int64
        L.orc_owned_pack.argtypes = [C.c_int64, _d, _d, _d, _i64, _i32, _i32, _d, _d, _d, _i32, _i64,
                                     C.c_int64, _d, _i32, _d]
        L.orc_owned_pack.restype = C.c_int64
        L.orc_nlist_brute.argtypes = [C.c_int64, _d, C.c_int64, _d, C.c_double, C.c_double,
                                      C.c_int64, _i64, _u32]
        L.orc_nlist_brute.restype = C.c_int64
        L.orc_nlist_cell.argtypes = [C.c_int64, _d, C.c_int64, _d, C.POINTER(Box), C.c_double,
                                     C.c_double, C.c_int64, _i64, _u32]
        L.orc_nlist_cell.restype = C.c_int64
        L.orc_lj.argtypes = [C.c_int64, _d, _i32, _i64, C.c_int64, _d, _i32, _i64, _i64, _u32,
                             C.POINTER(LjParams), _d, _d, _d]
        L.orc_lj.restype = None
        L.orc_lj_pair_energy.argtypes = [C.c_double] * 4
        L.orc_lj_pair_energy.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [C.c_int64, _d, _i32, _d, C.c_int]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_sim_create.argtypes = [C.POINTER(SimParams), C.c_int64, _d, _d, _i32, _i64]
        L.orc_sim_create.restype = _vp
        L.orc_sim_destroy.argtypes = [_vp]
        L.orc_sim_setup.argtypes = [_vp]
        L.orc_sim_step.argtypes = [_vp, _d]
        L.orc_sim_step.restype = C.c_int
        L.orc_sim_get.argtypes = [_vp, _d, _d, _d]
        L.orc_sim_nghost.argtypes = [_vp]
        L.orc_sim_nghost.restype = C.c_int64
        L.orc_sim_npairs.argtypes = [_vp]
        L.orc_sim_npairs.restype = C.c_int64
        L.orc_sim_thermo.argtypes = [_vp, _d, _d, _d, _d]
        L.orc_nve_lambda.argtypes = [C.c_double, C.c_int] + [C.c_double] * 5
        L.orc_nve_lambda.restype = C.c_double
        L.orc_sim_scale_velocities.argtypes = [_vp, C.c_double]
        L.orc_sim_set_molecules.argtypes = [_vp, _i32, C.c_int]
        _orc = L
    return _orc


# ---------------------------------------------------------------- functions
def wrap(pos, box):
    """Domain::fix_position on every atom; returns (pos, wrapped_flag)."""
    p = _f64(pos).copy()
    f = oracle_lib().orc_wrap(len(p), _p(p, _d), C.byref(box), None)
    return p, bool(f)


def ghosts(pos, box, ghost_cutoff):
    """Reference-ordered ghost images: (gpos[G,3], owner[G], code[G])."""
    p = _f64(pos)
    L = oracle_lib()
    n = len(p)
    cap = max(16, n * 2)
    while True:
        g = np.empty((cap, 3)); o = np.empty(cap, np.int64); c = np.empty(cap, np.int32)
        k = L.orc_ghosts(n, _p(p, _d), C.byref(box), ghost_cutoff, cap, _p(g, _d), _p(o, _i64), _p(c, _i32))
        if k <= cap:
            return g[:k].copy(), o[:k].copy(), c[:k].copy()
        cap = k


def owned_pack(pos, vel, acc, id_, grid_index, nprocs, lo_local, hi_local, size_global, bc):
    """Owned-atom exchange packets (exchange_owned_mpi_shared_atoms.cpp):
    (counts[27], packets concatenated in direction order, stay[n] as bool,
    positions after the shifts back into the domain [n, 3])."""
    p = _f64(pos); v = _f64(vel); a = _f64(acc)
    n = len(p)
    ids = np.ascontiguousarray(id_, np.int64)
    a32 = lambda x: np.ascontiguousarray(x, np.int32)
    gi, npr, bcs = a32(grid_index), a32(nprocs), a32(bc)
    lo, hi, sg = _f64(lo_local), _f64(hi_local), _f64(size_global)
    counts = np.zeros(27, np.int64)
    L = oracle_lib()
    args = (n, _p(p, _d), _p(v, _d), _p(a, _d), _p(ids, _i64), _p(gi, _i32), _p(npr, _i32),
            _p(lo, _d), _p(hi, _d), _p(sg, _d), _p(bcs, _i32), _p(counts, _i64))
    need = L.orc_owned_pack(*args, 0, None, None, None)
    out = np.empty(max(need, 1))
    stay = np.zeros(max(n, 1), np.int32)
    after = np.zeros((max(n, 1), 3))
    L.orc_owned_pack(*args, need, _p(out, _d), _p(stay, _i32), _p(after, _d))
    return counts, out[:need], stay[:n].astype(bool), after[:n]


def halo_pack(pos, vel, id_, type_, grid_index, nprocs, lo_local, hi_local, size_global, bc,
              ghost_cutoff, with_vel=False):
    """MPI halo packets in the reference's layout: (counts[27], packets
    concatenated in direction order d = i*9 + j*3 + k)."""
    p = _f64(pos)
    v = _f64(vel) if vel is not None else np.zeros_like(p)
    n = len(p)
    ids = np.ascontiguousarray(id_, np.int64)
    ty = np.ascontiguousarray(type_, np.int32)
    a32 = lambda x: np.ascontiguousarray(x, np.int32)
    gi, npr, bcs = a32(grid_index), a32(nprocs), a32(bc)
    lo, hi, sg = _f64(lo_local), _f64(hi_local), _f64(size_global)
    counts = np.zeros(27, np.int64)
    L = oracle_lib()
    args = (n, _p(p, _d), _p(v, _d), _p(ids, _i64), _p(ty, _i32), _p(gi, _i32), _p(npr, _i32),
            _p(lo, _d), _p(hi, _d), _p(sg, _d), _p(bcs, _i32), ghost_cutoff, int(with_vel),
            _p(counts, _i64))
    need = L.orc_halo_pack(*args, 0, None)
    out = np.empty(max(need, 1))
    L.orc_halo_pack(*args, need, _p(out, _d))
    return counts, out[:need]


def _csr(fn, n, *args):
    cap = max(1024, 80 * n)
    while True:
        offs = np.zeros(n + 1, np.int64); nbr = np.empty(cap, np.uint32)
        k = fn(*args, cap, _p(offs, _i64), _p(nbr, _u32))
        if k <= cap:
            return offs, nbr[:k].copy()
        cap = k


def nlist_brute(pos, gpos, cutoff, coef):
    p = _f64(pos); g = _f64(gpos) if gpos is not None and len(gpos) else np.zeros((0, 3))
    return _csr(oracle_lib().orc_nlist_brute, len(p), len(p), _p(p, _d), len(g), _p(g, _d),
                cutoff, coef)


def nlist_cell(pos, gpos, box, cutoff, coef):
    p = _f64(pos); g = _f64(gpos) if gpos is not None and len(gpos) else np.zeros((0, 3))
    return _csr(oracle_lib().orc_nlist_cell, len(p), len(p), _p(p, _d), len(g), _p(g, _d),
                C.byref(box), cutoff, coef)


def lj(pos, type_, id_, gpos, gtype, gid, offs, nbr, eps, sigma, cutoff, mass,
       cutoff_list=None, jump_tol=0.0, mol=None, gmol=None):
    """Reference LJ pass: returns (acc[N,3], virial, pe).  With mol (owned
    molecule indices) and gmol (the ghosts'), Lj::ignore_intra_molecule is on."""
    p = _f64(pos); n = len(p)
    t = np.ascontiguousarray(type_, np.int32); i = np.ascontiguousarray(id_, np.int64)
    g = _f64(gpos) if gpos is not None and len(gpos) else np.zeros((0, 3))
    gt = np.ascontiguousarray(gtype if gtype is not None else [], np.int32)
    gi = np.ascontiguousarray(gid if gid is not None else [], np.int64)
    eps = _f64(np.atleast_2d(eps)); sig = _f64(np.atleast_2d(sigma))
    cl = _f64(np.atleast_2d(cutoff_list)) if cutoff_list is not None else None
    minv = _f64(1.0 / np.atleast_1d(np.asarray(mass, np.float64)))
    m = np.ascontiguousarray(mol, np.int32) if mol is not None else None
    gm = np.ascontiguousarray(gmol if gmol is not None else np.zeros(len(g)), np.int32)
    prm = LjParams(eps.shape[0], _p(eps, _d), _p(sig, _d), _p(cl, _d), cutoff, _p(minv, _d), jump_tol,
                   1 if m is not None else 0, _p(m, _i32), _p(gm, _i32))
    acc = np.zeros((n, 3)); vir = np.zeros(1); pe = np.zeros(1)
    oracle_lib().orc_lj(n, _p(p, _d), _p(t, _i32), _p(i, _i64), len(g), _p(g, _d), _p(gt, _i32),
                        _p(gi, _i64), _p(np.ascontiguousarray(offs, np.int64), _i64),
                        _p(np.ascontiguousarray(nbr, np.uint32), _u32), C.byref(prm),
                        _p(acc, _d), _p(vir, _d), _p(pe, _d))
    return acc, float(vir[0]), float(pe[0])


def pair_energy(r2, eps, sigma, rc2):
    return oracle_lib().orc_lj_pair_energy(r2, eps, sigma, rc2)


def nve_lambda(ke_now, dof, energy_per_dof=0.0, energy_tot=0.0, kbt=0.0, kb=0.0, temperature=0.0):
    """Velocity factor of constraint::Nve::apply_thermostat (nve.cpp:147-184)."""
    return oracle_lib().orc_nve_lambda(float(ke_now), int(dof), float(energy_per_dof), float(energy_tot),
                                       float(kbt), float(kb), float(temperature))


def kinetic_energy(vel, type_, mass, n_r_df):
    v = _f64(vel); t = np.ascontiguousarray(type_, np.int32); m = _f64(np.atleast_1d(mass))
    return oracle_lib().orc_kinetic_energy(len(v), _p(v, _d), _p(t, _i32), _p(m, _d), n_r_df)


# ---------------------------------------------------------------- simulators
class _SimBase:
    n: int

    def close(self):
        pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class OracleSim(_SimBase):
    """Whole-step restatement of Md_simulator (md_simulator.cpp:445-869)."""

    def __init__(self, sysdef: dict):
        s = sysdef
        self._keep = []
        eps = _f64(np.atleast_2d(s["eps"])); sig = _f64(np.atleast_2d(s["sigma"]))
        cl = _f64(np.atleast_2d(s["cutoff_list"])) if s.get("cutoff_list") is not None else None
        mass = _f64(np.atleast_1d(s["mass"]))
        self._keep += [eps, sig, cl, mass]
        prm = SimParams()
        prm.box = make_box(s["lo"], s["hi"], s.get("bc", (1, 1, 1)))
        prm.ghost_cutoff = s["ghost_cutoff"]; prm.nl_cutoff = s["nl_cutoff"]; prm.nl_coef = s["nl_coef"]
        prm.use_cell_list = int(s.get("use_cell_list", 0)); prm.rebuild_every_step = int(s.get("rebuild_test", 0))
        prm.ntypes = eps.shape[0]; prm.eps = _p(eps, _d); prm.sigma = _p(sig, _d)
        prm.cutoff_list = _p(cl, _d); prm.mass = _p(mass, _d); prm.lj_cutoff = s["lj_cutoff"]
        prm.integrator = int(s.get("integrator", 0)); prm.dt = s["dt"]
        prm.langevin_gamma = s.get("gamma", 0.0); prm.langevin_kbt = s.get("kbt", 0.0)
        prm.k_b = s.get("k_b", 1.0); prm.n_r_df = int(s.get("n_r_df", 3))
        pos = _f64(s["pos"]); vel = _f64(s["vel"])
        typ = np.ascontiguousarray(s.get("type", np.zeros(len(pos))), np.int32)
        ids = np.arange(len(pos), dtype=np.int64)
        self.n = len(pos)
        self.L = oracle_lib()
        self.h = self.L.orc_sim_create(C.byref(prm), self.n, _p(pos, _d), _p(vel, _d), _p(typ, _i32), _p(ids, _i64))

    def setup(self):
        self.L.orc_sim_setup(self.h)

    def set_molecules(self, mol, ignore=True):
        m = np.ascontiguousarray(mol, np.int32)
        self.L.orc_sim_set_molecules(self.h, _p(m, _i32), int(ignore))

    def step(self, eta=None):
        e = _f64(eta) if eta is not None else None
        return bool(self.L.orc_sim_step(self.h, _p(e, _d)))

    def state(self):
        p = np.empty((self.n, 3)); v = np.empty((self.n, 3)); a = np.empty((self.n, 3))
        self.L.orc_sim_get(self.h, _p(p, _d), _p(v, _d), _p(a, _d))
        return p, v, a

    def thermo(self):
        out = [np.zeros(1) for _ in range(4)]
        self.L.orc_sim_thermo(self.h, *[_p(o, _d) for o in out])
        return dict(ke=out[0][0], temperature=out[1][0], virial=out[2][0], pe=out[3][0])

    def nghost(self):
        return self.L.orc_sim_nghost(self.h)

    def scale_velocities(self, lam):
        self.L.orc_sim_scale_velocities(self.h, float(lam))

    def close(self):
        if self.h:
            self.L.orc_sim_destroy(self.h)
            self.h = None

    __del__ = close


_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    """Load the compiled reference harness (built by ``make -C oracle ref``)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(f"{REF_SO} missing: build it with `make -C oracle ref`")
        L = C.CDLL(REF_SO)
        L.ref_sim_create.argtypes = [_d, _d, C.POINTER(C.c_int), C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, _d, _d, _d, _d, C.c_double, C.c_int,
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_int64, _d, _d, _i32]
        L.ref_sim_create.restype = _vp
        for name in ("ref_sim_destroy", "ref_sim_setup"):
            getattr(L, name).argtypes = [_vp]
        L.ref_sim_step.argtypes = [_vp]; L.ref_sim_step.restype = C.c_int
        L.ref_sim_natoms.argtypes = [_vp]; L.ref_sim_natoms.restype = C.c_int64
        L.ref_sim_get.argtypes = [_vp, _d, _d, _d]
        L.ref_sim_thermo.argtypes = [_vp, _d, _d, _d]
        L.ref_sim_eta.argtypes = [_vp, _d]
        L.ref_sim_nghost.argtypes = [_vp]; L.ref_sim_nghost.restype = C.c_int64
        L.ref_sim_ghosts.argtypes = [_vp, _d, _i64]
        L.ref_sim_nlist.argtypes = [_vp, C.c_int64, _i64, _u32]; L.ref_sim_nlist.restype = C.c_int64
        L.ref_sim_set_molecules.argtypes = [_vp, _i32, C.c_int]
        L.ref_random_velocity.argtypes = [C.c_int64, C.c_uint, C.c_double, _d]
        L.ref_lj_tables.argtypes = ([C.c_int64, _i32, C.c_int, _d, C.c_int, _d, C.c_int, _i32, _d, C.c_int,
                                     _i32, _d, C.c_int] + [C.c_int, _i32, _d] * 3 + [C.c_int, _d, _d, _d])
        _ref = L
    return _ref


class RefSim(_SimBase):
    """The compiled reference driven through oracle/ref_driver.cpp."""

    def __init__(self, sysdef: dict):
        s = sysdef
        self.L = ref_lib()
        eps = _f64(np.atleast_2d(s["eps"])); sig = _f64(np.atleast_2d(s["sigma"]))
        cl = _f64(np.atleast_2d(s["cutoff_list"])) if s.get("cutoff_list") is not None else None
        mass = _f64(np.atleast_1d(s["mass"]))
        lo = _f64(s["lo"]); hi = _f64(s["hi"])
        bc = (C.c_int * 3)(*[int(x) for x in s.get("bc", (1, 1, 1))])
        pos = _f64(s["pos"]); vel = _f64(s["vel"])
        typ = np.ascontiguousarray(s.get("type", np.zeros(len(pos))), np.int32)
        self.h = self.L.ref_sim_create(
            _p(lo, _d), _p(hi, _d), bc, s["ghost_cutoff"], s["nl_cutoff"], s["nl_coef"],
            int(s.get("use_cell_list", 0)), int(s.get("rebuild_test", 0)), eps.shape[0],
            _p(eps, _d), _p(sig, _d), _p(cl, _d), _p(mass, _d), s["lj_cutoff"],
            int(s.get("integrator", 0)), s["dt"], s.get("gamma", 0.0), s.get("kbt", 0.0),
            s.get("k_b", 1.0), int(s.get("n_r_df", 3)), len(pos), _p(pos, _d), _p(vel, _d),
            _p(typ, _i32))
        self.n = self.L.ref_sim_natoms(self.h)

    def setup(self):
        self.L.ref_sim_setup(self.h)

    def set_molecules(self, mol, ignore=True):
        m = np.ascontiguousarray(mol, np.int32)
        self.L.ref_sim_set_molecules(self.h, _p(m, _i32), int(ignore))

    def step(self):
        return bool(self.L.ref_sim_step(self.h))

    def eta(self):
        e = np.empty((self.n, 3)); self.L.ref_sim_eta(self.h, _p(e, _d)); return e

    def state(self):
        p = np.empty((self.n, 3)); v = np.empty((self.n, 3)); a = np.empty((self.n, 3))
        self.L.ref_sim_get(self.h, _p(p, _d), _p(v, _d), _p(a, _d))
        return p, v, a

    def thermo(self):
        out = [np.zeros(1) for _ in range(3)]
        self.L.ref_sim_thermo(self.h, *[_p(o, _d) for o in out])
        return dict(ke=out[0][0], temperature=out[1][0], virial=out[2][0])

    def ghosts(self):
        ng = self.L.ref_sim_nghost(self.h)
        g = np.empty((ng, 3)); gid = np.empty(ng, np.int64)
        self.L.ref_sim_ghosts(self.h, _p(g, _d), _p(gid, _i64))
        return g, gid

    def nlist(self):
        k = self.L.ref_sim_nlist(self.h, 0, None, None)
        offs = np.zeros(self.n + 1, np.int64); nbr = np.empty(max(k, 1), np.uint32)
        self.L.ref_sim_nlist(self.h, k, _p(offs, _i64), _p(nbr, _u32))
        return offs, nbr[:k].copy()

    def close(self):
        if getattr(self, "h", None):
            self.L.ref_sim_destroy(self.h)
            self.h = None

    __del__ = close


def ref_lj_tables(types, k, eps_atom=(), sigma_atom=(), eps_set=(), sigma_set=(), wca=False,
                  cutoff_list=None, lambda_s=None, lambda_e=None):
    """Lj::verify_settings (lj.cpp:180-391) of the compiled reference on the
    given inputs; the epsilon, sigma and cutoff-list tables it forms as dense
    k x k arrays.  eps_set / sigma_set: (a, b, value) set_epsilon / set_sigma
    calls; cutoff_list, lambda_s, lambda_e: ragged lists of rows."""
    def rag(m):
        if m is None:
            return 0, None, None
        ln = np.array([len(r) for r in m], np.int32)
        v = _f64(np.array([x for r in m for x in r], np.float64)) if ln.sum() else _f64(np.zeros(1))
        return len(m), ln, v
    ty = np.ascontiguousarray(types, np.int32)
    ea = _f64(np.array(eps_atom, np.float64)) if len(eps_atom) else None
    sa = _f64(np.array(sigma_atom, np.float64)) if len(sigma_atom) else None
    eab = np.array([(a, b) for a, b, _ in eps_set], np.int32).reshape(-1, 2)
    ev = _f64(np.array([v for _, _, v in eps_set], np.float64))
    sab = np.array([(a, b) for a, b, _ in sigma_set], np.int32).reshape(-1, 2)
    sv = _f64(np.array([v for _, _, v in sigma_set], np.float64))
    out = [np.zeros((k, k)) for _ in range(3)]
    args = [len(ty), _p(ty, _i32), len(eps_atom), _p(ea, _d), len(sigma_atom), _p(sa, _d),
            len(eps_set), _p(np.ascontiguousarray(eab), _i32), _p(ev, _d),
            len(sigma_set), _p(np.ascontiguousarray(sab), _i32), _p(sv, _d), int(bool(wca))]
    keep = []
    for m in (cutoff_list, lambda_s, lambda_e):
        nr, ln, v = rag(m)
        keep += [ln, v]
        args += [nr, _p(ln, _i32), _p(v, _d)]
    args += [k] + [_p(o, _d) for o in out]
    ref_lib().ref_lj_tables(*args)
    return out


def ref_random_velocity(n, seed, amplitude):
    v = np.empty((n, 3))
    ref_lib().ref_random_velocity(n, seed, amplitude, _p(v, _d))
    return v
