"""ctypes binding of the C ABI in include/cv2_md.h (lib/libcv2md.so).

There is no CPU fallback: if the HIP library is missing or no GPU is
visible, constructing a Context raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libcv2md.so")

CV2_PREC_FP64 = 0
CV2_PREC_FP32_FIXED = 1
CV2_NL_REBUILD_EVERY_STEP = 1
CV2_NL_REFERENCE_REBUILDS = 2
CV2_NL_VERIFY = 4

ERRORS = {1: "HIP", 2: "BAD_ARG", 3: "NAN", 4: "FIXED_OVERFLOW", 5: "CAPACITY", 6: "STATE"}

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)


class Stats(C.Structure):
    _fields_ = [("natoms", C.c_int64), ("nghost", C.c_int64), ("steps", C.c_int64),
                ("rebuilds", C.c_int64), ("list_entries", C.c_int64), ("max_neighbors", C.c_int32),
                ("capacity", C.c_int32), ("ncell", C.c_int32 * 3), ("precision", C.c_int32),
                ("build_kind", C.c_int32), ("force_tasks", C.c_int32), ("verified_builds", C.c_int64),
                ("verify_mismatches", C.c_int64), ("verify_sum", C.c_uint64 * 2),
                ("verify_pairs", C.c_int64 * 2)]


class BuildFilter(C.Structure):
    _fields_ = [("qunit", C.c_double * 3), ("lo", C.c_double * 3), ("ce", C.c_double),
                ("cell_width", C.c_double * 3), ("cthr_in", C.c_float), ("cthr_out", C.c_float),
                ("coarse", C.c_int32), ("ncell", C.c_int32 * 3)]


class Timing(C.Structure):
    _fields_ = [("force_ms", C.c_double), ("force_calls", C.c_int64), ("build_ms", C.c_double),
                ("integrate_ms", C.c_double)]


class HaloGeom(C.Structure):
    _fields_ = [("grid_index", C.c_int32 * 3), ("nprocs", C.c_int32 * 3), ("lo_local", C.c_double * 3),
                ("hi_local", C.c_double * 3), ("size_global", C.c_double * 3), ("bc", C.c_int32 * 3),
                ("ghost_cutoff", C.c_double), ("with_velocity", C.c_int32)]


class EnsTuning(C.Structure):
    _fields_ = [("force_block", C.c_int32), ("build_block", C.c_int32), ("rebuild_all", C.c_int32),
                ("ms_per_step", C.c_double), ("cell_scale", C.c_double)]


class EnsStats(C.Structure):
    _fields_ = [("natoms", C.c_int64), ("nsys", C.c_int32), ("capacity", C.c_int32),
                ("steps", C.c_int64), ("build_steps", C.c_int64), ("system_builds", C.c_int64),
                ("list_entries", C.c_int64), ("max_neighbors", C.c_int32), ("graph_nodes", C.c_int32)]


class Cv2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cv2 error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


_lib = None


def lib():
    """Load libcv2md.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (run __graft_entry__.build() or make -C caviar-v2_amd)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "cv2_create": ([C.POINTER(vp), C.c_int], C.c_int),
            "cv2_destroy": ([vp], C.c_int),
            "cv2_last_error": ([vp], C.c_char_p),
            "cv2_set_box": ([vp, _d, _d, _i32, C.c_double], C.c_int),
            "cv2_set_lj": ([vp, C.c_int, _d, _d, _d, C.c_double, _d], C.c_int),
            "cv2_set_neighbor": ([vp, C.c_double, C.c_double, C.c_uint], C.c_int),
            "cv2_set_jump_tol": ([vp, C.c_double], C.c_int),
            "cv2_scale_velocities": ([vp, C.c_double], C.c_int),
            "cv2_set_precision": ([vp, C.c_int], C.c_int),
            "cv2_upload_atoms": ([vp, C.c_int64, _d, _d, _i32, _i64], C.c_int),
            "cv2_set_ghosts": ([vp, C.c_int64, _d, _i32, _i64], C.c_int),
            "cv2_set_molecules": ([vp, C.c_int64, _i32, C.c_int], C.c_int),
            "cv2_owned_pack": ([vp, C.POINTER(HaloGeom), _i64, C.c_int64, _d, _i64], C.c_int),
            "cv2_owned_unpack": ([vp, C.c_int32, _i64, _d], C.c_int),
            "cv2_set_type_table": ([vp, C.c_int64, _i32], C.c_int),
            "cv2_get_ids": ([vp, _i64], C.c_int),
            "cv2_halo_pack": ([vp, C.POINTER(HaloGeom), _i64, C.c_int64, _d, _i64], C.c_int),
            "cv2_halo_unpack": ([vp, C.c_int32, _i64, _d, C.c_int32], C.c_int),
            "cv2_halo_unpack_at": ([vp, C.c_int32, _i64, _i64, _d, C.c_int32], C.c_int),
            "cv2_host_alloc": ([C.c_int64, C.POINTER(vp)], C.c_int),
            "cv2_host_free": ([vp], C.c_int),
            "cv2_halo_pack_noalloc": ([vp, C.POINTER(HaloGeom), _i64, C.c_int64, _d, _i64], C.c_int),
            "cv2_build_neighbors": ([vp, C.c_int, C.POINTER(C.c_int)], C.c_int),
            "cv2_compute_forces": ([vp], C.c_int),
            "cv2_compute_accelerations": ([vp], C.c_int),
            "cv2_run_vv": ([vp, C.c_int64, C.c_double], C.c_int),
            "cv2_step_begin": ([vp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_uint64], C.c_int),
            "cv2_step_end": ([vp, C.c_int], C.c_int),
            "cv2_run_langevin": ([vp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64], C.c_int),
            "cv2_thermo": ([vp, C.c_int, _d, _d, _d], C.c_int),
            "cv2_download": ([vp, _d, _d, _d], C.c_int),
            "cv2_get_pairs": ([vp, C.c_int64, _i64, _i64], C.c_int),
            "cv2_get_stats": ([vp, C.POINTER(Stats)], C.c_int),
            "cv2_get_build_filter": ([vp, C.POINTER(BuildFilter)], C.c_int),
            "cv2_synchronize": ([vp], C.c_int),
            "cv2_set_timing": ([vp, C.c_int], C.c_int),
            "cv2_get_timing": ([vp, C.POINTER(Timing)], C.c_int),
            "cv2_ens_create": ([C.POINTER(vp), C.c_int], C.c_int),
            "cv2_ens_destroy": ([vp], C.c_int),
            "cv2_ens_last_error": ([vp], C.c_char_p),
            "cv2_ens_set_lj": ([vp, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
            "cv2_ens_set_neighbor": ([vp, C.c_double, C.c_double], C.c_int),
            "cv2_ens_upload": ([vp, C.c_int32, _i64, _d, _d, _d, _d], C.c_int),
            "cv2_ens_set_tuning": ([vp, C.POINTER(EnsTuning)], C.c_int),
            "cv2_ens_tune": ([vp, C.c_double, C.c_int32, C.POINTER(EnsTuning)], C.c_int),
            "cv2_ens_run_vv": ([vp, C.c_int64, C.c_double], C.c_int),
            "cv2_ens_build_neighbors": ([vp], C.c_int),
            "cv2_ens_thermo": ([vp, _d, _d, _d], C.c_int),
            "cv2_ens_download": ([vp, _d, _d, _d], C.c_int),
            "cv2_ens_get_pairs": ([vp, C.c_int32, C.c_int64, _i64, _i64], C.c_int),
            "cv2_ens_get_stats": ([vp, C.POINTER(EnsStats)], C.c_int),
            "cv2_ens_use_graph": ([vp, C.c_int], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def pinned_empty(n: int) -> np.ndarray:
    """float64 array of n elements in page-locked host memory
    (cv2_host_alloc), released when the array is collected."""
    import weakref

    L = lib()
    ptr = C.c_void_p()
    rc = L.cv2_host_alloc(8 * max(int(n), 1), C.byref(ptr))
    if rc:
        raise Cv2Error(rc, "cv2_host_alloc failed")
    a = np.ctypeslib.as_array(C.cast(ptr, _d), shape=(max(int(n), 1),))
    weakref.finalize(a, L.cv2_host_free, ptr)
    return a[:int(n)] if n else a[:0]


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device-resident MD system (see include/cv2_md.h)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.cv2_create(C.byref(h), device)
        if rc:
            raise Cv2Error(rc, "cv2_create failed (no GPU?)")
        self.h = h
        self.n = 0

    def _chk(self, rc):
        if rc:
            raise Cv2Error(rc, self.L.cv2_last_error(self.h).decode())

    # ---- configuration
    def set_box(self, lo, hi, bc=(1, 1, 1), ghost_cutoff=0.0):
        lo = _f64(lo); hi = _f64(hi); b = np.ascontiguousarray(bc, np.int32)
        self._chk(self.L.cv2_set_box(self.h, _p(lo, _d), _p(hi, _d), _p(b, _i32), float(ghost_cutoff)))

    def set_lj(self, eps, sigma, cutoff, mass, cutoff_list=None):
        e = _f64(np.atleast_2d(eps)); s = _f64(np.atleast_2d(sigma))
        cl = _f64(np.atleast_2d(cutoff_list)) if cutoff_list is not None else None
        m = _f64(np.atleast_1d(mass))
        self._chk(self.L.cv2_set_lj(self.h, e.shape[0], _p(e, _d), _p(s, _d), _p(cl, _d), float(cutoff), _p(m, _d)))

    def set_jump_tol(self, jump_tol):
        """jump_fix: pairs closer than sqrt(jump_tol) take the force of r^2 = jump_tol."""
        self._chk(self.L.cv2_set_jump_tol(self.h, float(jump_tol)))

    def scale_velocities(self, lam):
        """Velocity rescale of the NVE constraint: every owned velocity times lam."""
        self._chk(self.L.cv2_scale_velocities(self.h, float(lam)))

    def set_neighbor(self, cutoff, extra_coef, flags=0):
        self._chk(self.L.cv2_set_neighbor(self.h, float(cutoff), float(extra_coef), int(flags)))

    def set_precision(self, prec):
        self._chk(self.L.cv2_set_precision(self.h, int(prec)))

    def upload(self, pos, vel=None, type_=None, id_=None):
        p = _f64(pos); v = _f64(vel)
        t = np.ascontiguousarray(type_, np.int32) if type_ is not None else None
        i = np.ascontiguousarray(id_, np.int64) if id_ is not None else None
        self.n = len(p)
        self._chk(self.L.cv2_upload_atoms(self.h, self.n, _p(p, _d), _p(v, _d), _p(t, _i32), _p(i, _i64)))

    def set_molecules(self, mol, ignore=True):
        """Owned molecule indices (upload order) and Lj::ignore_intra_molecule."""
        m = np.ascontiguousarray(mol, np.int32) if mol is not None else None
        n = len(m) if m is not None else self.n
        self._chk(self.L.cv2_set_molecules(self.h, n, _p(m, _i32), int(ignore)))

    def set_ghosts(self, pos, type_=None, id_=None):
        p = _f64(pos)
        t = np.ascontiguousarray(type_, np.int32) if type_ is not None else None
        i = np.ascontiguousarray(id_, np.int64) if id_ is not None else None
        self._chk(self.L.cv2_set_ghosts(self.h, len(p), _p(p, _d), _p(t, _i32), _p(i, _i64)))

    @classmethod
    def from_sysdef(cls, s: dict, precision=CV2_PREC_FP64, device=0, rebuild_every_step=False,
                    reference_rebuilds=False, verify=False):
        c = cls(device)
        c.set_box(s["lo"], s["hi"], s.get("bc", (1, 1, 1)), s["ghost_cutoff"])
        c.set_lj(s["eps"], s["sigma"], s["lj_cutoff"], s["mass"], s.get("cutoff_list"))
        flags = (CV2_NL_REBUILD_EVERY_STEP if rebuild_every_step else 0) | \
            (CV2_NL_REFERENCE_REBUILDS if reference_rebuilds else 0) | (CV2_NL_VERIFY if verify else 0)
        c.set_neighbor(s["nl_cutoff"], s["nl_coef"], flags)
        c.set_precision(precision)
        c.upload(s["pos"], s.get("vel"), s.get("type"))
        return c

    # ---- work
    def halo_pack(self, grid_index, nprocs, lo_local, hi_local, size_global, bc, ghost_cutoff,
                  with_velocity=False, noalloc=False):
        """MPI halo packets of the owned atoms in the reference's layout:
        (counts[27], packets concatenated in direction order d = i*9+j*3+k).
        noalloc: through cv2_halo_pack_noalloc (the library's static scratch)."""
        fn = self.L.cv2_halo_pack_noalloc if noalloc else self.L.cv2_halo_pack
        g = HaloGeom()
        for a in range(3):
            g.grid_index[a] = int(grid_index[a]); g.nprocs[a] = int(nprocs[a]); g.bc[a] = int(bc[a])
            g.lo_local[a] = float(lo_local[a]); g.hi_local[a] = float(hi_local[a])
            g.size_global[a] = float(size_global[a])
        g.ghost_cutoff = float(ghost_cutoff)
        g.with_velocity = int(bool(with_velocity))
        counts = np.zeros(27, np.int64)
        need = C.c_int64()
        self._chk(fn(self.h, C.byref(g), _p(counts, _i64), 0, None, C.byref(need)))
        out = np.empty(max(need.value, 1))
        self._chk(fn(self.h, C.byref(g), _p(counts, _i64), need.value, _p(out, _d), C.byref(need)))
        return counts, out[:need.value]

    def owned_pack(self, grid_index, nprocs, lo_local, hi_local, size_global, bc, cap=None):
        """Owned-atom exchange, sending half: (counts[27], packets concatenated
        in direction order).  The leavers leave this context.  cap: packet
        doubles offered in one call (None: size first, then pack)."""
        g = HaloGeom()
        for a in range(3):
            g.grid_index[a] = int(grid_index[a]); g.nprocs[a] = int(nprocs[a]); g.bc[a] = int(bc[a])
            g.lo_local[a] = float(lo_local[a]); g.hi_local[a] = float(hi_local[a])
            g.size_global[a] = float(size_global[a])
        counts = np.zeros(27, np.int64)
        need = C.c_int64()
        if cap is None:
            self._chk(self.L.cv2_owned_pack(self.h, C.byref(g), _p(counts, _i64), 0, None, C.byref(need)))
            cap = need.value
        out = np.empty(max(int(cap), 1))
        self._chk(self.L.cv2_owned_pack(self.h, C.byref(g), _p(counts, _i64), int(cap), _p(out, _d),
                                        C.byref(need)))
        self.n = self.stats()["natoms"]
        return counts, out[:need.value]

    def owned_unpack(self, counts, packets):
        """Owned-atom exchange, receiving half: the atoms of the packets join."""
        cnt = np.ascontiguousarray(counts, np.int64)
        pk = np.ascontiguousarray(packets, np.float64)
        self._chk(self.L.cv2_owned_unpack(self.h, len(cnt), _p(cnt, _i64), _p(pk, _d) if pk.size else None))
        self.n = self.stats()["natoms"]

    def set_type_table(self, type_by_id):
        t = np.ascontiguousarray(type_by_id, np.int32)
        self._chk(self.L.cv2_set_type_table(self.h, len(t), _p(t, _i32)))

    def ids(self):
        """Ids of the owned atoms in owned order (the order of download())."""
        out = np.empty(max(self.n, 1), np.int64)
        self._chk(self.L.cv2_get_ids(self.h, _p(out, _i64)))
        return out[:self.n]

    def halo_pack_into(self, out, grid_index, nprocs, lo_local, hi_local, size_global, bc, ghost_cutoff,
                       with_velocity=False):
        """cv2_halo_pack into a caller's float64 buffer in one call: returns
        (counts, need); the packets are out[:need] when need <= len(out),
        otherwise out is untouched and the caller grows it."""
        g = HaloGeom()
        for a in range(3):
            g.grid_index[a] = int(grid_index[a]); g.nprocs[a] = int(nprocs[a]); g.bc[a] = int(bc[a])
            g.lo_local[a] = float(lo_local[a]); g.hi_local[a] = float(hi_local[a])
            g.size_global[a] = float(size_global[a])
        g.ghost_cutoff = float(ghost_cutoff)
        g.with_velocity = int(bool(with_velocity))
        counts = np.zeros(27, np.int64)
        need = C.c_int64()
        assert out.dtype == np.float64 and out.flags.c_contiguous
        self._chk(self.L.cv2_halo_pack(self.h, C.byref(g), _p(counts, _i64), len(out), _p(out, _d),
                                       C.byref(need)))
        return counts, need.value

    def halo_unpack_at(self, counts, offsets, buf, with_velocity=False):
        """Ghosts from packets at offsets (doubles) of one buffer, appended in
        the order given (cv2_halo_unpack_at)."""
        cnt = np.ascontiguousarray(counts, np.int64)
        off = np.ascontiguousarray(offsets, np.int64)
        assert buf.dtype == np.float64 and buf.flags.c_contiguous and len(off) == len(cnt)
        self._chk(self.L.cv2_halo_unpack_at(self.h, len(cnt), _p(cnt, _i64), _p(off, _i64),
                                            _p(buf, _d) if buf.size else None, int(bool(with_velocity))))

    def halo_unpack(self, counts, packets, with_velocity=False):
        """Ghosts from received packets, appended in the order given."""
        cnt = np.ascontiguousarray(counts, np.int64)
        pk = np.ascontiguousarray(packets, np.float64)
        self._chk(self.L.cv2_halo_unpack(self.h, len(cnt), _p(cnt, _i64), _p(pk, _d) if pk.size else None,
                                         int(bool(with_velocity))))

    def build(self, force=True):
        r = C.c_int(0)
        self._chk(self.L.cv2_build_neighbors(self.h, int(force), C.byref(r)))
        return bool(r.value)

    def compute_forces(self):
        self._chk(self.L.cv2_compute_forces(self.h))

    def compute_accelerations(self):
        """Accelerations through the kernel the step loop runs (no energy pass)."""
        self._chk(self.L.cv2_compute_accelerations(self.h))

    def run_vv(self, nsteps, dt):
        self._chk(self.L.cv2_run_vv(self.h, int(nsteps), float(dt)))

    def run_langevin(self, nsteps, dt, gamma, kbt, seed):
        self._chk(self.L.cv2_run_langevin(self.h, int(nsteps), float(dt), float(gamma), float(kbt), int(seed)))

    def step_begin(self, dt, langevin=False, gamma=0.0, kbt=0.0, seed=0):
        """First half of one step (velocities, positions); the halo may be
        exchanged before step_end."""
        self._chk(self.L.cv2_step_begin(self.h, float(dt), int(bool(langevin)), float(gamma), float(kbt),
                                        int(seed)))

    def step_end(self, energy=True):
        """List build when needed, forces, second velocity half."""
        self._chk(self.L.cv2_step_end(self.h, int(bool(energy))))

    def synchronize(self):
        self._chk(self.L.cv2_synchronize(self.h))

    # ---- results
    def thermo(self, n_r_df=3):
        ke = C.c_double(); pe = C.c_double(); vir = C.c_double()
        self._chk(self.L.cv2_thermo(self.h, int(n_r_df), C.byref(ke), C.byref(pe), C.byref(vir)))
        return dict(ke=ke.value, pe=pe.value, virial=vir.value)

    def download(self):
        p = np.empty((self.n, 3)); v = np.empty((self.n, 3)); a = np.empty((self.n, 3))
        self._chk(self.L.cv2_download(self.h, _p(p, _d), _p(v, _d), _p(a, _d)))
        return p, v, a

    def pairs(self):
        """(i, j, image_code) rows in the reference's half-list convention."""
        k = C.c_int64()
        self._chk(self.L.cv2_get_pairs(self.h, 0, C.byref(k), None))
        out = np.empty((max(k.value, 1), 3), np.int64)
        self._chk(self.L.cv2_get_pairs(self.h, k.value, C.byref(k), _p(out, _i64)))
        return out[: k.value]

    def build_filter(self):
        """Fixed-point unit and coarse-filter thresholds of the list build."""
        f = BuildFilter()
        self._chk(self.L.cv2_get_build_filter(self.h, C.byref(f)))
        return dict(qunit=np.array(f.qunit[:]), lo=np.array(f.lo[:]), ce=f.ce,
                    cell_width=np.array(f.cell_width[:]), cthr_in=np.float32(f.cthr_in),
                    cthr_out=np.float32(f.cthr_out), coarse=f.coarse, ncell=tuple(f.ncell[:]))

    def stats(self):
        s = Stats()
        self._chk(self.L.cv2_get_stats(self.h, C.byref(s)))
        return dict(natoms=s.natoms, nghost=s.nghost, steps=s.steps, rebuilds=s.rebuilds,
                    list_entries=s.list_entries, max_neighbors=s.max_neighbors, capacity=s.capacity,
                    ncell=tuple(s.ncell), precision=s.precision,
                    build_kind=('cell_wave' if s.build_kind == 1 else 'per_atom'),
                    force_tasks=s.force_tasks, verified_builds=s.verified_builds,
                    verify_mismatches=s.verify_mismatches, verify_sum=tuple(s.verify_sum),
                    verify_pairs=tuple(s.verify_pairs))

    def set_timing(self, enable=True, period=1):
        """HIP-event timing: every kernel group (period 1), or only the force
        launch of every period-th force call."""
        self._chk(self.L.cv2_set_timing(self.h, int(period) if enable else 0))

    def timing(self):
        t = Timing()
        self._chk(self.L.cv2_get_timing(self.h, C.byref(t)))
        return dict(force_ms=t.force_ms, force_calls=t.force_calls, build_ms=t.build_ms,
                    integrate_ms=t.integrate_ms)

    def close(self):
        if getattr(self, "h", None):
            self.L.cv2_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Many independent periodic LJ systems advanced together (batched
    ensemble of include/cv2_md.h; FP32 pair arithmetic with 64-bit
    fixed-point accumulation, velocity Verlet, one captured graph per step)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.cv2_ens_create(C.byref(h), device)
        if rc:
            raise Cv2Error(rc, "cv2_ens_create failed (no GPU?)")
        self.h = h
        self.natoms = np.zeros(0, np.int64)

    def _chk(self, rc):
        if rc:
            raise Cv2Error(rc, self.L.cv2_ens_last_error(self.h).decode())

    @classmethod
    def from_systems(cls, systems, eps=1.0, sigma=1.0, rc=2.5, mass=1.0, nl_cutoff=2.5, nl_coef=1.12,
                     device=0):
        """systems: list of dicts with pos (n,3), vel (n,3), lo (3,), hi (3,)."""
        e = cls(device)
        e._chk(e.L.cv2_ens_set_lj(e.h, eps, sigma, rc, mass))
        e._chk(e.L.cv2_ens_set_neighbor(e.h, nl_cutoff, nl_coef))
        e.upload(systems)
        return e

    def upload(self, systems):
        nat = np.array([len(s["pos"]) for s in systems], np.int64)
        lo = _f64(np.concatenate([np.asarray(s["lo"], np.float64) for s in systems]))
        hi = _f64(np.concatenate([np.asarray(s["hi"], np.float64) for s in systems]))
        pos = _f64(np.concatenate([s["pos"] for s in systems]))
        vel = _f64(np.concatenate([s["vel"] for s in systems]))
        self._chk(self.L.cv2_ens_upload(self.h, len(systems), _p(nat, _i64), _p(lo, _d), _p(hi, _d),
                                        _p(pos, _d), _p(vel, _d)))
        self.natoms = nat
        self.offsets = np.concatenate([[0], np.cumsum(nat)])

    def set_tuning(self, force_block=256, build_block=128, rebuild_all=0, cell_scale=1.0):
        t = EnsTuning(force_block, build_block, rebuild_all, 0.0, float(cell_scale))
        self._chk(self.L.cv2_ens_set_tuning(self.h, C.byref(t)))

    def tune(self, dt, trial_steps=20):
        t = EnsTuning()
        self._chk(self.L.cv2_ens_tune(self.h, dt, int(trial_steps), C.byref(t)))
        return dict(force_block=t.force_block, build_block=t.build_block, rebuild_all=t.rebuild_all,
                    ms_per_step=t.ms_per_step, cell_scale=t.cell_scale)

    def use_graph(self, enable=True):
        self._chk(self.L.cv2_ens_use_graph(self.h, int(bool(enable))))

    def run_vv(self, nsteps, dt):
        self._chk(self.L.cv2_ens_run_vv(self.h, int(nsteps), float(dt)))

    def build(self):
        """Rebuild every system's list now (and recompute forces)."""
        self._chk(self.L.cv2_ens_build_neighbors(self.h))

    def thermo(self):
        k = len(self.natoms)
        ke, pe, vir = np.empty(k), np.empty(k), np.empty(k)
        self._chk(self.L.cv2_ens_thermo(self.h, _p(ke, _d), _p(pe, _d), _p(vir, _d)))
        return ke, pe, vir

    def download(self):
        n = int(self.natoms.sum())
        p = np.empty((n, 3)); v = np.empty((n, 3)); a = np.empty((n, 3))
        self._chk(self.L.cv2_ens_download(self.h, _p(p, _d), _p(v, _d), _p(a, _d)))
        return p, v, a

    def pairs(self, sys):
        k = C.c_int64()
        self._chk(self.L.cv2_ens_get_pairs(self.h, int(sys), 0, C.byref(k), None))
        out = np.empty((max(k.value, 1), 3), np.int64)
        self._chk(self.L.cv2_ens_get_pairs(self.h, int(sys), k.value, C.byref(k), _p(out, _i64)))
        return out[: k.value]

    def stats(self):
        s = EnsStats()
        self._chk(self.L.cv2_ens_get_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in EnsStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.L.cv2_ens_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
