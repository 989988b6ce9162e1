"""The larger configurations on the device.

256,000-atom FP64 LJ fluid (40^3 FCC cells, rho* = 0.8442, T* = 1.44, rc 2.5,
skin 0.3, velocity Verlet dt = 0.005, 10^4 steps):

- Every list build of the run is checked on the device (CV2_NL_VERIFY: an
  all-pairs FP64 recomputation of the pair set, compared by checksum and
  count; the check itself is pinned against the oracle in
  test_gpu_verify.py): no build of the 10^4 steps differs.
- Pair sets are bit-exact against the oracle at the device's positions of
  each of the first rebuilds, and after the 10^4 steps (a build forced at
  the final positions).  Trajectories of the two diverge, so each check is a
  function of the positions the device built from, as in the reference's
  own predicate.
- The total energy is sampled every 1000 steps; its per-atom excursion stays
  within the bound the 4000-atom run keeps.

The oracle's cell-list build takes about 3 s per check at this size, so the
rebuilds checked are the first four and the last.

1,000,000-atom Kob-Andersen mixture as one MPI domain's share (rho* = 1.2,
80:20, rc_ij = 2.5 sigma_ij, 2.8 sigma halo supplied by the host, FP32 +
fixed point): the pair set with the uploaded ghosts bit-exact against the
oracle's cell list and the forces within the FP32+fixed-point tolerance of
the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O
from cv2md import device as D
from cv2md.systems import ka_mixture, lj_fluid
from helpers import force_error, oracle_forces, oracle_pairs, sort_rows

pytestmark = pytest.mark.gpu

DT = 0.005
STEPS = 10000


def check_pairs(c, s, tag):
    pos = c.download()[0]
    got = sort_rows(c.pairs())
    exp = oracle_pairs(pos, dict(s, pos=pos))
    print(f"[c2] {tag}: {len(got)} pairs", flush=True)
    assert len(got) == len(exp), tag
    assert np.array_equal(got, exp), tag


def energy(c):
    t = c.thermo()
    return t["ke"] + t["pe"]


def test_256k_pairs_at_rebuilds_and_energy():
    s = lj_fluid(40, temperature=1.44, seed=1002)
    n = len(s["pos"])
    assert n == 256000
    with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP64, verify=True) as c:
        c.compute_forces()
        check_pairs(c, s, "setup")
        e = [energy(c)]
        step, checked = 0, 0
        # velocity Verlet drifts positions before the build of a step, so
        # after a step that rebuilt the positions are the build positions
        while checked < 4:
            r0 = c.stats()["rebuilds"]
            c.run_vv(1, DT)
            step += 1
            if c.stats()["rebuilds"] != r0:
                check_pairs(c, s, f"rebuild at step {step}")
                checked += 1
            assert step < 300, "fewer rebuilds than expected"
        while step < STEPS:
            k = min(1000, STEPS - step)
            c.run_vv(k, DT)
            step += k
            e.append(energy(c))
        rebuilds = c.stats()["rebuilds"]
        c.build(True)
        check_pairs(c, s, f"after {STEPS} steps ({rebuilds} rebuilds)")
        st = c.stats()
    print(f"[c2] {st['verified_builds']} of {st['rebuilds']} builds checked on the device, "
          f"{st['verify_mismatches']} differed", flush=True)
    assert st["verified_builds"] == st["rebuilds"] == rebuilds + 1
    assert st["verify_mismatches"] == 0
    e = np.asarray(e)
    de = np.abs(e - e[0]).max() / n
    print(f"[c2] E0 {e[0]:.4f}, max |E - E0| per atom {de:.3e}", flush=True)
    assert rebuilds > 100
    assert de < 2e-3


def test_1m_ka_host_halo():
    s = ka_mixture(1000000, seed=1005)
    n = len(s["pos"])
    box = O.make_box(s["lo"], s["hi"], s["bc"])
    p, _ = O.wrap(s["pos"], box)
    gpos, owner, _ = O.ghosts(p, box, s["ghost_cutoff"])
    t = np.asarray(s["type"], np.int32)
    print(f"[c5] {n} owned, {len(gpos)} ghosts", flush=True)
    with D.Context.from_sysdef(dict(s, pos=p), precision=D.CV2_PREC_FP32_FIXED) as c:
        c.set_ghosts(gpos, t[owner], owner)
        c.compute_forces()
        acc = c.download()[2]
        pairs = sort_rows(c.pairs())
        th = c.thermo(n_r_df=0)
    offs, nbr = O.nlist_cell(p, gpos, box, s["nl_cutoff"], s["nl_coef"])
    exp = sort_rows(np.stack([np.repeat(np.arange(n), np.diff(offs)), nbr.astype(np.int64),
                              np.full(len(nbr), 13)], 1))
    print(f"[c5] {len(pairs)} pairs", flush=True)
    assert len(pairs) == len(exp)
    assert np.array_equal(pairs, exp)
    _, a_ref, _vir, _pe = oracle_forces(s["pos"], s)
    emax, erms = force_error(acc, a_ref)
    print(f"[c5] force error max {emax:.2e} rms {erms:.2e}", flush=True)
    assert emax < 2e-5 and erms < 5e-6
    assert np.isfinite(th["pe"])
