"""The coarse distance filter of the wave-per-cell list build, with the
library's own fixed-point units and thresholds (cv2_get_build_filter), and
the FP64 band path of the list kernel driven on the device.

- The thresholds the library set up equal the restatement that
  tests/test_coarse_filter.py checks on the CPU, and with them the device's
  float arithmetic (int32 differences from a reference atom, float32 squared
  distance) never accepts a pair outside the list cutoff nor rejects one
  inside it, over 4e5 pairs within 3e-5 of the cutoff per box shape.
- Pairs placed at ce (1 + d) with |d| from 1e-15 to 3e-5, most of them
  inside the band between the two thresholds (the FP64 predicate decides
  there), give the oracle's pair set bit for bit on the device, in both
  precision forms.
"""
import numpy as np
import pytest

import oracle as O
from cv2md import device as D
from helpers import oracle_pairs, sort_rows
from test_coarse_filter import quantize, qunit, thresholds

pytestmark = pytest.mark.gpu

CE = 2.5 * 1.12


def sysdef(L, pos):
    n = len(pos)
    return dict(lo=np.zeros(3), hi=np.asarray(L, float), bc=(1, 1, 1), pos=pos, vel=np.zeros((n, 3)),
                type=np.zeros(n, np.int32), eps=np.ones((1, 1)), sigma=np.ones((1, 1)), cutoff_list=None,
                mass=np.ones(1), lj_cutoff=2.5, nl_cutoff=2.5, nl_coef=1.12, ghost_cutoff=CE, dt=0.005)


@pytest.mark.parametrize("L", [(134.3677,) * 3, (20.0, 30.0, 400.0), (14.5, 14.5, 14.5)])
def test_library_thresholds_are_sound(L):
    Ls = np.asarray(L)
    rng = np.random.default_rng(7)
    with D.Context.from_sysdef(sysdef(Ls, rng.uniform(0, 1, (64, 3)) * Ls)) as c:
        f = c.build_filter()
    assert f["coarse"] == 1 and abs(f["ce"] - CE) < 1e-12
    w = Ls / np.asarray(f["ncell"])
    # the CPU test's restatement is the library's setup
    u = qunit(Ls, 2.0 * max(w.max(), CE))
    tin, tout, _ = thresholds(u, CE, w.max())
    assert np.array_equal(f["qunit"], u)
    assert f["cthr_in"] == tin and f["cthr_out"] == tout
    unit = f["qunit"].astype(np.float32)
    n = 400000
    ref = rng.uniform(0, 1, (n, 3)) * Ls
    pi = ref + rng.uniform(-1, 1, (n, 3)) * w
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = CE * (1 + rng.uniform(-3e-5, 3e-5, n))
    pj = pi + d * r[:, None]
    pi, pj, ref = (np.mod(x, Ls) for x in (pi, pj, ref))
    dx = pj - pi
    dx -= Ls * np.round(dx / Ls)
    exact = ((dx[:, 0] * dx[:, 0] + dx[:, 1] * dx[:, 1]) + dx[:, 2] * dx[:, 2]) < CE * CE
    qo, qi, qj = (quantize(x - f["lo"], f["qunit"]) for x in (ref, pi, pj))
    rel = lambda q: (q - qo).view(np.int32).astype(np.float32) * unit
    fd = rel(qj) - rel(qi)
    r2c = fd[:, 0] * fd[:, 0] + fd[:, 1] * fd[:, 1] + fd[:, 2] * fd[:, 2]
    assert not np.any((r2c < f["cthr_in"]) & ~exact)
    assert not np.any((r2c >= f["cthr_out"]) & exact)


@pytest.mark.parametrize("prec", [D.CV2_PREC_FP64, D.CV2_PREC_FP32_FIXED])
def test_band_pairs_bit_exact_on_device(prec):
    rng = np.random.default_rng(21)
    L = np.full(3, 20.0)
    base = rng.uniform(0, 1, (2500, 3)) * L
    deltas = np.array([1e-15, 1e-13, 1e-11, 1e-9, 1e-8, 1e-7, 1e-6, 3e-6, 1e-5, 3e-5])
    k = 1500
    a = rng.integers(0, len(base), k)
    d = rng.normal(size=(k, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    delta = rng.choice(deltas, k) * rng.choice([-1.0, 1.0], k)
    partners = np.mod(base[a] + d * (CE * (1 + delta))[:, None], L)
    pos = np.concatenate([base, partners])
    s = sysdef(L, pos)
    # pairs whose exact squared distance lies between the thresholds take the
    # FP64 predicate on the device
    with D.Context.from_sysdef(s, precision=prec) as c:
        f = c.build_filter()
        c.build(True)
        got = sort_rows(c.pairs())
        assert c.stats()["build_kind"] == "cell_wave"
    dx = pos[len(base):] - base[a]
    dx -= L * np.round(dx / L)
    r2 = (dx[:, 0] * dx[:, 0] + dx[:, 1] * dx[:, 1]) + dx[:, 2] * dx[:, 2]
    in_band = (r2 >= f["cthr_in"]) & (r2 < f["cthr_out"])
    print(f"[band] prec {prec}: {int(in_band.sum())} of {k} placed pairs in the band, "
          f"{int((r2 < CE * CE).sum())} inside the cutoff", flush=True)
    assert in_band.sum() > 500
    assert np.array_equal(got, oracle_pairs(pos, s))
