"""Soundness of the coarse distance filter of the wave-per-cell list build.

The build accepts a candidate without the FP64 predicate when its float
squared distance (from the int32 fixed-point coordinates, relative to a
reference atom of the home cell) is below cthr_in, and rejects it when at or
above cthr_out.  This test reproduces the device arithmetic in float32 with
numpy on pairs placed within 3e-5 of the list cutoff, in boxes of several
shapes, with the thresholds restated from the library's setup path, and
checks that neither shortcut ever contradicts the FP64 predicate
(dx*dx + dy*dy) + dz*dz < ce^2.  tests/test_gpu_coarse_filter.py checks
that this restatement equals the values the library sets up
(cv2_get_build_filter), repeats the check with those values, and drives
the band path on the device.
"""
import math

import numpy as np
import pytest


def qunit(L, span):
    # set_qscale on a fully periodic box: u_d = L_d / 2^(32 + k_d), k_d the
    # largest shift with span < L_d / 2^(k_d + 1) (0.1% margin)
    u = []
    for Ld in L:
        k = 0
        while k < 30 and math.ldexp(span * 1.001, k + 2) < Ld:
            k += 1
        u.append(math.ldexp(Ld, -(32 + k)))
    return np.asarray(u)


def thresholds(u, ce, wmax):
    ef = 8.0 * 5.9604644775390625e-08 * (2.0 * wmax + ce) * math.sqrt(3.0)
    e = 1.01 * math.sqrt(float((u * u).sum())) + ef + 1e-6 * ce
    tin = np.nextafter(np.float32((ce - e) * (ce - e) * (1.0 - 4e-6)), np.float32(0))
    tout = np.nextafter(np.float32((ce + e) * (ce + e) * (1.0 + 4e-6)), np.float32(1e30))
    return tin, tout, u.astype(np.float32)


def quantize(x, u):
    # pack_q: llrint((x - lo) / u) kept modulo 2^32 (lo = 0 here)
    q = np.rint(x * (1.0 / u)).astype(np.int64)
    return (q & 0xFFFFFFFF).astype(np.uint32)


@pytest.mark.parametrize("L", [(134.3677,) * 3, (20.0, 30.0, 400.0), (14.5, 14.5, 14.5)])
def test_coarse_filter_never_contradicts_fp64(L):
    rng = np.random.default_rng(7)
    ce = 2.5 * 1.12
    Ls = np.asarray(L)
    nc = np.floor(Ls / (ce * (1 + 1e-12)))
    w = Ls / nc
    u = qunit(Ls, 2.0 * max(w.max(), ce))
    tin, tout, unit = thresholds(u, ce, w.max())
    n = 400000
    ref = rng.uniform(0, 1, (n, 3)) * Ls             # reference atom of the home cell
    pi = ref + rng.uniform(-1, 1, (n, 3)) * w         # home atom near it
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = ce * (1 + rng.uniform(-3e-5, 3e-5, n))
    pj = pi + d * r[:, None]
    wrapf = lambda x: np.mod(x, Ls)
    pi, pj, ref = wrapf(pi), wrapf(pj), wrapf(ref)
    # FP64 predicate on the minimum image
    dx = pj - pi
    dx -= Ls * np.round(dx / Ls)
    r2 = (dx[:, 0] * dx[:, 0] + dx[:, 1] * dx[:, 1]) + dx[:, 2] * dx[:, 2]
    exact = r2 < ce * ce
    # device arithmetic: int32 differences from the reference atom, float32
    qo, qi, qj = quantize(ref, u), quantize(pi, u), quantize(pj, u)
    rel = lambda q: (q - qo).view(np.int32).astype(np.float32) * unit
    fd = rel(qj) - rel(qi)
    r2c = fd[:, 0] * fd[:, 0] + fd[:, 1] * fd[:, 1] + fd[:, 2] * fd[:, 2]
    assert not np.any((r2c < tin) & ~exact)
    assert not np.any((r2c >= tout) & exact)
    # the band that goes to the FP64 predicate stays narrow
    assert (tout - tin) / (ce * ce) < 1e-4
