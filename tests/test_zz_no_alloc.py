"""cv2_halo_pack_noalloc (the halo pack that allocates no device memory: its
scratch is the library's static per-device array) against cv2_halo_pack, bit
for bit, on two calls in a row on one GPU -- the second with a smaller
system, so stale scratch from the first would show -- with and without
velocities, plus the limit checks."""
import numpy as np
import pytest

from cv2md import device as D
from cv2md.systems import lj_fluid

pytestmark = pytest.mark.gpu

GC = 2.8


def context(ncell, seed, grid, gi):
    """A domain's share of an LJ fluid: the owned atoms of cell gi of a
    process grid over the box (periodic grid edges)."""
    s = lj_fluid(ncell, seed=seed)
    rng = np.random.default_rng(seed)
    L = s["hi"] - s["lo"]
    pos = np.mod(s["pos"] + rng.normal(0.0, 0.05, s["pos"].shape), L)
    vel = rng.normal(0.0, 1.0, pos.shape)
    w = L / np.asarray(grid)
    lo = s["lo"] + w * np.asarray(gi)
    hi = lo + w
    own = np.nonzero(np.all((pos >= lo) & (pos < hi), axis=1))[0]
    c = D.Context()
    c.set_box(lo, hi, (0, 0, 0), GC)
    c.set_lj(s["eps"], s["sigma"], s["lj_cutoff"], s["mass"])
    c.set_neighbor(s["nl_cutoff"], s["nl_coef"], 0)
    c.upload(pos[own], vel[own], np.zeros(len(own), np.int32), own.astype(np.int64))
    geom = (gi, grid, lo, hi, L, (1, 1, 1), GC)
    return c, geom, len(own)


@pytest.mark.parametrize("with_velocity", [False, True])
def test_noalloc_pack_matches_pack_twice(with_velocity):
    sizes = []
    for ncell, seed, grid, gi in ((12, 51, (2, 2, 1), (1, 0, 0)), (6, 52, (2, 1, 1), (0, 0, 0))):
        c, geom, n = context(ncell, seed, grid, gi)
        with c:
            ref_counts, ref_out = c.halo_pack(*geom, with_velocity=with_velocity)
            got_counts, got_out = c.halo_pack(*geom, with_velocity=with_velocity, noalloc=True)
        print(f"[noalloc] n={n} with_velocity={with_velocity}: {int(ref_counts.sum())} sends, "
              f"{len(ref_out)} doubles", flush=True)
        assert ref_counts.sum() > 0
        assert np.array_equal(got_counts, ref_counts)
        assert len(got_out) == len(ref_out)
        assert np.array_equal(got_out.view(np.int64), ref_out.view(np.int64))
        sizes.append(n)
    assert sizes[1] < sizes[0]


def test_noalloc_pack_sizing_call_and_limits():
    c, geom, n = context(6, 53, (2, 2, 2), (1, 1, 1))
    with c:
        ref_counts, ref_out = c.halo_pack(*geom)
        # a sizing call (cap 0) reports the counts and the doubles needed only
        g = D.HaloGeom()
        gi, grid, lo, hi, L, bc, gc = geom
        for a in range(3):
            g.grid_index[a], g.nprocs[a], g.bc[a] = gi[a], grid[a], bc[a]
            g.lo_local[a], g.hi_local[a], g.size_global[a] = lo[a], hi[a], L[a]
        g.ghost_cutoff, g.with_velocity = gc, 0
        counts = np.zeros(27, np.int64)
        need = D.C.c_int64()
        rc = c.L.cv2_halo_pack_noalloc(c.h, D.C.byref(g), D._p(counts, D._i64), 0, None, D.C.byref(need))
        assert rc == 0 and need.value == len(ref_out) and np.array_equal(counts, ref_counts)
        # a bad process grid is rejected before any device work
        g.grid_index[0] = 5
        rc = c.L.cv2_halo_pack_noalloc(c.h, D.C.byref(g), D._p(counts, D._i64), 0, None, D.C.byref(need))
        assert rc == 2 and b"bad process grid" in c.L.cv2_last_error(c.h)
