"""MPI halo packets on the device (cv2_halo_pack / cv2_halo_unpack) against the
oracle's restatement of exchange_ghost_mpi_shared_atoms.cpp, and a whole
exchange between the domains of a 2 x 2 x 1 process grid run in one process.

- Each domain's 26 direction packets (counts, ids, types, shifted positions,
  velocities, atom order) equal the oracle's bit for bit.
- The packets are delivered as MPI would (a packet sent in direction d
  arrives in the receiver's slot 26 - d), each domain appends the received
  packets in direction order and then its own periodic images, as the
  reference does, and unpacks them on the device.  Every domain's pair set
  then equals the oracle's list over the same owned and ghost arrays, and its
  FP64 accelerations equal the single-box periodic answer for its atoms.
"""
import numpy as np
import pytest

import oracle as O
from cv2md import device as D
from cv2md.systems import lj_fluid
from helpers import oracle_forces, sort_rows

pytestmark = pytest.mark.gpu

GRID = (2, 2, 1)
GC = 2.8


def ranks():
    return [(i, j, k) for i in range(GRID[0]) for j in range(GRID[1]) for k in range(GRID[2])]


def neighbour(gi, d):
    v = (d // 9 - 1, (d // 3) % 3 - 1, d % 3 - 1)
    return tuple((gi[a] + v[a]) % GRID[a] for a in range(3))


def split(counts, out, total):
    res, at = {}, 0
    for d in range(27):
        n = int(counts[d])
        res[d] = (n, out[at:at + total * n])
        at += total * n
    return res


@pytest.fixture(scope="module")
def domains():
    # domains at least twice the ghost cutoff wide (the halo rule gives an atom
    # one direction per axis, so a narrower domain misses images)
    s = lj_fluid(8, seed=31)
    box = O.make_box(s["lo"], s["hi"])
    rng = np.random.default_rng(5)
    pos, _ = O.wrap(s["pos"] + rng.normal(0.0, 0.05, s["pos"].shape), box)
    vel = rng.normal(0.0, 1.0, pos.shape)
    L = s["hi"] - s["lo"]
    w = L / np.asarray(GRID)
    out = {}
    for gi in ranks():
        lo = s["lo"] + w * np.asarray(gi)
        hi = lo + w
        own = np.nonzero(np.all((pos >= lo) & (pos < hi), axis=1))[0]
        out[gi] = dict(lo=lo, hi=hi, own=own)
    return s, pos, vel, L, out


def make_ctx(s, dom, pos, vel):
    c = D.Context()
    c.set_box(dom["lo"], dom["hi"], (0, 0, 0), GC)  # a domain's share: neighbours come from the halo
    c.set_lj(s["eps"], s["sigma"], s["lj_cutoff"], s["mass"], s.get("cutoff_list"))
    c.set_neighbor(s["nl_cutoff"], s["nl_coef"], 0)
    c.set_precision(D.CV2_PREC_FP64)
    own = dom["own"]
    c.upload(pos[own], vel[own], np.zeros(len(own), np.int32), own.astype(np.int64))
    return c


def test_halo_pack_matches_oracle(domains):
    s, pos, vel, L, doms = domains
    for gi, dom in doms.items():
        own = dom["own"]
        for with_vel in (False, True):
            exp_c, exp = O.halo_pack(pos[own], vel[own], own, np.zeros(len(own), np.int32), gi, GRID,
                                     dom["lo"], dom["hi"], L, (1, 1, 1), GC, with_vel)
            with make_ctx(s, dom, pos, vel) as c:
                c.set_ghosts(np.zeros((0, 3)))  # one domain's share, no halo yet
                c.build(True)  # the device order is now the cell order
                got_c, got = c.halo_pack(gi, GRID, dom["lo"], dom["hi"], L, (1, 1, 1), GC, with_vel)
            print(f"[halo] rank {gi}: {len(own)} owned, {int(got_c.sum())} sent", flush=True)
            assert np.array_equal(got_c, exp_c), gi
            assert np.array_equal(got, exp), (gi, with_vel)


def test_halo_exchange_forces_and_pairs(domains):
    s, pos, vel, L, doms = domains
    packets = {}
    for gi, dom in doms.items():
        with make_ctx(s, dom, pos, vel) as c:
            cnt, pk = c.halo_pack(gi, GRID, dom["lo"], dom["hi"], L, (1, 1, 1), GC)
        packets[gi] = split(cnt, pk, 5)
    _, a_ref, _vir, _pe = oracle_forces(pos, s)
    for gi, dom in doms.items():
        # received packets in direction order, then the domain's own images
        recv = []
        for d in range(27):
            src = neighbour(gi, d)
            if d != 13 and src != gi:
                recv.append(packets[src][26 - d])
        recv += [packets[gi][d] for d in range(27) if d != 13 and neighbour(gi, d) == gi]
        counts = [n for n, _ in recv]
        flat = np.concatenate([p for _, p in recv]) if recv else np.zeros(0)
        gpos = np.concatenate([p[2 * n:5 * n].reshape(n, 3) for n, p in recv])
        own = dom["own"]
        with make_ctx(s, dom, pos, vel) as c:
            c.halo_unpack(counts, flat)
            c.compute_forces()
            acc = c.download()[2]
            pairs = sort_rows(c.pairs())
        offs, nbr = O.nlist_brute(pos[own], gpos, s["nl_cutoff"], s["nl_coef"])
        n = len(own)
        exp = sort_rows(np.stack([np.repeat(np.arange(n), np.diff(offs)), nbr.astype(np.int64),
                                  np.full(len(nbr), 13)], 1))
        print(f"[halo] rank {gi}: {sum(counts)} ghosts, {len(pairs)} pairs", flush=True)
        assert np.array_equal(pairs, exp), gi
        ref = a_ref[own]
        assert np.abs(acc - ref).max() < 1e-12 * max(np.abs(ref).max(), 1.0), gi


def test_unpack_at_offsets_equals_concatenated(domains):
    """cv2_halo_unpack_at with packets left where they lie in one page-locked
    buffer (received in reverse direction order) gives the ghosts, pairs and
    forces of cv2_halo_unpack over the same packets concatenated."""
    s, pos, vel, L, doms = domains
    gi, dom = next(iter(doms.items()))
    with make_ctx(s, dom, pos, vel) as c:
        cnt, pk = c.halo_pack(gi, GRID, dom["lo"], dom["hi"], L, (1, 1, 1), GC)
    starts = np.concatenate([[0], np.cumsum(5 * cnt)[:-1]])
    order = np.arange(26, -1, -1)
    flat = np.concatenate([pk[starts[d]:starts[d] + 5 * cnt[d]] for d in order])
    buf = D.pinned_empty(len(pk) + 7)
    buf[:len(pk)] = pk
    out = []
    for use_at in (False, True):
        with make_ctx(s, dom, pos, vel) as c:
            if use_at:
                c.halo_unpack_at(cnt[order], starts[order], buf)
            else:
                c.halo_unpack(cnt[order], flat)
            c.compute_forces()
            out.append((sort_rows(c.pairs()), c.download()[2], c.stats()["nghost"]))
    assert out[0][2] == out[1][2] == int(cnt.sum())
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
