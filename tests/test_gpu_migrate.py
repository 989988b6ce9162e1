"""Owned-atom exchange on the device (cv2_owned_pack / cv2_owned_unpack)
against the oracle's restatement of exchange_owned_mpi_shared_atoms.cpp, and
a whole decomposed run: the four domains of a 2 x 2 x 1 process grid, run in
one process, exchange owned atoms and ghost halos on every step and follow
the single periodic box.

- Each domain's packets (counts, ids, shifted positions, velocities,
  accelerations, molecule and bond fields, atom order) equal the oracle's
  bit for bit, whatever the device's cell order; the staying atoms keep
  their (shifted) positions in owned order.
- After the exchange each domain owns exactly the atoms the oracle's
  exchange gives it, in ascending id order, with their state bit for bit.
- 40 velocity-Verlet steps of the four domains, each step opened, the owned
  atoms exchanged, the halo exchanged, the step closed, equal 40 steps of
  the whole periodic box (FP64, positions and velocities within 1e-10).
"""
import numpy as np
import pytest

import oracle as O
from cv2md import device as D
from cv2md.systems import lj_fluid

pytestmark = pytest.mark.gpu

GRID = (2, 2, 1)
GC = 2.8


def ranks():
    return [(i, j, k) for i in range(GRID[0]) for j in range(GRID[1]) for k in range(GRID[2])]


def neighbour(gi, d):
    v = (d // 9 - 1, (d // 3) % 3 - 1, d % 3 - 1)
    return tuple((gi[a] + v[a]) % GRID[a] for a in range(3))


def split(counts, out, per_atom):
    res, at = {}, 0
    for d in range(27):
        n = int(counts[d])
        res[d] = (n, out[at:at + per_atom * n])
        at += per_atom * n
    return res


def received(gi, packets):
    """Packets arriving at domain gi in its direction order (a packet sent
    towards d arrives in slot 26 - d), then for the ghost halo the domain's
    own images (self directions)."""
    recv = []
    for d in range(27):
        src = neighbour(gi, d)
        if d != 13 and src != gi:
            recv.append(packets[src][26 - d])
    return recv


@pytest.fixture(scope="module")
def setup():
    # domains at least twice the ghost cutoff wide: the halo rule gives an atom
    # one direction per axis (below lo + g, else above hi - g), so a narrower
    # domain would not send the images of an atom near both of its faces
    s = lj_fluid(8, seed=41)
    L = s["hi"] - s["lo"]
    rng = np.random.default_rng(9)
    box = O.make_box(s["lo"], s["hi"])
    x0, _ = O.wrap(s["pos"] + rng.normal(0.0, 0.05, s["pos"].shape), box)
    w = L / np.asarray(GRID)
    doms = {}
    for gi in ranks():
        lo = s["lo"] + w * np.asarray(gi)
        doms[gi] = dict(lo=lo, hi=lo + w, own=np.nonzero(np.all((x0 >= lo) & (x0 < lo + w), axis=1))[0])
    return s, L, x0, doms


def make_ctx(s, dom, pos, vel, acc=None):
    c = D.Context()
    c.set_box(dom["lo"], dom["hi"], (0, 0, 0), GC)
    c.set_lj(s["eps"], s["sigma"], s["lj_cutoff"], s["mass"], s.get("cutoff_list"))
    c.set_neighbor(s["nl_cutoff"], s["nl_coef"], 0)
    c.set_precision(D.CV2_PREC_FP64)
    own = dom["own"]
    c.upload(pos[own], vel[own], np.zeros(len(own), np.int32), own.astype(np.int64))
    c.set_type_table(np.zeros(len(pos), np.int32))
    return c


def test_owned_pack_matches_oracle(setup):
    """Displaced atoms leave in the reference's packets; the device's cell
    order (a build before the exchange) does not change them."""
    s, L, x0, doms = setup
    rng = np.random.default_rng(11)
    x1 = x0 + rng.uniform(-0.5, 0.5, x0.shape)
    vel = rng.normal(size=x0.shape)
    for gi, dom in doms.items():
        own = dom["own"]
        with make_ctx(s, dom, x1, vel) as c:
            c.set_ghosts(np.zeros((0, 3)))
            c.compute_forces()  # device order: cell order; accelerations to carry
            acc = c.download()[2]
            exp_c, exp, stay, after = O.owned_pack(x1[own], vel[own], acc, own, gi, GRID, dom["lo"], dom["hi"],
                                                   L, (1, 1, 1))
            got_c0, _ = c.owned_pack(gi, GRID, dom["lo"], dom["hi"], L, (1, 1, 1), cap=0)  # sizing only
            assert c.n == len(own)
            got_c, got = c.owned_pack(gi, GRID, dom["lo"], dom["hi"], L, (1, 1, 1))
            print(f"[migrate] rank {gi}: {len(own)} owned, {int(got_c.sum())} leave", flush=True)
            assert np.array_equal(got_c0, exp_c) and np.array_equal(got_c, exp_c), gi
            assert np.array_equal(got, exp), gi
            assert got_c.sum() > 0
            ids = c.ids()
            p, _, _ = c.download()
        assert np.array_equal(ids, own[stay])
        assert np.array_equal(p, after[stay])


def test_exchange_places_every_atom(setup):
    s, L, x0, doms = setup
    rng = np.random.default_rng(12)
    x1 = x0 + rng.uniform(-0.5, 0.5, x0.shape)
    vel, acc = rng.normal(size=x0.shape), rng.normal(size=x0.shape)
    ctxs, packets, expect = {}, {}, {}
    try:
        for gi, dom in doms.items():
            ctxs[gi] = c = make_ctx(s, dom, x1, vel)
            c.build(True)
            cnt, pk = c.owned_pack(gi, GRID, dom["lo"], dom["hi"], L, (1, 1, 1))
            packets[gi] = split(cnt, pk, 12)
        x1w = np.where(x1 < s["lo"], x1 + L, np.where(x1 > s["hi"], x1 - L, x1))
        for gi, dom in doms.items():
            recv = received(gi, packets)
            counts = [n for n, _ in recv]
            flat = np.concatenate([p for _, p in recv]) if recv else np.zeros(0)
            c = ctxs[gi]
            c.owned_unpack(counts, flat)
            ids = c.ids()
            p, v, _ = c.download()
            lo, hi = dom["lo"], dom["hi"]
            inside = np.nonzero(np.all((x1w >= lo) & (x1w <= hi), axis=1))[0]
            print(f"[migrate] rank {gi}: owns {len(ids)} after the exchange", flush=True)
            assert np.array_equal(ids, inside), gi  # ascending id order, every atom of the domain
            assert np.array_equal(p, x1w[ids]) and np.array_equal(v, vel[ids]), gi
            expect[gi] = ids
        allids = np.concatenate(list(expect.values()))
        assert np.array_equal(np.sort(allids), np.arange(len(x0)))
    finally:
        for c in ctxs.values():
            c.close()


def halo_exchange(ctxs, doms, L):
    packets = {}
    for gi, c in ctxs.items():
        cnt, pk = c.halo_pack(gi, GRID, doms[gi]["lo"], doms[gi]["hi"], L, (1, 1, 1), GC)
        packets[gi] = split(cnt, pk, 5)
    for gi, c in ctxs.items():
        recv = received(gi, packets)
        recv += [packets[gi][d] for d in range(27) if d != 13 and neighbour(gi, d) == gi]
        c.halo_unpack([n for n, _ in recv], np.concatenate([p for _, p in recv]))


def owned_exchange(ctxs, doms, L):
    packets = {}
    for gi, c in ctxs.items():
        cnt, pk = c.owned_pack(gi, GRID, doms[gi]["lo"], doms[gi]["hi"], L, (1, 1, 1))
        packets[gi] = split(cnt, pk, 12)
    moved = 0
    for gi, c in ctxs.items():
        recv = received(gi, packets)
        moved += sum(n for n, _ in recv)
        if recv:
            c.owned_unpack([n for n, _ in recv], np.concatenate([p for _, p in recv]))
    return moved


def test_decomposed_run_follows_single_box(setup):
    s, L, x0, doms = setup
    vel = np.random.default_rng(13).normal(0.0, 1.2, x0.shape)
    vel -= vel.mean(0)
    dt, nsteps = 0.005, 40
    single = dict(s, pos=x0, vel=vel)
    with D.Context.from_sysdef(single) as c:
        c.compute_forces()
        c.run_vv(nsteps, dt)
        p_ref, v_ref, _ = c.download()
    ctxs = {gi: make_ctx(s, dom, x0, vel) for gi, dom in doms.items()}
    try:
        halo_exchange(ctxs, doms, L)
        for c in ctxs.values():
            c.compute_forces()
        moved = 0
        for k in range(nsteps):
            for c in ctxs.values():
                c.step_begin(dt)
            moved += owned_exchange(ctxs, doms, L)
            halo_exchange(ctxs, doms, L)
            for c in ctxs.values():
                c.step_end(energy=False)
        print(f"[migrate] {nsteps} steps, {moved} atoms moved between domains", flush=True)
        assert moved > 0
        ids, pos, vv = [], [], []
        for c in ctxs.values():
            ids.append(c.ids())
            p, v, _ = c.download()
            pos.append(p)
            vv.append(v)
        ids = np.concatenate(ids)
        assert np.array_equal(np.sort(ids), np.arange(len(x0)))
        pos = np.concatenate(pos)
        vv = np.concatenate(vv)
        d = pos - p_ref[ids]
        d -= L * np.round(d / L)
        assert np.abs(d).max() < 1e-10
        assert np.abs(vv - v_ref[ids]).max() < 1e-10
    finally:
        for c in ctxs.values():
            c.close()


def test_unpack_needs_the_type_table(setup):
    s, L, x0, doms = setup
    gi, dom = next(iter(doms.items()))
    c = D.Context()
    try:
        c.set_box(dom["lo"], dom["hi"], (0, 0, 0), GC)
        c.set_lj(s["eps"], s["sigma"], s["lj_cutoff"], s["mass"])
        c.set_neighbor(s["nl_cutoff"], s["nl_coef"], 0)
        own = dom["own"]
        c.upload(x0[own], None, np.zeros(len(own), np.int32), own.astype(np.int64))
        pk = np.array([5.0, 1.0, 1.0, 1.0, 0, 0, 0, 0, 0, 0, -1.0, 0.0])
        with pytest.raises(D.Cv2Error) as e:
            c.owned_unpack([1], pk)
        assert e.value.code == 6  # CV2_ERR_STATE
        c.set_type_table(np.zeros(3, np.int32))
        with pytest.raises(D.Cv2Error) as e:  # id 5 is outside the table
            c.owned_unpack([1], pk)
        assert e.value.code == 2  # CV2_ERR_BAD_ARG
    finally:
        c.close()
