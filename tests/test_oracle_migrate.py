"""Pins for the oracle's owned-atom exchange (orc_owned_pack, the restatement
of exchange_owned_mpi_shared_atoms.cpp).  The reference's MPI path does not
build here (nothing defines CAVIAR_WITH_MPI), so the restatement is pinned
by a hand-worked known answer and by what an exchange must conserve:

- known answer: five atoms of one domain of a 2 x 1 x 1 grid, each case of
  the direction rule (leaving across a periodic grid edge, leaving to an
  inner neighbour, leaving along a one-process axis, inside, outside a
  non-periodic face) with its packet entries written out by hand;
- an exchange over the 4 domains of a 2 x 2 x 1 grid: every atom ends owned
  by exactly one domain, the one whose box holds its position, the position
  is the displaced one wrapped once into the global box (x + L below, x - L
  above), and id, velocity and acceleration travel unchanged.
"""
import numpy as np

import oracle as O

L = np.array([10.0, 10.0, 10.0])


def test_known_answer_2x1x1():
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([5.0, 10.0, 10.0])
    pos = np.array([[5.5, 1.0, 1.0],     # +x to rank 1 (inner edge: no shift), d = 22
                    [-0.5, 2.0, 2.0],    # -x across the periodic edge: x + 10, d = 4
                    [1.0, 10.5, 3.0],    # +y, one process in y: stays, y - 10
                    [2.0, 3.0, 4.0],     # inside
                    [5.5, -0.5, 3.0],    # +x -y: d = 19, y + 10 (periodic edge in y)
                    [2.0, 3.0, -0.1]])   # below an open z face: zeroed, stays unshifted
    n = len(pos)
    vel = np.arange(3 * n, dtype=float).reshape(n, 3) + 0.25
    acc = -np.arange(3 * n, dtype=float).reshape(n, 3) - 0.5
    ids = np.array([10, 11, 12, 13, 14, 15])
    counts, out, stay, after = O.owned_pack(pos, vel, acc, ids, (0, 0, 0), (2, 1, 1), lo, hi, L, (1, 1, 0))
    exp_counts = np.zeros(27, np.int64)
    exp_counts[[4, 19, 22]] = 1
    assert np.array_equal(counts, exp_counts)
    assert np.array_equal(stay, [False, False, True, True, False, True])
    pk4 = [11, 9.5, 2.0, 2.0, *vel[1], *acc[1], -1, 0]
    pk19 = [14, 5.5, 9.5, 3.0, *vel[4], *acc[4], -1, 0]
    pk22 = [10, 5.5, 1.0, 1.0, *vel[0], *acc[0], -1, 0]
    assert np.array_equal(out, np.array(pk4 + pk19 + pk22))
    assert np.array_equal(after[2], [1.0, 0.5, 3.0])
    assert np.array_equal(after[3], pos[3]) and np.array_equal(after[5], pos[5])


GRID = (2, 2, 1)


def neighbour(gi, d):
    v = (d // 9 - 1, (d // 3) % 3 - 1, d % 3 - 1)
    return tuple((gi[a] + v[a]) % GRID[a] for a in range(3))


def domain(gi):
    w = L / np.asarray(GRID)
    lo = w * np.asarray(gi)
    return lo, lo + w


def exchange(states):
    """One owned exchange over the grid with the oracle: states[gi] = (ids,
    pos, vel, acc) in owned order; returns the new states, arrivals appended
    in receive order (direction order of the receiver)."""
    packets, kept = {}, {}
    for gi, (ids, pos, vel, acc) in states.items():
        lo, hi = domain(gi)
        counts, out, stay, after = O.owned_pack(pos, vel, acc, ids, gi, GRID, lo, hi, L, (1, 1, 1))
        base = np.concatenate([[0], np.cumsum(12 * counts)[:-1]])
        packets[gi] = {d: (int(counts[d]), out[base[d]:base[d] + 12 * counts[d]]) for d in range(27)}
        kept[gi] = (ids[stay], after[stay], vel[stay], acc[stay])
    new = {}
    for gi in states:
        ids, pos, vel, acc = (list(x) for x in kept[gi])
        for d in range(27):
            src = neighbour(gi, d)
            if d == 13 or src == gi:
                continue
            n, p = packets[src][26 - d]
            for c in range(n):
                ids.append(int(p[c]))
                pos.append(p[n + 3 * c:n + 3 * c + 3])
                vel.append(p[4 * n + 3 * c:4 * n + 3 * c + 3])
                acc.append(p[7 * n + 3 * c:7 * n + 3 * c + 3])
        new[gi] = (np.asarray(ids, np.int64), np.asarray(pos).reshape(-1, 3), np.asarray(vel).reshape(-1, 3),
                   np.asarray(acc).reshape(-1, 3))
    return new


def test_exchange_conserves_and_places_every_atom():
    rng = np.random.default_rng(3)
    n = 4000
    x0 = rng.uniform(0.0, 10.0, (n, 3))
    disp = rng.uniform(-0.6, 0.6, (n, 3))
    vel, acc = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    states = {}
    for i in range(GRID[0]):
        for j in range(GRID[1]):
            gi = (i, j, 0)
            lo, hi = domain(gi)
            own = np.nonzero(np.all((x0 >= lo) & (x0 < hi), axis=1))[0]
            states[gi] = (own.astype(np.int64), x0[own] + disp[own], vel[own], acc[own])
    new = exchange(states)
    x1 = x0 + disp
    wrapped = np.where(x1 < 0.0, x1 + 10.0, np.where(x1 > 10.0, x1 - 10.0, x1))
    seen = np.zeros(n, int)
    for gi, (ids, pos, v, a) in new.items():
        lo, hi = domain(gi)
        seen[ids] += 1
        assert np.all((pos >= lo) & (pos <= hi)), gi
        assert np.array_equal(pos, wrapped[ids])
        assert np.array_equal(v, vel[ids]) and np.array_equal(a, acc[ids])
    assert np.all(seen == 1)
    moved = sum(int(np.sum(~np.all((x1[s[0]] >= domain(g)[0]) & (x1[s[0]] <= domain(g)[1]), axis=1)))
                for g, s in states.items())
    assert moved > 500
