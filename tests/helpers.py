"""Shared helpers for the parity tests: the oracle's view of a configuration
(pair keys, accelerations, energy) on the same inputs the device sees."""
import numpy as np

import oracle as O


def oracle_pairs(pos, s, cell=True):
    """Reference pair set as sorted int64 rows (i, j, image_code); ghost
    entries are keyed by their owner and image (the device stores no ghosts)."""
    box = O.make_box(s["lo"], s["hi"], s.get("bc", (1, 1, 1)))
    p, _ = O.wrap(pos, box)
    gpos, owner, code = O.ghosts(p, box, s["ghost_cutoff"])
    if cell:
        offs, nbr = O.nlist_cell(p, gpos, box, s["nl_cutoff"], s["nl_coef"])
    else:
        offs, nbr = O.nlist_brute(p, gpos, s["nl_cutoff"], s["nl_coef"])
    n = len(p)
    i = np.repeat(np.arange(n), np.diff(offs))
    j = nbr.astype(np.int64)
    g = j >= n
    gi = np.clip(j - n, 0, max(len(owner) - 1, 0))  # ghost index (owned entries: unused)
    jj = np.where(g, owner[gi] if len(owner) else j, j)
    cc = np.where(g, code[gi] if len(code) else 13, 13)
    return sort_rows(np.stack([i, jj, cc], 1))


def pair_keys(rows, n):
    """Sorted int64 keys (i * n + j) * 27 + image_code of (i, j, code) rows:
    one sort of a flat array instead of a row sort, for lists of 10^7-10^8
    pairs."""
    a = np.asarray(rows, np.int64)
    return np.sort((a[:, 0] * n + a[:, 1]) * 27 + a[:, 2])


def oracle_pair_keys(pos, s):
    """oracle_pairs as sorted pair_keys (cell-list path)."""
    box = O.make_box(s["lo"], s["hi"], s.get("bc", (1, 1, 1)))
    p, _ = O.wrap(pos, box)
    gpos, owner, code = O.ghosts(p, box, s["ghost_cutoff"])
    offs, nbr = O.nlist_cell(p, gpos, box, s["nl_cutoff"], s["nl_coef"])
    del gpos
    n = len(p)
    j = nbr.astype(np.int64)
    del nbr
    g = j >= n
    gi = np.where(g, j - n, 0)
    cc = np.where(g, code[gi], 13).astype(np.int64)
    j = np.where(g, owner[gi], j)
    del g, gi
    k = np.repeat(np.arange(n, dtype=np.int64), np.diff(offs)) * n
    k += j
    k *= 27
    k += cc
    return np.sort(k)


def sort_rows(a):
    a = np.asarray(a, np.int64).reshape(-1, 3)
    return a[np.lexsort((a[:, 2], a[:, 1], a[:, 0]))]


def oracle_forces(pos, s):
    """Reference accelerations, virial and (added) potential energy."""
    box = O.make_box(s["lo"], s["hi"], s.get("bc", (1, 1, 1)))
    p, _ = O.wrap(pos, box)
    gpos, owner, _ = O.ghosts(p, box, s["ghost_cutoff"])
    offs, nbr = O.nlist_cell(p, gpos, box, s["nl_cutoff"], s["nl_coef"])
    t = np.asarray(s["type"], np.int32)
    mol = s.get("mol")  # ignore_intra_molecule on: ghosts carry their owners' indices
    acc, vir, pe = O.lj(p, t, np.arange(len(p)), gpos, t[owner], owner, offs, nbr, s["eps"],
                        s["sigma"], s["lj_cutoff"], s["mass"], s.get("cutoff_list"),
                        s.get("jump_tol", 0.0), mol=mol,
                        gmol=None if mol is None else np.asarray(mol)[owner])
    return p, acc, vir, pe


def force_error(a, ref):
    """max_i |a_i - ref_i| / max(|ref_i|, rms|ref|), and the rms of |a - ref| / rms|ref|."""
    d = np.linalg.norm(a - ref, axis=1)
    m = np.linalg.norm(ref, axis=1)
    rms = np.sqrt((m ** 2).mean())
    return float((d / np.maximum(m, rms)).max()), float(np.sqrt((d ** 2).mean()) / rms)


_U = np.uint64


def splitmix64(z):
    """The splitmix64 output function of state + 0x9e3779b97f4a7c15
    (Steele, Lea and Flood, OOPSLA 2014), on uint64 arrays, wrapping."""
    z = np.asarray(z, np.uint64) + _U(0x9E3779B97F4A7C15)
    z = (z ^ (z >> _U(30))) * _U(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> _U(27))) * _U(0x94D049BB133111EB)
    return z ^ (z >> _U(31))


def pair_checksum(rows, ids=None):
    """(sum mod 2^64, count) over the ordered pairs of a half-list pair set
    (i, j, image_code) rows, as CV2_NL_VERIFY reports them: owned pairs
    (code 13) count in both directions, image pairs as listed; a pair adds
    splitmix64((id_i << 38) ^ (id_j << 6) ^ image) with image = (vx + 1) |
    (vy + 1) << 2 | (vz + 1) << 4 and code = (vx + 1) + 3 (vy + 1) + 9 (vz + 1)."""
    a = np.asarray(rows, np.int64).reshape(-1, 3)
    own = a[:, 2] == 13
    i = np.concatenate([a[:, 0], a[own, 1]])
    j = np.concatenate([a[:, 1], a[own, 0]])
    c = np.concatenate([a[:, 2], a[own, 2]])
    if ids is not None:
        i, j = np.asarray(ids, np.int64)[i], np.asarray(ids, np.int64)[j]
    img = (c % 3) | ((c // 3 % 3) << 2) | ((c // 9) << 4)
    key = (i.astype(np.uint64) << _U(38)) ^ (j.astype(np.uint64) << _U(6)) ^ img.astype(np.uint64)
    return int(np.add.reduce(splitmix64(key), dtype=np.uint64)) if len(key) else 0, len(key)
