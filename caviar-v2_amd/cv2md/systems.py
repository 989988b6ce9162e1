"""Synthetic LJ system builders for the MD-step workloads.

Reduced LJ units (epsilon = sigma = mass = k_B = 1).  The builders return a
plain dict ("sysdef") that both the device runtime and the test oracle
accept:  lo, hi, bc, pos[N,3], vel[N,3], type[N], eps/sigma/mass tables,
lj_cutoff, nl_cutoff, nl_coef, ghost_cutoff, dt, integrator parameters.
"""
from __future__ import annotations

import numpy as np

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])


def fcc_lattice(ncell: int, rho: float):
    """4*ncell^3 atoms on an FCC lattice filling a cubic box [0, L)^3."""
    a = (4.0 / rho) ** (1.0 / 3.0)
    g = np.stack(np.meshgrid(np.arange(ncell), np.arange(ncell), np.arange(ncell), indexing="ij"), -1)
    g = g.reshape(-1, 3).astype(np.float64)
    pos = ((g[:, None, :] + FCC_BASIS[None, :, :]) * a).reshape(-1, 3)
    return pos, a * ncell


def maxwell_velocities(n: int, temperature: float, seed: int, mass: np.ndarray | None = None):
    """Gaussian velocities, zero total momentum, rescaled so that
    T = sum m v^2 / (3N - 3) (kinetic_energy.cpp:158-198 with n_r_df = 3)."""
    rng = np.random.default_rng(seed)
    m = np.ones(n) if mass is None else np.asarray(mass, np.float64)
    v = rng.normal(0.0, 1.0, (n, 3)) / np.sqrt(m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    t_now = (m[:, None] * v * v).sum() / (3 * n - 3)
    return v * np.sqrt(temperature / t_now)


def lj_fluid(ncell: int, rho: float = 0.8442, temperature: float = 1.44, seed: int = 1001,
             rc: float = 2.5, skin: float = 0.3, dt: float = 0.005, integrator: int = 0,
             gamma: float = 0.0):
    """Single-species LJ fluid on an FCC start (the reference's
    tests/t002_lj.cpp set-up, by default at reduced density 0.8442)."""
    pos, L = fcc_lattice(ncell, rho)
    n = len(pos)
    vel = maxwell_velocities(n, temperature, seed)
    ce = rc + skin
    return dict(lo=np.zeros(3), hi=np.full(3, L), bc=(1, 1, 1), pos=pos, vel=vel,
                type=np.zeros(n, np.int32), eps=np.ones((1, 1)), sigma=np.ones((1, 1)),
                cutoff_list=None, mass=np.ones(1), lj_cutoff=rc, nl_cutoff=rc,
                nl_coef=ce / rc, ghost_cutoff=ce, dt=dt, integrator=integrator, gamma=gamma,
                kbt=temperature, k_b=1.0, n_r_df=0 if integrator == 1 else 3,
                temperature=temperature, seed=seed)


KA_EPS = np.array([[1.0, 1.5], [1.5, 0.5]])
KA_SIGMA = np.array([[1.0, 0.8], [0.8, 0.88]])


def ka_mixture(n: int, rho: float = 1.2, frac_a: float = 0.8, seed: int = 1005,
               temperature: float = 0.5, skin: float = 0.3, dt: float = 0.003, gamma: float = 0.5,
               jitter: float = 0.05):
    """Kob-Andersen 80:20 binary LJ mixture, rc_ij = 2.5 sigma_ij.
    Positions are a jittered simple-cubic fill of the box (so that no two
    atoms overlap), types assigned by a seeded permutation."""
    rng = np.random.default_rng(seed)
    L = (n / rho) ** (1.0 / 3.0)
    m = int(np.ceil(n ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    g = g[rng.permutation(len(g))[:n]]
    a = L / m
    pos = (g + 0.5 + rng.uniform(-jitter, jitter, (n, 3))) * a
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: n - int(round(frac_a * n))]] = 1
    vel = maxwell_velocities(n, temperature, seed + 1)
    cl = 2.5 * KA_SIGMA
    rc_max = float(cl.max())
    ce = rc_max + skin
    return dict(lo=np.zeros(3), hi=np.full(3, L), bc=(1, 1, 1), pos=pos, vel=vel, type=types,
                eps=KA_EPS.copy(), sigma=KA_SIGMA.copy(), cutoff_list=cl, mass=np.ones(2),
                lj_cutoff=rc_max, nl_cutoff=rc_max, nl_coef=ce / rc_max, ghost_cutoff=ce, dt=dt,
                integrator=1, gamma=gamma, kbt=temperature, k_b=1.0, n_r_df=0,
                temperature=temperature, seed=seed)


def ensemble_systems(nsys: int = 512, seed: int = 1004, nmin: int = 500, nmax: int = 20000,
                     rho=(0.60, 0.90), temperature=(0.8, 2.0), jitter: float = 0.05):
    """Batched-ensemble workload: N_s log-uniform in [nmin, nmax], density and
    temperature uniform in the given ranges; half the systems start on a
    random FCC lattice (the smallest 4k^3 >= N_s lattice with random
    vacancies), half on the same lattice jittered by `jitter` sigma.
    Returns a list of dicts (lo, hi, pos, vel, rho, temperature)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nsys):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        rho_s = float(rng.uniform(*rho))
        t_s = float(rng.uniform(*temperature))
        k = int(np.ceil((n / 4.0) ** (1.0 / 3.0) - 1e-9))
        L = (n / rho_s) ** (1.0 / 3.0)
        a = L / k
        g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1)
        sites = ((g.reshape(-1, 3)[:, None, :] + FCC_BASIS[None, :, :]) * a).reshape(-1, 3)
        pos = sites[np.sort(rng.choice(len(sites), n, replace=False))]
        if rng.random() < 0.5:
            pos = pos + rng.normal(0.0, jitter, pos.shape)
        pos = np.mod(pos, L)
        vel = maxwell_velocities(n, t_s, int(rng.integers(1 << 31)))
        out.append(dict(lo=np.zeros(3), hi=np.full(3, L), pos=pos, vel=vel, rho=rho_s, temperature=t_s))
    return out
