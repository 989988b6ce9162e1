"""One MPI domain's share on one GPU, with its ghost halo exchanged through
the host every step (the binary Lennard-Jones workload of the benchmark).

The share is its own neighbour in all 26 directions: a 1x1x1 periodic
process grid, as caviar2 runs a single domain with MPI.  Each step is the
reference's order of work (Md_simulator::integrate_* with
boundary_condition, md_simulator.cpp:422-437,731-869):

  cv2_step_begin   velocity half kick (+ Langevin noise), drift
  cv2_halo_pack    the owned atoms within the ghost cutoff of a face, as the
                   26 direction packets of exchange_ghost_mpi_shared_atoms
                   (device -> host)
  (host)           the packet sent in direction d arrives from direction
                   26 - d; the share appends the received packets in the
                   order they arrive
  cv2_halo_unpack_at  the packets become the ghosts (host -> device, from
                   the same page-locked buffer; a changed ghost count forces
                   a list rebuild, neighborlist.cpp:117-120)
  cv2_step_end     list build when needed, forces, second velocity half

Owned atoms that leave the box are wrapped back by cv2_step_begin, before
the halo is packed (the share is periodic on every axis): the owned-atom
migration of a 1x1x1 grid, which precedes the ghost exchange
(Md_simulator::boundary_condition, md_simulator.cpp:547-557).  Nothing here
runs on the CPU except the forwarding of packets.
"""
from __future__ import annotations

import time

import numpy as np

from . import device as D
from .systems import KA_EPS, KA_SIGMA, maxwell_velocities


def ka_random(n: int, rho: float = 1.2, frac_a: float = 0.8, seed: int = 1005, temperature: float = 0.5):
    """Kob-Andersen 80:20 mixture at uniformly random positions (atoms may
    overlap: the start of the relaxation), types by a seeded permutation,
    rc_ij = 2.5 sigma_ij, list cutoff 2.5 + 0.3 skin, ghost cutoff 2.8."""
    rng = np.random.default_rng(seed)
    L = (n / rho) ** (1.0 / 3.0)
    pos = rng.uniform(0.0, L, (n, 3))
    types = np.zeros(n, np.int32)
    types[rng.permutation(n)[: n - int(round(frac_a * n))]] = 1
    vel = maxwell_velocities(n, temperature, seed + 1)
    cl = 2.5 * KA_SIGMA
    return dict(lo=np.zeros(3), hi=np.full(3, L), bc=(1, 1, 1), pos=pos, vel=vel, type=types,
                eps=KA_EPS.copy(), sigma=KA_SIGMA.copy(), cutoff_list=cl, mass=np.ones(2),
                lj_cutoff=float(cl.max()), nl_cutoff=float(cl.max()), nl_coef=2.8 / float(cl.max()),
                ghost_cutoff=2.8, dt=0.003, integrator=1, gamma=0.5, kbt=temperature, k_b=1.0,
                n_r_df=0, temperature=temperature, seed=seed)


class Share:
    """A context holding one domain's share, its halo refreshed each step
    through the device halo packets and the host."""

    def __init__(self, s: dict, precision=D.CV2_PREC_FP64, device: int = 0):
        self.s = s
        self.precision = precision
        self.L = np.asarray(s["hi"], float) - np.asarray(s["lo"], float)
        c = D.Context(device)
        c.set_box(s["lo"], s["hi"], (1, 1, 1), s["ghost_cutoff"])
        c.set_lj(s["eps"], s["sigma"], s["lj_cutoff"], s["mass"], s.get("cutoff_list"))
        c.set_neighbor(s["nl_cutoff"], s["nl_coef"], 0)
        c.set_precision(precision)
        c.upload(s["pos"], s.get("vel"), s.get("type"))
        self.c = c
        self.exchange_seconds = 0.0  # pack (with the device work queued before it) + unpack
        self.pack_seconds = 0.0
        self.unpack_seconds = 0.0
        self.exchanges = 0
        self.halo_atoms = 0
        # the packets: one page-locked buffer, sent and received in place
        self.buf = D.pinned_empty(0)
        self.exchange()

    def exchange(self):
        """One halo exchange: pack on the device into the host buffer, then
        hand the packets back in receive order (slot k receives the packet
        sent towards 26 - k) without moving them on the host."""
        t0 = time.perf_counter()
        s = self.s
        geom = ((0, 0, 0), (1, 1, 1), s["lo"], s["hi"], self.L, (1, 1, 1), s["ghost_cutoff"])
        counts, need = self.c.halo_pack_into(self.buf, *geom)
        if need > len(self.buf):
            self.buf = D.pinned_empty(int(need * 1.25) + 1024)
            counts, need = self.c.halo_pack_into(self.buf, *geom)
        t1 = time.perf_counter()
        per = 5  # doubles per atom: id, type, x, y, z
        starts = np.concatenate([[0], np.cumsum(per * counts)[:-1]])
        order = np.arange(26, -1, -1)
        rc, ro = counts[order], starts[order]
        self.c.halo_unpack_at(rc, ro, self.buf)
        self._recv = (rc, ro)
        self.halo_atoms = int(counts.sum())
        t2 = time.perf_counter()
        self.pack_seconds += t1 - t0
        self.unpack_seconds += t2 - t1
        self.exchange_seconds += t2 - t0
        self.exchanges += 1

    def step(self, dt: float, langevin: bool = True, gamma: float = 0.0, kbt: float = 0.0, seed: int = 0,
             energy: bool = True):
        self.c.step_begin(dt, langevin, gamma, kbt, seed)
        self.exchange()
        self.c.step_end(energy)

    def run(self, nsteps: int, dt: float, langevin: bool = True, gamma: float = 0.0, kbt: float = 0.0,
            seed: int = 0):
        """nsteps steps; the energy and virial are summed on the last one
        only, as cv2_run_* do."""
        for k in range(nsteps):
            self.step(dt, langevin, gamma, kbt, seed, energy=k == nsteps - 1)

    def ghosts(self):
        """The ghosts of the last exchange as the device holds them (ghost g
        is atom n + g of the pair list): positions (G, 3), types, ids."""
        rc, ro = self._recv
        pos, typ, ids = [], [], []
        for N, at in zip((int(x) for x in rc), (int(x) for x in ro)):
            pk = self.buf[at:at + 5 * N]
            ids.append(pk[:N].copy())
            typ.append(pk[N:2 * N].copy())
            pos.append(pk[2 * N:5 * N].reshape(N, 3).copy())
        return (np.concatenate(pos), np.concatenate(typ).astype(np.int32),
                np.concatenate(ids).astype(np.int64))

    def relax(self, schedule=((0.81, 400), (0.49, 300), (0.36, 150)), dt: float = 0.001,
              gamma: float = 20.0, seed: int = 1):
        """FP64 relaxation of the random start: damped dynamics at zero
        temperature with jump_fix (lj.cpp:459-468: a pair closer than
        sqrt(jump_tol) takes the force of r^2 = jump_tol), in stages of
        decreasing jump_tol, then jump_fix off.  The precision set at
        construction is restored afterwards."""
        prec = self.precision
        self.c.set_precision(D.CV2_PREC_FP64)
        for jt, steps in schedule:
            self.c.set_jump_tol(jt)
            self.run(steps, dt, True, gamma, 0.0, seed)
        self.c.set_jump_tol(0.0)
        self.c.set_precision(prec)

    def close(self):
        self.c.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
