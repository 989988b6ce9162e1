"""Batched ensemble (many independent systems in one set of arrays, one
captured graph per step) against the oracle and the single-box context.

- pair sets of every system are bit-exact against the oracle, at the start
  and after dynamics;
- accelerations match the FP64 oracle within the FP32+fixed-point tolerance;
- each system's trajectory is bit-identical to the same system advanced
  alone by the single-box context in FP32+fixed-point (fixed-point sums do
  not depend on list order, so the two must agree exactly);
- replays of the captured step graph, interleaved with eager work and
  synchronisations, give the same state as eager launches;
- the autotuner leaves the uploaded state unchanged.
"""
import numpy as np
import pytest

from cv2md import device as D
from cv2md.systems import ensemble_systems
from helpers import force_error, oracle_forces, oracle_pairs, sort_rows

pytestmark = pytest.mark.gpu

RC, NL_COEF = 2.5, 1.12
DT = 0.005


def sysdef(sy):
    n = len(sy["pos"])
    ce = RC * NL_COEF
    return dict(lo=sy["lo"], hi=sy["hi"], bc=(1, 1, 1), pos=sy["pos"], vel=sy["vel"],
                type=np.zeros(n, np.int32), eps=np.ones((1, 1)), sigma=np.ones((1, 1)),
                cutoff_list=None, mass=np.ones(1), lj_cutoff=RC, nl_cutoff=RC, nl_coef=NL_COEF,
                ghost_cutoff=ce, dt=DT, integrator=0, gamma=0.0, kbt=1.0, k_b=1.0, n_r_df=3)


def small_ensemble(seed=21, nsys=6):
    return ensemble_systems(nsys, seed=seed, nmin=500, nmax=3000)


def split(e, a):
    return [a[e.offsets[s]:e.offsets[s + 1]] for s in range(len(e.natoms))]


def test_ensemble_pairs_bit_exact_and_forces():
    systems = small_ensemble()
    print(f"[ensemble] sizes {[len(s['pos']) for s in systems]}", flush=True)
    with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
        for rnd in range(2):
            if rnd:
                e.build()  # the list of the current positions (as Neighborlist::build(true))
            pos, _, acc = e.download()
            for s, (sy, p, a) in enumerate(zip(systems, split(e, pos), split(e, acc))):
                sd = sysdef(dict(sy, pos=p))
                got = sort_rows(e.pairs(s))
                exp = oracle_pairs(p, sd)
                assert len(got) == len(exp), (rnd, s)
                assert np.array_equal(got, exp), (rnd, s)
                _p, ref, _vir, _pe = oracle_forces(p, sd)
                emax, erms = force_error(a, ref)
                assert emax < 2e-5 and erms < 5e-6, (rnd, s, emax, erms)
            e.run_vv(300, DT)
            if not rnd:
                st = e.stats()
        assert st["system_builds"] > len(systems)  # systems were rebuilt during the run
        assert st["graph_nodes"] > 0


def lattice_system(L, rho, seed, temperature=1.5):
    """Jittered simple-cubic fill of an (Lx, Ly, Lz) periodic box."""
    rng = np.random.default_rng(seed)
    L = np.asarray(L, float)
    a = (1.0 / rho) ** (1.0 / 3.0)
    k = np.maximum(1, np.round(L / a)).astype(int)
    g = np.stack(np.meshgrid(*[np.arange(m) for m in k], indexing="ij"), -1).reshape(-1, 3)
    pos = np.mod((g + 0.5) * (L / k) + rng.normal(0.0, 0.04, g.shape), L)
    vel = rng.normal(0.0, np.sqrt(temperature), pos.shape)
    return dict(lo=np.zeros(3), hi=L, pos=pos, vel=vel - vel.mean(0))


def test_small_boxes_pairs_bit_exact():
    """Boxes with 2, 3 and 4 list-cutoff-wide cells on some axis (the wave
    build meets a cell under two images when an axis has 2; relative
    coordinates come from the FP64 positions below 5 cells), next to a
    larger one: pair sets bit-exact against the oracle before and after
    dynamics, forces within the FP32+fixed-point tolerance."""
    boxes = [(6.0, 6.0, 6.0), (8.6, 8.6, 8.6), (11.5, 11.5, 11.5), (6.2, 9.1, 14.6), (15.0, 15.0, 15.0)]
    systems = [lattice_system(L, 0.8, 90 + i) for i, L in enumerate(boxes)]
    print(f"[ensemble] small boxes {boxes}", flush=True)
    with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
        for rnd in range(2):
            if rnd:
                e.run_vv(200, DT)
                e.build()
            pos, _, acc = e.download()
            for s, (sy, p, a) in enumerate(zip(systems, split(e, pos), split(e, acc))):
                sd = sysdef(dict(sy, pos=p))
                got = sort_rows(e.pairs(s))
                exp = oracle_pairs(p, sd)
                assert len(got) == len(exp), (rnd, s)
                assert np.array_equal(got, exp), (rnd, s)
                _p, ref, _vir, _pe = oracle_forces(p, sd)
                emax, erms = force_error(a, ref)
                assert emax < 2e-5 and erms < 5e-6, (rnd, s, emax, erms)


def test_ensemble_matches_single_box_context():
    systems = small_ensemble(seed=22, nsys=4)
    with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
        e.run_vv(200, DT)
        pos, vel, acc = e.download()
    for s, sy in enumerate(systems):
        o0 = int(np.sum([len(x["pos"]) for x in systems[:s]]))
        o1 = o0 + len(sy["pos"])
        with D.Context.from_sysdef(sysdef(sy), precision=D.CV2_PREC_FP32_FIXED) as c:
            c.compute_forces()
            c.run_vv(200, DT)
            p1, v1, a1 = c.download()
        # same pair forces in 64-bit fixed point: identical trajectories
        assert np.array_equal(np.mod(p1, sy["hi"][0]), np.mod(pos[o0:o1], sy["hi"][0])), s
        assert np.array_equal(v1, vel[o0:o1]), s
        assert np.array_equal(a1, acc[o0:o1]), s


def test_graph_replays_match_eager_launches():
    systems = small_ensemble(seed=23, nsys=5)
    states = []
    for graph in (True, False):
        with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
            e.use_graph(graph)
            ke_hist = []
            for _ in range(4):  # replays with eager work and synchronisations between them
                e.run_vv(60, DT)
                ke_hist.append(e.thermo()[0].copy())
            states.append((e.download(), ke_hist, e.stats()["graph_nodes"]))
    (pg, kg, ng), (pe, ke, ne) = states
    assert ng > 0 and ne == 0
    for a, b in zip(pg, pe):
        assert np.array_equal(a, b)
    for a, b in zip(kg, ke):
        assert np.array_equal(a, b)


def test_energy_conserved_per_system():
    systems = small_ensemble(seed=24, nsys=4)
    with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
        ke0, pe0, _ = e.thermo()
        e.run_vv(2000, DT)
        ke1, pe1, _ = e.thermo()
    n = e.natoms
    drift = np.abs((ke1 + pe1) - (ke0 + pe0)) / n
    print(f"per-atom energy change {drift}")
    assert np.all(drift < 5e-3)


def test_tuner_keeps_state():
    systems = small_ensemble(seed=25, nsys=3)
    with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
        before = e.download()
        best = e.tune(DT, trial_steps=5)
        after = e.download()
        assert best["force_block"] in (64, 128, 256) and best["build_block"] in (64, 128, 256)
        assert best["rebuild_all"] in (0, 1) and best["ms_per_step"] > 0
        assert best["cell_scale"] in (1.0, 1.15, 1.3)
        for a, b in zip(before, after):
            assert np.array_equal(a, b)
        e.run_vv(50, DT)


@pytest.mark.parametrize("scale", [1.15, 1.3])
def test_cell_width_gives_the_same_bits(scale):
    """The tuner's cell-width axis: wider cells (fewer cells, more
    candidates per home atom) give the same pair sets, and with fixed-point
    sums the same trajectory bit for bit, as cells one list cutoff wide."""
    systems = small_ensemble(seed=26, nsys=5)
    out = []
    for sc in (1.0, scale):
        with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
            e.set_tuning(cell_scale=sc)
            e.run_vv(150, DT)
            e.build()
            out.append((e.download(), [sort_rows(e.pairs(k)) for k in range(len(systems))],
                        e.stats()["system_builds"]))
    (d0, p0, b0), (d1, p1, b1) = out
    assert b0 == b1 and b0 > len(systems)
    for a, b in zip(d0, d1):
        assert np.array_equal(a, b)
    for a, b in zip(p0, p1):
        assert np.array_equal(a, b)


def test_c4_workload_512_systems():
    """The benchmark's ensemble at its real shape: 512 systems of 500 to
    20,000 atoms (seed 1004, about 2.7M atoms), tuned by the autotuner, then
    200 velocity-Verlet steps through the captured step graph and a list
    rebuild of every system.  For a seeded sample of 40 systems that
    includes the smallest and the largest: the pair set is bit-exact against
    the oracle's cell list at the device's positions, and the FP32 +
    fixed-point accelerations lie within 1e-6 of the FP64 oracle on the
    metric of bench.py (per atom |a32 - a64| / max(|a64|, rms|a64|))."""
    systems = ensemble_systems(512, seed=1004)
    sizes = np.array([len(s["pos"]) for s in systems])
    rng = np.random.default_rng(1004)
    sample = sorted({int(sizes.argmin()), int(sizes.argmax())} |
                    set(int(x) for x in rng.choice(len(systems), 38, replace=False)))
    with D.Ensemble.from_systems(systems, rc=RC, nl_cutoff=RC, nl_coef=NL_COEF) as e:
        tuned = e.tune(DT, trial_steps=10)
        e.run_vv(200, DT)
        st = e.stats()
        e.build()
        pos, _, acc = e.download()
        pp, aa = split(e, pos), split(e, acc)
        print(f"[c4] {len(systems)} systems, {int(sizes.sum())} atoms, tuning {tuned}, "
              f"system builds {st['system_builds']}, graph nodes {st['graph_nodes']}", flush=True)
        worst = 0.0
        for s in sample:
            sd = sysdef(dict(systems[s], pos=pp[s]))
            got = sort_rows(e.pairs(s))
            exp = oracle_pairs(pp[s], sd)
            assert np.array_equal(got, exp), s
            _p, ref, _vir, _pe = oracle_forces(pp[s], sd)
            emax, erms = force_error(aa[s], ref)
            worst = max(worst, emax)
            assert emax <= 1e-6 and erms <= 2.5e-7, (s, sizes[s], emax, erms)
        print(f"[c4] sample of {len(sample)} systems ({sizes[sample].min()}..{sizes[sample].max()} "
              f"atoms): pairs bit-exact, worst force error {worst:.3e}", flush=True)
    assert st["graph_nodes"] > 0 and st["system_builds"] > len(systems)
