"""CV2_NL_VERIFY: the device's check of every list build.

With the flag on, each list build is followed on the device by an all-pairs
FP64 recomputation of the pair set (no cells, coarse filter or fixed point),
and the two are compared by a 64-bit checksum and a count.  These tests pin
the check against the oracle: at the positions of a build, the checksum and
count the device reports for its list and for its recomputation both equal
the sum taken in numpy over the oracle's pair set (helpers.pair_checksum).
A recomputation that dropped or added a pair, or labelled an image wrongly,
would differ from the oracle's sum, so a check that reports no mismatch is a
check that could have seen one.
"""
import numpy as np
import pytest

from cv2md import device as D
from cv2md.systems import ka_mixture, lj_fluid
from helpers import oracle_pairs, pair_checksum, sort_rows
from test_gpu_parity import CASES

pytestmark = pytest.mark.gpu

PRECISIONS = {"fp64": D.CV2_PREC_FP64, "fp32": D.CV2_PREC_FP32_FIXED}


def expect_check(c, s, tag):
    """The last check's sums against the oracle at the device's positions."""
    st = c.stats()
    pos = c.download()[0]
    exp = pair_checksum(oracle_pairs(pos, dict(s, pos=pos)))
    print(f"[verify] {tag}: {exp[1]} ordered pairs, checked {st['verified_builds']}", flush=True)
    assert st["verify_pairs"] == (exp[1], exp[1]), tag
    assert st["verify_sum"] == (exp[0], exp[0]), tag
    assert st["verify_mismatches"] == 0, tag
    return st


@pytest.mark.parametrize("prec", sorted(PRECISIONS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_check_matches_oracle(name, prec):
    s = CASES[name]()
    with D.Context.from_sysdef(s, precision=PRECISIONS[prec], verify=True) as c:
        c.build(True)
        st = expect_check(c, s, f"{name} {prec}")
        assert st["verified_builds"] == st["rebuilds"] == 1
        # the list the device built is the pair set the oracle gives
        assert np.array_equal(sort_rows(c.pairs()), oracle_pairs(s["pos"], s))


@pytest.mark.parametrize("prec", sorted(PRECISIONS))
def test_every_rebuild_of_a_run_checked(prec):
    # 4000-atom LJ fluid, velocity Verlet dt 0.005: every build of 400 steps
    # is checked; after each of the first rebuilds the positions are those the
    # build saw, so the sums are compared with the oracle's there
    s = lj_fluid(10, temperature=1.44, seed=1101)
    with D.Context.from_sysdef(s, precision=PRECISIONS[prec], verify=True) as c:
        c.compute_forces()
        expect_check(c, s, f"setup {prec}")
        step, seen = 0, 0
        while seen < 3:
            r0 = c.stats()["rebuilds"]
            c.run_vv(1, 0.005)
            step += 1
            if c.stats()["rebuilds"] != r0:
                expect_check(c, s, f"{prec} rebuild at step {step}")
                seen += 1
            assert step < 200, "fewer rebuilds than expected"
        c.run_vv(400 - step, 0.005)
        st = c.stats()
    print(f"[verify] {prec}: {st['rebuilds']} rebuilds, {st['verified_builds']} checked", flush=True)
    assert st["rebuilds"] > 10
    assert st["verified_builds"] == st["rebuilds"]
    assert st["verify_mismatches"] == 0


def test_mixture_ids_and_switching():
    # caller-given atom ids enter the checksum; the flag can be dropped and
    # set again on a live context
    s = ka_mixture(2000, seed=1102)
    n = len(s["pos"])
    ids = np.random.default_rng(7).permutation(n).astype(np.int64) + 1000
    with D.Context(0) as c:
        c.set_box(s["lo"], s["hi"], s["bc"], s["ghost_cutoff"])
        c.set_lj(s["eps"], s["sigma"], s["lj_cutoff"], s["mass"], s.get("cutoff_list"))
        c.set_neighbor(s["nl_cutoff"], s["nl_coef"], D.CV2_NL_VERIFY)
        c.set_precision(D.CV2_PREC_FP32_FIXED)
        c.upload(s["pos"], s["vel"], s["type"], ids)
        c.compute_forces()
        st = c.stats()
        exp = pair_checksum(oracle_pairs(s["pos"], s), ids)
        assert st["verify_sum"] == (exp[0], exp[0])
        assert st["verify_pairs"] == (exp[1], exp[1])
        c.set_neighbor(s["nl_cutoff"], s["nl_coef"], 0)
        c.run_vv(20, 0.002)
        assert c.stats()["verified_builds"] == 0
        c.set_neighbor(s["nl_cutoff"], s["nl_coef"], D.CV2_NL_VERIFY)
        c.run_vv(20, 0.002)
        st = c.stats()
        assert st["verified_builds"] >= 1 and st["verify_mismatches"] == 0


def test_host_halo_refused():
    s = lj_fluid(6, seed=1103)
    with D.Context.from_sysdef(dict(s, bc=(0, 0, 0)), verify=True) as c:
        with pytest.raises(D.Cv2Error) as e:
            c.set_ghosts(np.zeros((1, 3)) + 1.0, np.zeros(1, np.int32), np.zeros(1, np.int64))
        assert e.value.code == 6  # CV2_ERR_STATE


def test_c3_every_build_checked():
    # the 2,048,000-atom FP32 + fixed-point Langevin workload (80^3 FCC start,
    # T* = 1, gamma = 1, dt = 0.005): every list build of 300 steps, the
    # lattice melting included, checked on the device (the coarse-filter wave
    # build in its 32-bit form)
    s = lj_fluid(80, temperature=1.0, seed=1003, integrator=1, gamma=1.0)
    with D.Context.from_sysdef(s, precision=D.CV2_PREC_FP32_FIXED, verify=True) as c:
        c.compute_forces()
        c.run_langevin(300, s["dt"], 1.0, 1.0, 1003)
        st = c.stats()
    print(f"[verify] C3: {st['rebuilds']} rebuilds, {st['verified_builds']} checked, "
          f"{st['verify_mismatches']} differed, last {st['verify_pairs'][0]} ordered pairs", flush=True)
    assert st["build_kind"] == "cell_wave"
    assert st["rebuilds"] > 30
    assert st["verified_builds"] == st["rebuilds"]
    assert st["verify_mismatches"] == 0
