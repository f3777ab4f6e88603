"""The CPU oracle's two-component model (mcmp2D / mcmp3D and the film forms
mcmp2DFilm / mcmp3DFilm, restated in oracle/scmp_oracle.c) against fixtures
written by the reference's own classes (tests/golden/mcmp*.npz, via
oracle/_ref/ref_harness_mcmp): bit for bit, every array, every dumped step
(NaN where the reference has NaN: mcmp3DFilm's wall rows);
plus the invariants the model fixes whatever its rounding: each component's
mass is conserved, the two cross forces cancel in total (momentum), the
common velocity is the same in both components' arrays after an mcmp3D
step, and with G = 0 each component relaxes as an independent BGK fluid
(identical to the single-component oracle at G = 0)."""
import numpy as np
import pytest

import oracle
from conftest import MC_FIXTURES, load_fixture, same_bits

KEYS = ("rho", "u", "v", "w", "fx", "fy", "fz")


@pytest.mark.parametrize("name", MC_FIXTURES)
def test_mc_oracle_matches_reference_fixture_bitwise(name):
    meta, data = load_fixture(name)
    q, shape = meta["q"], tuple(meta["shape"])
    o = oracle.MCOracle(q, shape, meta["tauA"], meta["tauB"], meta["G"], nthreads=3, film=meta.get("film"))
    o.set_rho(data["rhoA0"], data["rhoB0"])
    o.init()
    cur = 0
    for s in sorted(meta["steps"]):
        o.step(s - cur)
        cur = s
        for comp, c in ((0, "A"), (1, "B")):
            got = o.fields(comp)
            for k in KEYS:
                key = f"{k}_{c}_{s}"
                if key in data:
                    assert same_bits(got[k].ravel(), data[key]), (name, s, key)
            if f"f_{c}_{s}" in data:
                assert same_bits(got["f"].ravel(), data[f"f_{c}_{s}"]), (name, s, c)
    o.close()


@pytest.mark.parametrize("q,shape,order", [(9, (20, 16), 2), (19, (8, 7, 6), 3), (9, (12, 10), 3),
                                           (19, (6, 5, 7), 2)])
def test_mc_invariants(q, shape, order):
    rA = oracle.noise_rho(shape, 3, mean=0.5, amp=0.1)
    rB = oracle.noise_rho(shape, 4, mean=0.5, amp=0.1)
    o = oracle.MCOracle(q, shape, 1.0, 0.8, 3.0, order=order)
    o.set_rho(rA, rB)
    o.init()
    o.step(25)
    for comp, r0 in ((0, rA), (1, rB)):
        m = o.moments(comp)["rho"].sum()
        assert abs(m / r0.sum() - 1.0) < 1e-12
    fa, fb = o.fields(0), o.fields(1)
    for k in ("fx", "fy", "fz"):
        # sum over the lattice of F_A + F_B = -G sum_x sum_n w e (psiA(x) psiB(x+e) - ...)
        # which telescopes to zero on a periodic lattice
        assert abs((fa[k] + fb[k]).sum()) < 1e-12
    if order == 3:
        for k in ("u", "v", "w"):
            assert same_bits(fa[k], fb[k])
    o.close()


@pytest.mark.parametrize("q,shape,order", [(9, (14, 12), 2), (19, (7, 6, 5), 3), (9, (10, 9), 3)])
def test_mc_without_coupling_sums_to_one_bgk_fluid(q, shape, order):
    # G = 0 and tauA = tauB: the common velocity is the mixture's j / rho and
    # feq is linear in the density, so f_A + f_B evolves exactly as a single
    # BGK fluid of density rho_A + rho_B at G = 0 (up to rounding)
    tau = 0.8
    rA = oracle.noise_rho(shape, 5, mean=0.5, amp=0.1)
    rB = oracle.noise_rho(shape, 6, mean=0.4, amp=0.2)
    mc = oracle.MCOracle(q, shape, tau, tau, 0.0, order=order)
    mc.set_rho(rA, rB)
    mc.init()
    one = oracle.Oracle(q, shape, tau, 0.0, order=order)
    one.set_rho(rA + rB)
    one.init()
    mc.step(30)
    one.step(30)
    ftot = mc.fields(0)["f"] + mc.fields(1)["f"]
    ref = one.fields()["f"]
    assert np.max(np.abs(ftot - ref)) < 1e-13 * np.max(np.abs(ref))
    # and a wrongly weighted common velocity (tau ignored) would not do this:
    # with unequal taus the sum departs from the one-fluid result
    mc2 = oracle.MCOracle(q, shape, tau, 1.4, 0.0, order=order)
    mc2.set_rho(rA, rB)
    mc2.init()
    mc2.step(30)
    assert np.max(np.abs(mc2.fields(0)["f"] + mc2.fields(1)["f"] - ref)) > 1e-8
    mc.close()
    mc2.close()
    one.close()
