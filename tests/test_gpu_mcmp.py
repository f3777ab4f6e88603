"""Two-component Shan-Chen (mcmp2D / mcmp3D, and the film forms mcmp2DFilm /
mcmp3DFilm) on the GPU, FP64, through the C ABI: bit for bit against
fixtures written by the reference's own classes (every dumped array of both
components: r, u, v, w, the force arrays and f), and against the oracle on
further lattices, taus, wall densities and both step orders, with restarts
through set_f."""
import numpy as np
import pytest

import mesobasic_amd as mb
import oracle
from conftest import MC_FIXTURES, load_fixture, rel_l2, same_bits

pytestmark = pytest.mark.gpu


def _mc_lattice(q, shape, tauA, tauB, G, order=None, film=None):
    return mb.Lattice(q, shape, tauA, G, order=order, precision=mb.FP64, components=2, tau_b=tauB,
                      walls=film is not None, rho_sol=film if film is not None else (0.5, 0.5))


@pytest.mark.parametrize("name", MC_FIXTURES)
def test_mc_fixture_bitwise(name):
    meta, data = load_fixture(name)
    q, shape = meta["q"], tuple(meta["shape"])
    print(f"mcmp fixture {name}", flush=True)
    lat = _mc_lattice(q, shape, meta["tauA"], meta["tauB"], meta["G"], film=meta.get("film"))
    assert lat.kernel == mb.KERNEL_TWO_PASS
    lat.set_rho(data["rhoA0"], comp=0)
    lat.set_rho(data["rhoB0"], comp=1)
    lat.init()
    cur = 0
    for s in sorted(meta["steps"]):
        lat.step(s - cur)
        cur = s
        for comp, c in ((0, "A"), (1, "B")):
            got = lat.fields(mb.FIELDS_REFERENCE, comp=comp)
            for k in ("rho", "u", "v", "w"):
                key = f"{k}_{c}_{s}"
                if key in data:
                    assert same_bits(got[k].ravel(), data[key]), (name, s, key, rel_l2(got[k], data[key]))
            fx, fy, fz = lat.force(comp=comp)
            for k, arr in (("fx", fx), ("fy", fy), ("fz", fz)):
                key = f"{k}_{c}_{s}"
                if key in data:
                    assert same_bits(arr.ravel(), data[key]), (name, s, key)
            if f"f_{c}_{s}" in data:
                assert same_bits(lat.get_f(comp=comp).ravel(), data[f"f_{c}_{s}"]), (name, s, c)
    lat.close()


MC_CASES = [
    # q, shape, tauA, tauB, G, order, steps[, film]
    (9, (31, 27), 1.0, 0.7, 3.2, 2, 60),
    (9, (20, 64), 0.8, 1.3, 2.6, 3, 40),
    (19, (9, 8, 7), 0.9, 1.1, 3.0, 3, 30),
    (19, (6, 10, 33), 1.0, 1.0, 3.4, 2, 25),
    (9, (26, 19), 0.9, 1.2, 3.0, 2, 50, (0.2, 0.8)),
    (19, (7, 9, 12), 1.0, 0.8, 2.8, 3, 30, (0.7, 0.3)),
    (19, (5, 6, 9), 1.1, 1.0, 3.0, 2, 25, (0.5, 0.5)),
]


@pytest.mark.parametrize("case", MC_CASES)
def test_mc_matches_oracle_bitwise(case):
    q, shape, tauA, tauB, G, order, steps = case[:7]
    film = case[7] if len(case) > 7 else None
    print(f"mcmp case {case}", flush=True)
    rA = mb.noise_rho(shape, 31, mean=0.5, amp=0.1)
    rB = mb.noise_rho(shape, 32, mean=0.5, amp=0.1)
    o = oracle.MCOracle(q, shape, tauA, tauB, G, order=order, film=film)
    o.set_rho(rA, rB)
    o.init()
    lat = _mc_lattice(q, shape, tauA, tauB, G, order=order, film=film)
    lat.set_rho(rA, comp=0)
    lat.set_rho(rB, comp=1)
    lat.init()
    for chunk in (1, 1, steps - 2):
        o.step(chunk)
        lat.step(chunk)
        for comp in (0, 1):
            ref = o.fields(comp)
            got = lat.fields(mb.FIELDS_REFERENCE, comp=comp)
            for k in ("rho", "u", "v", "w"):
                assert same_bits(got[k], ref[k]), (comp, k, rel_l2(got[k], ref[k]))
            fx, fy, fz = lat.force(comp=comp)
            assert same_bits(fx, ref["fx"]) and same_bits(fy, ref["fy"]) and same_bits(fz, ref["fz"]), comp
            assert same_bits(lat.get_f(comp=comp), ref["f"]), comp
            cur = lat.fields(mb.FIELDS_CURRENT, comp=comp)
            mom = o.moments(comp)
            for k in ("rho", "u", "v", "w"):
                assert same_bits(cur[k], mom[k]), (comp, k)
    lat.close()
    o.close()


@pytest.mark.parametrize("q,shape,order", [(9, (24, 18), 2), (19, (8, 6, 10), 3)])
def test_mc_restart_and_conservation(q, shape, order):
    rA = mb.noise_rho(shape, 41, mean=0.5, amp=0.1)
    rB = mb.noise_rho(shape, 42, mean=0.5, amp=0.1)
    a = _mc_lattice(q, shape, 1.0, 0.8, 3.0, order=order)
    a.set_rho(rA, comp=0)
    a.set_rho(rB, comp=1)
    a.init()
    a.step(9)
    snap = [a.get_f(comp=c) for c in (0, 1)]
    a.step(11)
    # the same run restarted from the snapshot taken after 9 steps
    b = _mc_lattice(q, shape, 1.0, 0.8, 3.0, order=order)
    b.set_rho(rA, comp=0)
    b.set_rho(rB, comp=1)
    b.init()
    b.step(3)                       # post-collision state in the other buffer
    b.set_f(snap[0], comp=0)
    b.set_f(snap[1], comp=1)
    b.step(11)
    for c in (0, 1):
        assert same_bits(a.get_f(comp=c), b.get_f(comp=c)), c
        m = a.conservation(comp=c)
        assert abs(m[0] / (rA if c == 0 else rB).sum() - 1.0) < 1e-12
    a.close()
    b.close()


def test_mc_descriptions_refused():
    for kw in (dict(precision=mb.FP16N), dict(kernel=mb.KERNEL_FUSED)):
        with pytest.raises(mb.LbmError):
            mb.Lattice(19, (8, 8, 64), 1.0, 3.0, components=2, tau_b=1.0, **kw)
    lat = mb.Lattice(9, (8, 8), 1.0, 3.0, components=2, tau_b=1.0)
    with pytest.raises(mb.LbmError):
        lat.set_rho(np.ones((8, 8)), comp=2)
    lat.close()
    one = mb.Lattice(9, (8, 8), 1.0, -5.0)
    with pytest.raises(mb.LbmError):
        one.set_rho(np.ones((8, 8)), comp=1)
    one.close()
