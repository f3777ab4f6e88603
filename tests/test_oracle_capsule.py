"""The CPU oracle's immersed-boundary capsule model (capsule2D with
IBcapsules2D, restated in oracle/scmp_oracle.c) against fixtures written by
the reference's own classes (tests/golden/capsule*.npz, via
oracle/_ref/ref_harness_capsule): bit for bit, every array (fluid fields,
force arrays, populations, IB node positions, velocities and forces, and the
force left on the ghost rows), every dumped step, NaN where the reference
has NaN; plus what the model fixes whatever its rounding: a capsule at rest
in fluid at rest stays exactly where it is, a node's membrane force is
-kstiff times its displacement, and the spread puts each node's force on the
fluid with weights that sum to one."""
import numpy as np
import pytest

import oracle
from conftest import CAPSULE_FIXTURES, load_fixture, same_bits

FLUID = ("rho", "u", "v", "fx", "fy")
NODES = (("x", "x"), ("y", "y"), ("vx", "vx"), ("vy", "vy"), ("nfx", "fx"), ("nfy", "fy"))


def _oracle_for(meta, data):
    o = oracle.CapsuleOracle(tuple(meta["shape"]), meta["tau"], meta["nnodes"], meta["kstiff"],
                             meta["dt"], tuple(meta["body"]))
    o.set_rho(data["rho0"])
    o.circle(*meta["circle"])
    o.init()
    return o


@pytest.mark.parametrize("name", CAPSULE_FIXTURES)
def test_capsule_oracle_matches_reference_fixture_bitwise(name):
    meta, data = load_fixture(name)
    o = _oracle_for(meta, data)
    cur = 0
    for s in sorted(meta["steps"]):
        o.step(s - cur)
        cur = s
        fl, nd = o.fields(), o.nodes()
        for k in FLUID:
            assert same_bits(fl[k].ravel(), data[f"{k}_{s}"]), (name, s, k)
        if f"f_{s}" in data:
            assert same_bits(fl["f"].ravel(), data[f"f_{s}"]), (name, s, "f")
        for key, k in NODES:
            assert same_bits(nd[k], data[f"{key}_{s}"]), (name, s, key)
        gfx, gfy = o.ghost_force()
        assert same_bits(gfx, data[f"gfx_{s}"]) and same_bits(gfy, data[f"gfy_{s}"]), (name, s)
    o.close()


def test_wall_fixture_reaches_the_ghost_row():
    """The wall case exercises the ghost-row path: nodes whose footprint
    takes in the row below the lower wall read r = 0 there and turn
    non-finite (the reference then skips them); the others go on."""
    meta, data = load_fixture("capsule_30x16_wall")
    s = max(meta["steps"])
    assert np.isnan(data[f"x_{s}"]).any() and np.isfinite(data[f"x_{s}"]).any()
    assert np.isnan(data["vx_1"]).any()


def test_capsule_at_rest_stays_at_rest():
    shape, nn = (24, 20), 12
    o = oracle.CapsuleOracle(shape, 0.9, nn, kstiff=3.0, body=(0.0, 0.0))
    o.set_rho(np.ones(shape))
    o.circle(11.3, 9.6, 4.0)
    x0 = o.nodes()["x"].copy()
    y0 = o.nodes()["y"].copy()
    o.init()
    o.step(25)
    nd = o.nodes()
    assert np.array_equal(nd["x"], x0) and np.array_equal(nd["y"], y0)
    assert not np.any(nd["vx"]) and not np.any(nd["vy"])
    assert not np.any(nd["fx"]) and not np.any(nd["fy"])
    fl = o.fields()
    assert np.max(np.abs(fl["rho"] - 1.0)) < 1e-14 and not np.any(fl["u"]) and not np.any(fl["v"])
    o.close()


def test_membrane_force_and_spread_weights():
    """One node displaced from rest in fluid at rest: its force is
    -kstiff * displacement, and the first step's force arrays hold that
    force on its four footprint nodes with the bilinear weights (exact in
    binary here: 0.75/0.25 in x, 0.5/0.5 in y) and zero elsewhere."""
    shape = (16, 12)
    k = 2.5
    o = oracle.CapsuleOracle(shape, 1.0, 1, kstiff=k, body=(0.0, 0.0))
    o.set_rho(np.ones(shape))
    o.set_nodes(np.array([7.0]), np.array([5.0]))
    o.move_nodes(np.array([7.25]), np.array([5.5]))
    o.init()
    o.step(1)
    nd = o.nodes()
    Fx, Fy = -k * 0.25, -k * 0.5
    assert nd["fx"][0] == Fx and nd["fy"][0] == Fy
    fl = o.fields()
    want_x, want_y = np.zeros(shape), np.zeros(shape)
    for px, wx in ((7, 0.75), (8, 0.25)):
        for py, wy in ((5, 0.5), (6, 0.5)):
            want_x[px, py] = wx * wy * Fx
            want_y[px, py] = wx * wy * Fy
    assert np.array_equal(fl["fx"], want_x) and np.array_equal(fl["fy"], want_y)
    assert fl["fx"].sum() == Fx and fl["fy"].sum() == Fy
    o.close()


def test_body_force_rows():
    """The body force acts on the rows between the walls only (j = 2 ..
    ny-1 in ghost-shell indices: every row but the first and the last)."""
    shape = (10, 8)
    o = oracle.CapsuleOracle(shape, 1.0, 1, kstiff=0.0, body=(0.000008, 0.0))
    o.set_rho(np.ones(shape))
    o.set_nodes(np.array([-5.0]), np.array([3.0]))     # outside the lattice: skipped
    o.init()
    o.step(1)
    fl = o.fields()
    assert np.all(fl["fx"][:, 1:-1] == 0.000008) and not np.any(fl["fx"][:, [0, -1]])
    assert not np.any(fl["fy"])
    o.close()
