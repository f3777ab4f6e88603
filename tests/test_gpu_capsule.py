"""Immersed-boundary capsule (capsule2D) on the GPU, FP64, through the C ABI:
bit for bit against fixtures written by the reference's own capsule2D /
IBcapsules2D classes (fluid r, u, v, force arrays and f; IB node positions,
velocities and membrane forces; the ghost-row path), and against the oracle
on further lattices, stiffnesses, body forces and node counts above one LDS
chunk (256 nodes), with the force arrays also read through a snapshot."""
import numpy as np
import pytest

import mesobasic_amd as mb
import oracle
from conftest import CAPSULE_FIXTURES, load_fixture, same_bits

pytestmark = pytest.mark.gpu


def _gpu(shape, tau, nnodes, kstiff, dt=1.0, body=(0.000008, 0.0)):
    lat = mb.Lattice(9, shape, tau, 0.0, order=2, walls=True, precision=mb.FP64)
    lat.attach_capsule(nnodes, kstiff, dt, body)
    return lat


def _circle(shape, tau, nnodes, kstiff, xo, yo, rad):
    """The circle's node positions as initCircularCapsule computes them (the
    oracle's restatement, bit-identical to the reference's)."""
    o = oracle.CapsuleOracle(shape, tau, nnodes, kstiff)
    o.circle(xo, yo, rad)
    nd = o.nodes()
    o.close()
    return nd["x"], nd["y"]


@pytest.mark.parametrize("name", CAPSULE_FIXTURES)
def test_capsule_fixture_bitwise(name):
    meta, data = load_fixture(name)
    shape = tuple(meta["shape"])
    print(f"capsule fixture {name}", flush=True)
    lat = _gpu(shape, meta["tau"], meta["nnodes"], meta["kstiff"], meta["dt"], tuple(meta["body"]))
    assert lat.kernel == mb.KERNEL_TWO_PASS
    x, y = _circle(shape, meta["tau"], meta["nnodes"], meta["kstiff"], *meta["circle"])
    lat.set_capsule_nodes(x, y)
    lat.set_rho(data["rho0"])
    lat.init()
    cur = 0
    for s in sorted(meta["steps"]):
        lat.step(s - cur)
        cur = s
        got = lat.fields(mb.FIELDS_REFERENCE)
        for k in ("rho", "u", "v"):
            assert same_bits(got[k].ravel(), data[f"{k}_{s}"]), (name, s, k)
        fx, fy, fz = lat.force()
        assert same_bits(fx.ravel(), data[f"fx_{s}"]) and same_bits(fy.ravel(), data[f"fy_{s}"]), (name, s)
        assert not np.any(fz)
        if f"f_{s}" in data:
            assert same_bits(lat.get_f().ravel(), data[f"f_{s}"]), (name, s, "f")
        nd = lat.capsule_nodes()
        for key, k in (("x", "x"), ("y", "y"), ("vx", "vx"), ("vy", "vy"), ("nfx", "fx"), ("nfy", "fy")):
            assert same_bits(nd[k], data[f"{key}_{s}"]), (name, s, key)
    lat.close()


@pytest.mark.parametrize("shape,tau,nnodes,kstiff,body,circle,steps", [
    # node counts above one LDS chunk, with stiffnesses the dense membranes
    # stay stable at (a stiff dense membrane blows up, in the oracle as here)
    ((96, 64), 0.7, 600, 0.02, (0.00002, -0.000005), (40.0, 31.5, 20.0), 120),
    ((130, 37), 1.3, 40, 1.0, (0.000008, 0.0), (20.5, 18.0, 9.0), 200),
    ((64, 48), 1.0, 300, 0.05, (0.0, 0.0001), (31.0, 23.0, 14.0), 150),
])
def test_capsule_matches_oracle(shape, tau, nnodes, kstiff, body, circle, steps):
    print(f"capsule vs oracle {shape} nnodes={nnodes}", flush=True)
    rho0 = oracle.noise_rho(shape, 7, mean=0.5, amp=0.1)
    o = oracle.CapsuleOracle(shape, tau, nnodes, kstiff, body=body)
    o.set_rho(rho0)
    o.circle(*circle)
    x, y = o.nodes()["x"].copy(), o.nodes()["y"].copy()
    o.init()
    lat = _gpu(shape, tau, nnodes, kstiff, body=body)
    lat.set_capsule_nodes(x, y)
    lat.set_rho(rho0)
    lat.init()
    for chunk in (1, steps // 2 - 1, steps - steps // 2):
        o.step(chunk)
        lat.step(chunk)
        want = o.fields()
        got = lat.fields(mb.FIELDS_REFERENCE)
        for k in ("rho", "u", "v"):
            assert same_bits(got[k], want[k]), k
        fx, fy, _ = lat.force()
        assert same_bits(fx, want["fx"]) and same_bits(fy, want["fy"])
        nd, wn = lat.capsule_nodes(), o.nodes()
        for k in ("x", "y", "vx", "vy", "fx", "fy"):
            assert same_bits(nd[k], wn[k]), k
    # the force arrays through an asynchronous snapshot
    lat.snapshot_begin()
    _, snap = lat.snapshot_wait()
    assert same_bits(snap["fx"], want["fx"]) and same_bits(snap["fy"], want["fy"])
    assert np.isfinite(nd["x"]).all() and np.max(np.abs(nd["x"] - x)) > 0.0
    lat.close()
    o.close()


def test_capsule_errors():
    with pytest.raises(mb.LbmError):
        mb.Lattice(19, (8, 8, 16), 1.0, 0.0).attach_capsule(10)            # D3Q19
    with pytest.raises(mb.LbmError):
        mb.Lattice(9, (16, 12), 1.0, 0.0, order=3).attach_capsule(10)      # scmp3D step order
    with pytest.raises(mb.LbmError):
        mb.Lattice(9, (16, 12), 1.0, 0.0, precision=mb.FP32).attach_capsule(10)
    lat = _gpu((16, 12), 1.0, 4, 1.0)
    with pytest.raises(mb.LbmError):
        lat.attach_capsule(4)                                              # twice
    with pytest.raises(mb.LbmError):
        lat.set_capsule_nodes(np.array([3.0, 4.0, 5.0, 6.0]), np.array([2.0, 2.0, 12.5, 2.0]))  # y >= ny
    lat.set_rho(np.full((16, 12), 0.5))
    lat.init()
    with pytest.raises(mb.LbmError):
        lat.step(1)                                                        # nodes not set
    lat.set_capsule_nodes(np.array([3.0, 4.0, 5.0, 6.0]), np.array([2.0, 3.0, 4.0, 5.0]))
    lat.step(2)
    assert lat.steps_done == 2
    late = mb.Lattice(9, (16, 12), 1.0, 0.0, walls=True)
    late.set_rho(np.full((16, 12), 0.5))
    late.init()
    with pytest.raises(mb.LbmError):
        late.attach_capsule(4)                                             # after init
