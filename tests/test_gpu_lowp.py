"""Reduced-precision forms (FP32 compute; FP32, FP16 or BF16 storage of the
shifted populations f_i - w_i*rho_shift) against the FP64 oracle, by
tolerance.  Tolerances are relative L2 of the density field after the stated
steps on a noisy start (rho = 0.7 +- 0.01, G = -5.0, early spinodal
growth); the FP16 bound is the 1e-3 target of the precision form."""
import numpy as np
import pytest

import mesobasic_amd as mb
import oracle
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = {mb.FP32: 2e-5, mb.FP16S: 1e-3, mb.BF16S: 1e-2}


@pytest.mark.parametrize("kernel", [mb.KERNEL_TWO_PASS, mb.KERNEL_FUSED])
@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16S, mb.BF16S])
def test_d3q19_lowp_vs_oracle(prec, kernel):
    shape, steps = (32, 32, 32), 200
    print(f"lowp d3q19 prec {prec} kernel {kernel}", flush=True)
    rho0 = mb.noise_rho(shape, 31)
    o = oracle.Oracle(19, shape, 1.0, -5.0, nthreads=4)
    o.set_rho(rho0); o.init(); o.step(steps)
    ref = o.moments()
    lat = mb.Lattice(19, shape, 1.0, -5.0, precision=prec, rho_shift=0.7, kernel=kernel)
    assert lat.kernel == kernel
    lat.set_rho(rho0); lat.init(); lat.step(steps)
    got = lat.fields(mb.FIELDS_CURRENT)
    err = rel_l2(got["rho"], ref["rho"])
    print(f"  rel L2 rho = {err:.3e}", flush=True)
    assert err < TOL[prec]
    c = lat.conservation()
    assert abs(c[0] - ref["rho"].sum()) / ref["rho"].sum() < 10 * TOL[prec]


@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16S, mb.BF16S])
def test_d2q9_lowp_vs_oracle(prec):
    shape, steps = (96, 80), 300
    rho0 = mb.noise_rho(shape, 32)
    o = oracle.Oracle(9, shape, 1.0, -5.0)
    o.set_rho(rho0); o.init(); o.step(steps)
    ref = o.moments()
    lat = mb.Lattice(9, shape, 1.0, -5.0, precision=prec, rho_shift=0.7)
    lat.set_rho(rho0); lat.init(); lat.step(steps)
    got = lat.fields(mb.FIELDS_CURRENT)
    err = rel_l2(got["rho"], ref["rho"])
    print(f"lowp d2q9 prec {prec}: rel L2 rho = {err:.3e}", flush=True)
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16S])
def test_lowp_walls_droplet(prec):
    shape, steps = (24, 32, 32), 100
    rho0 = mb.droplet_rho(shape, 6.0, rl=1.9, rv=0.2)
    o = oracle.Oracle(19, shape, 0.9, -5.5, walls=True, nthreads=4)
    o.set_rho(rho0); o.init(); o.step(steps)
    ref = o.moments()
    lat = mb.Lattice(19, shape, 0.9, -5.5, walls=True, precision=prec, rho_shift=0.7)
    lat.set_rho(rho0); lat.init(); lat.step(steps)
    err = rel_l2(lat.fields(mb.FIELDS_CURRENT)["rho"], ref["rho"])
    print(f"walls droplet prec {prec}: rel L2 rho = {err:.3e}", flush=True)
    assert err < 5 * TOL[prec]
