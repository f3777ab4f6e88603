"""Accuracy of the reduced-precision forms at full size, against the FP64
GPU path on the same seeded inputs.  The FP64 path is bit-identical to the
reference (tests/test_fullsize.py checks the reference's own hash of this
256^3 droplet run), so it stands in for the FP64 oracle at sizes the CPU
restatement cannot reach in a test.

Droplet (D3Q19 256^3, tau 1, G -5.5, R0 48, W 2, 2.0 / 0.15, 2,000 steps):
density L2, radius, Laplace pressure jump and coexistence densities.

Spinodal (D3Q19 512^3, tau 1, G -5.0, rho 0.7 +- 0.01, seed 1, 5,000 steps):
spinodal decomposition grows any difference exponentially, so even FP32
storage ends 5e-3 off FP64 at 5,000 steps and no reduced form can hold the
pointwise L2 to 1e-3 there.  The L2 is checked at 500 steps (FP32 and FP16
node shift below 1e-3; FP32 stays below it through 1,000 steps, FP16 with
the node shift crosses it between 500 and 1,000), and at 5,000 steps the
size-independent results: bulk coexistence densities (phase medians) and
total-mass drift."""
import numpy as np
import pytest

import mesobasic_amd as mb
from mesobasic_amd import precision_report as pr

pytestmark = pytest.mark.gpu

CHECKPOINTS = (500, 1000, 5000)


@pytest.fixture(scope="module")
def droplet_fp64():
    print("droplet 256^3 FP64, 2000 steps", flush=True)
    rho = pr.run_droplet(mb.FP64, 256, 2000)
    return rho, pr.droplet_metrics(rho, -5.5)


# measured: FP16 node shift 1.17e-5, BF16 node shift 2.45e-4 (FP32 1.42e-4)
@pytest.mark.parametrize("prec,tol", [(mb.FP16N, 1e-4), (mb.BF16N, 1e-3), (mb.FP32, 1e-3)])
def test_droplet_256(droplet_fp64, prec, tol):
    ref, m = droplet_fp64
    print(f"droplet 256^3 precision {prec}, 2000 steps", flush=True)
    got = pr.run_droplet(prec, 256, 2000)
    g = pr.droplet_metrics(got, -5.5)
    err = pr.rel_l2(got, ref)
    print(f"precision {prec}: rel L2 {err:.3e}, R {g['R']:.4f} vs {m['R']:.4f}, dp {g['dp']:.5e} vs "
          f"{m['dp']:.5e}, rho_l {g['rho_l']:.5f} vs {m['rho_l']:.5f}, rho_v {g['rho_v']:.5f} vs "
          f"{m['rho_v']:.5f}", flush=True)
    assert err <= tol
    assert abs(g["R"] - m["R"]) <= 0.01
    assert abs(g["dp"] - m["dp"]) <= 0.01 * abs(m["dp"])
    assert abs(g["rho_l"] / m["rho_l"] - 1.0) <= 1e-4
    assert abs(g["rho_v"] / m["rho_v"] - 1.0) <= 1e-3


def test_droplet_constant_shift_is_outside_the_bar(droplet_fp64):
    # the constant shift (0.7) cannot resolve 2.0 / 0.15 in 16 bits: the
    # reason the node shift exists and is the serving default
    ref, _ = droplet_fp64
    got = pr.run_droplet(mb.FP16S, 256, 2000)
    assert pr.rel_l2(got, ref) > 1e-2


@pytest.fixture(scope="module")
def spinodal_fp64():
    print("spinodal 512^3 FP64, 5000 steps", flush=True)
    fields, drift, _ = pr.run_spinodal(mb.FP64, 512, 5000, CHECKPOINTS)
    return fields, drift


def test_spinodal_fp64_conserves_mass(spinodal_fp64):
    _, drift = spinodal_fp64
    assert abs(drift) <= 1e-12


# measured rel L2 at 500 / 1000 steps: FP32 2.2e-4 / 5.9e-4, FP16 node
# shift 7.9e-4 / 1.5e-3, BF16 node shift 7.3e-3 / 1.7e-2; mass drift at
# 5000 steps: FP32 7.7e-5, FP16N 2.9e-6, BF16N -1.5e-6
@pytest.mark.parametrize("prec,l2_500,l2_1000,drift_tol", [
    (mb.FP32, 1e-3, 1e-3, 2e-4),
    (mb.FP16N, 1e-3, None, 1e-5),
    (mb.BF16N, None, None, 1e-5),
])
def test_spinodal_512(spinodal_fp64, prec, l2_500, l2_1000, drift_tol):
    ref, _ = spinodal_fp64
    print(f"spinodal 512^3 precision {prec}, 5000 steps", flush=True)
    got, drift, _ = pr.run_spinodal(prec, 512, 5000, CHECKPOINTS)
    errs = {c: pr.rel_l2(got[c], ref[c]) for c in CHECKPOINTS}
    rl, rv = pr.coexistence(ref[5000])
    gl, gv = pr.coexistence(got[5000])
    print(f"precision {prec}: rel L2 {errs}, liquid {gl:.5f} vs {rl:.5f}, vapour {gv:.5f} vs {rv:.5f}, "
          f"mass drift {drift:.2e}", flush=True)
    if l2_500 is not None:
        assert errs[500] <= l2_500
    if l2_1000 is not None:
        assert errs[1000] <= l2_1000
    assert abs(gl / rl - 1.0) <= 1e-4
    assert abs(gv / rv - 1.0) <= 1e-3
    assert abs(drift) <= drift_tol
