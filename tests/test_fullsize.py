"""Full-size parity: the drop-in configurations at their stated sizes and
step counts, against sha256 hashes of the reference's own run (written by
tests/golden/make_fullsize.py from the compiled reference classes).  The
GPU runs through the C ABI; the oracle run is the CPU restatement."""
import hashlib
import json
import os

import numpy as np
import pytest

import mesobasic_amd as mb
import oracle
from conftest import FULL_FIXTURES, GOLDEN


def _meta(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return json.loads(bytes(z["meta"]).decode()), z


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def _rho0(meta):
    shape = tuple(meta["shape"])
    init = meta["init"]
    return oracle.noise_rho(shape, init[1]) if init[0] == "noise" else oracle.droplet_rho(shape, init[1])


def test_oracle_full_d2q9_spinodal_matches_reference_hash():
    name = "full_d2q9_256_spinodal_10000"
    if name not in FULL_FIXTURES:
        pytest.skip("fixture not generated")
    meta, z = _meta(name)
    o = oracle.Oracle(9, tuple(meta["shape"]), meta["tau"], meta["G"], nthreads=8)
    o.set_rho(_rho0(meta)); o.init(); o.step(meta["steps"])
    fl = o.fields()
    assert _sha(fl["rho"].ravel()) == meta["sha256_rho"]
    assert _sha(fl["u"].ravel()) == meta["sha256_u"]
    assert _sha(fl["f"].ravel()) == meta["sha256_f"]


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [mb.KERNEL_TWO_PASS, mb.KERNEL_FUSED])
@pytest.mark.parametrize("name", FULL_FIXTURES)
def test_gpu_full_size_matches_reference_hash(name, kernel):
    meta, z = _meta(name)
    q, shape = meta["q"], tuple(meta["shape"])
    print(f"full-size {name} kernel {kernel}", flush=True)
    try:
        lat = mb.Lattice(q, shape, meta["tau"], meta["G"], kernel=kernel)
    except mb.LbmError:
        if kernel == mb.KERNEL_FUSED:
            pytest.skip("fused kernel does not take this lattice")
        raise
    lat.set_rho(_rho0(meta)); lat.init(); lat.step(meta["steps"])
    fl = lat.fields(mb.FIELDS_REFERENCE)
    step = 8 if q == 19 else 4
    sub = fl["rho"][::step, ::step, ::step] if q == 19 else fl["rho"][::step, ::step]
    rel = np.linalg.norm(sub - z["rho_sub"]) / np.linalg.norm(z["rho_sub"])
    print(f"  rel L2 of the subsampled density: {rel:.3e}", flush=True)
    assert _sha(fl["rho"].ravel()) == meta["sha256_rho"]
    assert _sha(fl["u"].ravel()) == meta["sha256_u"]
    assert _sha(fl["v"].ravel()) == meta["sha256_v"]
    if q == 19:
        assert _sha(fl["w"].ravel()) == meta["sha256_w"]
    assert _sha(lat.get_f().ravel()) == meta["sha256_f"]
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("block", [8, 11, 12, 13])
def test_gpu_full_d2q9_temporal_blocks_match_reference_hash(block):
    """The 10,000-step 256^2 spinodal with K steps per launch (temporal
    blocks of 16x16 tiles x 2 or 4 steps, 8x8 x 4; the default here) and
    with the single-step row march (shape 8) gives the reference's bits."""
    meta, z = _meta("full_d2q9_256_spinodal_10000")
    shape = tuple(meta["shape"])
    print(f"full-size d2q9 block {block}", flush=True)
    lat = mb.Lattice(9, shape, meta["tau"], meta["G"], kernel=mb.KERNEL_FUSED)
    lat.set_tuning(block)
    lat.set_rho(_rho0(meta)); lat.init(); lat.step(meta["steps"])
    fl = lat.fields(mb.FIELDS_REFERENCE)
    assert _sha(fl["rho"].ravel()) == meta["sha256_rho"]
    assert _sha(fl["u"].ravel()) == meta["sha256_u"]
    assert _sha(fl["v"].ravel()) == meta["sha256_v"]
    assert _sha(lat.get_f().ravel()) == meta["sha256_f"]
    lat.close()
