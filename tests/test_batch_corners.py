"""The batched D2Q9 benchmark draws tau in [0.6, 1.5] and G in [-6, -4.5];
some of those lattices end non-finite in FP16 storage on the GPU.  The FP64
oracle shows that the Shan-Chen model itself is unstable there (not the
number format): those jobs blow up in FP64 too, while stable jobs stay
finite and bounded."""
import numpy as np
import pytest

import mesobasic_amd as mb
import oracle
from mesobasic_amd import serve

# job ids of the 512-job benchmark set (serve.make_batched_jobs(512)):
# non-finite after 2,000 FP16 steps on MI355X, and two that stayed finite
UNSTABLE = [8, 11, 13, 20]
STABLE = [0, 1]


def _run(job, steps):
    o = oracle.Oracle(9, job["shape"], job["tau"], job["G"], order=2, nthreads=8)
    o.set_rho(mb.noise_rho(job["shape"], job["seed"]))
    o.init()
    blow = None
    for s in range(50, steps + 1, 50):
        o.step(50)
        r = o.fields()["rho"]
        if not np.isfinite(r).all():
            blow = s
            break
    o.close()
    return blow, r


@pytest.mark.parametrize("k", UNSTABLE)
def test_unstable_corner_blows_up_in_fp64(k):
    job = serve.make_batched_jobs(512)[k]
    assert job["tau"] < 0.9
    blow, _ = _run(job, 1000)
    assert blow is not None and blow <= 500, (k, blow)


@pytest.mark.parametrize("k", STABLE)
def test_stable_job_stays_finite_in_fp64(k):
    job = serve.make_batched_jobs(512)[k]
    blow, r = _run(job, 2000)
    assert blow is None
    assert 0.0 < r.min() and r.max() < 5.0


def test_serving_droplet_corner_blows_up_in_fp64():
    # the serving stream's D3Q19 droplets (2.0 / 0.15, width 2) at strong
    # coupling: the batch engine retires request 26 (G = -5.95) at its first
    # mass check; the FP64 oracle on the same inputs is non-finite by step 10
    reqs = serve.make_requests(200, 200.0, 7)
    r = reqs[26]
    assert r["q"] == 19 and r["init"] == "droplet" and r["G"] < -5.9
    o = oracle.Oracle(19, r["shape"], r["tau"], r["G"], nthreads=8)
    o.set_rho(serve.initial_rho(r))
    o.init()
    o.step(10)
    assert not np.isfinite(o.fields()["rho"]).all()
    o.close()
