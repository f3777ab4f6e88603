"""Failure handling of the batched / serving engine on the GPU, through the C
ABI: a poisoned initial density (NaN, Inf), a run that diverges, a
cancellation and an injected device-allocation failure each end only the
job concerned -- the engine's FP64 mass check retires a non-finite job at
the end of the tick that sees it, and every other job's result stays
bit-identical to its run alone."""
import numpy as np
import pytest

import mesobasic_amd as mb
from mesobasic_amd import lbm, serve
from mesobasic_amd.lbm import Batch
from conftest import same_bits

pytestmark = pytest.mark.gpu

NORMAL = [dict(q=9, shape=(40, 36), tau=1.0, G=-5.0, steps=30, seed=1),
          dict(q=19, shape=(12, 10, 14), tau=0.9, G=-5.5, steps=24, seed=2),
          dict(q=19, shape=(6, 9, 70), tau=1.0, G=-5.2, steps=26, seed=3),
          dict(q=9, shape=(70, 130), tau=1.2, G=-4.8, steps=18, seed=4)]


def _rho0(j):
    return j["rho0"] if "rho0" in j else mb.noise_rho(j["shape"], j["seed"])


def _run(batch, jobs, cancel_after_first_tick=()):
    ids = {}
    for k, j in enumerate(jobs):
        jid = batch.submit(j["q"], j["shape"], j["tau"], j["G"], j["steps"], _rho0(j))
        assert jid is not None
        ids[jid] = k
    out, first = {}, True
    while len(out) < len(jobs):
        batch.tick()
        if first:
            for jid, k in list(ids.items()):
                if k in cancel_after_first_tick:
                    batch.cancel(jid)
                    out[k] = None
            first = False
        for jid in batch.finished():
            out[ids[jid]] = batch.collect(jid)
    return out


def _alone(prec, j, k=6):
    return _run(Batch(precision=prec, steps_per_tick=k), [j])[0]


@pytest.mark.parametrize("prec", [mb.FP16S, mb.FP16N, mb.FP32])
def test_poisoned_jobs_fail_alone(prec):
    bad_nan = dict(q=9, shape=(32, 40), tau=1.0, G=-5.0, steps=40, seed=5)
    bad_nan["rho0"] = mb.noise_rho(bad_nan["shape"], 5)
    bad_nan["rho0"][7, 9] = np.nan
    bad_inf = dict(q=19, shape=(10, 12, 16), tau=1.0, G=-5.0, steps=40, seed=6)
    bad_inf["rho0"] = mb.noise_rho(bad_inf["shape"], 6)
    bad_inf["rho0"][3, 4, 5] = np.inf
    jobs = NORMAL[:2] + [bad_nan] + NORMAL[2:] + [bad_inf]
    print(f"poisoned jobs, precision {prec}", flush=True)
    got = _run(Batch(precision=prec, steps_per_tick=6), jobs)
    for k in (2, 5):
        r = got[k]
        assert r["failed"], k
        # caught by the first tick's mass check, retired before running on
        assert r["steps_done"] == 6 and r["steps_finite"] == 0, r
        assert "non-finite mass" in r["error"]
    for k, j in enumerate(jobs):
        if k in (2, 5):
            continue
        assert not got[k]["failed"] and got[k]["steps_done"] == j["steps"]
        alone = _alone(prec, j)
        for f in ("rho", "u", "v", "w"):
            assert same_bits(alone[f], got[k][f]), (k, f)


def test_diverging_job_is_retired_and_others_unchanged():
    # job 8 of the 512-job batched set sits in an unstable corner of the
    # model (tau 0.6-0.86): it diverges in FP64 as well, by step 500
    corner = serve.make_batched_jobs(512)[8]
    job = dict(q=9, shape=corner["shape"], tau=corner["tau"], G=corner["G"], steps=2000,
               seed=corner["seed"])
    jobs = [NORMAL[0], job, NORMAL[3]]
    got = _run(Batch(precision=mb.FP16S, steps_per_tick=20), jobs)
    r = got[1]
    print(f"diverging job {job['shape']}: failed {r['failed']} at step {r['steps_done']}, "
          f"finite to {r['steps_finite']}", flush=True)
    assert r["failed"]
    assert r["steps_done"] <= 1000 and r["steps_done"] % 20 == 0
    assert r["steps_finite"] == r["steps_done"] - 20
    for k in (0, 2):
        alone = _alone(mb.FP16S, jobs[k], 20)
        assert same_bits(alone["rho"], got[k]["rho"]) and same_bits(alone["u"], got[k]["u"]), k


def test_mass_check_tracks_the_total_mass():
    # periodic lattices conserve mass: the engine's FP64 check after a tick
    # equals the initial total to the FP32 storage rounding
    for q, shape in ((9, (64, 48)), (19, (16, 12, 20))):
        rho0 = mb.droplet_rho(shape, min(shape) / 4.0)
        b = Batch(precision=mb.FP32, steps_per_tick=10)
        jid = b.submit(q, shape, 1.0, -5.5, 40, rho0)
        b.tick()
        st = b.status(jid)
        assert st["state"] == lbm.JOB_RUNNING and st["steps_done"] == 10
        assert abs(st["mass"] / rho0.sum() - 1.0) < 1e-5, (q, st["mass"], rho0.sum())
        b.close()


def test_cancel_frees_memory_and_leaves_others_unchanged():
    prec = mb.FP16N
    jobs = [dict(NORMAL[0], steps=60), dict(NORMAL[1], steps=60), dict(NORMAL[3], steps=60)]
    b = Batch(precision=prec, steps_per_tick=6)
    need = b.job_bytes(jobs[1]["q"], jobs[1]["shape"])
    got = _run(b, jobs, cancel_after_first_tick=(1,))
    assert got[1] is None and b.bytes_in_use == 0 and b.active == 0
    for k in (0, 2):
        alone = _alone(prec, jobs[k])
        assert same_bits(alone["rho"], got[k]["rho"]), k
    # cancelling a finished, uncollected job frees it too; an unknown id is an error
    jid = b.submit(9, (16, 16), 1.0, -5.0, 6, mb.noise_rho((16, 16), 1))
    b.tick()
    assert b.finished() == [jid] and b.bytes_in_use > 0
    b.cancel(jid)
    assert b.bytes_in_use == 0 and need > 0
    with pytest.raises(mb.LbmError):
        b.cancel(jid)


def test_injected_allocation_failure():
    prec = mb.FP16N
    b = Batch(precision=prec, steps_per_tick=6)
    try:
        j0 = b.submit(9, (40, 36), 1.0, -5.0, 30, mb.noise_rho((40, 36), 1))
        used = b.bytes_in_use
        # the first, then the third device allocation of a submission fails:
        # nothing it allocated stays behind, and the next submission works
        for after in (0, 2):
            lbm.fail_alloc(after)
            assert b.submit(19, (12, 10, 14), 0.9, -5.5, 24, mb.noise_rho((12, 10, 14), 2)) is None
            assert b.bytes_in_use == used
        j1 = b.submit(19, (12, 10, 14), 0.9, -5.5, 24, mb.noise_rho((12, 10, 14), 2))
        assert j1 is not None
        # a lattice handle whose second allocation fails reports LBM_ENOMEM
        lbm.fail_alloc(1)
        with pytest.raises(mb.LbmError) as e:
            mb.Lattice(19, (8, 8, 8), 1.0, -5.0)
        assert e.value.code == lbm.ENOMEM
    finally:
        lbm.fail_alloc(-1)
    out = {}
    while len(out) < 2:
        b.tick()
        for jid in b.finished():
            out[jid] = b.collect(jid)
    assert same_bits(out[j0]["rho"], _alone(prec, NORMAL[0])["rho"])
    assert same_bits(out[j1]["rho"], _alone(prec, NORMAL[1])["rho"])


def test_serving_deadline_cancels_a_long_request():
    reqs = [dict(id=0, arrival=0.0, q=9, shape=(48, 40), steps=40, tau=1.0, G=-5.0, init="noise", seed=3),
            dict(id=1, arrival=0.0, q=19, shape=(24, 24, 24), steps=10 ** 7, tau=1.0, G=-5.0,
                 init="droplet", seed=4, deadline=0.5),
            dict(id=2, arrival=0.001, q=9, shape=(30, 64), steps=50, tau=1.0, G=-5.5, init="droplet", seed=5)]
    summary, res = serve.run_stream(reqs, precision=mb.FP16N, budget_bytes=1 << 30, steps_per_tick=10,
                                    keep_results=True)
    assert summary["outcomes"] == dict(ok=2, failed=0, cancelled=1, rejected=0)
    for r in (reqs[0], reqs[2]):
        b = Batch(precision=mb.FP16N, steps_per_tick=4)
        jid = b.submit(r["q"], r["shape"], r["tau"], r["G"], r["steps"], serve.initial_rho(r))
        while not b.finished():
            b.tick()
        assert same_bits(b.collect(jid)["rho"], res[r["id"]]["rho"]), r


def test_served_droplets_in_node_shift_storage_stay_within_1e3_of_fp64():
    # the serving default (FP16, node shift) against the bit-exact FP64 path
    reqs = []
    for k, (q, shape, steps) in enumerate([(9, (160, 140), 600), (19, (64, 72, 80), 500),
                                           (9, (128, 200), 800), (19, (70, 64, 64), 400)]):
        reqs.append(dict(id=k, arrival=0.0, q=q, shape=shape, steps=steps, tau=1.0, G=-5.5,
                         init="droplet", seed=k))
    summary, res = serve.run_stream(reqs, budget_bytes=8 << 30, keep_results=True)
    err = serve.check_against_fp64(reqs, res, [r["id"] for r in reqs])
    print("rel L2 vs FP64:", err, flush=True)
    assert max(err.values()) <= 1e-3
