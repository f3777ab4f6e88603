"""Batched / serving form: many jobs of different sizes, stencils and
parameters in one batch.  Checks that a job's result does not depend on the
other jobs, on graph capture, on the launch configuration the tuner picks,
or on where a tick boundary falls; that it stops exactly at its step count;
that admission respects the HBM budget; and that results agree with the
FP64 oracle to the 16-bit storage tolerance."""
import numpy as np
import pytest

import mesobasic_amd as mb
from mesobasic_amd.lbm import Batch
import oracle
from conftest import rel_l2, same_bits

pytestmark = pytest.mark.gpu


def _jobs(rng, n, q9_share=0.7):
    out = []
    for _ in range(n):
        if rng.random() < q9_share:
            shape = (int(rng.integers(20, 140)), int(rng.integers(20, 140)))
            q = 9
        else:
            shape = tuple(int(v) for v in rng.integers(8, 28, 3))
            q = 19
        out.append(dict(q=q, shape=shape, tau=float(rng.uniform(0.6, 1.5)), G=float(rng.uniform(-6.0, -4.5)),
                        steps=int(rng.integers(5, 60)), seed=int(rng.integers(1 << 30))))
    return out


def _run(batch, jobs):
    ids = {}
    for k, j in enumerate(jobs):
        rho0 = mb.noise_rho(j["shape"], j["seed"])
        jid = batch.submit(j["q"], j["shape"], j["tau"], j["G"], j["steps"], rho0)
        assert jid is not None
        ids[jid] = k
    out = {}
    while len(out) < len(jobs):
        batch.tick()
        for jid in batch.finished():
            out[ids[jid]] = batch.collect(jid)
    return out


@pytest.mark.parametrize("prec", [mb.FP16S, mb.FP16N])
def test_alone_equals_in_the_mix_bitwise(prec):
    rng = np.random.default_rng(5)
    jobs = _jobs(rng, 24)
    mixed = _run(Batch(precision=prec, steps_per_tick=8), jobs)
    for k in (0, 3, 7, 11, 19):
        alone = _run(Batch(precision=prec, steps_per_tick=6), [jobs[k]])[0]
        for f in ("rho", "u", "v", "w"):
            assert same_bits(alone[f], mixed[k][f]), (k, f, jobs[k])


@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16S, mb.BF16S, mb.FP16N, mb.BF16N])
def test_graph_tuner_and_tick_length_do_not_change_results(prec):
    rng = np.random.default_rng(6)
    jobs = _jobs(rng, 10)
    a = _run(Batch(precision=prec, steps_per_tick=4, use_graph=False), jobs)
    b = _run(Batch(precision=prec, steps_per_tick=10, use_graph=True, autotune=True), jobs)
    for k in range(len(jobs)):
        assert same_bits(a[k]["rho"], b[k]["rho"]), (k, jobs[k])
        assert same_bits(a[k]["u"], b[k]["u"]), (k, jobs[k])


@pytest.mark.parametrize("prec,launch", [(mb.FP16S, None), (mb.FP16N, None), (mb.FP16S, (16, 2, 1)),
                                         (mb.FP32, (16, 0, 2))])
def test_concurrent_classes_equal_one_stream_bitwise(prec, launch):
    # kernel classes on streams of their own (the default) or one after
    # another on the batch's stream, replayed as graphs or launched directly:
    # the same bits.  The mix spans D2Q9 row widths 64..1024, two-pass D2Q9
    # and D3Q19 jobs, D3Q19 march jobs and (launch (16, 2, 1)) temporal blocks.
    rng = np.random.default_rng(11)
    jobs = _jobs(rng, 12)
    for shape in ((40, 700), (24, 1500), (6, 16, 64), (5, 9, 70)):
        jobs.append(dict(q=19 if len(shape) == 3 else 9, shape=shape, tau=1.0, G=-5.0,
                         steps=int(rng.integers(9, 45)), seed=int(rng.integers(1 << 30))))
    runs = []
    for conc, graph in ((True, True), (False, True), (True, False)):
        b = Batch(precision=prec, steps_per_tick=6, use_graph=graph, concurrent=conc)
        if launch is not None:
            b.set_launch(*launch)
        runs.append(_run(b, jobs))
        b.close()
    for k in range(len(jobs)):
        for f in ("rho", "u", "v", "w"):
            assert same_bits(runs[0][k][f], runs[1][k][f]), (k, f, jobs[k])
            assert same_bits(runs[0][k][f], runs[2][k][f]), (k, f, jobs[k])


def test_overlapped_ticks_of_two_batches_equal_their_own_ticks_bitwise():
    # two batches of different forms ticked together (both launched with
    # tick_begin, then both completed with tick_end) give the bits each gives
    # ticked alone; between begin and end a batch refuses changes
    rng = np.random.default_rng(21)
    jobs_a, jobs_b = _jobs(rng, 8), _jobs(rng, 8, q9_share=0.3)
    alone_a = _run(Batch(precision=mb.FP32, steps_per_tick=6), jobs_a)
    alone_b = _run(Batch(precision=mb.FP16N, steps_per_tick=6), jobs_b)
    a, b = Batch(precision=mb.FP32, steps_per_tick=6), Batch(precision=mb.FP16N, steps_per_tick=6)
    ids = {}
    for batch, jobs, tag in ((a, jobs_a, "a"), (b, jobs_b, "b")):
        for k, j in enumerate(jobs):
            jid = batch.submit(j["q"], j["shape"], j["tau"], j["G"], j["steps"], mb.noise_rho(j["shape"], j["seed"]))
            ids[(tag, jid)] = k
    out = {"a": {}, "b": {}}
    checked = False
    while len(out["a"]) < len(jobs_a) or len(out["b"]) < len(jobs_b):
        a.tick_begin()
        b.tick_begin()
        if not checked:
            j = jobs_a[0]
            for call in (lambda: a.submit(j["q"], j["shape"], j["tau"], j["G"], 4, mb.noise_rho(j["shape"], 1)),
                         lambda: a.set_work_limit(100), lambda: a.tick_begin()):
                with pytest.raises(mb.lbm.LbmError):
                    call()
            checked = True
        a.tick_end()
        b.tick_end()
        for batch, tag in ((a, "a"), (b, "b")):
            for jid in batch.finished():
                out[tag][ids[(tag, jid)]] = batch.collect(jid)
    with pytest.raises(mb.lbm.LbmError):
        a.tick_end()            # nothing in flight
    a.close()
    b.close()
    for got, ref, jobs in ((out["a"], alone_a, jobs_a), (out["b"], alone_b, jobs_b)):
        for k in range(len(jobs)):
            for f in ("rho", "u", "v", "w"):
                assert same_bits(got[k][f], ref[k][f]), (k, f, jobs[k])


@pytest.mark.parametrize("q,shape,steps", [(9, (64, 48), 37), (9, (40, 200), 20), (19, (20, 18, 16), 23)])
def test_job_matches_fp64_oracle(q, shape, steps):
    tau, G = 0.9, -5.0
    rho0 = mb.noise_rho(shape, 3)
    o = oracle.Oracle(q, shape, tau, G)
    o.set_rho(rho0); o.init(); o.step(steps)
    ref = o.moments()
    for prec, tol in ((mb.FP32, 2e-6), (mb.FP16S, 1e-4), (mb.FP16N, 1e-4), (mb.BF16N, 5e-4)):
        res = _run(Batch(precision=prec, steps_per_tick=6), [dict(q=q, shape=shape, tau=tau, G=G,
                                                                  steps=steps, seed=3)])[0]
        err = rel_l2(res["rho"], ref["rho"])
        print(f"q {q} {shape} prec {prec}: rel L2 rho {err:.2e}", flush=True)
        assert err < tol


def test_admission_by_budget():
    b = Batch(precision=mb.FP16S, steps_per_tick=4, hbm_budget=1)
    assert b.submit(9, (32, 32), 1.0, -5.0, 4, np.full((32, 32), 0.7)) is None
    need = Batch(precision=mb.FP16S).job_bytes(9, (32, 32))
    b = Batch(precision=mb.FP16S, steps_per_tick=4, hbm_budget=int(2.5 * need))
    ids = [b.submit(9, (32, 32), 1.0, -5.0, 4, mb.noise_rho((32, 32), s)) for s in range(3)]
    assert ids[0] is not None and ids[1] is not None and ids[2] is None
    assert b.bytes_in_use == 2 * need
    b.tick()
    for jid in b.finished():
        b.collect(jid)
    assert b.bytes_in_use == 0
    assert b.submit(9, (32, 32), 1.0, -5.0, 4, mb.noise_rho((32, 32), 2)) is not None


@pytest.mark.parametrize("prec", [mb.FP16S, mb.FP16N])
def test_row_items_of_every_length_give_the_same_bits(prec):
    # D2Q9 row items hold at most the launch's chunk of x rows, fewer where
    # a row-width class holds too few rows in all to fill two rounds of
    # workgroups (the small mixes of the other tests get 2).  Here the class of rows of
    # up to 128 nodes holds ~77,000 rows, so caps 1, 5 and 64 give items of
    # 1, 5 and ~18 rows, with ragged last items (random nx); the node-shift
    # form also writes each row's reference density from exactly one item.
    rng = np.random.default_rng(9)
    jobs = [dict(q=9, shape=(int(rng.integers(400, 700)), 100), tau=float(rng.uniform(0.8, 1.3)),
                 G=float(rng.uniform(-5.5, -4.5)), steps=7, seed=int(rng.integers(1 << 30)))
            for _ in range(140)]
    jobs += [dict(q=9, shape=(300, 50), tau=1.0, G=-5.0, steps=7, seed=3),
             dict(q=9, shape=(90, 300), tau=0.9, G=-5.2, steps=7, seed=4)]
    runs = []
    for cap in (1, 5, 64):
        b = Batch(precision=prec, steps_per_tick=8)
        b.set_launch(cap, 1, 1)
        assert b.tuning["d2q9_rows_kernel"] == 1
        runs.append(_run(b, jobs))
    for k in range(len(jobs)):
        for r in runs[1:]:
            assert same_bits(runs[0][k]["rho"], r[k]["rho"]), (k, jobs[k])
            assert same_bits(runs[0][k]["u"], r[k]["u"]), (k, jobs[k])


def test_large_rows_take_the_two_pass_path():
    # rows wider than the largest row-kernel workgroup (1024) go to the
    # two-pass kernels; both paths must agree with the oracle
    shape = (12, 1100)
    rho0 = mb.noise_rho(shape, 8)
    o = oracle.Oracle(9, shape, 1.0, -5.0)
    o.set_rho(rho0); o.init(); o.step(10)
    res = _run(Batch(precision=mb.FP32, steps_per_tick=4), [dict(q=9, shape=shape, tau=1.0, G=-5.0,
                                                                steps=10, seed=8)])[0]
    assert rel_l2(res["rho"], o.moments()["rho"]) < 2e-6


@pytest.mark.parametrize("work_limit,forms", [(0, None), (3000, None), (3000, {9: mb.FP32, 19: mb.FP16N})])
def test_serving_stream_admits_as_memory_frees_and_matches_alone(work_limit, forms):
    # forms: one storage form for every request (FP16 constant shift), or one
    # per stencil, each on an engine of its own sharing the budget
    from mesobasic_amd import serve
    rng = np.random.default_rng(9)
    reqs = []
    for k in range(14):
        q = 9 if k % 3 else 19
        shape = (int(rng.integers(24, 90)), int(rng.integers(24, 90))) if q == 9 else \
            tuple(int(v) for v in rng.integers(10, 24, 3))
        reqs.append(dict(id=k, arrival=0.004 * max(0, k - 3), q=q, shape=shape, steps=int(rng.integers(20, 90)),
                         tau=1.0, G=float(rng.uniform(-6.0, -4.5)),
                         init="droplet" if k % 2 else "noise", seed=100 + k))
    form = (lambda r: forms[r["q"]]) if forms else (lambda r: mb.FP16S)
    need = max(Batch(precision=form(r)).job_bytes(r["q"], r["shape"]) for r in reqs)
    summary, res = serve.run_stream(reqs, precision=forms or mb.FP16S, budget_bytes=3 * need, steps_per_tick=8,
                                    work_limit=work_limit, keep_results=True)
    print(summary, flush=True)
    # the first four arrive together: at least three fit the budget at once
    assert summary["requests"] == len(reqs) and summary["max_active"] >= 3
    assert summary["peak_hbm_bytes"] <= 3 * need
    assert summary["p50_s"] <= summary["p99_s"]
    for r in reqs:
        b = Batch(precision=form(r), steps_per_tick=4)
        jid = b.submit(r["q"], r["shape"], r["tau"], r["G"], r["steps"], serve.initial_rho(r))
        while not b.finished():
            b.tick()
        alone = b.collect(jid)
        assert same_bits(alone["rho"], res[r["id"]]["rho"]), r
        assert same_bits(alone["u"], res[r["id"]]["u"]), r


@pytest.mark.parametrize("q,shape", [(9, (96, 80)), (19, (24, 24, 32))])
def test_node_shift_resolves_droplet_jobs(q, shape):
    # a droplet (2.0 / 0.15) is far from any single shift: the node-shifted
    # batch forms stay close to the FP64 oracle, the constant shift does not
    steps, tau, G = 200, 1.0, -5.5
    rho0 = mb.droplet_rho(shape, min(shape) * 3.0 / 16.0)
    o = oracle.Oracle(q, shape, tau, G)
    o.set_rho(rho0); o.init(); o.step(steps)
    ref = o.moments()["rho"]
    err = {}
    for prec in (mb.FP16S, mb.FP16N, mb.BF16N):
        b = Batch(precision=prec, steps_per_tick=8)
        jid = b.submit(q, shape, tau, G, steps, rho0)
        while not b.finished():
            b.tick()
        err[prec] = rel_l2(b.collect(jid)["rho"], ref)
    print(f"q {q} {shape}: rel L2 FP16S {err[mb.FP16S]:.2e} FP16N {err[mb.FP16N]:.2e} "
          f"BF16N {err[mb.BF16N]:.2e}", flush=True)
    assert err[mb.FP16N] < 3e-4 and err[mb.BF16N] < 3e-3
    assert err[mb.FP16N] < 0.1 * err[mb.FP16S]


# D3Q19 jobs with nx >= 2, ny >= 8 and nz >= 64 take the fused march kernel
# (8 x 64 tiles); sides that are not multiples of the tile exercise the
# wrapped partial tiles
MARCH_SHAPES = [(9, 20, 100), (5, 8, 64), (12, 33, 70), (3, 9, 129), (6, 16, 128)]
MARCH_TOL = {mb.FP32: 2e-6, mb.FP16S: 1e-4, mb.BF16S: 1e-3, mb.FP16N: 1e-4, mb.BF16N: 5e-4}


@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16S, mb.BF16S, mb.FP16N, mb.BF16N])
def test_d3q19_march_agrees_with_two_pass_and_oracle(prec):
    # the two D3Q19 kernels differ in FP32 rounding only: both within the
    # storage tolerance of the FP64 oracle, and of each other
    jobs = [dict(q=19, shape=sh, tau=t, G=G, steps=st, seed=11 + k)
            for k, (sh, t, G, st) in enumerate(zip(MARCH_SHAPES, (1.0, 0.7, 1.3, 0.9, 1.0),
                                                   (-5.0, -5.5, -4.8, -6.0, -5.2), (9, 14, 7, 12, 10)))]
    runs = []
    for march in (0, 2):   # two-pass; the march for every job that fits
        b = Batch(precision=prec, steps_per_tick=4)
        b.set_launch(32, 1, march)
        assert b.tuning["d3q19_march"] == march
        runs.append(_run(b, jobs))
    for k, j in enumerate(jobs):
        o = oracle.Oracle(19, j["shape"], j["tau"], j["G"])
        o.set_rho(mb.noise_rho(j["shape"], j["seed"])); o.init(); o.step(j["steps"])
        ref = o.moments()["rho"]
        e_two, e_march = rel_l2(runs[0][k]["rho"], ref), rel_l2(runs[1][k]["rho"], ref)
        print(f"prec {prec} job {j['shape']}: two-pass {e_two:.2e} march {e_march:.2e}", flush=True)
        assert e_march < MARCH_TOL[prec] and e_two < MARCH_TOL[prec]
        assert rel_l2(runs[1][k]["rho"], runs[0][k]["rho"]) < MARCH_TOL[prec]


def test_d3q19_march_default_alone_equals_in_the_mix():
    # by default node-shift batches run D3Q19 jobs that fill 90% of the
    # march tiles on the march ((6, 16, 128)), the others on the two passes
    # ((10, 12, 100) fills 59%); a job's bits do not depend on the other jobs
    # or on the tuner
    for prec, shape in ((mb.FP16N, (10, 12, 100)), (mb.BF16N, (10, 12, 100)), (mb.FP16N, (6, 16, 128)),
                        (mb.FP16S, (10, 16, 64))):
        job = dict(q=19, shape=shape, tau=0.9, G=-5.0, steps=25, seed=4)
        b = Batch(precision=prec, steps_per_tick=6)
        assert b.tuning["d3q19_march"] == 1
        alone = _run(b, [job])[0]
        rng = np.random.default_rng(8)
        mix = _jobs(rng, 6) + [job] + [dict(q=19, shape=(6, 9, 80), tau=1.1, G=-5.3, steps=17, seed=2)]
        got = _run(Batch(precision=prec, steps_per_tick=4, autotune=True), mix)[6]
        for f in ("rho", "u", "v", "w"):
            assert same_bits(alone[f], got[f]), (prec, shape, f)


@pytest.mark.parametrize("max_wait", [0, 3])
def test_wait_bound_keeps_a_large_job_from_starving(max_wait):
    # shortest remaining work first under a node limit: a steady supply of
    # small jobs (one per tick, below the limit's capacity of 1.5 per tick)
    # keeps a large one waiting for ever, unless waiting is bounded
    big = dict(q=9, shape=(64, 64), tau=1.0, G=-5.0, steps=24, seed=3)
    b = Batch(precision=mb.FP16S, steps_per_tick=4, work_limit=3 * 16 * 16, max_wait=max_wait)
    big_id = b.submit(9, big["shape"], 1.0, -5.0, big["steps"], mb.noise_rho(big["shape"], 3))
    got, seed = None, 100
    for tick in range(120):
        b.submit(9, (16, 16), 1.0, -5.0, 8, mb.noise_rho((16, 16), seed))
        seed += 1
        b.tick()
        for jid in b.finished():
            out = b.collect(jid)
            if jid == big_id:
                got = (tick, out)
        if got:
            break
    print(f"max_wait {max_wait}: large job done at tick {got[0] if got else None}", flush=True)
    if max_wait == 0:
        assert got is None and b.status(big_id)["steps_done"] == 0
    else:
        assert got is not None and got[0] < 60
        alone = Batch(precision=mb.FP16S, steps_per_tick=6)
        jid = alone.submit(9, big["shape"], 1.0, -5.0, big["steps"], mb.noise_rho(big["shape"], 3))
        while not alone.finished():
            alone.tick()
        assert same_bits(alone.collect(jid)["rho"], got[1]["rho"])


BLOCK_JOBS = [dict(q=9, shape=(64, 64), tau=1.0, G=-5.0, steps=13, seed=41),
              dict(q=9, shape=(33, 70), tau=0.7, G=-5.5, steps=10, seed=42),
              dict(q=9, shape=(100, 37), tau=1.3, G=-4.8, steps=7, seed=43),
              dict(q=9, shape=(20, 20), tau=0.9, G=-5.2, steps=29, seed=44),    # smaller than a tile
              dict(q=9, shape=(3, 5), tau=1.1, G=-5.0, steps=5, seed=45),       # smaller than the halo
              dict(q=9, shape=(97, 130), tau=0.8, G=-5.0, steps=44, seed=46),
              dict(q=19, shape=(12, 10, 16), tau=1.0, G=-5.5, steps=9, seed=47)]


@pytest.mark.parametrize("prec", [mb.FP32, mb.FP16S, mb.BF16S])
@pytest.mark.parametrize("spt", [4, 20])
def test_temporal_blocks_equal_the_row_kernel_bitwise(prec, spt):
    # D2Q9 jobs in temporal blocks (3 steps per launch, 1 when fewer are
    # left) against the row kernel: ragged sizes, lattices smaller than a
    # tile and than the halo, odd and even step counts, the LBfluid2D-order
    # first step inside a block, a D3Q19 job alongside; tick lengths 4 and 20
    rows = Batch(precision=prec, steps_per_tick=spt)
    rows.set_launch(32, 1, 1)
    blocks = Batch(precision=prec, steps_per_tick=spt)
    blocks.set_launch(32, 2, 1)
    assert blocks.tuning["d2q9_rows_kernel"] == 2
    print(f"blocks vs rows precision {prec} steps/tick {spt}", flush=True)
    a = _run(rows, BLOCK_JOBS)
    b = _run(blocks, BLOCK_JOBS)
    for k in range(len(BLOCK_JOBS)):
        for f in ("rho", "u", "v", "w"):
            assert same_bits(a[k][f], b[k][f]), (k, f, BLOCK_JOBS[k])
        assert a[k]["steps_done"] == b[k]["steps_done"] == BLOCK_JOBS[k]["steps"]


def test_temporal_blocks_read_extend_and_node_shift_fallback():
    # a block job stopped at an odd count, read, extended, equals one run
    # through; the node-shift forms keep the row kernel under the same setting
    shape, tau, G = (48, 40), 0.9, -5.5
    rho0 = mb.noise_rho(shape, 9)
    ref = Batch(precision=mb.FP16S, steps_per_tick=6)
    ref.set_launch(32, 1, 1)
    want = _run(ref, [dict(q=9, shape=shape, tau=tau, G=G, steps=17, seed=9)])[0]
    b = Batch(precision=mb.FP16S, steps_per_tick=6)
    b.set_launch(32, 2, 1)
    jid = b.submit(9, shape, tau, G, 7, rho0)
    while jid not in b.finished():
        b.tick()
    assert b.status(jid)["steps_done"] == 7
    b.extend(jid, 10)
    while jid not in b.finished():
        b.tick()
    got = b.collect(jid)
    for f in ("rho", "u", "v"):
        assert same_bits(got[f], want[f]), f
    for prec in (mb.FP16N, mb.BF16N):
        r = Batch(precision=prec, steps_per_tick=6)
        r.set_launch(32, 1, 1)
        n = Batch(precision=prec, steps_per_tick=6)
        n.set_launch(32, 2, 1)
        x = _run(r, BLOCK_JOBS[:3])
        y = _run(n, BLOCK_JOBS[:3])
        for k in range(3):
            assert same_bits(x[k]["rho"], y[k]["rho"]), (prec, k)
