"""Serving form: continuous batching of independent Shan-Chen simulations on
one GPU, on top of the batched C ABI (lbm_batch_*).

    python -m mesobasic_amd.serve [--requests 200] [--rate 200] [--seed 7]
           [--budget-gb 24] [--max-active 0] [--steps-per-tick 20]
    python -m mesobasic_amd.serve --batched [--jobs 512] [--steps 2000]

Request stream: seeded Poisson arrivals at `rate` requests/s; 70% D2Q9 with
side lengths uniform in [128, 1024], 30% D3Q19 with side lengths uniform in
[64, 192] (each side drawn on its own, so lattices need not be square);
step counts uniform in [500, 5000]; G uniform in [-6.0, -4.5]; tau 1.0;
half the requests start from a droplet (R = 3/16 of the shortest side,
width 2, 2.0 / 0.15), half from noisy-uniform rho = 0.7 +- 0.01.

Loop (host side; all lattice state stays on the device): at each tick
boundary, requests that have arrived are admitted in arrival order while
the HBM budget (and the optional cap on active jobs) allows -- a request
that does not fit waits, and later ones wait behind it, so large requests
are not starved; then one tick (steps_per_tick steps of every active job,
one graph replay) runs, and finished jobs are read back and freed, which
makes room for the next admissions.  Latency = completion - arrival, on
the host's monotonic clock; aggregate MLUPS = sum(nodes * steps) / wall
time from the first arrival to the last completion.  The command line
serves D2Q9 requests in FP32 storage and D3Q19 requests in FP16 with the
node shift, on two engines that share the budget: droplet requests
(densities 2.0 / 0.15) then stay within 1e-3 of the FP64 path.  The D3Q19
node-shift form is closer to FP64 than FP32 storage on droplets (1e-5
against 5e-5 at 128^3), the D2Q9 one is not (up to 3e-3 after 3,000 steps
at G = -5.7, against 2e-4), and the constant shift is 0.2-0.7 off.

Failure handling: the engine retires a job whose FP64 mass total goes
non-finite at the end of the tick that sees it (outcome "failed", its work
counted up to its last finite check); a request past its optional deadline
is cancelled; a request larger than the whole budget, or whose allocation
keeps failing with nothing else resident, is rejected.  None of these
change the other requests' results.

--batched runs the batched configuration instead: `jobs` D2Q9 lattices
with NX and NY uniform in {64..512}, G uniform in [-6.0, -4.5], tau uniform
in [0.6, 1.5], noisy rho = 0.7 +- 0.01, `steps` steps each, all submitted
before the clock starts, the tuner on; reports aggregate MLUPS.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from .lbm import FP16N, FP16S, PRECISIONS, Batch, droplet_rho, noise_rho


def make_requests(n, rate=200.0, seed=7):
    """The seeded request stream (list of dicts sorted by arrival time, s)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    for k in range(n):
        if rng.random() < 0.7:
            q, shape = 9, (int(rng.integers(128, 1025)), int(rng.integers(128, 1025)))
        else:
            q, shape = 19, tuple(int(v) for v in rng.integers(64, 193, size=3))
        reqs.append(dict(id=k, arrival=float(t[k]), q=q, shape=shape,
                         steps=int(rng.integers(500, 5001)), tau=1.0,
                         G=float(rng.uniform(-6.0, -4.5)),
                         init="droplet" if rng.random() < 0.5 else "noise",
                         seed=int(rng.integers(1 << 31))))
    return reqs


def tick_all(engines):
    """One tick of every engine: all are launched first (tick_begin), then
    completed (tick_end), so that their kernels share the device; an engine
    without the split interface ticks on its own.  A tick that was launched
    is always completed, even when a later launch fails."""
    begun = []
    try:
        for b in engines:
            if hasattr(b, "tick_begin"):
                b.tick_begin()
                begun.append(b)
            else:
                b.tick()
    finally:
        first = None
        for b in begun:
            try:
                b.tick_end()
            except Exception as e:   # completes the others, then reports the first
                first = first or e
        if first is not None:
            raise first


def initial_rho(req):
    if req["init"] == "droplet":
        return droplet_rho(req["shape"], 3.0 * min(req["shape"]) / 16.0)
    return noise_rho(req["shape"], req["seed"])


def run_stream(reqs, precision=FP16N, budget_bytes=24 << 30, max_active=0, steps_per_tick=20,
               work_limit=0, keep_results=False, clock=time.perf_counter, engine=None,
               alloc_retries=3, keep_ids=(), max_wait=50, split="global"):
    """Serves `reqs` in real time; returns a summary dict (and, when
    keep_results, the per-request fields of every request, else of the ids in
    keep_ids, as a second value).

    precision: the storage form of every request, or a dict {q: form} that
    chooses it by stencil; each form runs on a batch engine of its own, the
    engines share the HBM budget and each loop turn ticks every engine that
    holds jobs.  work_limit > 0: each tick advances the admitted jobs with
    the least remaining work, up to that many nodes (split evenly between
    the engines); a job left
    waiting for max_wait ticks goes first (no starvation of large requests).
    split: how the node limit is shared between the engines of several
    forms -- "global" (each turn, the live requests of all engines are
    ranked by remaining work, nodes x steps left, and each engine gets the
    nodes of its requests in the prefix that fits the limit; an engine with
    none waits that turn) or "even" (each engine gets limit / number of
    forms).
    A request may carry a "deadline" (seconds after its arrival): a request
    still queued or running past it is cancelled and its memory freed.
    Outcomes: "ok"; "failed" (the engine retired the job because its FP64
    mass total went non-finite; its work counts up to the last finite
    check); "cancelled"; "rejected" (larger than the whole budget, or its
    allocation kept failing with nothing else resident).  `engine` replaces
    the GPU batch for every request (an object with the Batch methods; tests
    pass a host-only stand-in)."""
    forms = dict(precision) if isinstance(precision, dict) else None
    engines = {}                    # form -> engine
    # the node limit per tick is shared by the engines (each turn ticks them all)
    n_forms = len(set(forms.values())) if forms is not None else 1
    limit = -(-work_limit // n_forms) if work_limit else 0

    def engine_of(r):
        key = None if engine is not None else (forms[r["q"]] if forms is not None else precision)
        if key not in engines:
            engines[key] = engine if engine is not None else Batch(
                precision=key, hbm_budget=budget_bytes, steps_per_tick=steps_per_tick, use_graph=True,
                autotune=False, work_limit=limit, max_wait=max_wait if limit else 0)
        return engines[key]

    def in_use():
        return sum(e.bytes_in_use for e in engines.values())

    pending = list(reqs)            # not yet admitted, arrival order
    live = {}                       # (engine, job id) -> request
    skipped = {}                    # engine -> turns in a row without a tick (split="global")
    outcome, latency, fail_info = {}, {}, {}
    results = {}
    keep_ids = set(keep_ids)
    work = 0
    ticks = 0
    max_live = 0
    peak_bytes = 0
    retries = 0
    t0 = clock()

    def finish(r, how, now):
        outcome[r["id"]] = how
        latency[r["id"]] = now - r["arrival"]

    while pending or live:
        now = clock() - t0
        # deadlines: queued requests past theirs are dropped, running ones cancelled
        for (b, jid), r in list(live.items()):
            if r.get("deadline") is not None and now - r["arrival"] > r["deadline"]:
                b.cancel(jid)
                del live[(b, jid)]
                finish(r, "cancelled", now)
        while pending and pending[0].get("deadline") is not None and \
                now - pending[0]["arrival"] > pending[0]["deadline"]:
            finish(pending.pop(0), "cancelled", now)
        # admissions, in arrival order, no overtaking
        while pending and pending[0]["arrival"] <= now and (max_active <= 0 or len(live) < max_active):
            r = pending[0]
            b = engine_of(r)
            need = b.job_bytes(r["q"], r["shape"], r["tau"], r["G"], r["steps"])
            if need > budget_bytes:
                finish(pending.pop(0), "rejected", now)
                continue
            # the engines share the budget
            jid = None if (len(engines) > 1 and in_use() + need > budget_bytes) else \
                b.submit(r["q"], r["shape"], r["tau"], r["G"], r["steps"], initial_rho(r))
            if jid is None:
                if live:
                    break           # waits for memory to free up
                # nothing else holds memory: an allocation failure, retried a
                # few times before the request is turned away
                retries += 1
                if retries > alloc_retries:
                    finish(pending.pop(0), "rejected", now)
                    retries = 0
                continue
            retries = 0
            live[(b, jid)] = r
            pending.pop(0)
        max_live = max(max_live, len(live))
        peak_bytes = max(peak_bytes, in_use())
        if not live:
            if pending:
                time.sleep(max(0.0, pending[0]["arrival"] - (clock() - t0)))
            continue
        busy = {b for (b, _) in live}
        if limit and len(engines) > 1 and split == "global":
            # the global shortest-remaining-work prefix, shared out by engine
            rank = []
            for (b, jid), r in live.items():
                left = r["steps"] - b.status(jid)["steps_done"]
                rank.append((int(np.prod(r["shape"])) * left, int(np.prod(r["shape"])), id(b), b))
            rank.sort(key=lambda t: (t[0], t[2]))
            share, total = {}, 0
            for _, n, _, b in rank:
                if total and total + n > work_limit:
                    break
                total += n
                share[b] = share.get(b, 0) + n
            # an engine left out max_wait turns in a row runs its smallest
            # job at least (no starvation across engines)
            for b in {b for (b, _) in live}:
                skipped[b] = 0 if b in share else skipped.get(b, 0) + 1
                if max_wait and skipped[b] >= max_wait:
                    share[b] = max(share.get(b, 0), min(n for _, n, _, e in rank if e is b))
                    skipped[b] = 0
            busy = set(share)
            for b in busy:
                b.set_work_limit(share[b])
        tick_all(busy)
        ticks += 1
        for b in {b for (b, _) in live}:
            for jid in b.finished():
                r = live.pop((b, jid))
                out = b.collect(jid)
                now = clock() - t0
                nodes = int(np.prod(r["shape"]))
                if out["failed"]:
                    finish(r, "failed", now)
                    fail_info[r["id"]] = dict(steps_done=out["steps_done"], steps_finite=out["steps_finite"])
                    work += nodes * out["steps_finite"]
                else:
                    finish(r, "ok", now)
                    work += nodes * r["steps"]
                if keep_results or r["id"] in keep_ids:
                    results[r["id"]] = out
    end = clock() - t0
    wall = end - reqs[0]["arrival"]
    ok = [r["id"] for r in reqs if outcome[r["id"]] == "ok"]
    lat = np.array([latency[i] for i in ok]) if ok else np.array([np.nan])
    counts = {k: sum(1 for v in outcome.values() if v == k) for k in ("ok", "failed", "cancelled", "rejected")}
    summary = dict(requests=len(reqs), p50_s=float(np.percentile(lat, 50)),
                   p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()),
                   aggregate_mlups=work / wall / 1e6, wall_s=wall, ticks=ticks,
                   max_active=max_live, peak_hbm_bytes=int(peak_bytes), outcomes=counts,
                   failed=fail_info)
    if engine is None:
        for b in engines.values():
            b.close()
    return (summary, results) if (keep_results or keep_ids) else summary


def check_against_fp64(reqs, results, ids, fp32_floor=False):
    """Relative L2 of the served density against the FP64 single-lattice
    path (the bit-exact drop-in) on the same inputs, for the requests `ids`
    (served ok, fields kept).  Returns {id: rel_l2}; with fp32_floor,
    {id: (rel_l2, rel_l2 of an FP32-storage run)} -- the error that FP32
    rounding alone leaves, which on spinodal (noise) requests grows with the
    instability, not with the storage format."""
    from .lbm import FIELDS_CURRENT, FP32, FP64, Lattice
    by_id = {r["id"]: r for r in reqs}

    def run(r, prec):
        lat = Lattice(r["q"], r["shape"], r["tau"], r["G"], precision=prec)
        lat.set_rho(initial_rho(r))
        lat.init()
        lat.step(r["steps"])
        rho = lat.fields(FIELDS_CURRENT)["rho"]
        lat.close()
        return rho

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    out = {}
    for i in ids:
        r = by_id[i]
        ref = run(r, FP64)
        e = rel(results[i]["rho"], ref)
        out[i] = (e, rel(run(r, FP32), ref)) if fp32_floor else e
    return out


def sample_ids(reqs, k, seed=5, max_work=2e10):
    """k requests drawn (seeded) from `reqs`, half of each init type where
    possible, each of at most max_work lattice updates (so the FP64 check
    stays short)."""
    rng = np.random.default_rng(seed)
    picks = []
    for init in ("droplet", "noise"):
        pool = [r["id"] for r in reqs
                if r["init"] == init and np.prod(r["shape"]) * r["steps"] <= max_work]
        if pool:
            picks += [int(v) for v in rng.choice(pool, size=min(len(pool), (k + 1) // 2), replace=False)]
    return sorted(picks[:k])


def make_batched_jobs(n=512, seed=11):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        jobs.append(dict(id=k, q=9, shape=(int(rng.integers(64, 513)), int(rng.integers(64, 513))),
                         tau=float(rng.uniform(0.6, 1.5)), G=float(rng.uniform(-6.0, -4.5)),
                         seed=int(rng.integers(1 << 31))))
    return jobs


def run_batched(jobs, steps=2000, precision=FP16S, steps_per_tick=20, autotune=True,
                clock=time.perf_counter, launch=None):
    """All jobs submitted, then ticked to completion; returns a summary.
    Jobs the engine retires as failed (non-finite mass) count their work up
    to their last finite mass check only."""
    b = Batch(precision=precision, steps_per_tick=steps_per_tick, use_graph=True,
              autotune=autotune and launch is None)
    if launch is not None:
        b.set_launch(*launch)   # (d2q9_chunk, d2q9_rows_kernel, d3q19_march); the tuner stays off
    ids = []
    for j in jobs:
        ids.append(b.submit(9, j["shape"], j["tau"], j["G"], steps, noise_rho(j["shape"], j["seed"])))
    assert all(i is not None for i in ids)
    nodes = sum(int(np.prod(j["shape"])) for j in jobs)
    t0 = clock()
    n_ticks = 0
    while b.active:
        b.tick()
        n_ticks += 1
        if len(b.finished()) == len(jobs):
            break
    wall = clock() - t0
    index = {jid: k for k, jid in enumerate(ids)}
    work = 0
    failed = {}
    nonfinite_ok = []
    for jid in b.finished():
        k = index[jid]
        out = b.collect(jid)
        n = int(np.prod(jobs[k]["shape"]))
        if out["failed"]:
            failed[k] = out["steps_done"]
            work += n * out["steps_finite"]
        else:
            work += n * steps
            if not np.isfinite(out["rho"]).all():
                nonfinite_ok.append(k)
    summary = dict(jobs=len(jobs), steps=steps, nodes=nodes, wall_s=wall, ticks=n_ticks,
                   aggregate_mlups=work / wall / 1e6, tuning=b.tuning,
                   completed=len(jobs) - len(failed), failed_jobs=sorted(failed),
                   failed_at_step={str(k): v for k, v in sorted(failed.items())},
                   nonfinite_completed=nonfinite_ok)
    b.close()
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--budget-gb", type=float, default=24.0)
    ap.add_argument("--max-active", type=int, default=0)
    ap.add_argument("--steps-per-tick", type=int, default=20)
    ap.add_argument("--work-limit", type=int, default=0,
                    help="nodes advanced per step, least remaining work first (0: all)")
    ap.add_argument("--precision", default=None,
                    help="storage form, fp32 / fp16 / bf16 / fp16n / bf16n, or FORM2D,FORM3D for the "
                         "stream's D2Q9 and D3Q19 requests (default: fp32,fp16n for the stream, fp16 for "
                         "--batched)")
    ap.add_argument("--d2q9-kernel", default="auto", choices=["auto", "rows", "two_pass", "blocks"],
                    help="--batched: D2Q9 kernel (auto: the tuner picks)")
    args = ap.parse_args()
    launch = {"auto": None, "rows": (32, 1, 1), "two_pass": (32, 0, 1), "blocks": (32, 2, 1)}[args.d2q9_kernel]
    spec = args.precision or ("fp16" if args.batched else "fp32,fp16n")
    if "," in spec and not args.batched:
        p2, p3 = spec.split(",")
        prec = {9: PRECISIONS[p2], 19: PRECISIONS[p3]}
    else:
        prec = PRECISIONS[spec]
    if args.batched:
        s = run_batched(make_batched_jobs(args.jobs), args.steps, prec, args.steps_per_tick, launch=launch)
    else:
        s = run_stream(make_requests(args.requests, args.rate, args.seed), prec,
                       int(args.budget_gb * (1 << 30)), args.max_active, args.steps_per_tick,
                       args.work_limit)
    print(json.dumps(s), flush=True)


if __name__ == "__main__":
    main()
