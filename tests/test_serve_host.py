"""Host-side logic of the serving loop (mesobasic_amd.serve.run_stream),
driven by a stand-in engine with the Batch interface, so it runs without a
GPU: requests larger than the budget are rejected, a transient allocation
failure is retried, a job the engine retires as failed is reported as failed
with its work counted up to its last finite check, a request past its
deadline is cancelled (queued or running), and none of this changes what the
other requests get."""
import numpy as np

from mesobasic_amd import serve


class FakeBatch:
    """Stand-in engine: a job's 'result' is a function of its request only;
    jobs with G == -99 go non-finite after `bad_after` steps."""

    def __init__(self, budget, steps_per_tick=10, fail_submits=0, bad_after=25):
        self.budget, self.k = budget, steps_per_tick
        self.fail_submits = fail_submits
        self.bad_after = bad_after
        self.jobs, self.next_id, self.in_use = {}, 1, 0
        self.log = []

    def job_bytes(self, q, shape, tau=1.0, G=-5.0, steps=1):
        return int(np.prod(shape)) * 10

    def submit(self, q, shape, tau, G, steps, rho0):
        need = self.job_bytes(q, shape)
        if self.fail_submits > 0:
            self.fail_submits -= 1
            self.log.append("alloc-fail")
            return None
        if self.in_use + need > self.budget:
            return None
        jid = self.next_id
        self.next_id += 1
        self.in_use += need
        self.jobs[jid] = dict(need=need, steps=steps, done=0, finite=0, failed=False, G=G,
                              rho=np.asarray(rho0, dtype=float) * 2.0)
        return jid

    @property
    def bytes_in_use(self):
        return self.in_use

    def tick(self):
        for j in self.jobs.values():
            if j["done"] >= j["steps"] or j["failed"]:
                continue
            j["done"] = min(j["steps"], j["done"] + self.k)
            if j["G"] == -99 and j["done"] > self.bad_after:
                j["failed"] = True
            else:
                j["finite"] = j["done"]

    def finished(self):
        return [i for i, j in self.jobs.items() if j["failed"] or j["done"] >= j["steps"]]

    def collect(self, jid):
        j = self.jobs.pop(jid)
        self.in_use -= j["need"]
        rho = np.full_like(j["rho"], np.nan) if j["failed"] else j["rho"]
        return dict(rho=rho, failed=j["failed"], steps_done=j["done"], steps_finite=j["finite"])

    def cancel(self, jid):
        j = self.jobs.pop(jid)
        self.in_use -= j["need"]
        self.log.append(f"cancel {jid}")


def _req(k, arrival, shape=(8, 8), steps=40, G=-5.0, deadline=None):
    return dict(id=k, arrival=arrival, q=9, shape=shape, steps=steps, tau=1.0, G=G, init="noise",
                seed=10 + k, deadline=deadline)


def test_outcomes_and_unaffected_results():
    reqs = [_req(0, 0.0), _req(1, 0.0, G=-99), _req(2, 0.0, shape=(100, 100)),
            _req(3, 0.001, steps=10 ** 7, deadline=0.05), _req(4, 0.002), _req(5, 0.003, shape=(9, 7))]
    eng = FakeBatch(budget=4000)
    summary, res = serve.run_stream(reqs, budget_bytes=4000, engine=eng, keep_results=True)
    assert summary["outcomes"] == dict(ok=3, failed=1, cancelled=1, rejected=1)
    # the failed job counts its work up to its last finite check only
    assert summary["failed"] == {1: dict(steps_done=30, steps_finite=20)}
    assert res[1]["failed"] and not np.isfinite(res[1]["rho"]).any()
    for k in (0, 4, 5):
        assert not res[k]["failed"]
        assert np.array_equal(res[k]["rho"], 2.0 * serve.initial_rho(reqs[k]))
    assert any(e.startswith("cancel") for e in eng.log)
    assert eng.in_use == 0 and not eng.jobs


def test_transient_allocation_failure_is_retried():
    reqs = [_req(0, 0.0), _req(1, 0.0)]
    eng = FakeBatch(budget=10 ** 6, fail_submits=2)
    summary, res = serve.run_stream(reqs, budget_bytes=10 ** 6, engine=eng, keep_results=True)
    assert eng.log.count("alloc-fail") == 2
    assert summary["outcomes"]["ok"] == 2
    assert np.array_equal(res[1]["rho"], 2.0 * serve.initial_rho(reqs[1]))


def test_persistent_allocation_failure_rejects_only_that_request():
    reqs = [_req(0, 0.0), _req(1, 0.0)]
    eng = FakeBatch(budget=10 ** 6, fail_submits=4)
    summary, res = serve.run_stream(reqs, budget_bytes=10 ** 6, engine=eng, keep_results=True,
                                    alloc_retries=3)
    assert summary["outcomes"] == dict(ok=1, failed=0, cancelled=0, rejected=1)
    assert np.array_equal(res[1]["rho"], 2.0 * serve.initial_rho(reqs[1]))


def test_queued_request_past_its_deadline_is_dropped():
    # request 1 cannot be admitted while request 0 holds the budget, and its
    # deadline passes while it waits
    reqs = [_req(0, 0.0, shape=(20, 20), steps=10 ** 9), _req(1, 0.0, shape=(20, 20), deadline=0.0),
            _req(2, 0.0, shape=(4, 4))]
    reqs[0]["deadline"] = 0.2
    eng = FakeBatch(budget=4500)
    summary = serve.run_stream(reqs, budget_bytes=4500, engine=eng)
    assert summary["outcomes"] == dict(ok=1, failed=0, cancelled=2, rejected=0)


def test_sample_ids_takes_both_init_types():
    reqs = serve.make_requests(60, seed=3)
    ids = serve.sample_ids(reqs, 6, max_work=1e12)
    inits = {reqs[i]["init"] for i in ids}
    assert len(ids) == 6 and inits == {"droplet", "noise"}


class FakeLimitedBatch(FakeBatch):
    """Stand-in engine with the work limit of the real one: a tick advances
    the unfinished jobs in order of remaining work (nodes x steps left), the
    prefix that fits the limit (at least one job)."""

    engines = []

    def __init__(self, precision=None, hbm_budget=0, steps_per_tick=10, work_limit=0, max_wait=0, **_):
        super().__init__(hbm_budget, steps_per_tick=steps_per_tick)
        self.form, self.limit, self.ticks = precision, work_limit, []
        FakeLimitedBatch.engines.append(self)

    def set_work_limit(self, n):
        self.limit = n

    def status(self, jid):
        return dict(steps_done=self.jobs[jid]["done"])

    def tick(self):
        live = [(int(j["rho"].size) * (j["steps"] - j["done"]), i) for i, j in self.jobs.items()
                if j["done"] < j["steps"]]
        live.sort()
        picked, nodes = [], 0
        for _, i in live:
            n = int(self.jobs[i]["rho"].size)
            if picked and self.limit and nodes + n > self.limit:
                break
            picked.append(i)
            nodes += n
        for i in picked:
            j = self.jobs[i]
            j["done"] = min(j["steps"], j["done"] + self.k)
            j["finite"] = j["done"]
        self.ticks.append((tuple(picked), nodes))

    def close(self):
        pass


def test_global_split_shares_the_limit_and_starves_no_engine(monkeypatch):
    # one large D3Q19 request among a stream of small D2Q9 ones: under the
    # global split the small ones take the node limit, yet the large one
    # still advances at least once every max_wait + 1 turns
    FakeLimitedBatch.engines = []
    monkeypatch.setattr(serve, "Batch", FakeLimitedBatch)
    reqs = [dict(id=0, arrival=0.0, q=19, shape=(8, 8, 8), steps=200, tau=1.0, G=-5.0, init="noise", seed=1)]
    reqs += [dict(id=k, arrival=0.0, q=9, shape=(8, 8), steps=40, tau=1.0, G=-5.0, init="noise", seed=k)
             for k in range(1, 40)]
    summary, res = serve.run_stream(reqs, precision={9: "A", 19: "B"}, budget_bytes=10 ** 9,
                                    steps_per_tick=10, work_limit=4 * 64, max_wait=3, keep_results=True)
    assert summary["outcomes"]["ok"] == len(reqs)
    by_form = {e.form: e for e in FakeLimitedBatch.engines}
    a, b = by_form["A"], by_form["B"]
    # the small jobs went first: A ran every one of its ticks at the limit
    # while B was held back, and B was never left out more than 3 turns
    assert all(n <= 4 * 64 for _, n in a.ticks)
    assert summary["ticks"] > len(b.ticks)
    assert len(b.ticks) >= summary["ticks"] // 4
    # results are the requests' own (the stand-in's result depends on nothing else)
    for r in reqs:
        np.testing.assert_array_equal(res[r["id"]]["rho"], serve.initial_rho(r) * 2.0)
