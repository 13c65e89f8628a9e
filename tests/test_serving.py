"""Scheduling of the serving loop on the CPU with simulated jobs and clock:
admission in arrival order within the slot count and the HBM budget, every
job completes, and one-at-a-time serving never overlaps jobs."""
import numpy as np

from supernormal_amd import serving


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-3)


class FakeJob:
    """Each iteration costs `it_s` of simulated time shared by the running
    jobs; the Chamfer distance halves every `eval_every` iterations."""
    running = []

    def __init__(self, spec, clock, it_s=0.002):
        self.spec, self.clock, self.it_s = spec, clock, it_s
        self.iters = 0
        self.chamfer = None
        self.history = []
        self.samples = 0.0

    def advance(self, n):
        n = min(n, self.spec.max_iters - self.iters)
        self.iters += n
        self.samples += 1000.0 * n
        self.clock.t += n * self.it_s

    def due_for_eval(self):
        return self.iters >= self.spec.max_iters or self.iters % self.spec.eval_every == 0

    def evaluate(self):
        self.chamfer = 0.02 * 0.5 ** (self.iters // self.spec.eval_every)
        self.history.append((self.iters, self.chamfer, 0.0))
        return self.chamfer

    def finished(self):
        return self.iters >= self.spec.max_iters or self.chamfer <= self.spec.chamfer_tol

    def finish(self):
        self.clock.t += 0.05

    def release(self):
        pass


def _run(max_running, budget, n=12):
    specs = serving.sample_stream(n_jobs=n, rate=2.0, seed=3, views=(8, 50),
                                  resolutions=(256, 512), log2_sizes=(17, 19), max_iters=400,
                                  chamfer_tol=3e-3, eval_every=50)
    clock = FakeClock()
    seen = []

    def make(spec):
        j = FakeJob(spec, clock)
        seen.append(spec.job_id)
        return j

    srv = serving.Server(device="cpu", hbm_budget_bytes=budget, max_running=max_running,
                         chunk=10, make_job=make, clock=clock, sleep=clock.sleep)
    return specs, seen, srv.run(specs)


def test_admission_order_and_completion():
    specs, seen, out = _run(4, 10 ** 13)
    assert seen == sorted(seen)                     # arrival order
    assert len(out["records"]) == len(specs)        # every job completes
    assert out["peak_running"] <= 4
    for r in out["records"]:
        assert r["admit_s"] >= r["arrival_s"] - 1e-9
        assert r["done_s"] > r["admit_s"]
        assert r["chamfer"] <= 3e-3 or r["iters"] == 400
    assert out["p99_s"] >= out["p50_s"] > 0


def test_budget_limits_concurrency():
    specs = serving.sample_stream(n_jobs=12, rate=2.0, seed=3, resolutions=(256, 512),
                                  log2_sizes=(17, 19))
    biggest = max(serving.job_bytes(s) for s in specs)
    _, _, out = _run(8, int(biggest * 2.5))
    assert out["peak_running"] <= 2


def test_one_at_a_time_never_overlaps():
    _, _, solo = _run(1, 10 ** 13)
    recs = sorted(solo["records"], key=lambda r: r["admit_s"])
    for a, b in zip(recs, recs[1:]):
        assert b["admit_s"] >= a["done_s"] - 1e-9
    assert solo["peak_running"] == 1


def test_job_bytes_grows_with_table_and_views():
    s = serving.JobSpec(0, 1, 8, 256, 17, 0.0)
    t = serving.JobSpec(0, 1, 8, 256, 21, 0.0)
    v = serving.JobSpec(0, 1, 50, 1024, 17, 0.0)
    assert serving.job_bytes(t) > serving.job_bytes(s)
    assert serving.job_bytes(v) > serving.job_bytes(s)
    assert serving.hash_entries(16, 19, 16, 1.38) == 6098120
