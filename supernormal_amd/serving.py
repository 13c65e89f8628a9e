"""A request loop for reconstruction jobs on one GPU (continuous batching).

Jobs (a scene's normal maps + masks to reconstruct) arrive over time.  A job
is admitted as soon as a slot is free and its resident state fits the HBM
budget; its state (hash table, MLP, Adam moments, occupancy grid, capacity-
sized step buffers and a captured HIP graph of its training iteration) stays
on the device until its mesh is out.  Every admitted job owns a HIP stream:
the loop issues a few graph replays per job in turn, so the kernels of all
running jobs overlap on the device, and newly arrived jobs join between
rounds without waiting for the others to finish.  A job ends when its
Chamfer distance to the ground-truth visible points falls under the job's
tolerance (checked every `eval_every` iterations) or at `max_iters`; then its
mesh is extracted and its memory released.

Each job draws its patches and occupancy samples from a generator of its own,
so its result does not depend on what else runs beside it (up to the order
of fp32 atomics in the hash-grid backward).

time-to-mesh = mesh ready - arrival (queueing + data preparation + training +
mesh extraction).  `run_stream` reports p50 / p99 and the aggregate sample
throughput, and `one_at_a_time=True` gives the comparison the same jobs run
back to back.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Callable, List, Optional

import numpy as np
import torch


@dataclasses.dataclass
class JobSpec:
    job_id: int
    seed: int
    n_views: int
    resolution: int          # square images
    log2_hashmap_size: int
    arrival_s: float
    batch_size: int = 2048
    max_iters: int = 2000
    chamfer_tol: float = 5e-3
    eval_every: int = 250
    mesh_resolution: int = 128


def sample_stream(n_jobs=64, rate=0.5, seed=0, views=(8, 50), resolutions=(256, 512, 768, 1024),
                  log2_sizes=(17, 18, 19, 20, 21), **kw) -> List[JobSpec]:
    """Poisson arrivals at `rate` jobs/s; per job: views U[views], a resolution
    and a hash-table size drawn uniformly from the lists, a seeded scene."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for j in range(n_jobs):
        t += float(rng.exponential(1.0 / rate))
        out.append(JobSpec(job_id=j, seed=int(rng.integers(1 << 30)),
                           n_views=int(rng.integers(views[0], views[1] + 1)),
                           resolution=int(rng.choice(resolutions)),
                           log2_hashmap_size=int(rng.choice(log2_sizes)), arrival_s=t, **kw))
    return out


def hash_entries(n_levels, log2_T, base, scale):
    """Table entries under tcnn's per-level rule (same as the kernels)."""
    tot = 0
    for l in range(n_levels):
        s = base * scale ** l - 1.0
        res = int(math.ceil(s)) + 1
        dense = res ** 3
        size = min((dense + 7) // 8 * 8, 1 << log2_T)
        tot += size
    return tot


def job_bytes(spec: JobSpec, max_samples_per_ray=1024, n_levels=16, F=2, patch=9) -> int:
    """Resident device bytes of a job: parameters with gradient and Adam
    moments, the step buffers at capacity, the hash-grid backward's record
    pool (tcnn.bucket_workspace_bytes), the dataset and the occupancy grid."""
    E = n_levels * F
    params = hash_entries(n_levels, spec.log2_hashmap_size, 16, 1.38) * F + 8000
    c0 = spec.batch_size * max_samples_per_ray
    c1 = 2 * c0 * patch
    step = c1 * (3 + 2 * E + 2) * 4 + c0 * (2 * 3 + E + 1) * 4 * 2 + c0 * 40
    data = spec.n_views * spec.resolution ** 2 * (12 + 4 + 36)
    occ = 128 ** 3 * 5
    rec = 2 + 4 * F
    hash_ws = min(8 * n_levels * c1, 196608 * 747, (4 << 30) // rec) * rec
    return int(params * 16 + step + hash_ws + data + occ)


def job_data(spec: JobSpec, device):
    """A job's input: the normal maps, masks and cameras of its views (here
    rendered from its seeded synthetic scene, standing in for the images a
    request brings) and the ground-truth surface points for its Chamfer."""
    from .scene import Dataset, random_primitives_scene
    scene = random_primitives_scene(spec.seed)
    ds = Dataset(scene, spec.n_views, spec.resolution, spec.resolution, 0.02,
                 torch.device(device), seed=spec.seed,
                 elevations=(0.0, 20.0, -20.0, 40.0, -40.0))
    return ds, ds.surface_points(20000, seed=spec.seed)


class Job:
    """A running reconstruction: Runner (graph mode) on a stream of its own.
    `data` = job_data(spec) when the request's input is already resident."""

    def __init__(self, spec: JobSpec, device, precision="fp16", data=None):
        from .runner import Runner, synthetic_conf
        self.spec = spec
        self.device = torch.device(device)
        self.stream = torch.cuda.Stream(device=self.device)
        self.iters = 0
        self.samples = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.chamfer = None
        self.history = []
        with torch.cuda.stream(self.stream):
            if data is None:
                data = job_data(spec, self.device)
            ds, self.gt = data
            conf = synthetic_conf("prims1024", end_iter=spec.max_iters,
                                  batch_size=spec.batch_size, max_samples_per_ray=1024,
                                  mlp_precision=precision, step_mode="graph")
            conf["model.encoding"]["log2_hashmap_size"] = spec.log2_hashmap_size
            conf["model.ray_marching"]["occ_resolution"] = 128
            self.runner = Runner(conf, fused=True, dataset=ds, device=self.device,
                                 seed=spec.seed % (1 << 31), rng_seed=spec.seed)
            self.runner.update_learning_rate()
        self.ready = torch.cuda.Event()
        self.ready.record(self.stream)

    def advance(self, n):
        """Issue n iterations on the job's stream (asynchronous)."""
        with torch.cuda.stream(self.stream):
            for _ in range(n):
                if self.iters >= self.spec.max_iters:
                    break
                self.runner.train_iteration(self.iters)
                self.samples += self.runner.sf_step.ns
                self.iters += 1

    def due_for_eval(self):
        return self.iters >= self.spec.max_iters or (
            self.iters > 0 and self.iters % self.spec.eval_every == 0)

    def evaluate(self):
        """Chamfer of the learned surface's visible points against the ground truth."""
        from .mesh import chamfer_distance_and_f1_score, visible_surface_points
        with torch.cuda.stream(self.stream):
            net = self.runner.sdf_network
            pts = visible_surface_points(lambda x: net.fused_sdf(x), self.runner.dataset,
                                         max(1, 20000 // len(self.runner.dataset.train_images)))
            cd, fs = chamfer_distance_and_f1_score(self.gt, pts, f_threshold=0.005)
        self.chamfer = cd
        self.history.append((self.iters, cd, fs))
        return cd

    def finished(self):
        return self.iters >= self.spec.max_iters or (
            self.chamfer is not None and self.chamfer <= self.spec.chamfer_tol)

    def extract_mesh(self):
        from .mesh import extract_geometry
        with torch.cuda.stream(self.stream):
            net = self.runner.sdf_network
            r = self.spec.mesh_resolution
            v, f = extract_geometry(np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, 1.0]), r,
                                    0.0, lambda x: -net.fused_sdf(x.contiguous()))
        return v, f

    def finish(self):
        """Mesh extraction; returns when the job's work is done."""
        self.mesh = self.extract_mesh()
        with torch.cuda.stream(self.stream):
            self.runner.sf_step.drain()
        self.stream.synchronize()

    def release(self):
        self.runner = None
        self.gt = None


class Server:
    """Admission by HBM budget and slot count; round-robin issue of graph
    replays over the running jobs' streams."""

    def __init__(self, device="cuda", hbm_budget_bytes=None, max_running=8, chunk=10,
                 precision="fp16", make_job: Optional[Callable] = None,
                 clock: Callable[[], float] = time.perf_counter, sleep=time.sleep):
        self.device = torch.device(device)
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget_bytes = int(free * 0.8)
        self.budget = int(hbm_budget_bytes)
        self.max_running = int(max_running)
        self.chunk = int(chunk)
        self.precision = precision
        self.make_job = make_job or (lambda spec: Job(spec, self.device, precision))
        self.clock, self.sleep = clock, sleep

    def run(self, specs: List[JobSpec], log=None, on_done=None):
        """Serve `specs`; `on_done(job)` sees each finished job before its state
        is released."""
        queue = sorted(specs, key=lambda s: s.arrival_s)
        running, records = [], []
        used = 0
        t0 = self.clock()
        peak_running = 0
        # host seconds per phase of the loop (job set-up, issue, evaluation, finish)
        phase = dict(admit=0.0, issue=0.0, evaluate=0.0, finish=0.0)
        while queue or running:
            now = self.clock() - t0
            # admit, in arrival order, while slots and memory allow
            while (queue and queue[0].arrival_s <= now and len(running) < self.max_running
                   and used + job_bytes(queue[0]) <= self.budget):
                spec = queue.pop(0)
                t_admit = self.clock() - t0
                job = self.make_job(spec)
                phase["admit"] += self.clock() - t0 - t_admit
                job.t_arrival, job.t_admit = spec.arrival_s, t_admit
                job.bytes = job_bytes(spec)
                used += job.bytes
                running.append(job)
            peak_running = max(peak_running, len(running))
            if not running:
                if queue:
                    self.sleep(max(0.0, queue[0].arrival_s - (self.clock() - t0)))
                continue
            tp = self.clock()
            for job in running:
                job.advance(self.chunk)
            phase["issue"] += self.clock() - tp
            done = []
            for job in running:
                if job.due_for_eval():
                    tp = self.clock()
                    job.evaluate()
                    phase["evaluate"] += self.clock() - tp
                    if job.finished():
                        tp = self.clock()
                        job.finish()
                        job.t_done = self.clock() - t0
                        phase["finish"] += job.t_done + t0 - tp
                        done.append(job)
            for job in done:
                running.remove(job)
                used -= job.bytes
                rec = dict(job_id=job.spec.job_id, arrival_s=job.t_arrival,
                           admit_s=job.t_admit, done_s=job.t_done,
                           time_to_mesh_s=job.t_done - job.t_arrival, iters=job.iters,
                           chamfer=job.chamfer, samples=float(job.samples),
                           history=job.history)
                records.append(rec)
                if log:
                    log(rec)
                if on_done:
                    on_done(job)
                job.release()
        t_end = self.clock() - t0
        ttm = np.array([r["time_to_mesh_s"] for r in records])
        return dict(records=records, p50_s=float(np.percentile(ttm, 50)),
                    p99_s=float(np.percentile(ttm, 99)), makespan_s=t_end,
                    samples_per_s=float(sum(r["samples"] for r in records)) / max(t_end, 1e-9),
                    peak_running=peak_running,
                    host_s={k: round(v, 3) for k, v in phase.items()})


def run_stream(specs, device="cuda", one_at_a_time=False, **kw):
    max_running = kw.pop("max_running", 8)
    server = Server(device=device, max_running=1 if one_at_a_time else max_running, **kw)
    return server.run(specs)


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="serving stream: p50/p99 time-to-mesh")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--rate", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-running", type=int, default=8)
    ap.add_argument("--max-iters", type=int, default=2000)
    ap.add_argument("--tol", type=float, default=5e-3)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--resolutions", default="256,512,768,1024")
    ap.add_argument("--views", default="8,50")
    ap.add_argument("--one-at-a-time", action="store_true")
    a = ap.parse_args(argv)
    specs = sample_stream(a.jobs, a.rate, a.seed,
                          views=tuple(int(v) for v in a.views.split(",")),
                          resolutions=tuple(int(r) for r in a.resolutions.split(",")),
                          max_iters=a.max_iters, chamfer_tol=a.tol, batch_size=a.batch)
    def log(rec):  # one line per finished job, as it finishes
        print(json.dumps({k: v for k, v in rec.items() if k != "history"}), flush=True)

    # the requests' inputs (synthetic views) are prepared before the clock
    # starts, as a request would bring its images
    t = time.perf_counter()
    data = {}
    for sp in specs:
        data[sp.job_id] = job_data(sp, "cuda")
        print(json.dumps({"input_ready": sp.job_id, "s": round(time.perf_counter() - t, 2)}),
              flush=True)
    max_running = 1 if a.one_at_a_time else a.max_running
    server = Server(max_running=max_running, precision="fp16",
                    make_job=lambda sp: Job(sp, "cuda", "fp16", data=data.pop(sp.job_id)))
    out = server.run(specs, log=log)
    out.pop("records")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
