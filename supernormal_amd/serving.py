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

Shared launches (the default, `fused=True`): after a job's first iteration
(its own graphs, which also warm its kernels up), the running jobs' iterations
are captured together into ONE HIP graph per set of jobs -- each job's
sampling, march, rendering, backward and step guard on a stream forked from
the capture stream, then a single Adam launch over every job's parameters
(sn_adam_multi, one segment per job) -- and a round is one replay of it.  The
graph is captured again when a job joins or leaves.  With `fused=False` each
job replays its own graphs on its own stream.

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
    mesh_resolution: int = 256


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
        self.mesh_dev = None
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

    def begin_joint(self):
        """Host part of the job's next iteration inside a joint round (on the
        current stream); returns the step count token for end_joint()."""
        r = self.runner
        step = r.begin_iteration(self.iters)
        return r.sf_step.prepare(r.lr, step)

    def end_joint(self, t):
        r = self.runner
        r.sf_step.finished(t)
        r.end_iteration()
        self.samples += r.sf_step.ns
        self.iters += 1

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
        """Chamfer of the learned surface against the ground truth, the
        reference's way (exp_runner.py:579-630): marching cubes of the SDF at
        the job's mesh resolution, first hits of training-pixel rays with the
        mesh; the mesh is kept as the job's result."""
        from .mesh import (chamfer_distance_and_f1_score, extract_fields_device,
                           marching_cubes_device, visible_mesh_points)
        # after the iterations issued on the current stream (joint rounds)
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            net = self.runner.sdf_network
            r = self.spec.mesh_resolution
            lo, hi = [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]
            u = extract_fields_device(lo, hi, r, lambda x: -net.fused_sdf(x.contiguous()),
                                      self.device)
            self.mesh_dev = marching_cubes_device(-u, 0.0, lo, hi)
            del u
            pts = visible_mesh_points(self.mesh_dev, self.runner.dataset,
                                      max(1, 20000 // len(self.runner.dataset.train_images)))
            cd, fs = chamfer_distance_and_f1_score(self.gt, pts, f_threshold=0.005)
        # later iterations (next rounds) overwrite what the evaluation read
        torch.cuda.current_stream(self.device).wait_stream(self.stream)
        self.chamfer = cd
        self.history.append((self.iters, cd, fs))
        return cd

    def finished(self):
        return self.iters >= self.spec.max_iters or (
            self.chamfer is not None and self.chamfer <= self.spec.chamfer_tol)

    def extract_mesh(self):
        """The mesh of the last evaluation (a job finishes right after one) as
        PyMCubes returns it: vertices, triangles."""
        with torch.cuda.stream(self.stream):
            return self.mesh_dev.indexed()

    def finish(self):
        """Mesh extraction; returns when the job's work is done."""
        self.mesh = self.extract_mesh()
        with torch.cuda.stream(self.stream):
            self.runner.sf_step.drain()
        self.stream.synchronize()

    def release(self):
        self.runner = None
        self.gt = None
        self.mesh_dev = None


class JointRound:
    """One HIP graph holding one iteration of every job in `jobs`: each job's
    iteration up to its step guard on a stream of its own forked from the
    capture stream, joined, then one sn_adam_multi launch over all the jobs'
    parameters.  replay() runs one iteration of all of them."""

    def __init__(self, jobs, device):
        import ctypes as C
        from . import _lib
        self.jobs = list(jobs)
        self.device = torch.device(device)
        lib = _lib.lib()
        segs = (_lib.SnAdamSegment * len(self.jobs))()
        block0 = 0
        for i, job in enumerate(self.jobs):
            sf = job.runner.sf_step
            nb = int(lib.sn_adam_multi_blocks(sf.flat.n))
            segs[i] = _lib.SnAdamSegment(*sf.adam_segment(), block0, nb)
            block0 += nb
        self.total_blocks = block0
        raw = bytes(bytearray(segs))
        self.segments = torch.tensor(list(raw), dtype=torch.uint8).to(self.device)
        self.graph = self._capture()

    def _capture(self):
        from ._lib import call, ptr, stream_of
        dev = self.device
        torch.cuda.synchronize(dev)
        g = torch.cuda.CUDAGraph()
        for job in self.jobs:
            gen = job.runner.sf_step.gen
            if gen is not None:
                g.register_generator_state(gen)
        streams = [torch.cuda.Stream(device=dev) for _ in self.jobs]
        with torch.cuda.graph(g):
            cur = torch.cuda.current_stream(dev)
            for job, st in zip(self.jobs, streams):
                st.wait_stream(cur)
                with torch.cuda.stream(st):
                    job.runner.sf_step.capture_parts()
            for st in streams:
                cur.wait_stream(st)
            call("sn_adam_multi", ptr(self.segments), len(self.jobs), self.total_blocks, 0.9,
                 0.999, 1e-8, stream_of(dev))
        self._streams = streams
        return g

    def replay(self):
        tokens = [job.begin_joint() for job in self.jobs]
        self.graph.replay()
        for job, t in zip(self.jobs, tokens):
            job.end_joint(t)


class Server:
    """Admission by HBM budget and slot count; each round advances every
    running job by `chunk` iterations: replays of one joint graph over the
    running jobs (fused=True), or each job's own graphs on its own stream."""

    def __init__(self, device="cuda", hbm_budget_bytes=None, max_running=8, chunk=10,
                 precision="fp16", make_job: Optional[Callable] = None,
                 clock: Callable[[], float] = time.perf_counter, sleep=time.sleep,
                 fused=True):
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
        self.fused = bool(fused)
        self.captures = 0
        self._joint = None
        self._joint_key = None

    def _issue(self, running):
        """One round: `chunk` iterations of every running job."""
        joint = [j for j in running if self.fused and hasattr(j, "begin_joint")]
        for job in running:
            if job not in joint:
                job.advance(self.chunk)
        # a job's first iteration runs on its own graphs (no march left pending)
        for job in joint:
            if job.iters == 0:
                with torch.cuda.stream(job.stream):
                    job.runner.train_iteration(0, ahead=False)
                    job.samples += job.runner.sf_step.ns
                    job.iters += 1
                torch.cuda.current_stream(job.device).wait_stream(job.stream)
        members = [j for j in joint if _joint_ready(j)]
        if not members:
            return
        key = tuple(id(j) for j in members)
        if self._joint_key != key:
            self._joint = JointRound(members, members[0].device)
            self._joint_key = key
            self.captures += 1
        for _ in range(self.chunk):
            self._joint.replay()
            if any(j.due_for_eval() for j in members):
                break

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
            self._issue(running)
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
        self._joint = None
        self._joint_key = None
        return dict(records=records, p50_s=float(np.percentile(ttm, 50)),
                    p99_s=float(np.percentile(ttm, 99)), makespan_s=t_end,
                    joint_captures=self.captures, fused=self.fused,
                    samples_per_s=float(sum(r["samples"] for r in records)) / max(t_end, 1e-9),
                    peak_running=peak_running,
                    host_s={k: round(v, 3) for k, v in phase.items()})


def _joint_ready(job):
    """A job takes part in joint rounds once its own first iteration ran (its
    kernels warmed up, nothing pending) and while iterations remain."""
    return job.iters > 0 and job.iters < job.spec.max_iters


def run_stream(specs, device="cuda", one_at_a_time=False, **kw):
    max_running = kw.pop("max_running", 8)
    server = Server(device=device, max_running=1 if one_at_a_time else max_running, **kw)
    return server.run(specs)


def main(argv=None):
    import argparse
    import faulthandler
    import json
    faulthandler.enable()  # a crash prints the Python stack of every thread
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
    ap.add_argument("--no-fuse", action="store_true",
                    help="each job replays its own graphs on its own stream")
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
    server = Server(max_running=max_running, precision="fp16", fused=not a.no_fuse,
                    make_job=lambda sp: Job(sp, "cuda", "fp16", data=data.pop(sp.job_id)))
    out = server.run(specs, log=log)
    out.pop("records")
    out["max_running"] = max_running
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
