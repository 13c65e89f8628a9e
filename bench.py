"""Training throughput of SuperNormal's hot path on one MI355X.

Workload (the full-size single-object configuration): seeded smooth union of
a torus (R 0.4, r 0.12), a 0.3-side box and 3 spheres; 20 ring + elevation
cameras at 612x512 with N(0, 0.02) normal noise + masks; hash grid 16 levels
x 2 features, T = 2^19, base 16, x1.38; SDF MLP 64 wide, 2 hidden layers;
2048 patches of 3x3 (18 432 rays) per iteration; occupancy grid 128^3;
marching with a capacity of 1024 samples per centre ray (the reference is
unbounded; a cap of 128 truncates rays before the surface late in training and
every path diverges, --max-samples 128 measures it); 5000 iterations with the
reference's schedules.  A "step" is
one training iteration (occupancy update when due, patch sampling, marching,
rendering, losses, backward, Adam) on the fused HIP path: by default with no
host round trip and replayed as one HIP graph (--mode graph), the MLP on the
fp16/bf16 matrix cores with fp32 accumulation and fp32 master weights
(--precision fp16).

Timed region: W untimed warm-up iterations, then K iterations bracketed by
device synchronisation; the default W + K = 5000 covers the whole schedule,
so `value` is the run-average iterations/s.  Beside it:
  * plain_gpu: the reference-structured path (torch autograd over the drop-in
    kernels, torch Linear MLP) timed from the same trained state;
  * roofline: the dominant kernel's algorithmic bytes per launch / its mean
    duration, from HIP events on its launch stream around every launch of the
    timed region (in graph mode the iteration is four graphs with the
    hash-grid backward alone in the third, so the events sit between
    replays); `traffic` = the HBM bytes of one late launch from rocprofv3
    FETCH_SIZE / WRITE_SIZE counter runs (--traffic-probe), scaled by what
    the counters read for a copy of known size;
  * cpu_baseline: the CPU oracle (torch restatement of the reference
    iteration) on the same workload, a bounded number of iterations;
  * accuracy: Chamfer distance / F-score of the visible surface points of the
    trained field against the analytic scene's.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def hash_bwd_bytes(n, n_active, grid):
    """Algorithmic bytes of one hash-grid backward launch over n points:
    positions (12 B) and dL/denc (L*F*4 B) read per point, plus an fp32
    read-modify-write of the 8 corners x F features of every active level."""
    L, F = grid["n_levels"], grid["n_features"]
    return n * (12 + L * F * 4 + n_active * 8 * F * 4 * 2)


def hash_fwd_bytes(n, n_active, grid):
    """Positions read, 8 corners x F fp32 gathered per active level, L*F written."""
    L, F = grid["n_levels"], grid["n_features"]
    return n * (12 + n_active * 8 * F * 4 + L * F * 4)


ROOFLINE_KERNELS = {"sn_hashgrid_bwd": hash_bwd_bytes, "sn_hashgrid_fwd": hash_fwd_bytes}

# HBM bytes of one late hash-grid backward launch on the tbs612 workload, from
# two rocprofv3 counter runs over --traffic-probe (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, the factors being what a 64 MiB copy reads on each
# counter), with that launch's point count and algorithmic bytes; the
# commands and per-kernel figures are in profiles/hash_bwd_tbs612_late_traffic.md
HASH_BWD_LATE_TRAFFIC = {"points": 1473021, "traffic_bytes": 2740882158,
                         "algorithmic_bytes": 3222969948}


def cpu_baseline(runner, iters=2):
    """The oracle's CPU iteration on the same workload, bounded to `iters`
    iterations after one untimed iteration."""
    from oracle.train_oracle import OracleTrainer
    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    ds = runner.dataset
    tr = OracleTrainer(runner.sdf_network, float(runner.deviation_network.variance),
                       runner.occ_resolution, {"lr": runner.learning_rate, "loss_type": "l2"})
    tr.binary = runner.renderer.occupancy_grid.binary.cpu().numpy().astype(np.uint8)
    tr.occs = runner.renderer.occupancy_grid.occs.cpu().numpy()
    g = torch.Generator(device=ds.device).manual_seed(123)
    times = []
    step = float(runner.renderer.sampling_step_size)
    for i in range(iters + 1):
        o, d, pn, vi, nt, m = ds.gen_random_patches(runner.batch_size, generator=g)
        near, far = ds.near_far_from_sphere(o[:, 1, 1], d[:, 1, 1])
        cpu = [t.cpu() for t in (o, d, pn, vi, nt, m, near, far)]
        t0 = time.perf_counter()
        tr.step(*cpu, step_size=step, lr=runner.lr, cap=runner.max_samples_per_ray,
                method=runner.gradient_method)
        if i > 0:
            times.append(time.perf_counter() - t0)
    sec = float(np.median(times))
    return {"value": 1.0 / sec, "unit": "train iters/s", "cores": threads, "kind": "port",
            "sample": f"{iters} iterations of the CPU oracle (torch fp32 restatement of the "
                      f"reference iteration) at 2048 patches from the trained state, median"}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4980)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="tbs612")
    ap.add_argument("--method", default="dfd", choices=["dfd", "ad"])
    ap.add_argument("--roofline-kernel", default="sn_hashgrid_bwd")
    ap.add_argument("--plain-iters", type=int, default=20)
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--mode", default="graph", choices=["fused", "sync_free", "graph"])
    ap.add_argument("--precision", default="fp16", choices=["fp32", "fp16"])
    ap.add_argument("--seed", type=int, default=0,
                    help="seed of the initial weights and of the synthetic views' noise")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="after the timed run: 5 hash-grid backward calls on the last "
                         "iteration's points and 3 copies of 64 MiB (calibration), for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--max-samples", type=int, default=None,
                    help="samples per centre ray capacity (default: the workload's)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); this build runs on a "
              "single GPU and produces no multi-GPU figure.", file=sys.stderr)
        return 2
    from supernormal_amd import _lib
    from supernormal_amd.runner import Runner, synthetic_conf

    dev = torch.device("cuda:0")
    mode = a.mode if a.method == "dfd" else "fused"
    conf = synthetic_conf(a.workload, gradient_method=a.method,
                          end_iter=max(a.steps + a.warmup, 1), mlp_precision=a.precision,
                          step_mode=mode, max_samples_per_ray=a.max_samples)
    runner = Runner(conf, fused=True, device=dev, seed=a.seed)
    sf = runner.sf_step
    grid = runner.sdf_network.encoding.grid
    runner.update_learning_rate()
    torch.cuda.synchronize()
    for i in range(a.warmup):
        runner.train_iteration(i)
    torch.cuda.synchronize()
    timer = _lib.time_calls(a.roofline_kernel) if sf is None else None
    if sf is not None:
        # events around the hash-grid backward's replay (its own graph), read at the end
        sf.time_hash_bwd = True
        sf.hash_times = []
    samples = 0
    samples_dev = torch.zeros(1, dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(a.warmup, a.warmup + a.steps):
        runner.train_iteration(i)
        if sf is not None:
            samples_dev += sf.ns
        else:
            samples += getattr(runner, "last_samples", 0)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    if sf is not None:
        samples = float(samples_dev)
    byte_fn = ROOFLINE_KERNELS[a.roofline_kernel]
    if sf is None:
        _lib.stop_timing(a.roofline_kernel)
        durs = timer.durations_ms()
        # (x, n, params, levels, F, log2 T, base, scale, n_active, ...)
        nbytes = [byte_fn(int(args[1]), int(args[8]), grid) for _, _, args in timer.records]
    else:
        # every launch of the timed region (device count copied next to it)
        sf.time_hash_bwd = False
        timed = sf.hash_bwd_timings()
        durs = [ms for ms, _, _ in timed]
        nbytes = [byte_fn(n, la, grid) for _, n, la in timed]
    ms_per_step = ms / max(a.steps, 1)
    value = 1000.0 / ms_per_step
    roof = None
    if durs:
        avg_ms = float(np.mean(durs))
        avg_bytes = float(np.mean(nbytes))
        achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": a.roofline_kernel, "launches_timed": len(durs),
                "avg_launch_ms": round(avg_ms, 4), "avg_algorithmic_bytes": int(avg_bytes)}
        if a.roofline_kernel == "sn_hashgrid_bwd" and a.workload == "tbs612":
            t = HASH_BWD_LATE_TRAFFIC
            roof["traffic"] = t["traffic_bytes"]
            roof["traffic_launch"] = {"points": t["points"],
                                      "algorithmic_bytes": t["algorithmic_bytes"],
                                      "traffic_over_algorithmic":
                                          round(t["traffic_bytes"] / t["algorithmic_bytes"], 3)}
        if len(durs) >= 1000:
            # the last 1000 launches (all levels active, the most points)
            late_ms = float(np.mean(durs[-1000:]))
            late_bytes = float(np.mean(nbytes[-1000:]))
            roof["last_1000"] = {"avg_launch_ms": round(late_ms, 4),
                                 "achieved": round(late_bytes / (late_ms * 1e-3) / 1e9, 1),
                                 "avg_algorithmic_bytes": int(late_bytes)}
    out = {
        "metric": "train iters/s at 18k rays vs plain GPU @ equal Chamfer; p50/p99 job time-to-mesh",
        "value": round(value, 2), "unit": "train iters/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if a.precision == "fp32" else "fp16/bf16 MFMA + f32",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: torus+box+3 spheres, 20 views 612x512, hash 16x2 "
                               f"T=2^19 base 16 x1.38, MLP 64x2, 2048 3x3 patches (18432 rays), "
                               f"occupancy 128^3, {runner.max_samples_per_ray} samples/ray "
                               f"capacity, gradient {a.method}, "
                               f"mode {mode}, mlp {a.precision}",
                   "global_batch": runner.batch_size * 9, "iterations": a.warmup + a.steps,
                   "parallelism": "single GPU"},
        "samples_per_iter": round(samples / max(a.steps, 1), 1),
        "wall_s": round(wall, 2),
        "roofline": roof,
    }
    if a.traffic_probe and sf is not None:
        sf.drain()
        torch.cuda.synchronize()
        for _ in range(5):
            sf._part_hash()
        torch.cuda.synchronize()
        xc = torch.rand(16 << 20, device=dev)
        yc = torch.empty_like(xc)
        for _ in range(3):
            torch.mul(xc, 1.0, out=yc)
        torch.cuda.synchronize()
        out["traffic_probe"] = {"hash_bwd_points": int(sf.ne), "copy_bytes_each_way": 64 << 20}
    if not a.no_accuracy:
        from supernormal_amd.mesh import chamfer_distance_and_f1_score, visible_surface_points
        net = runner.sdf_network
        pts = visible_surface_points(lambda x: net.fused_sdf(x), runner.dataset, 4000)
        gt = runner.dataset.surface_points(4000 * len(runner.dataset.train_images))
        cd, fs = chamfer_distance_and_f1_score(gt, pts, f_threshold=0.005)
        out["accuracy"] = {"chamfer": round(cd, 6), "fscore@0.005": round(fs, 4),
                           "n_eval_points": int(len(pts))}
    if not a.no_plain:
        runner.fused = False
        torch.cuda.synchronize()
        it0 = runner.iter_step
        runner.train_iteration(it0)  # untimed
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(a.plain_iters):
            runner.train_iteration(it0 + 1 + i)
        torch.cuda.synchronize()
        plain = a.plain_iters / (time.perf_counter() - t)
        runner.fused = True
        t = time.perf_counter()
        for i in range(a.plain_iters):
            runner.train_iteration(it0 + 1 + a.plain_iters + i, ahead=i + 1 < a.plain_iters)
        torch.cuda.synchronize()
        fused_same = a.plain_iters / (time.perf_counter() - t)
        out["plain_gpu"] = {"value": round(plain, 2), "unit": "train iters/s",
                            "path": "reference-structured (autograd over drop-in kernels, "
                                    "torch Linear MLP)",
                            "fused_same_state": round(fused_same, 2),
                            "speedup": round(fused_same / plain, 2)}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(runner, a.cpu_iters)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
