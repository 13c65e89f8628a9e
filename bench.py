"""Training throughput of SuperNormal's hot path on one MI355X.

Workload (the full-size single-object configuration, tbs612): seeded smooth
union of a torus (R 0.4, r 0.12), a 0.3-side box and 3 spheres; 20 ring +
elevation cameras at 612x512 with N(0, 0.02) normal noise + masks; hash grid
16 levels x 2 features, T = 2^19, base 16, x1.38; SDF MLP 64 wide, 2 hidden
layers; 2048 patches of 3x3 (18 432 rays) per iteration; occupancy grid
128^3; marching with a capacity of 1024 samples per centre ray (the reference
is unbounded; rays that fill it are counted and reported); the reference's
schedules over 5000 iterations: step size 1e-2 -> 1e-3, learning rate warm-up
50 + cosine, bandwidth (active hash levels) +1 every 250 iterations, occupancy
update every 8.  A "step" is one training iteration (occupancy update when
due, patch sampling, marching, visibility pre-pass, rendering, losses,
backward, Adam) on the fused HIP path: with no host round trip, replayed as
HIP graphs (--mode graph), the MLP on fp16/bf16 matrix cores with fp32
accumulation and fp32 master weights (--precision fp16).

The schedule always runs whole (5000 iterations), whatever --steps is: the
first 5000 - W - K iterations are an untimed prefix, then W untimed warm-up
iterations, then the K timed iterations -- the end of training, where all 16
levels are active, the step size is 1e-3 and the samples per iteration are at
their largest.  `value` = K / (their time), bracketed by device
synchronisation.  `full_schedule` = 5000 / the time of the whole schedule in
the same run.  Beside it:
  * counts of the timed window: marched / visible centre samples, patch
    samples, SDF evaluations (patch points, centre points of the visibility
    pre-pass, occupancy cells), rays that filled the sample capacity, skipped
    iterations; samples/s, evaluations/s, MLP TFLOP/s and the MFMA
    utilisation against the dense fp16/bf16 peak, the step's algorithmic
    HBM GB/s;
  * roofline: the hash-grid backward (the largest kernel time of the late
    iterations), algorithmic bytes of the active levels per launch / its mean
    duration from HIP events on its launch stream around every launch of the
    timed window (in graph mode it is the third of the iteration's four
    graphs, so the events sit between replays); `traffic` = the HBM bytes of a
    late launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs over
    --traffic-probe, per point, times the timed launches' mean points;
  * fp32: the same schedule with the fp32 MLP and fp32 table (--no-fp32 skips);
  * method_ad: the same schedule with the analytic gradient (method ad: the
    rendered normal is the SDF's x-gradient, trained through its double
    backward) (--no-ad skips);
  * accuracy: Chamfer distance / F-score of the trained field's visible
    surface points against the analytic scene's, for each leg;
  * plain_gpu: the reference-structured path (torch autograd over the drop-in
    kernels, torch Linear MLP, fp32) timed from the trained state, and with
    --plain-full trained over the whole schedule from the same seed (its
    Chamfer beside the device path's: the equal-Chamfer comparison);
  * cpu_baseline: the CPU oracle (torch restatement of the reference
    iteration) on a bounded sample of this workload from the trained state,
    plus the small CPU workload (sphere, 8 views 128^2, 256 patches, 200
    iterations) timed as median of 20 iterations after 5 and trained whole,
    with the device path on the same small workload beside it.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "tbs612": "torus+box+3 spheres, 20 views 612x512, hash 16x2 T=2^19 base 16 x1.38, MLP 64x2, "
              "2048 3x3 patches (18432 rays), occupancy 128^3",
    "prims1024": "8 random primitives with 0.005 sinusoidal ridges, 50 views 1024x1024, hash "
                 "16x2 T=2^21 base 16 x1.38, MLP 64x2, 8192 3x3 patches (73728 rays), "
                 "occupancy 256^3",
    "sphere128": "sphere r 0.5, 8 views 128x128, hash 8x2 T=2^15 base 16 x1.5, MLP 32x1, 256 3x3 "
                 "patches, occupancy 64^3",
}
MFMA_PEAK = {"fp16": 2.5e15, "fp32": 157.3e12}   # dense fp16/bf16, fp32 MFMA (FLOP/s)


def hash_bwd_bytes(n, n_active, grid):
    """Algorithmic bytes of one hash-grid backward launch over n points:
    positions (12 B) and dL/denc of the active levels (F fp32 each) read per
    point, plus an fp32 read-modify-write of the 8 corners x F features of
    every active level."""
    F = grid["n_features"]
    return n * (12 + n_active * F * 4 + n_active * 8 * F * 4 * 2)


def hash_fwd_bytes(n, n_active, grid, table_bytes=4):
    """Positions read, 8 corners x F gathered per active level, L*F fp32 written."""
    L, F = grid["n_levels"], grid["n_features"]
    return n * (12 + n_active * 8 * F * table_bytes + L * F * 4)


# HBM bytes of a late hash-grid backward launch on this workload (all 16 levels
# active), per point: two rocprofv3 counter runs over --traffic-probe (5 launches
# on the last iteration's points: 1225053 in the FETCH_SIZE run, 1225593 in the
# WRITE_SIZE run), FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 -- the factors are what a
# 64 MiB copy reads on each counter: 1.4195e9 / 1225053 + 0.8553e9 / 1225593.
HASH_BWD_TRAFFIC_PER_POINT = {"bytes_per_point": 1856.6, "probe_points": 1225053,
                              "active_levels": 16}


def mlp_flops_fwd(cfg):
    """2 * sum(d_in * d_out) of the SDF MLP (35 -> 64 -> 64 -> 1 here)."""
    dims = [3 + cfg.E] + [cfg.d_hidden] * cfg.n_layers + [1]
    return 2 * sum(a * b for a, b in zip(dims[:-1], dims[1:]))


def iteration_bytes(c, grid, n_params, table_bytes, res3, B, P):
    """Algorithmic HBM bytes of one iteration from its counts (per-iteration
    means): marched / visible centre samples n0 / ns, patch points ne, centre
    points ne0, occupancy cells evaluated occ (formulas per kernel, read once /
    written once)."""
    L, F, La = grid["n_levels"], grid["n_features"], c["levels"]
    E = L * F
    ne, ne0, occ = c["ne"], c["ne0"], c["occ"]
    b = 0.0
    b += hash_fwd_bytes(ne + ne0 + occ, La, grid, table_bytes)     # hash forward
    b += (ne + ne0 + occ) * (12 + E * 4 + 4)                        # MLP forward in/out
    b += ne * (12 + E * 4 + 4 + E * 4)                               # MLP backward
    b += hash_bwd_bytes(ne, La, grid)                                # hash backward
    b += B * 40 + c["n0"] * 16 + res3                                # march (+ grid bytes)
    b += c["n0"] * 13                                                # visibility
    b += c["ns"] * P * (8 + 16 + 32)                                 # weights fwd/bwd, accumulate
    b += occ * 21                                                    # occupancy EMA
    b += n_params * 28 + (n_params * 2 if table_bytes == 2 else 0)  # Adam (+ fp16 table copy)
    return b


def run_schedule(dev, workload, precision, method, mode, seed, W, K, max_samples,
                 time_hash=True, log=None):
    """Trains the workload's whole schedule; times the last K iterations after W
    untimed ones.  Returns (runner, result dict)."""
    from supernormal_amd.runner import Runner, synthetic_conf
    conf = synthetic_conf(workload, gradient_method=method, mlp_precision=precision,
                          step_mode=mode, max_samples_per_ray=max_samples)
    runner = Runner(conf, fused=True, device=dev, seed=seed)
    E = runner.end_iter
    if W + K > E:
        raise SystemExit(f"--warmup {W} + --steps {K} exceed the schedule's {E} iterations")
    sf = runner.sf_step
    grid = runner.sdf_network.encoding.grid
    occ = runner.renderer.occupancy_grid
    runner.update_learning_rate()
    t_first = E - K
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for i in range(t_first):
        # no march of the first timed iteration runs ahead of the window
        runner.train_iteration(i, ahead=False if i == t_first - 1 else None)
        if log and (i + 1) % 1000 == 0:
            log(f"{workload} {precision}: iteration {i + 1}")
    torch.cuda.synchronize()
    from supernormal_amd import _lib
    timer = None
    if sf is not None:
        sf.time_hash_bwd = time_hash
        sf.hash_times = []
        acc0 = sf.counts_total()
        capped0, skipped0 = sf.capped_rays, sf.skipped_iterations
    else:
        timer = _lib.time_calls("sn_hashgrid_bwd") if time_hash else None
        samples = 0
    occ0 = occ.points_evaluated
    levels = runner.sdf_network.active_levels()
    step_size = float(runner.renderer.sampling_step_size)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(t_first, E):
        runner.train_iteration(i)
        if sf is None:
            samples += getattr(runner, "last_samples", 0)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_wall = time.perf_counter() - t_all
    ms = e0.elapsed_time(e1)
    res = {"value": K / (ms * 1e-3), "ms_per_step": ms / K, "wall_s": wall,
           "full_schedule": E / total_wall, "iterations": E, "levels": levels,
           "step_size": step_size}
    counts = {"levels": levels, "occ": (occ.points_evaluated - occ0) / K}
    if sf is not None:
        sf.time_hash_bwd = False
        acc1 = sf.counts_total()
        n0, ns, ne0, ne = ((a - b) / K for a, b in zip(acc1, acc0))
        counts.update(n0=n0, ns=ns, ne0=ne0, ne=ne)
        res["capped_rays"] = sf.capped_rays - capped0
        res["skipped_iterations"] = sf.skipped_iterations - skipped0
        timed = sf.hash_bwd_timings() if time_hash else []
        res["hash_launches"] = [(ms_, n, la) for ms_, n, la in timed]
    else:
        counts.update(n0=float("nan"), ns=samples / K, ne0=float("nan"), ne=float("nan"))
        if timer is not None:
            _lib.stop_timing("sn_hashgrid_bwd")
            durs = timer.durations_ms()
            res["hash_launches"] = [(d, int(a[1]), int(a[8])) for d, (_, _, a) in
                                    zip(durs, timer.records)]
    res["counts"] = counts
    return runner, res


def accuracy(runner, resolution=512, rays_per_view=4000, net=None):
    """Chamfer distance / F-score@0.005 of a trained field the reference's way
    (exp_runner.py:579-630): marching cubes of the SDF on a resolution^3 grid
    of the object's bounding box (renderer.py:27-34), first hits of masked
    training-pixel rays with that mesh, against the true surface's first hits
    along the training pixels (cd_and_fscore.py, exact fp64 nearest
    neighbours)."""
    from supernormal_amd.mesh import (chamfer_distance_and_f1_score, extract_fields_device,
                                      marching_cubes_device, visible_mesh_points)
    net = net or runner.sdf_network
    ds = runner.dataset
    bmin, bmax = list(ds.object_bbox_min), list(ds.object_bbox_max)
    u = extract_fields_device(bmin, bmax, resolution, lambda x: -net.fused_sdf(x), ds.device)
    mesh = marching_cubes_device(-u, 0.0, bmin, bmax, vertex_ids=False)
    del u
    pts = visible_mesh_points(mesh, ds, rays_per_view=rays_per_view)
    gt = ds.surface_points(rays_per_view * len(ds.train_images))
    cd, fs = chamfer_distance_and_f1_score(gt, pts, f_threshold=0.005)
    return {"chamfer": round(cd, 7), "fscore@0.005": round(fs, 4), "n_eval_points": int(len(pts)),
            "mesh_triangles": int(mesh.tris.shape[0]), "mesh_resolution": resolution}


def throughput_fields(runner, res, precision):
    """samples/s, SDF evaluations/s, MLP FLOP rate and MFMA utilisation, step
    GB/s of the timed window."""
    c = res["counts"]
    sec = res["ms_per_step"] * 1e-3
    cfg = runner.sdf_network.sdf_config()
    grid = runner.sdf_network.encoding.grid
    P = runner.patch_size ** 2
    evals = c["ne"] + c["ne0"] + c["occ"]
    flops = mlp_flops_fwd(cfg) * (c["ne0"] + c["occ"] + 3 * c["ne"])
    n_params = runner.optimizer.n
    res3 = int(np.prod(runner.renderer.occupancy_grid.binary.shape))
    tb = 2 if precision == "fp16" else 4
    nbytes = iteration_bytes(c, grid, n_params, tb, res3, runner.batch_size, P)
    return {
        "marched_centre_samples_per_iter": round(c["n0"], 1),
        "visible_centre_samples_per_iter": round(c["ns"], 1),
        "patch_samples_per_iter": round(c["ns"] * P, 1),
        "sdf_evals_per_iter": round(evals, 1),
        "samples_per_s": round(c["ns"] * P / sec, 1),
        "sdf_evals_per_s": round(evals / sec, 1),
        "mlp_tflops": round(flops / sec / 1e12, 3),
        "mfma_utilisation": round(flops / sec / MFMA_PEAK[precision], 5),
        "mfma_peak_tflops": MFMA_PEAK[precision] / 1e12,
        "step_algorithmic_gb": round(nbytes / 1e9, 4),
        "step_hbm_gbs": round(nbytes / sec / 1e9, 1),
    }


def roofline(res, grid):
    la = res.get("hash_launches") or []
    if not la:
        return None
    durs = [d for d, _, _ in la]
    nbytes = [hash_bwd_bytes(n, a, grid) for _, n, a in la]
    pts = float(np.mean([n for _, n, _ in la]))
    avg_ms = float(np.mean(durs))
    avg_bytes = float(np.mean(nbytes))
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    t = HASH_BWD_TRAFFIC_PER_POINT
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": int(t["bytes_per_point"] * pts),
            "kernel": "sn_hashgrid_bwd (bucket records + fixed-point bucket sums)",
            "launches_timed": len(durs), "avg_launch_ms": round(avg_ms, 4),
            "avg_points": int(pts), "active_levels": int(la[-1][2]),
            "avg_algorithmic_bytes": int(avg_bytes),
            "traffic_source": {"probe_points": t["probe_points"],
                               "probe_active_levels": t["active_levels"],
                               "bytes_per_point": round(t["bytes_per_point"], 1)}}


def cpu_threads():
    """Host threads for the CPU legs: the cores this process may run on, capped
    by OMP_NUM_THREADS when it is set (the host share of one GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(env)) if env and env.isdigit() and int(env) > 0 else n


def cpu_sample(runner, iters=2):
    """The CPU oracle's iteration at this workload's size from the trained
    state: `iters` timed iterations after one untimed, median."""
    from oracle.train_oracle import OracleTrainer
    threads = cpu_threads()
    torch.set_num_threads(threads)
    ds = runner.dataset
    tr = OracleTrainer(runner.sdf_network, float(runner.deviation_network.variance.detach()),
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
    return {"value": round(1.0 / sec, 4), "unit": "train iters/s", "cores": threads,
            "kind": "port",
            "sample": f"{iters} iterations (median, after 1 untimed) of the CPU oracle (torch "
                      f"fp32 restatement of the reference iteration) at 2048 patches from the "
                      f"trained state of the same workload"}


def cpu_small_workload(seed=0):
    """The plain method on the CPU (torch fp32 restatement of the reference
    iteration) on the small workload (sphere r 0.5, 8 views 128^2, hash 8x2 T
    2^15, MLP 32x1, 256 patches, occupancy 64^3, <= 32 samples/ray, lr 1e-2):
    median time of iterations 5..24 (after 5 warm-up)."""
    from oracle.train_oracle import OracleTrainer
    from supernormal_amd.fields import SDFNetwork, SingleVarianceNetwork
    from supernormal_amd.runner import synthetic_conf
    from supernormal_amd.scene import config_dataset
    threads = cpu_threads()
    torch.set_num_threads(threads)
    conf = synthetic_conf("sphere128")
    torch.manual_seed(seed)
    net = SDFNetwork(**conf["model.sdf_network"], encoding_config=conf["model.encoding"])
    var = SingleVarianceNetwork(**conf["model.variance_network"])
    ds = config_dataset("sphere128", torch.device("cpu"), seed)
    tr = OracleTrainer(net, float(var.variance.detach()), 64,
                       {"lr": 1e-2, "loss_type": "l2"})
    g = torch.Generator().manual_seed(seed + 99)
    rm = conf["model.ray_marching"]
    start, slope = rm["start_step_size"], \
        (np.log10(rm["start_step_size"]) - np.log10(rm["end_step_size"])) / 200
    times = []
    for i in range(25):
        tr.occupancy_update(i, n=8, thre=0.1)
        if i % 25 == 0:
            tr.field.bindwidth += 1
        o, d, pn, vi, nt, m = ds.gen_random_patches(256, generator=g)
        near, far = ds.near_far_from_sphere(o[:, 1, 1], d[:, 1, 1])
        t0 = time.perf_counter()
        tr.step(o, d, pn, vi, nt, m, near, far, step_size=float(10 ** (np.log10(start) - slope * i)),
                lr=1e-2, cap=32)
        times.append(time.perf_counter() - t0)
    sec = float(np.median(times[5:25]))
    return {"value": round(1.0 / sec, 3), "unit": "train iters/s", "cores": threads,
            "kind": "port",
            "sample": "small CPU workload (plain method, torch fp32): sphere r 0.5, 8 views "
                      "128^2, hash 8x2 T 2^15, MLP 32x1, 256 patches, occupancy 64^3, <= 32 "
                      "samples/ray; median iteration time of iterations 5..24 after 5 warm-up "
                      "(occupancy updates outside the timed calls)"}


def fp64_gate(dev, seed=0, precisions=("fp16", "fp32"), resolution=256):
    """The device path against the FP64 oracle on the small workload (the size
    at which an FP64 run is feasible on the host: 200 iterations), on identical
    inputs: every iteration's patches, stratified start and occupancy cells /
    jitter are drawn once and fed to the CPU oracle in float64
    (oracle/train_oracle.py) and to the device iteration (graph mode, given
    inputs) in each precision.  Returns each run's Chamfer / F-score (the
    reference's evaluation, `accuracy`) and its ratio to the oracle's."""
    from oracle.train_oracle import OracleTrainer
    from supernormal_amd.runner import Runner, synthetic_conf
    torch.set_num_threads(cpu_threads())
    runs = {}
    for p in precisions:
        conf = synthetic_conf("sphere128", mlp_precision=p, step_mode="graph")
        runs[p] = Runner(conf, fused=True, device=dev, seed=seed, rng_seed=seed)
    r0 = runs[precisions[0]]
    tr = OracleTrainer(r0.sdf_network, float(r0.deviation_network.variance.detach()),
                       r0.occ_resolution, {"lr": r0.learning_rate, "loss_type": r0.loss_type},
                       dtype=torch.float64)
    ds = r0.dataset
    g = torch.Generator(device=dev).manual_seed(seed + 99)
    rng = np.random.default_rng(seed + 7)
    E = r0.end_iter
    t0 = time.perf_counter()
    for i in range(E):
        step = float(10 ** (np.log10(r0.start_step_size) - r0.slop_step * i))
        if i % r0.occ_update_freq == 0:
            idx, jit = tr.occupancy_cells(i)
            tr.occupancy_update(i, n=r0.occ_update_freq, thre=r0.occ_threshold,
                                cells=(idx, jit))
            for r in runs.values():
                if r.sf_step is not None:
                    r.sf_step.drain()
                r.renderer.occupancy_grid.update_cells(
                    torch.from_numpy(idx), torch.from_numpy(jit), r.occ_eval_fn(),
                    r.occ_threshold, 0.95)
        if i % r0.increase_bindwidth_every == 0:
            tr.field.bindwidth += 1
            for r in runs.values():
                r.sdf_network.increase_bandwidth()
        o, d, pn, vi, nt, m = ds.gen_random_patches(r0.batch_size, generator=g)
        near, far = ds.near_far_from_sphere(o[:, 1, 1], d[:, 1, 1])
        u = torch.from_numpy(rng.random(near.shape[0], dtype=np.float32)).to(dev)
        t_min = near + u * step
        for r in runs.values():
            r.iter_step = i
            r.update_learning_rate()
            r.sf_step.step(r.lr, step, use_graph=True,
                           inputs=(o, d, pn, vi, nt, m, t_min, far))
        cpu = [t.cpu() for t in (o, d, pn, vi, nt, m, near, far)]
        tr.step(*cpu, step_size=step, lr=r0.lr, cap=r0.max_samples_per_ray, t_min=t_min.cpu())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {}
    for p, r in runs.items():
        out[p] = accuracy(r, resolution=resolution)
    net = r0.sdf_network
    with torch.no_grad():
        net.encoding.params.copy_(tr.field.params.detach().float())
        for l in range(net.num_layers - 1):
            lin = getattr(net, f"lin{l}")
            lin.weight_v.copy_(tr.field.v[l].detach().float())
            lin.weight_g.copy_(tr.field.g[l].detach().float())
            lin.bias.copy_(tr.field.b[l].detach().float())
    net.bindwidth = tr.field.bindwidth
    out["fp64_oracle"] = accuracy(r0, resolution=resolution)
    ref = out["fp64_oracle"]["chamfer"]
    for p in precisions:
        out[p]["chamfer_ratio_to_fp64_oracle"] = round(out[p]["chamfer"] / ref, 4)
        out[p]["within_2pct"] = bool(abs(out[p]["chamfer"] / ref - 1.0) <= 0.02)
    out["workload"] = ("small workload (sphere r 0.5, 8 views 128^2, hash 8x2 T 2^15, MLP 32x1, "
                       f"256 patches, occupancy 64^3, <= 32 samples/ray, {E} iterations), "
                       f"identical inputs, mesh resolution {resolution}")
    out["wall_s"] = round(wall, 1)
    return out


def plain_from_state(runner, iters):
    """The reference-structured path (fused=False) from the runner's trained
    state, and the fused device path over the same number of iterations."""
    runner.fused = False
    torch.cuda.synchronize()
    it0 = runner.iter_step
    runner.train_iteration(it0)  # untimed
    torch.cuda.synchronize()
    t = time.perf_counter()
    for i in range(iters):
        runner.train_iteration(it0 + 1 + i)
    torch.cuda.synchronize()
    plain = iters / (time.perf_counter() - t)
    runner.fused = True
    t = time.perf_counter()
    for i in range(iters):
        runner.train_iteration(it0 + 1 + iters + i, ahead=i + 1 < iters)
    torch.cuda.synchronize()
    fused_same = iters / (time.perf_counter() - t)
    return {"value": round(plain, 2), "unit": "train iters/s",
            "path": "reference-structured (autograd over drop-in kernels, torch Linear MLP, "
                    "fp32), from the trained state",
            "device_same_state": round(fused_same, 2), "speedup": round(fused_same / plain, 2)}


def plain_full(dev, workload, seed, max_samples, log=None):
    """The reference-structured path trained over the whole schedule from the
    same seed (same initial weights and views): its iterations/s and Chamfer."""
    from supernormal_amd.runner import Runner, synthetic_conf
    conf = synthetic_conf(workload, mlp_precision="fp32", step_mode="fused",
                          max_samples_per_ray=max_samples)
    r = Runner(conf, fused=False, device=dev, seed=seed)
    r.update_learning_rate()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for i in range(r.end_iter):
        r.train_iteration(i)
        if log and (i + 1) % 1000 == 0:
            log(f"plain {workload}: iteration {i + 1}")
    torch.cuda.synchronize()
    sec = time.perf_counter() - t
    return {"full_schedule": round(r.end_iter / sec, 2), "wall_s": round(sec, 1),
            "accuracy": accuracy(r)}


def main(argv=None):
    import faulthandler
    faulthandler.enable()  # a crash prints the Python stack of every thread
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed iterations (end of schedule)")
    ap.add_argument("--warmup", type=int, default=20, help="untimed iterations before them")
    ap.add_argument("--workload", default="tbs612", choices=list(WORKLOADS))
    ap.add_argument("--method", default="dfd", choices=["dfd", "ad"])
    ap.add_argument("--mode", default="graph", choices=["fused", "sync_free", "graph"])
    ap.add_argument("--precision", default="fp16", choices=["fp32", "fp16"])
    ap.add_argument("--seed", type=int, default=0,
                    help="seed of the initial weights and of the synthetic views' noise")
    ap.add_argument("--max-samples", type=int, default=None,
                    help="samples per centre ray capacity (default: the workload's)")
    ap.add_argument("--plain-iters", type=int, default=20)
    ap.add_argument("--plain-full", action="store_true",
                    help="also train the reference-structured path over the whole schedule")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-small", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-ad", action="store_true",
                    help="skip the leg that trains the same schedule with the analytic "
                         "gradient (method ad)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-gate", action="store_true",
                    help="skip the small workload's run against the FP64 oracle")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="after the timed run: 5 hash-grid backward calls on the last "
                         "iteration's points and 3 copies of 64 MiB (calibration), for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); this build runs on a "
              "single GPU and produces no multi-GPU figure.", file=sys.stderr)
        return 2
    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    mode = a.mode
    t_bench = time.perf_counter()
    runner, res = run_schedule(dev, a.workload, a.precision, a.method, mode, a.seed,
                               a.warmup, a.steps, a.max_samples, log=log)
    log(f"device path: {res['value']:.1f} it/s in the timed window, "
        f"{res['full_schedule']:.1f} it/s over the schedule")
    grid = runner.sdf_network.encoding.grid
    prec_name = "fp16" if a.precision == "fp16" else "fp32"
    out = {
        "metric": "train iters/s at 18k rays vs plain GPU @ equal Chamfer; p50/p99 job time-to-mesh",
        "value": round(res["value"], 2), "unit": "train iters/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(res["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if a.precision == "fp32" else "fp16/bf16 MFMA + f32",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {WORKLOADS[a.workload]}, "
                               f"{runner.max_samples_per_ray} samples/ray "
                               f"capacity, gradient {a.method}, mode {mode}, mlp {a.precision}; "
                               f"timed: the last {a.steps} of {res['iterations']} iterations "
                               f"after {a.warmup} untimed",
                   "global_batch": runner.batch_size * runner.patch_size ** 2,
                   "iterations": res["iterations"], "parallelism": "single GPU"},
        "timed_window": {"first_iteration": res["iterations"] - a.steps,
                         "active_levels": res["levels"], "step_size": res["step_size"],
                         "capped_rays": res.get("capped_rays"),
                         "skipped_iterations": res.get("skipped_iterations")},
        "full_schedule": {"value": round(res["full_schedule"], 2), "unit": "train iters/s",
                          "iterations": res["iterations"]},
        "throughput": throughput_fields(runner, res, prec_name),
        "wall_s": round(res["wall_s"], 3),
        "roofline": roofline(res, grid),
    }
    if a.traffic_probe and runner.sf_step is not None:
        sf = runner.sf_step
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
        out["accuracy"] = accuracy(runner)
    if not a.no_plain:
        out["plain_gpu"] = plain_from_state(runner, a.plain_iters)
        log(f"plain path from the trained state: {out['plain_gpu']['value']} it/s")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_sample(runner, a.cpu_iters)
        log(f"CPU sample: {out['cpu_baseline']['value']} it/s")
    del runner
    torch.cuda.empty_cache()
    if not a.no_fp32 and a.precision != "fp32":
        r32, res32 = run_schedule(dev, a.workload, "fp32", a.method, mode, a.seed, a.warmup,
                                  a.steps, a.max_samples, log=log)
        out["fp32"] = {"value": round(res32["value"], 2), "unit": "train iters/s",
                       "ms_per_step": round(res32["ms_per_step"], 4),
                       "full_schedule": round(res32["full_schedule"], 2),
                       "throughput": throughput_fields(r32, res32, "fp32"),
                       "roofline": roofline(res32, r32.sdf_network.encoding.grid)}
        if not a.no_accuracy:
            out["fp32"]["accuracy"] = accuracy(r32)
        log(f"fp32 leg: {res32['value']:.1f} it/s")
        del r32
        torch.cuda.empty_cache()
    if not a.no_ad and a.method != "ad":
        rad, resad = run_schedule(dev, a.workload, a.precision, "ad", mode, a.seed, a.warmup,
                                  a.steps, a.max_samples, log=log)
        out["method_ad"] = {"value": round(resad["value"], 2), "unit": "train iters/s",
                            "ms_per_step": round(resad["ms_per_step"], 4),
                            "full_schedule": round(resad["full_schedule"], 2),
                            "note": "same schedule and seed, rendered normal = analytic "
                                    "x-gradient of the SDF (renderer.py:225-226), its double "
                                    "backward forward-over-reverse on the fp16/bf16 MLP "
                                    "kernels (fp32 ones with --precision fp32)"}
        if not a.no_accuracy:
            out["method_ad"]["accuracy"] = accuracy(rad)
        log(f"method ad leg: {resad['value']:.1f} it/s")
        del rad
        torch.cuda.empty_cache()
    if a.plain_full:
        pf = plain_full(dev, a.workload, a.seed, a.max_samples, log=log)
        out.setdefault("plain_gpu", {})["full_schedule_run"] = pf
        if "accuracy" in out:
            out["plain_gpu"]["device_vs_plain_full_schedule"] = round(
                res["full_schedule"] / pf["full_schedule"], 2)
            out["plain_gpu"]["chamfer_ratio_device_to_plain"] = round(
                out["accuracy"]["chamfer"] / pf["accuracy"]["chamfer"], 4)
    if not a.no_cpu and not a.no_cpu_small:
        out["cpu_baseline"]["small_workload"] = cpu_small_workload(a.seed)
        log(f"CPU small workload: {out['cpu_baseline']['small_workload']['value']} it/s")
    if not a.no_gate:
        out["fp64_gate"] = fp64_gate(dev, a.seed)
        log(f"FP64 gate: {json.dumps(out['fp64_gate'])}")
    out["bench_wall_s"] = round(time.perf_counter() - t_bench, 1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
