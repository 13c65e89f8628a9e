"""One training iteration of the device path (SyncFreeStep replayed as HIP
graphs, the path the throughput figures come from) against the CPU oracle's
iteration (oracle/train_oracle.py: the reference's iteration restated in torch,
field in fp64) on identical patches, stratified jitter, occupancy grid and
parameters, at the full-size workload's shapes: 2048 patches of 3x3 rays,
occupancy grid 128^3, hash grid 16 levels x 2 features (all active), T = 2^19,
MLP 64 x 2 hidden.  The state is the one the throughput figure is measured in:
the full 5000-iteration schedule trained on the device first.

Bit-exact (integer / index work and the fp32 march recurrence):
  the march: packed_info, ray indices, t_starts, t_ends (ray_marching.cu);
  the visible samples after the visibility filter: packed_info (pack.py),
  ray indices, t_starts, t_ends.
The keep decision of the visibility filter (T >= 1e-8, renderer.py:133) is
compared with the oracle's own; it may differ only for samples whose
transmittance lies within a stated band around 1e-8 (the SDF in fp16 / fp32
against fp64: |ln T - ln 1e-8| < 0.05 for fp32, < 4 for the fp16 MLP, whose
SDF errors of ~2e-4 are multiplied by inv_s in every alpha of the ray and add
up along it), and the oracle then continues from the device's decision.
Within stated tolerances (fp32 / fp16 MLP with fp16 table copy):
  the normal, mask and eikonal losses (relative);
  the gradient of every parameter tensor: relative 2-norm of the difference.
Measured relative gradient errors: fp32 1.5e-5 to 1e-4 per tensor; fp16
4e-2 to 1.0e-1 per MLP tensor and 3.8e-2 to 4.5e-2 for the table over two
trained states (the fp16 forward's 11-bit and the bf16 backward's 8-bit
significands, summed over ~1.2 M points whose terms largely cancel in the
weight gradients), hence 2e-1 for the fp16 MLP tensors; a wrong term moves a
tensor's gradient by O(1).
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

# (keep band in |ln T - ln 1e-8|, loss rtol, gradient relative-norm tolerance
#  for the hash table and for the MLP / variance tensors)
TOL = {"fp32": dict(band=0.05, loss=2e-3, table=2e-2, mlp=2e-2),
       "fp16": dict(band=4.0, loss=3e-2, table=1e-1, mlp=2e-1)}


def _trained_runner(dev, precision):
    from supernormal_amd.runner import Runner, synthetic_conf
    conf = synthetic_conf("tbs612", mlp_precision=precision, step_mode="graph")
    r = Runner(conf, fused=True, device=dev, seed=0, rng_seed=0)
    r.update_learning_rate()
    for i in range(r.end_iter):
        r.train_iteration(i, ahead=i + 1 < r.end_iter)
    torch.cuda.synchronize()
    assert r.sdf_network.active_levels() == 16
    return r


def _given_inputs(r, dev, step):
    g = torch.Generator(device=dev).manual_seed(1234)
    o, d, pn, vi, nt, m = r.dataset.gen_random_patches(r.batch_size, generator=g)
    near, far = r.dataset.near_far_from_sphere(o[:, 1, 1], d[:, 1, 1])
    u = torch.rand(near.shape, generator=g, device=dev, dtype=torch.float32)
    t_min = near + u * step
    return o, d, pn, vi, nt, m, t_min, far


def _device_grads(r):
    net, var = r.sdf_network, r.deviation_network
    out = {"params": net.encoding.params.grad}
    for l in range(net.num_layers - 1):
        lin = getattr(net, f"lin{l}")
        out[f"lin{l}.weight_v"] = lin.weight_v.grad
        out[f"lin{l}.weight_g"] = lin.weight_g.grad
        out[f"lin{l}.bias"] = lin.bias.grad
    out["variance"] = var.variance.grad
    return {k: v.detach().double().cpu().reshape(-1) for k, v in out.items()}


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_device_iteration_matches_oracle(dev, precision):
    from oracle.train_oracle import OracleTrainer
    tol = TOL[precision]
    r = _trained_runner(dev, precision)
    sf = r.sf_step
    step = float(r.renderer.sampling_step_size)
    inputs = _given_inputs(r, dev, step)
    tr = OracleTrainer(r.sdf_network, float(r.deviation_network.variance.detach()),
                       r.occ_resolution, {"lr": r.lr, "loss_type": r.loss_type},
                       dtype=torch.float64)
    tr.binary = r.renderer.occupancy_grid.binary.cpu().numpy().astype(np.uint8)
    sf.step(r.lr, step, use_graph=True, inputs=inputs)
    torch.cuda.synchronize()
    n0, ns = int(sf.n0), int(sf.ns)
    cpu = [t.cpu() for t in inputs]
    keep_dev = sf.keep[:n0].cpu().numpy().astype(bool)
    res = tr.forward_backward(*cpu, step, cap=r.max_samples_per_ray, keep=keep_dev)
    assert res is not None
    print(f"\n[{precision}] marched {n0}, visible {ns}, patch points {int(sf.ne)}", flush=True)
    # the march: bit-exact
    assert n0 == res["ri0"].size
    np.testing.assert_array_equal(sf.packed0.cpu().numpy(), res["packed0"])
    np.testing.assert_array_equal(sf.ri0[:n0].cpu().numpy(), res["ri0"])
    np.testing.assert_array_equal(sf.t00[:n0].cpu().numpy(), res["t00"][:, 0])
    np.testing.assert_array_equal(sf.t10[:n0].cpu().numpy(), res["t10"][:, 0])
    # the visibility decision: equal outside the band around T = 1e-8
    diff = keep_dev != res["keep_oracle"]
    lnT = np.log(np.maximum(res["T0"], 1e-300))
    near_thr = np.abs(lnT - math.log(1e-8)) < tol["band"]
    dist = np.abs(lnT - math.log(1e-8))
    print(f"[{precision}] keep decisions differing: {int(diff.sum())} of {n0} "
          f"({int(near_thr.sum())} samples in the band); largest |ln T - ln 1e-8| of a "
          f"differing decision {float(dist[diff].max()) if diff.any() else 0.0:.3f}", flush=True)
    assert not (diff & ~near_thr).any()
    # the visible samples: bit-exact
    assert ns == res["ri"].size
    np.testing.assert_array_equal(sf.packed.cpu().numpy(), res["packed"])
    np.testing.assert_array_equal(sf.ri[:ns].cpu().numpy(), res["ri"])
    np.testing.assert_array_equal(sf.t0[:ns].cpu().numpy(), res["t0"][:, 0])
    np.testing.assert_array_equal(sf.t1[:ns].cpu().numpy(), res["t1"][:, 0])
    # losses
    L = sf.losses.cpu().double().numpy()
    for k, name in enumerate(("normal", "mask", "eikonal")):
        ref = res["parts"][name]
        rel = abs(L[k] - ref) / abs(ref)
        print(f"[{precision}] loss {name}: device {L[k]:.8e} oracle {ref:.8e} rel {rel:.2e}",
              flush=True)
        assert rel <= tol["loss"], name
    # gradients of every parameter tensor
    gd, go = _device_grads(r), {k: v.double().reshape(-1) for k, v in tr.gradients().items()}
    for k in go:
        rel = float((gd[k] - go[k]).norm() / go[k].norm())
        print(f"[{precision}] grad {k}: |g| {float(go[k].norm()):.4e} rel err {rel:.2e}",
              flush=True)
        assert rel <= (tol["table"] if k == "params" else tol["mlp"]), k
