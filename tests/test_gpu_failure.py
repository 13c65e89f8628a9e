"""Failure handling and resume of the device iteration (SyncFreeStep, HIP
graphs), every condition forced from the host so that every kernel exits
normally:
  * kill + resume from a checkpoint => identical continuation: the run that
    saved and went on, and a fresh Runner that loaded the checkpoint, end with
    bit-identical parameters, moments and step count (exp_runner.py:298-315
    saves the keys; the occupancy grid, bandwidth and the samplers' RNG are
    saved too);
  * an empty occupancy grid => no sample => the iteration is skipped:
    parameters, moments and the Adam step count unchanged, the skip counted
    (the reference raises KeyError, renderer.py:137-140 vs exp_runner.py:182);
  * a NaN in the iteration's inputs => non-finite loss => skipped and counted,
    and the next iteration on clean inputs updates as usual.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _runner(dev, base, mode="graph", is_continue=False, precision="fp32"):
    from supernormal_amd.runner import Runner, synthetic_conf
    conf = synthetic_conf("sphere128", mlp_precision=precision, step_mode=mode)
    conf["val"]["save_freq"] = 10 ** 9
    return Runner(conf, fused=True, device=dev, seed=0, rng_seed=0, base_exp_dir=str(base),
                  is_continue=is_continue)


def _state(r):
    f = r.optimizer
    if r.sf_step is not None:
        r.sf_step.drain()
        r.sf_step.sync_step_count()
    torch.cuda.synchronize()
    return (f.data.clone(), f.exp_avg.clone(), f.exp_avg_sq.clone(), int(f.step_count),
            r.iter_step, r.sdf_network.bindwidth,
            r.renderer.occupancy_grid.binary.clone())


@pytest.mark.parametrize("mode,precision", [("graph", "fp16"), ("graph", "fp32"),
                                            ("fused", "fp32")])
def test_resume_continues_identically(dev, tmp_path, mode, precision):
    a = _runner(dev, tmp_path, mode, precision=precision)
    a.train(n_iters=45, log=None)       # past bandwidth steps, occupancy updates
    a.save_checkpoint()
    a.train(n_iters=37, log=None)
    sa = _state(a)
    names = os.listdir(tmp_path / "checkpoints")
    assert names == ["ckpt_000045.pth"]
    del a
    torch.cuda.empty_cache()
    b = _runner(dev, tmp_path, mode, is_continue=True, precision=precision)
    assert b.iter_step == 45
    b.train(n_iters=37, log=None)
    sb = _state(b)
    print(f"[{mode}/{precision}] iter {sa[4]}, adam t {sa[3]}, bandwidth {sa[5]}", flush=True)
    for x, y, name in zip(sa, sb, ("params", "exp_avg", "exp_avg_sq", "step", "iter",
                                   "bandwidth", "occupancy")):
        if isinstance(x, torch.Tensor):
            assert torch.equal(x, y), name
        else:
            assert x == y, name


def _given(r, dev, seed=5, nan=False):
    g = torch.Generator(device=dev).manual_seed(seed)
    o, d, pn, vi, nt, m = r.dataset.gen_random_patches(r.batch_size, generator=g)
    if nan:
        nt = nt.clone()
        nt[3, 1, 1, 0] = float("nan")
    return (o, d, pn, vi, nt, m)


def test_empty_grid_skips_the_step(dev, tmp_path):
    r = _runner(dev, tmp_path)
    r.train(n_iters=20, log=None)
    sf = r.sf_step
    sf.drain()
    before = _state(r)
    skipped0 = sf.skipped_iterations
    occ = r.renderer.occupancy_grid
    occ.binary.zero_()
    sf.step(r.lr, float(r.renderer.sampling_step_size), use_graph=True,
            inputs=_given(r, dev))
    torch.cuda.synchronize()
    assert int(sf.n0) == 0 and int(sf.ns) == 0
    after = _state(r)
    assert sf.skipped_iterations == skipped0 + 1
    assert sf.nonfinite_iterations == 0
    for x, y in zip(before[:4], after[:4]):
        assert (torch.equal(x, y) if isinstance(x, torch.Tensor) else x == y)


def test_nan_input_skips_the_step(dev, tmp_path):
    r = _runner(dev, tmp_path)
    r.train(n_iters=30, log=None)
    sf = r.sf_step
    sf.drain()
    before = _state(r)
    skipped0 = sf.skipped_iterations
    step = float(r.renderer.sampling_step_size)
    sf.step(r.lr, step, use_graph=True, inputs=_given(r, dev, nan=True))
    torch.cuda.synchronize()
    assert int(sf.ns) > 0
    assert not torch.isfinite(sf.losses[:3]).all()
    after = _state(r)
    assert sf.nonfinite_iterations == 1 and sf.skipped_iterations == skipped0 + 1
    for x, y in zip(before[:4], after[:4]):
        assert (torch.equal(x, y) if isinstance(x, torch.Tensor) else x == y)
    # the same inputs without the NaN: an ordinary update
    sf.step(r.lr, step, use_graph=True, inputs=_given(r, dev))
    torch.cuda.synchronize()
    assert torch.isfinite(sf.losses[:3]).all()
    again = _state(r)
    assert again[3] == before[3] + 1
    assert not torch.equal(again[0], before[0])
    assert torch.isfinite(again[0]).all()
    assert sf.nonfinite_iterations == 1
