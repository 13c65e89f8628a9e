"""The sync-free iteration (SyncFreeStep) against FusedStep on the same
patches, and its HIP-graph replays against eager runs of the same inputs.

FusedStep reads the sample counts back to the host; SyncFreeStep keeps them on
the device and sizes launches by capacity.  On identical inputs (unjittered
starts) both must select the same samples (bit-exact counts) and give the same
loss and gradients up to the order of fp32 atomics (rtol 1e-4).  Graph replays
are checked step by step against an eager rerun from the same state on the
replay's own inputs, with eager work and synchronisations between replays.
"""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _runner(dev, precision="fp32", step_mode="sync_free"):
    from supernormal_amd.runner import Runner, synthetic_conf
    conf = synthetic_conf("sphere128", max_samples_per_ray=32, mlp_precision=precision,
                          step_mode=step_mode)
    r = Runner(conf, fused=True, device=dev, seed=3)
    r.update_learning_rate()
    # a few ordinary iterations so that the occupancy grid and the field are not trivial
    for i in range(12):
        r.train_iteration(i, ahead=i < 11)  # no march left pending
    torch.cuda.synchronize()
    return r


def _state(flat):
    return [t.clone() for t in (flat.data, flat.exp_avg, flat.exp_avg_sq)]


def _device_packing(r, sf, monkeypatch):
    """FusedStep with the packed MLP weights and inv_s of the device kernels
    (sn_mlp_pack_fwd) in value, torch autograd for their backward: the dfd
    normals difference SDF values of neighbouring points, which turns
    one-ulp differences of the weights (another norm summation order) into
    1e-3 relative differences of some gradients."""
    from supernormal_amd._lib import call, ptr, stream_of
    net = r.sdf_network
    orig_pack, orig_inv = net.packed_mlp, r.fused_step.inv_s

    def dev_values():
        packed = torch.empty_like(sf.mlp_packed)
        inv = torch.empty(1, device=sf.dev)
        call("sn_mlp_pack_fwd", ptr(sf.flat.data), sf.n_pack, sf.pack_layers.ctypes.data,
             ptr(packed), ptr(sf.var_data), ptr(inv), stream_of(sf.dev))
        return packed, inv

    def packed_mlp():
        m = orig_pack()
        return m + (dev_values()[0] - m).detach()

    def inv_s():
        t = orig_inv()
        return t + (dev_values()[1] - t).detach()

    monkeypatch.setattr(net, "packed_mlp", packed_mlp)
    monkeypatch.setattr(r.fused_step, "inv_s", inv_s)


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_sync_free_matches_fused_step(dev, precision, monkeypatch):
    r = _runner(dev, precision)
    fs, flat = r.fused_step, r.optimizer
    # FusedStep gathers from the fp32 table: compare without the fp16 table copy
    from supernormal_amd.step import SyncFreeStep
    sf = SyncFreeStep(r.sdf_network, r.deviation_network, r.renderer.occupancy_grid, flat,
                      r.dataset, r.batch_size, r.patch_size, r.max_samples_per_ray,
                      r.normal_weight, r.eikonal_weight, r.mask_weight, r.loss_type,
                      precision=precision, device=dev, half_table=False)
    _device_packing(r, sf, monkeypatch)
    g = torch.Generator(device=dev).manual_seed(7)
    o, d, pn, vi, nt, m = r.dataset.gen_random_patches(r.batch_size, generator=g)
    near, far = r.dataset.near_far_from_sphere(o[:, 1, 1], d[:, 1, 1])
    step = 5e-3
    # FusedStep (host-synchronised sizes, autograd)
    flat.zero_grad()
    res = fs.forward_backward(o, d, pn, vi, nt, m, near, far, step, stratified=False)
    flat.collect()
    g_ref = flat.grad.clone()
    loss_ref = float(res.loss)
    # SyncFreeStep on the same patches (it also takes the Adam step)
    before = _state(flat)
    losses = sf.step(r.lr, step, use_graph=False, inputs=(o, d, pn, vi, nt, m))
    torch.cuda.synchronize()
    assert int(sf.ns) == res.n_samples
    assert int(sf.n0) == res.n_marched
    loss = float(losses[0] + losses[1] + losses[2])
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref) + 1e-7
    scale = float(g_ref.abs().max())
    torch.testing.assert_close(flat.grad, g_ref, rtol=1e-4, atol=1e-5 * scale)
    # the Adam step matches FlatParams.adam on the same gradient
    after = flat.data.clone()
    flat.data.copy_(before[0])
    flat.exp_avg.copy_(before[1])
    flat.exp_avg_sq.copy_(before[2])
    flat.step_count -= 1
    flat.adam(r.lr)
    torch.testing.assert_close(after, flat.data, rtol=1e-6, atol=1e-7)


def test_graph_replays_match_eager(dev):
    r = _runner(dev, "fp16", "graph")
    sf, flat = r.sf_step, r.optimizer
    lr, step = r.lr, 5e-3
    sf.step(lr, step, use_graph=True)          # capture + first replay
    for k in range(3):
        torch.cuda.synchronize()
        before = _state(flat)
        t_before = sf.adam_t.clone()
        losses_g = sf.step(lr, step, use_graph=True).clone()
        after_g = _state(flat)
        torch.cuda.synchronize()
        # eager work between replays, then the same step eagerly on the replay's inputs
        for t, v in zip((flat.data, flat.exp_avg, flat.exp_avg_sq), before):
            t.copy_(v)
        sf.refresh_copies()  # the fp16 table copy follows the restored parameters
        sf.adam_t.copy_(t_before)
        flat.step_count -= 1
        losses_e = sf.step(lr, step, use_graph=False, inputs="keep").clone()
        torch.cuda.synchronize()
        torch.testing.assert_close(losses_g, losses_e, rtol=1e-4, atol=1e-7)
        for a, b in zip(after_g, _state(flat)):
            torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-6)


def _seeded_runner(dev):
    from supernormal_amd.runner import Runner, synthetic_conf
    conf = synthetic_conf("sphere128", max_samples_per_ray=32, mlp_precision="fp16",
                          step_mode="graph")
    r = Runner(conf, fused=True, device=dev, seed=3, rng_seed=11)
    r.update_learning_rate()
    return r


def test_march_ahead_matches_in_order(dev):
    """Runs the same seeded training with the next iteration's march overlapped
    (graph replays on a side stream) and strictly in order.  Sampling draws
    from the run's own generator in the same order either way, so the samples
    are the same and losses / counts agree up to the order of fp32 atomics.
    Iterations 0..7 lie between two occupancy updates (0 and 8), so iterations
    1..7 are marched ahead and nothing is pending after iteration 7."""
    runs = []
    for ahead in (None, False):
        r = _seeded_runner(dev)
        losses, counts = [], []
        for i in range(8):
            losses.append(r.train_iteration(i, ahead=ahead).clone())
            counts.append(r.sf_step.ns.clone())
            assert r.sf_step.march_pending == (ahead is None and i < 7)
        torch.cuda.synchronize()
        assert not r.sf_step.march_pending  # none past the last iteration
        runs.append(([float(x) for x in losses], [int(x) for x in counts],
                     r.optimizer.data.clone()))
    (la, ca, pa), (lb, cb, pb) = runs
    for a, b in zip(la, lb):
        assert abs(a - b) <= 1e-3 * abs(b) + 1e-7
    for a, b in zip(ca, cb):
        assert abs(a - b) <= 1e-3 * b
    scale = float(pb.abs().max())
    assert float((pa - pb).abs().max()) <= 1e-3 * scale


def test_march_ahead_step_size_must_match(dev):
    r = _seeded_runner(dev)
    sf = r.sf_step
    r.update_learning_rate()
    sf.step(r.lr, 5e-3, use_graph=True, next_step_size=4e-3)
    assert sf.march_pending
    with pytest.raises(ValueError):
        sf.step(r.lr, 5e-3, use_graph=True)
    with pytest.raises(RuntimeError):
        sf.step(r.lr, 4e-3, use_graph=False)
    sf.step(r.lr, 4e-3, use_graph=True)
    torch.cuda.synchronize()
    assert not sf.march_pending


def _runner_ad(dev, step_mode):
    from supernormal_amd.runner import Runner, synthetic_conf
    conf = synthetic_conf("sphere128", gradient_method="ad", max_samples_per_ray=32,
                          mlp_precision="fp32", step_mode=step_mode)
    r = Runner(conf, fused=True, device=dev, seed=3)
    r.update_learning_rate()
    for i in range(12):
        r.train_iteration(i, ahead=i < 11)
    torch.cuda.synchronize()
    return r


def test_sync_free_ad_matches_fused_step(dev, monkeypatch):
    """'ad' (analytic normals, double backward through the x-gradient) on the
    device iteration against FusedStep's autograd over the same kernels."""
    r = _runner_ad(dev, "sync_free")
    fs, flat = r.fused_step, r.optimizer
    from supernormal_amd.step import SyncFreeStep
    sf = SyncFreeStep(r.sdf_network, r.deviation_network, r.renderer.occupancy_grid, flat,
                      r.dataset, r.batch_size, r.patch_size, r.max_samples_per_ray,
                      r.normal_weight, r.eikonal_weight, r.mask_weight, r.loss_type,
                      precision="fp32", device=dev, half_table=False, method="ad")
    _device_packing(r, sf, monkeypatch)
    g = torch.Generator(device=dev).manual_seed(8)
    o, d, pn, vi, nt, m = r.dataset.gen_random_patches(r.batch_size, generator=g)
    near, far = r.dataset.near_far_from_sphere(o[:, 1, 1], d[:, 1, 1])
    step = 5e-3
    flat.zero_grad()
    res = fs.forward_backward(o, d, pn, vi, nt, m, near, far, step, stratified=False)
    flat.collect()
    g_ref = flat.grad.clone()
    loss_ref = float(res.loss)
    losses = sf.step(r.lr, step, use_graph=False, inputs=(o, d, pn, vi, nt, m))
    torch.cuda.synchronize()
    assert int(sf.ns) == res.n_samples and int(sf.n0) == res.n_marched
    loss = float(losses[0] + losses[1] + losses[2])
    print(f"\nad: loss {loss:.8e} fused {loss_ref:.8e}", flush=True)
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref) + 1e-7
    scale = float(g_ref.abs().max())
    torch.testing.assert_close(flat.grad, g_ref, rtol=1e-4, atol=1e-5 * scale)


def test_graph_ad_replays_match_eager(dev):
    r = _runner_ad(dev, "graph")
    sf, flat = r.sf_step, r.optimizer
    lr, step = r.lr, 5e-3
    g = torch.Generator(device=dev).manual_seed(9)
    for k in range(3):
        inputs = r.dataset.gen_random_patches(r.batch_size, generator=g)
        before = _state(flat)
        t0 = int(flat.step_count)
        sf.step(lr, step, use_graph=True, inputs=inputs)
        torch.cuda.synchronize()
        after, g_graph = _state(flat), flat.grad.clone()
        lg = sf.losses.clone()
        for t, b in zip((flat.data, flat.exp_avg, flat.exp_avg_sq), before):
            t.copy_(b)
        flat.step_count = t0
        sf.adam_t.fill_(t0)
        sf._adam_t_host = t0
        sf.step(lr, step, use_graph=False, inputs=inputs)
        torch.cuda.synchronize()
        torch.testing.assert_close(sf.losses, lg, rtol=1e-5, atol=1e-7)
        scale = float(g_graph.abs().max())
        torch.testing.assert_close(flat.grad, g_graph, rtol=1e-4, atol=1e-5 * scale)
        for t, a in zip((flat.data, flat.exp_avg, flat.exp_avg_sq), after):
            torch.testing.assert_close(t, a, rtol=1e-4, atol=1e-6)
