"""GPU parity of the nerfacc-path kernels against the CPU oracle and the
reference's golden vectors.  Integer outputs bit-exact; fp32 forward outputs
bit-exact where the oracle runs the same fp32 sequence; gradients within the
stated tolerance of the fp64 oracle."""
import numpy as np
import pytest
import torch

from oracle import c_oracle

pytestmark = pytest.mark.gpu


def _t(a, dev, dt=None):
    return torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dt)


# ------------------------------------------------------------------ pack_info
def test_pack_info_golden(golden, dev):
    from supernormal_amd.nerfacc import pack_info
    g = golden["pack"]
    pi = pack_info(_t(g["ray_indices"], dev, torch.int64), n_rays=3)
    assert pi.dtype == torch.int32
    assert pi.cpu().tolist() == g["packed_info"]
    pi2 = pack_info(_t(g["ray_indices"], dev, torch.int64))
    assert pi2.cpu().tolist() == g["packed_info"]


@pytest.mark.parametrize("n,n_rays", [(0, 5), (1, 1), (1000, 37), (300000, 2048), (2_000_000, 70000)])
def test_pack_info_random_bitexact(dev, n, n_rays):
    from supernormal_amd.nerfacc import pack_info
    rng = np.random.default_rng(n)
    ri = np.sort(rng.integers(0, n_rays, n)).astype(np.int64)
    got = pack_info(_t(ri, dev), n_rays=n_rays).cpu().numpy()
    np.testing.assert_array_equal(got, c_oracle.pack_info(ri, n_rays))
    # unsorted indices: counts still exact
    rng.shuffle(ri)
    got = pack_info(_t(ri, dev), n_rays=n_rays).cpu().numpy()
    np.testing.assert_array_equal(got, c_oracle.pack_info(ri, n_rays))


# ------------------------------------------------------------------ weights
def test_weights_golden(golden, dev):
    from supernormal_amd.nerfacc import (render_weight_from_alpha,
                                         render_transmittance_from_alpha,
                                         render_weight_from_alpha_patch_based)
    g = golden["weights"]
    a = _t(np.array(g["alphas"], np.float32)[:, None], dev)
    ri = _t(g["ray_indices"], dev, torch.int64)
    w = render_weight_from_alpha(a, ray_indices=ri, n_rays=3)
    torch.testing.assert_close(w.cpu()[:, 0], torch.tensor(g["weights"]), rtol=1e-6, atol=1e-7)
    wp = render_weight_from_alpha_patch_based(a[:, :, None], ri, n_rays=3)
    assert torch.equal(wp[:, :, 0], w)
    d = golden["docstring"]
    a = _t(np.array(d["alphas"], np.float32)[:, None], dev)
    ri = _t(d["ray_indices"], dev, torch.int64)
    w = render_weight_from_alpha(a, ray_indices=ri)
    torch.testing.assert_close(w.cpu()[:, 0], torch.tensor(d["weights"]), rtol=1e-6, atol=1e-7)
    T = render_transmittance_from_alpha(a, ray_indices=ri)
    torch.testing.assert_close(T.cpu()[:, 0], torch.tensor(d["transmittance"]), rtol=1e-6,
                               atol=1e-7)


def test_weight_grads_golden(golden, dev):
    from supernormal_amd.nerfacc import render_weight_from_alpha
    g = golden["grads"]
    sig = _t(np.array(g["sigmas"], np.float32)[:, None], dev).requires_grad_(True)
    alphas = 1.0 - torch.exp(-sig * g["dt"])
    w = render_weight_from_alpha(alphas, packed_info=_t(g["packed_info"], dev, torch.int32))
    w.sum().backward()
    torch.testing.assert_close(w.detach().cpu()[:, 0], torch.tensor(g["weights"]), atol=1e-4,
                               rtol=0)
    torch.testing.assert_close(sig.grad.cpu()[:, 0], torch.tensor(g["sigmas_grad"]), atol=1e-4,
                               rtol=0)


def _rand_patches(rng, n_patches, P, max_steps, lo=0.0, hi=1.0):
    steps = rng.integers(0, max_steps + 1, n_patches)
    pi = np.stack([np.concatenate([[0], np.cumsum(steps)[:-1]]), steps], -1).astype(np.int32)
    ri = np.repeat(np.arange(n_patches), steps).astype(np.int64)
    a = rng.uniform(lo, hi, (int(steps.sum()), P)).astype(np.float32)
    return pi, ri, a


@pytest.mark.parametrize("n_patches,P,max_steps", [(1, 1, 1), (7, 9, 5), (2048, 9, 128), (300, 25, 40)])
def test_patch_weight_fwd_bitexact(dev, n_patches, P, max_steps):
    from supernormal_amd.nerfacc import render_weight_and_transmittance_from_alpha_patch_based
    rng = np.random.default_rng(n_patches * 31 + P)
    pi, ri, a = _rand_patches(rng, n_patches, P, max_steps)
    w, T = render_weight_and_transmittance_from_alpha_patch_based(
        _t(a[:, :, None], dev), _t(ri, dev), n_rays=n_patches)
    w_o, T_o = c_oracle.patch_weight_fwd(pi, a)
    np.testing.assert_array_equal(w.cpu().numpy()[:, :, 0], w_o)
    np.testing.assert_array_equal(T.cpu().numpy()[:, :, 0], T_o)


@pytest.mark.parametrize("hi", [0.95, 1.0])
def test_patch_weight_bwd_vs_fp64(dev, hi):
    """grad alpha of sum(gw w) + sum(gT T): |err| <= 1e-5 + 1e-4 |ref| against fp64;
    also within 1e-4 of the reference's divide-by-(1-alpha) formula where alpha <= 1-1e-3."""
    from supernormal_amd.nerfacc import render_weight_and_transmittance_from_alpha_patch_based
    rng = np.random.default_rng(11)
    pi, ri, a = _rand_patches(rng, 2048, 9, 64, hi=hi)
    if hi == 1.0:
        a[rng.random(a.shape) < 0.02] = 1.0
    gw = rng.normal(size=a.shape).astype(np.float32)
    gT = rng.normal(size=a.shape).astype(np.float32)
    at = _t(a[:, :, None], dev).requires_grad_(True)
    w, T = render_weight_and_transmittance_from_alpha_patch_based(at, _t(ri, dev), n_rays=2048)
    ((w * _t(gw[:, :, None], dev)).sum() + (T * _t(gT[:, :, None], dev)).sum()).backward()
    ga = at.grad.cpu().numpy()[:, :, 0]
    ref = c_oracle.patch_weight_bwd_f64(pi, a.astype(np.float64), gw.astype(np.float64),
                                        gT.astype(np.float64))
    assert np.isfinite(ga).all()
    np.testing.assert_allclose(ga, ref, rtol=1e-4, atol=1e-5)
    if hi < 1.0:
        w_o, _ = c_oracle.patch_weight_fwd(pi, a)
        ref_formula = c_oracle.patch_weight_bwd_ref(pi, a, w_o, gw)
        at2 = _t(a[:, :, None], dev).requires_grad_(True)
        from supernormal_amd.nerfacc import render_weight_from_alpha_patch_based
        w2 = render_weight_from_alpha_patch_based(at2, _t(ri, dev), n_rays=2048)
        (w2 * _t(gw[:, :, None], dev)).sum().backward()
        np.testing.assert_allclose(at2.grad.cpu().numpy()[:, :, 0], ref_formula, rtol=1e-3,
                                   atol=1e-4)


# ------------------------------------------------------------------ visibility
def test_visibility_golden(golden, dev):
    from supernormal_amd.nerfacc import render_visibility
    g = golden["visibility"]
    a = _t(np.array(g["alphas"], np.float32)[:, None], dev)
    ri = _t(g["ray_indices"], dev, torch.int64)
    for case in g["cases"]:
        vis = render_visibility(a, ray_indices=ri, n_rays=3,
                                early_stop_eps=case["early_stop_eps"],
                                alpha_thre=case["alpha_thre"])
        assert vis.cpu().tolist() == case["visibility"]


def test_visibility_random_bitexact(dev):
    from supernormal_amd.nerfacc import render_visibility
    rng = np.random.default_rng(12)
    steps = rng.integers(0, 300, 2048)
    pi = np.stack([np.concatenate([[0], np.cumsum(steps)[:-1]]), steps], -1).astype(np.int32)
    a = rng.uniform(0, 0.3, int(steps.sum())).astype(np.float32)
    for eps, thre in [(1e-8, 0.0), (1e-3, 0.05)]:
        vis = render_visibility(_t(a[:, None], dev), packed_info=_t(pi, dev), early_stop_eps=eps,
                                alpha_thre=thre)
        np.testing.assert_array_equal(vis.cpu().numpy(), c_oracle.visibility(pi, a, eps, thre))


# ------------------------------------------------------------------ accumulate
def test_accumulate_patch(dev):
    from supernormal_amd.nerfacc import accumulate_along_rays_patch_based
    rng = np.random.default_rng(13)
    pi, ri, a = _rand_patches(rng, 256, 9, 20)
    w = _t(a[:, :, None], dev).requires_grad_(True)
    v = _t(rng.normal(size=(a.shape[0], 9, 3)).astype(np.float32), dev).requires_grad_(True)
    out = accumulate_along_rays_patch_based(w, _t(ri, dev), values=v, n_patches=256)
    ref = torch.zeros(256, 9, 3, dtype=torch.float64)
    ref.index_add_(0, torch.from_numpy(ri), (w.detach().cpu().double() * v.detach().cpu().double()))
    torch.testing.assert_close(out.detach().cpu().double(), ref, rtol=1e-5, atol=1e-5)
    go = torch.randn_like(out)
    out.backward(go)
    torch.testing.assert_close(w.grad, (go[_t(ri, dev)] * v).sum(-1, keepdim=True), rtol=1e-5,
                               atol=1e-6)
    empty = accumulate_along_rays_patch_based(w[:0], _t(ri[:0], dev), n_patches=4)
    assert empty.shape == (4, 9, 1) and (empty == 0).all()


# ------------------------------------------------------------------ marching
def _scene_grid(res):
    c = (np.arange(res) + 0.5) / res * 2 - 1
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    return ((r < 0.55) & (np.abs(X) + 0.3 > np.abs(Y))).astype(np.uint8)


def _cam_rays(rng, n):
    o = np.repeat(np.array([[0.3, -0.2, -2.5]], np.float32), n, 0)
    d = rng.normal(size=(n, 3)) * 0.12 + np.array([-0.1, 0.08, 1.0])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    b = 2 * (o * d).sum(1)
    c = (o ** 2).sum(1) - 1
    disc = np.maximum(b * b - 4 * c, 0)
    near = ((-b - np.sqrt(disc)) / 2).astype(np.float32)
    far = ((-b + np.sqrt(disc)) / 2).astype(np.float32)
    return o, d, near, far


@pytest.mark.parametrize("res,step,cap", [(128, 1e-2, 0), (128, 1e-3, 0), (64, 2e-3, 128), (32, 5e-3, 7)])
def test_ray_marching_bitexact(dev, res, step, cap):
    from supernormal_amd.nerfacc import march_raw
    rng = np.random.default_rng(res + cap)
    o, d, near, far = _cam_rays(rng, 2048)
    near = near + rng.random(2048).astype(np.float32) * np.float32(step)
    grid = _scene_grid(res)
    roi = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    pi, ri, t0, t1 = march_raw(_t(o, dev), _t(d, dev), _t(near, dev), _t(far, dev), _t(roi, dev),
                               _t(grid.astype(bool), dev), step, 0.0, cap)
    pi_o, ri_o, t0_o, t1_o = c_oracle.ray_marching(o, d, near, far, roi, grid, step,
                                                   max_per_ray=cap)
    assert ri_o.size > 1000
    np.testing.assert_array_equal(pi.cpu().numpy(), pi_o)
    np.testing.assert_array_equal(ri.cpu().numpy(), ri_o)
    np.testing.assert_array_equal(t0.cpu().numpy(), t0_o)
    np.testing.assert_array_equal(t1.cpu().numpy(), t1_o)


ROIS = {"pow2": [-1, -1, -1, 1, 1, 1], "odd": [-1.05, -0.95, -1.1, 1.0, 1.15, 0.95]}


@pytest.mark.parametrize("res,step,cap,roi_name,bricks",
                         [(128, 1e-3, 128, "pow2", True), (128, 1e-3, 128, "pow2", False),
                          (64, 2e-3, 128, "odd", True), (64, 2e-3, 128, "odd", False),
                          (32, 5e-3, 7, "pow2", False), (32, 5e-3, 7, "odd", True)])
def test_ray_marching_padded_bitexact(dev, res, step, cap, roi_name, bricks):
    """One-pass padded march + gather == the oracle's count + fill, with the
    step size read from device memory; occupancy read from the byte grid or
    from 4x4x4 brick masks (sn_occ_pack_bricks), with a roi whose spans are
    powers of two (divisions done as exact multiplications) or not."""
    from supernormal_amd._lib import call, lib, ptr, stream_of
    rng = np.random.default_rng(res + cap + 1)
    n = 2048
    o, d, near, far = _cam_rays(rng, n)
    near = near + rng.random(n).astype(np.float32) * np.float32(step)
    grid = _scene_grid(res)
    roi = np.array(ROIS[roi_name], np.float32)
    od, dd, nd, fd, rd = (_t(a, dev) for a in (o, d, near, far, roi))
    gd = _t(grid.astype(np.uint8), dev)
    step_d = torch.tensor([step], dtype=torch.float32, device=dev)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    t0p = torch.full((n * cap,), -1.0, device=dev)
    t1p = torch.full((n * cap,), -1.0, device=dev)
    st = stream_of(dev)
    if bricks:
        bk = torch.empty((res // 4) ** 3, dtype=torch.int64, device=dev)
        call("sn_occ_pack_bricks", ptr(gd), res, res, res, ptr(bk), st)
        # every brick bit is its cell
        bits = ((bk.cpu().numpy().view(np.uint64)[:, None] >> np.arange(64, dtype=np.uint64)) & 1)
        b = bits.reshape(res // 4, res // 4, res // 4, 4, 4, 4)  # bx by bz | z y x
        np.testing.assert_array_equal(b.transpose(0, 5, 1, 4, 2, 3).reshape(res, res, res),
                                      grid.astype(np.uint64))
        call("sn_ray_march_padded_bricks", ptr(od), ptr(dd), ptr(nd), ptr(fd), n, ptr(rd),
             ptr(gd), ptr(bk), res, res, res, 1.0, ptr(step_d), 0.0, cap, ptr(counts), ptr(t0p),
             ptr(t1p), st)
    else:
        call("sn_ray_march_padded", ptr(od), ptr(dd), ptr(nd), ptr(fd), n, ptr(rd), ptr(gd), res,
             res, res, 1.0, ptr(step_d), 0.0, cap, ptr(counts), ptr(t0p), ptr(t1p), st)
    packed = torch.empty(n, 2, dtype=torch.int32, device=dev)
    total = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(lib().sn_scan_workspace_bytes(n), dtype=torch.uint8, device=dev)
    call("sn_counts_to_packed_info", ptr(counts), n, ptr(packed), ptr(total), ptr(ws),
         ws.numel(), st)
    ri = torch.empty(n * cap, dtype=torch.int64, device=dev)
    t0 = torch.empty(n * cap, device=dev)
    t1 = torch.empty(n * cap, device=dev)
    n_capped = torch.zeros(1, dtype=torch.int32, device=dev)
    call("sn_march_gather", ptr(packed), n, cap, ptr(t0p), ptr(t1p), ptr(ri), ptr(t0), ptr(t1),
         ptr(n_capped), st)
    m = int(total)
    pi_o, ri_o, t0_o, t1_o = c_oracle.ray_marching(o, d, near, far, roi, grid, step,
                                                   max_per_ray=cap)
    assert m == ri_o.size and m > 1000
    # rays whose count reached the cap (the ones it may have truncated)
    assert int(n_capped) == int((pi_o[:, 1] == cap).sum())
    np.testing.assert_array_equal(packed.cpu().numpy(), pi_o)
    np.testing.assert_array_equal(ri[:m].cpu().numpy(), ri_o)
    np.testing.assert_array_equal(t0[:m].cpu().numpy(), t0_o[:, 0] if t0_o.ndim == 2 else t0_o)
    np.testing.assert_array_equal(t1[:m].cpu().numpy(), t1_o[:, 0] if t1_o.ndim == 2 else t1_o)


def test_ray_marching_visibility_filter(dev):
    """ray_marching with alpha_fn: keep mask and compaction equal the oracle's."""
    from supernormal_amd.nerfacc import ray_marching, OccupancyGrid
    rng = np.random.default_rng(21)
    o, d, near, far = _cam_rays(rng, 512)
    grid = OccupancyGrid(roi_aabb=[-1, -1, -1, 1, 1, 1], resolution=64).to(dev)
    grid._binary.copy_(_t(_scene_grid(64).astype(bool), dev))

    def alpha_fn(t0, t1, ri):
        return torch.sigmoid((t0 - 2.3) * 20.0) * 0.2

    ri, t0, t1 = ray_marching(_t(o, dev), _t(d, dev), t_min=_t(near, dev), t_max=_t(far, dev),
                              grid=grid, render_step_size=5e-3, early_stop_eps=1e-2,
                              alpha_fn=alpha_fn)
    roi = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    pi_o, ri_o, t0_o, t1_o = c_oracle.ray_marching(o, d, near, far, roi, _scene_grid(64), 5e-3)
    a_o = alpha_fn(torch.from_numpy(t0_o), torch.from_numpy(t1_o), None).numpy()
    keep = c_oracle.visibility(pi_o, a_o, 1e-2, 0.0)
    assert 0 < keep.sum() < keep.size
    np.testing.assert_array_equal(ri.cpu().numpy(), ri_o[keep])
    np.testing.assert_array_equal(t0.cpu().numpy(), t0_o[keep])


def test_ray_marching_empty(dev):
    from supernormal_amd.nerfacc import ray_marching, OccupancyGrid
    grid = OccupancyGrid(roi_aabb=[-1, -1, -1, 1, 1, 1], resolution=16).to(dev)
    o = torch.zeros(8, 3, device=dev)
    o[:, 2] = -2
    d = torch.zeros(8, 3, device=dev)
    d[:, 2] = 1
    ri, t0, t1 = ray_marching(o, d, t_min=torch.ones(8, device=dev),
                              t_max=torch.full((8,), 3.0, device=dev), grid=grid,
                              render_step_size=1e-2, alpha_fn=lambda a, b, c: a)
    assert ri.numel() == 0 and t0.shape == (0, 1)


# ------------------------------------------------------------------ occupancy
def test_occupancy_update_bitexact(dev):
    from supernormal_amd.nerfacc import OccupancyGrid
    grid = OccupancyGrid(roi_aabb=[0, 0, 0, 1, 1, 1], resolution=32).to(dev)
    torch.manual_seed(0)
    captured = {}

    def occ_fn(x):
        captured["x"] = x.clone()
        captured["occ"] = ((x - 0.5).norm(dim=-1, keepdim=True) < 0.5).float()
        return captured["occ"]

    for step in range(3):
        before = grid.occs.cpu().numpy().copy()
        bin_before = grid.binary.cpu().numpy().copy()
        grid.every_n_step(step * 16 + (300 if step == 2 else 0), occ_fn, occ_thre=0.1,
                          warmup_steps=256, n=16)
        x = captured["x"].cpu().numpy()
        n_all = 32 ** 3
        if step < 2:
            idx = np.arange(n_all)
            # every sample lies in its own cell (cell points are checked bit-exactly
            # against the oracle with a known jitter in test_occ_kernels_vs_oracle)
            coord = np.stack(np.meshgrid(*[np.arange(32)] * 3, indexing="ij"), -1).reshape(-1, 3)
            assert ((x * 32 >= coord - 1e-4) & (x * 32 <= coord + 1 + 1e-4)).all()
            occ = captured["occ"].cpu().numpy()[:, 0]
            want = c_oracle.occ_ema_max(before, idx, occ, 0.95)
            np.testing.assert_array_equal(grid.occs.cpu().numpy(), want)
            b, t = c_oracle.occ_binarize(want, 0.1)
            np.testing.assert_array_equal(grid.binary.cpu().numpy().reshape(-1), b.astype(bool))
    # after warm-up: N/4 uniform + <= N/4 occupied samples, binary still consistent
    b, _ = c_oracle.occ_binarize(grid.occs.cpu().numpy(), 0.1)
    np.testing.assert_array_equal(grid.binary.cpu().numpy().reshape(-1), b.astype(bool))
    assert bin_before.sum() > 0


def test_occ_kernels_vs_oracle(dev):
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of, workspace
    rng = np.random.default_rng(14)
    res = (24, 20, 16)
    n_cells = int(np.prod(res))
    idx = rng.integers(0, n_cells, 50000).astype(np.int64)
    jit = rng.random((50000, 3), dtype=np.float32)
    roi = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    x = torch.empty(50000, 3, device=dev)
    # device copies are held in locals: a raw pointer does not keep a tensor alive
    idx_d, jit_d, roi_d = _t(idx, dev), _t(jit, dev), _t(roi, dev)
    call("sn_occ_cell_points", ptr(idx_d), 50000, ptr(jit_d), *res, ptr(roi_d), ptr(x),
         stream_of(dev))
    np.testing.assert_array_equal(x.cpu().numpy(), c_oracle.occ_cell_points(idx, jit, res, roi))
    occs0 = rng.random(n_cells, dtype=np.float32)
    occ = rng.random(50000, dtype=np.float32)
    occs = _t(occs0, dev)
    ws = workspace(_lib.lib().sn_occ_workspace_bytes(n_cells), dev)
    occ_d = _t(occ, dev)
    call("sn_occ_ema_max", ptr(occs), n_cells, ptr(idx_d), ptr(occ_d), 50000, 0.95, ptr(ws),
         ws.numel(), stream_of(dev))
    want = c_oracle.occ_ema_max(occs0, idx, occ, 0.95)
    np.testing.assert_array_equal(occs.cpu().numpy(), want)
    binary = torch.empty(n_cells, dtype=torch.uint8, device=dev)
    thr = torch.empty(1, device=dev)
    call("sn_occ_binarize", ptr(occs), n_cells, 0.9, ptr(binary), ptr(thr), ptr(ws), ws.numel(),
         stream_of(dev))
    b, t = c_oracle.occ_binarize(want, 0.9)
    assert thr.item() == t
    np.testing.assert_array_equal(binary.cpu().numpy(), b)
    out = torch.empty(n_cells, dtype=torch.int64, device=dev)
    cnt = torch.empty(1, dtype=torch.int32, device=dev)
    ws2 = workspace(_lib.lib().sn_compact_workspace_bytes(n_cells), dev)
    call("sn_occ_nonzero", ptr(binary), n_cells, ptr(out), ptr(cnt), ptr(ws2), ws2.numel(),
         stream_of(dev))
    nz = np.nonzero(b)[0]
    assert cnt.item() == nz.size
    np.testing.assert_array_equal(out[: nz.size].cpu().numpy(), nz)
