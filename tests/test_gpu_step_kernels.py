"""GPU parity of the device-count kernels of the sync-free iteration.

Every kernel runs with a capacity larger than the valid prefix (garbage
beyond it) and the valid count in device memory, and must reproduce the
plain entry points / the oracle on the prefix: integer outputs (diff slots,
packed_info, compaction, cell lists, visibility bits away from the threshold)
bit-exact; points bit-exact where the oracle runs the same fp32 sequence
(o + d*t, no FMA) and within 1e-6 relative where a 3-term dot product is
summed in another order.
"""
import numpy as np
import pytest
import torch

from oracle import c_oracle, render_oracle

pytestmark = pytest.mark.gpu

STEP = 2.0 ** -7


def _samples(n_rays, max_per_ray, seed):
    """Marching-like samples: sorted ray indices, t on a 2^-7 grid (exact in
    fp32) with random gaps, so contiguity is decided by exact equality."""
    rng = np.random.default_rng(seed)
    ri, t0 = [], []
    for r in range(n_rays):
        c = int(rng.integers(0, max_per_ray + 1))
        k = np.sort(rng.choice(np.arange(2 * max_per_ray + 2), c, replace=False))
        ri += [r] * c
        t0 += list(0.5 + k * STEP)
    ri = np.asarray(ri, np.int64)
    t0 = np.asarray(t0, np.float32)
    t1 = (t0 + np.float32(STEP)).astype(np.float32)
    return ri, t0, t1


def _pad(a, extra, fill):
    return np.concatenate([a, np.full((extra,) + a.shape[1:], fill, a.dtype)])


def _cap_tensors(dev, ri, t0, t1, extra=97):
    S = ri.shape[0]
    rit = torch.as_tensor(_pad(ri, extra, 5), device=dev)
    t0t = torch.as_tensor(_pad(t0, extra, 0.75), device=dev)
    t1t = torch.as_tensor(_pad(t1, extra, 0.75), device=dev)
    n_dev = torch.tensor([S], dtype=torch.int32, device=dev)
    return rit, t0t, t1t, n_dev, S + extra


def _diff_slots(ri, t0, t1):
    dm = render_oracle.diff_mask(torch.as_tensor(ri), torch.as_tensor(t0)[:, None],
                                 torch.as_tensor(t1)[:, None]).numpy()
    slots = np.full(ri.shape[0], -1, np.int32)
    slots[dm] = np.arange(int(dm.sum()), dtype=np.int32)
    return dm, slots


def _rays(B, Ph, Pw, seed):
    g = torch.Generator().manual_seed(seed)
    o = torch.randn(B, 3, generator=g)[:, None, None, :] * 0.1 + \
        torch.tensor([0.0, 0.0, -2.0]) + torch.randn(B, Ph, Pw, 3, generator=g) * 1e-3
    d = torch.tensor([0.0, 0.0, 1.0]) + torch.randn(B, Ph, Pw, 3, generator=g) * 0.05
    d = d / d.norm(dim=-1, keepdim=True)
    n = -d[:, Ph // 2, Pw // 2] + torch.randn(B, 3, generator=g) * 0.1
    n = n / n.norm(dim=-1, keepdim=True)
    return o.contiguous(), d.contiguous(), n.contiguous()


@pytest.mark.parametrize("B,seed", [(1, 0), (300, 1), (2048, 2)])
def test_centre_points(dev, B, seed):
    from supernormal_amd._lib import call, lib, ptr, stream_of
    ri, t0, t1 = _samples(B, 40, seed)
    S = ri.shape[0]
    o, d, _ = _rays(B, 3, 3, seed)
    rit, t0t, t1t, n_dev, cap = _cap_tensors(dev, ri, t0, t1)
    od, dd = o.to(dev), d.to(dev)
    ws = torch.empty(lib().sn_step_workspace_bytes(cap), dtype=torch.uint8, device=dev)
    didx = torch.full((cap,), 7, dtype=torch.int32, device=dev)
    pts = torch.zeros(2 * cap, 3, device=dev)
    n_diff = torch.zeros(1, dtype=torch.int32, device=dev)
    n_eval = torch.zeros(1, dtype=torch.int32, device=dev)
    call("sn_centre_points", ptr(rit), ptr(t0t), ptr(t1t), cap, ptr(n_dev),
         ptr(od[:, 1, 1]), ptr(dd[:, 1, 1]), 27, ptr(didx), ptr(n_diff), ptr(pts),
         ptr(n_eval), ptr(ws), ws.numel(), stream_of(dev))
    dm, slots = _diff_slots(ri, t0, t1)
    Nd = int(dm.sum())
    assert int(n_diff) == Nd and int(n_eval) == S + Nd
    np.testing.assert_array_equal(didx[:S].cpu().numpy(), slots)
    oc, dc = o[:, 1, 1], d[:, 1, 1]
    rt = torch.as_tensor(ri)
    xs = oc[rt] + dc[rt] * torch.as_tensor(t0)[:, None]
    xe = (oc[rt] + dc[rt] * torch.as_tensor(t1)[:, None])[torch.as_tensor(dm)]
    got = pts[:S + Nd].cpu()
    assert torch.equal(got[:S], xs)
    assert torch.equal(got[S:], xe)


@pytest.mark.parametrize("B,seed", [(5, 3), (700, 4)])
def test_alpha_visibility(dev, B, seed):
    from supernormal_amd._lib import call, ptr, stream_of
    ri, t0, t1 = _samples(B, 40, seed)
    S = ri.shape[0]
    dm, slots = _diff_slots(ri, t0, t1)
    Nd = int(dm.sum())
    rng = np.random.default_rng(seed)
    # SDF decreasing along each ray through zero: alphas from ~0 to ~1.
    sdf = np.concatenate([0.3 - (t0 - 0.5) * rng.uniform(0.5, 3.0),
                          rng.uniform(-0.2, 0.2, Nd)]).astype(np.float32)
    inv_s = np.float32(30.0)
    packed = c_oracle.pack_info(ri, B)
    rit, t0t, t1t, n_dev, cap = _cap_tensors(dev, ri, t0, t1)
    sdf_cap = torch.zeros(cap + Nd + 1, device=dev)
    # kernel reads end values at [n + slot]: lay the tail after the valid prefix
    sdf_cap[:S] = torch.as_tensor(sdf[:S], device=dev)
    sdf_cap[S:S + Nd] = torch.as_tensor(sdf[S:], device=dev)
    didx = torch.as_tensor(_pad(slots, cap - S, -1), device=dev)
    keep = torch.full((cap,), 9, dtype=torch.uint8, device=dev)
    keep[:S] = 0
    # (device tensors held in locals: a temporary's memory can be reused before the launch)
    packed_d = torch.as_tensor(packed, device=dev)
    inv_s_d = torch.tensor([inv_s], device=dev)
    call("sn_alpha_visibility", ptr(packed_d), B, ptr(sdf_cap), ptr(didx), cap, ptr(n_dev),
         ptr(inv_s_d), 1e-4, ptr(keep), stream_of(dev))
    # fp64 restatement: alpha from (start, end) SDF, keep where T_before >= eps
    fs = sdf[:S].astype(np.float64)
    fe = np.where(slots >= 0, sdf[S + np.maximum(slots, 0)], np.roll(sdf[:S], -1)).astype(np.float64)
    sig = lambda x: 1.0 / (1.0 + np.exp(-x))
    c, e = sig(fs * inv_s), sig(fe * inv_s)
    a = np.clip((c - e + 1e-5) / (c + 1e-5), 0.0, 1.0)
    want = np.zeros(S, bool)
    near = np.zeros(S, bool)
    for r in range(B):
        b, n = packed[r]
        T = 1.0
        for j in range(b, b + n):
            want[j] = T >= 1e-4
            near[j] = abs(T - 1e-4) < 1e-3 * 1e-4 + 1e-9
            T *= 1.0 - a[j]
    got = keep[:S].cpu().numpy().astype(bool)
    assert np.array_equal(got[~near], want[~near])
    assert (keep[S:].cpu().numpy() == 9).all()


@pytest.mark.parametrize("B,Ph,Pw,seed", [(64, 3, 3, 5), (1000, 3, 3, 6), (40, 5, 5, 7),
                                          (40, 2, 4, 8)])
def test_patch_points(dev, B, Ph, Pw, seed):
    from supernormal_amd._lib import call, lib, ptr, stream_of
    P = Ph * Pw
    ri, t0, t1 = _samples(B, 30, seed)
    S = ri.shape[0]
    o, d, n = _rays(B, Ph, Pw, seed)
    rit, t0t, t1t, n_dev, cap = _cap_tensors(dev, ri, t0, t1)
    ws = torch.empty(lib().sn_step_workspace_bytes(cap), dtype=torch.uint8, device=dev)
    didx = torch.full((cap,), 7, dtype=torch.int32, device=dev)
    pts = torch.zeros(2 * cap * P, 3, device=dev)
    n_diff = torch.zeros(1, dtype=torch.int32, device=dev)
    n_eval = torch.zeros(1, dtype=torch.int32, device=dev)
    od, dd, nd = o.to(dev), d.to(dev), n.to(dev)
    call("sn_patch_points", ptr(rit), ptr(t0t), ptr(t1t), cap, ptr(n_dev), ptr(od), ptr(dd),
         ptr(nd), Ph, Pw, ptr(didx), ptr(n_diff), ptr(pts), ptr(n_eval), ptr(ws), ws.numel(),
         stream_of(dev))
    dm, slots = _diff_slots(ri, t0, t1)
    Nd = int(dm.sum())
    assert int(n_diff) == Nd and int(n_eval) == (S + Nd) * P
    np.testing.assert_array_equal(didx[:S].cpu().numpy(), slots)
    rt = torch.as_tensor(ri)
    ts = render_oracle.expand_t(torch.as_tensor(t0)[:, None], rt, d, n)
    te = render_oracle.expand_t(torch.as_tensor(t1)[:, None], rt, d, n)
    ps = render_oracle.positions(o, d, rt, ts).reshape(-1, 3)
    pe = render_oracle.positions(o, d, rt, te)[torch.as_tensor(dm)].reshape(-1, 3)
    got = pts[:(S + Nd) * P].cpu()
    torch.testing.assert_close(got[:S * P], ps, rtol=1e-6, atol=1e-6)
    torch.testing.assert_close(got[S * P:], pe, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("n,n_rays", [(1, 1), (5000, 37), (300000, 2048)])
def test_pack_info_dev_prefix(dev, n, n_rays):
    from supernormal_amd._lib import call, lib, ptr, stream_of
    rng = np.random.default_rng(n)
    ri = np.sort(rng.integers(0, n_rays, n)).astype(np.int64)
    cap = n + 1000
    rit = torch.as_tensor(_pad(ri, 1000, 0), device=dev)
    n_dev = torch.tensor([n], dtype=torch.int32, device=dev)
    packed = torch.empty(n_rays, 2, dtype=torch.int32, device=dev)
    ws = torch.empty(lib().sn_pack_info_workspace_bytes(n_rays), dtype=torch.uint8, device=dev)
    call("sn_pack_info_dev", ptr(rit), cap, ptr(n_dev), n_rays, ptr(packed), None, ptr(ws),
         ws.numel(), stream_of(dev))
    np.testing.assert_array_equal(packed.cpu().numpy(), c_oracle.pack_info(ri, n_rays))


@pytest.mark.parametrize("n", [1, 4097, 250000])
def test_compact_samples_dev_prefix(dev, n):
    from supernormal_amd._lib import call, lib, ptr, stream_of
    rng = np.random.default_rng(n)
    keep = (rng.random(n + 500) < 0.6).astype(np.uint8)
    ri = np.sort(rng.integers(0, 999, n + 500)).astype(np.int64)
    t0 = rng.random(n + 500).astype(np.float32)
    t1 = rng.random(n + 500).astype(np.float32)
    cap = n + 500
    T = lambda a: torch.as_tensor(a, device=dev)
    n_dev = torch.tensor([n], dtype=torch.int32, device=dev)
    ori = torch.empty(cap, dtype=torch.int64, device=dev)
    o0 = torch.empty(cap, device=dev)
    o1 = torch.empty(cap, device=dev)
    n_out = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(lib().sn_compact_workspace_bytes(cap), dtype=torch.uint8, device=dev)
    kt, rt, t0t, t1t = T(keep), T(ri), T(t0), T(t1)
    # packed_info of the rays before compaction (sorted ray indices, 999 rays)
    n_rays = 999
    pin = torch.as_tensor(c_oracle.pack_info(ri[:n], n_rays), device=dev)
    pout = torch.full((n_rays, 2), -7, dtype=torch.int32, device=dev)
    call("sn_compact_samples_dev", ptr(kt), cap, ptr(n_dev), ptr(rt), ptr(t0t), ptr(t1t),
         ptr(ori), ptr(o0), ptr(o1), ptr(n_out), ptr(pin), n_rays, ptr(pout), ptr(ws),
         ws.numel(), stream_of(dev))
    k = keep[:n].astype(bool)
    m = int(k.sum())
    assert int(n_out) == m
    np.testing.assert_array_equal(pout.cpu().numpy(), c_oracle.pack_info(ri[:n][k], n_rays))
    np.testing.assert_array_equal(ori[:m].cpu().numpy(), ri[:n][k])
    np.testing.assert_array_equal(o0[:m].cpu().numpy(), t0[:n][k])
    np.testing.assert_array_equal(o1[:m].cpu().numpy(), t1[:n][k])


@pytest.mark.parametrize("n_cells,density", [(64 ** 3, 0.05), (64 ** 3, 0.6), (128 ** 3, 0.01),
                                             (8, 0.0)])
def test_occ_sample_cells(dev, n_cells, density):
    from supernormal_amd._lib import call, lib, ptr, stream_of
    rng = np.random.default_rng(n_cells + int(density * 100))
    binary = (rng.random(n_cells) < density).astype(np.uint8)
    q = n_cells // 4
    u1 = rng.random(q).astype(np.float32)
    u2 = rng.random(q).astype(np.float32)
    out = torch.full((2 * q,), -5, dtype=torch.int64, device=dev)
    n_out = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(lib().sn_occ_sample_workspace_bytes(n_cells), dtype=torch.uint8, device=dev)
    bt, u1t, u2t = (torch.as_tensor(a, device=dev) for a in (binary, u1, u2))
    call("sn_occ_sample_cells", ptr(bt), n_cells, ptr(u1t), ptr(u2t), ptr(out), ptr(n_out),
         ptr(ws), ws.numel(), stream_of(dev))
    occ = np.nonzero(binary)[0]
    c = occ.size
    uni = np.minimum((u1 * np.float32(n_cells)).astype(np.int64), n_cells - 1)
    if c > q:
        sel = occ[np.minimum((u2 * np.float32(c)).astype(np.int64), c - 1)]
    else:
        sel = occ
    want = np.concatenate([uni, sel])
    assert int(n_out) == want.size
    np.testing.assert_array_equal(out[:want.size].cpu().numpy(), want)
