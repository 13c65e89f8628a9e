"""GPU parity of the hash-grid encoder (tinycudann.Encoding drop-in) against the
fp64 oracle: forward, backward to params and x, and the double backward
through the x-gradient.  Tolerances: fp32 kernel vs fp64 oracle,
|err| <= atol + rtol |ref| with rtol 1e-5 (forward) / 1e-4 (gradients)."""
import numpy as np
import pytest
import torch

from oracle.hashgrid_oracle import encode, hashgrid_desc

pytestmark = pytest.mark.gpu

CFGS = {
    "cfg1": dict(n_levels=8, n_features_per_level=2, log2_hashmap_size=15, base_resolution=16,
                 per_level_scale=1.5),
    "cfg2": dict(n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
                 per_level_scale=1.38),
    "diligent": dict(n_levels=14, n_features_per_level=2, log2_hashmap_size=19,
                     base_resolution=32, per_level_scale=1.3195079107728942),
}


def _enc(cfg, dev, seed=0):
    from supernormal_amd.tcnn import Encoding
    e = Encoding(3, {"otype": "HashGrid", **cfg}).to(dev)
    g = torch.Generator(device="cpu").manual_seed(seed)
    with torch.no_grad():
        e.params.copy_((torch.rand(e.params.shape, generator=g) * 2 - 1).to(dev))
    return e


def _desc(cfg):
    return hashgrid_desc(cfg["n_levels"], cfg["log2_hashmap_size"], cfg["base_resolution"],
                         cfg["per_level_scale"])


@pytest.mark.parametrize("name", list(CFGS))
def test_layout_matches_oracle(name, dev):
    e = _enc(CFGS[name], dev)
    d = _desc(CFGS[name])
    assert e.level_offsets() == d["offset"].tolist()
    assert e.params.numel() == d["n_entries"] * 2


@pytest.mark.parametrize("name,n_active", [("cfg1", 8), ("cfg2", 16), ("cfg2", 3), ("diligent", 14)])
def test_forward_backward_vs_fp64(name, n_active, dev):
    cfg = CFGS[name]
    e = _enc(cfg, dev, seed=1)
    e.n_active_levels = n_active
    d = _desc(cfg)
    rng = np.random.default_rng(2)
    n = 3000
    x = (rng.random((n, 3)) * 2 - 1).astype(np.float32)  # [-1,1]: wrap semantics exercised
    xt = torch.from_numpy(x).to(dev).requires_grad_(True)
    out = e(xt)
    g = torch.from_numpy(rng.normal(size=out.shape).astype(np.float32)).to(dev)
    (out * g).sum().backward()

    x64 = torch.from_numpy(x).double().requires_grad_(True)
    p64 = e.params.detach().cpu().double().requires_grad_(True)
    ref = encode(x64, p64, d, n_active=n_active)
    (ref * g.cpu().double()).sum().backward()
    torch.testing.assert_close(out.detach().cpu().double(), ref.detach(), rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(e.params.grad.cpu().double(), p64.grad, rtol=1e-4, atol=1e-5)
    torch.testing.assert_close(xt.grad.cpu().double(), x64.grad, rtol=1e-4, atol=2e-4)


@pytest.mark.parametrize("name", ["cfg2", "cfg1"])
def test_backward_aggregation_matches_plain_atomics(name, dev):
    """The merged backwards (2: workgroup LDS table over all levels, 1:
    wavefront merge of coarse levels) equal plain atomics (0) within fp32
    reassociation, on clustered points like a patch's samples, with the
    x-gradient unchanged; 300 clusters of 300 points overflow nothing but
    exercise long probe chains on the coarse levels."""
    from supernormal_amd import _lib
    cfg = CFGS[name]
    e = _enc(cfg, dev, seed=5)
    g = torch.Generator().manual_seed(6)
    centers = torch.rand(300, 3, generator=g) * 1.6 - 0.8
    x = (centers[:, None, :] + torch.randn(300, 300, 3, generator=g) * 0.004).reshape(-1, 3)
    x = x.to(dev).requires_grad_(True)
    genc = torch.randn(x.shape[0], e.n_output_dims, generator=g).to(dev)
    grads, gxs = [], []
    try:
        for mode in (2, 1, 0):
            _lib.lib().sn_hashgrid_set_aggregation(mode)
            e.params.grad = None
            x.grad = None
            (e(x) * genc).sum().backward()
            grads.append(e.params.grad.clone())
            gxs.append(x.grad.clone())
    finally:
        _lib.lib().sn_hashgrid_set_aggregation(2)
    scale = float(grads[2].abs().max())
    for k in (0, 1):
        torch.testing.assert_close(grads[k], grads[2], rtol=1e-4, atol=1e-5 * scale)
        torch.testing.assert_close(gxs[k], gxs[2], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("log2_T", [19, 21])
@pytest.mark.parametrize("records", [65536, 7])
def test_bucket_backward_matches_atomics(records, log2_T, dev):
    """sn_hashgrid_bwd_ws_dev (bucket records, then per-bucket fixed-point
    sums) equals the atomic path within fp32 reassociation, adding onto a
    non-zero gradient and giving the same x-gradient, and repeats bit for bit;
    with 7 records per bucket almost every record takes the overflow atomics.
    Only the first n_dev of n_cap points count.  The workspace starts as
    garbage and is reused by a second call.  T = 2^19 gives 64 buckets per
    hashed level, T = 2^21 gives 256 (several buckets per lane of the
    record kernel's reservation)."""
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = dict(CFGS["cfg2"], log2_hashmap_size=log2_T)
    e = _enc(cfg, dev, seed=7)
    g = torch.Generator().manual_seed(8)
    centers = torch.rand(200, 3, generator=g) * 1.6 - 0.8
    x = (centers[:, None, :] + torch.randn(200, 200, 3, generator=g) * 0.003).reshape(-1, 3)
    x = x.contiguous().to(dev)
    n_cap, n_use = x.shape[0], 35000
    n_dev = torch.tensor([n_use], dtype=torch.int32, device=dev)
    genc = torch.randn(n_cap, e.n_output_dims, generator=g).to(dev)
    gargs = (16, 2, log2_T, 16, 1.38, 16)
    lib = _lib.lib()
    nb = lib.sn_hashgrid_bwd_workspace_bytes(*gargs[:5], records)
    assert nb > 0
    ws = torch.full((nb,), 255, dtype=torch.uint8, device=dev)
    base = torch.randn(e.params.shape, generator=g).to(dev) * 1e-3
    s = stream_of(dev)
    g_ref, gx_ref = base.clone(), torch.empty_like(x)
    call("sn_hashgrid_bwd_dev", ptr(x), n_cap, ptr(n_dev), ptr(e.params), *gargs, None, ptr(genc),
         ptr(g_ref), ptr(gx_ref), s)
    outs = []
    for _ in range(2):
        g_b, gx_b = base.clone(), torch.empty_like(x)
        call("sn_hashgrid_bwd_ws_dev", ptr(x), n_cap, ptr(n_dev), ptr(e.params), *gargs, None,
             ptr(genc), ptr(g_b), ptr(gx_b), ptr(ws), nb, s)
        torch.cuda.synchronize()
        scale = float((g_ref - base).abs().max())
        torch.testing.assert_close(g_b, g_ref, rtol=1e-4, atol=1e-5 * scale)
        torch.testing.assert_close(gx_b[:n_use], gx_ref[:n_use], rtol=1e-5, atol=1e-6)
        outs.append(g_b)
    if records >= 65536:  # no list overflows (35000 points: <= 4375 records per list)
        # integer (fixed-point) sums: the same bits on every run
        assert torch.equal(outs[0], outs[1])
    # a workspace too small for one record per bucket is refused
    with pytest.raises(RuntimeError):
        call("sn_hashgrid_bwd_ws_dev", ptr(x), n_cap, ptr(n_dev), ptr(e.params), *gargs, None,
             ptr(genc), ptr(g_b), None, ptr(ws), 64, s)


def test_bucket_backward_propagates_nan_inf(dev):
    """A NaN and an inf in dL/denc reach the table gradient through the bucket
    path as through the fp32 atomic path: the level maximum is taken over the
    bits of |value| (a NaN's bits lie above inf's), so both levels fall back to
    fp32 sums, the rows the two points touch end NaN / +inf in both paths, and
    every other row agrees within fp32 reassociation."""
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS["cfg2"]
    e = _enc(cfg, dev, seed=21)
    g = torch.Generator().manual_seed(22)
    x = (torch.rand(20000, 3, generator=g) * 1.6 - 0.8).contiguous().to(dev)
    n = x.shape[0]
    genc = torch.randn(n, e.n_output_dims, generator=g)
    genc[5, 2 * 3 + 0] = float("nan")    # level 3 (dense), feature 0
    genc[7, 2 * 10 + 1] = float("inf")   # level 10 (hashed), feature 1
    genc = genc.to(dev)
    gargs = (16, 2, 19, 16, 1.38, 16)
    lib = _lib.lib()
    nb = lib.sn_hashgrid_bwd_workspace_bytes(*gargs[:5], 65536)
    ws = torch.zeros((nb,), dtype=torch.uint8, device=dev)
    s = stream_of(dev)
    g_ref = torch.zeros_like(e.params)
    call("sn_hashgrid_bwd_dev", ptr(x), n, None, ptr(e.params), *gargs, None, ptr(genc),
         ptr(g_ref), None, s)
    g_b = torch.zeros_like(e.params)
    call("sn_hashgrid_bwd_ws_dev", ptr(x), n, None, ptr(e.params), *gargs, None, ptr(genc),
         ptr(g_b), None, ptr(ws), nb, s)
    torch.cuda.synchronize()
    nan_ref, inf_ref = torch.isnan(g_ref), torch.isinf(g_ref)
    assert int(nan_ref.sum()) > 0 and int(inf_ref.sum()) > 0
    assert torch.equal(torch.isnan(g_b), nan_ref)
    assert torch.equal(torch.isinf(g_b), inf_ref)
    assert torch.equal(torch.sign(g_b[inf_ref]), torch.sign(g_ref[inf_ref]))
    fin = ~(nan_ref | inf_ref)
    scale = float(g_ref[fin].abs().max())
    torch.testing.assert_close(g_b[fin], g_ref[fin], rtol=1e-4, atol=1e-5 * scale)


def test_bucket_pool_follows_previous_counts(dev):
    """The record pool is shared out by the counts of the previous call on the
    same workspace.  12000 uniform points leave ~96k records in the single
    bucket of level 0: the even split of a fresh workspace (16384 per bucket)
    overflows there (atomics), the second call's lists hold every record, and
    from then on calls repeat bit for bit; all agree with the atomic path."""
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS["cfg2"]
    e = _enc(cfg, dev, seed=11)
    g = torch.Generator().manual_seed(12)
    n = 12000
    x = (torch.rand(n, 3, generator=g) * 2 - 1).to(dev)
    genc = torch.randn(n, 32, generator=g).to(dev)
    gargs = (16, 2, 19, 16, 1.38, 16)
    lib = _lib.lib()
    b1 = lib.sn_hashgrid_bwd_workspace_bytes(*gargs[:5], 1)
    b2 = lib.sn_hashgrid_bwd_workspace_bytes(*gargs[:5], 1025)
    nbk = round((b2 - b1) / (1024 * 10))
    assert nbk == 747
    nbytes = lib.sn_hashgrid_bwd_workspace_bytes(*gargs[:5], 16384)
    ws = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    head = ws[:4 * (4 + 3 * nbk + 32)].view(torch.int32)
    cursor = lambda: head[4:4 + nbk].cpu()
    bcap = lambda: head[4 + 2 * nbk + 32:4 + 3 * nbk + 32].cpu()
    s = stream_of(dev)
    g_ref = torch.zeros_like(e.params)
    call("sn_hashgrid_bwd_dev", ptr(x), n, None, ptr(e.params), *gargs, None, ptr(genc),
         ptr(g_ref), None, s)
    scale = float(g_ref.abs().max())
    outs, over = [], []
    for _ in range(3):
        gp = torch.zeros_like(e.params)
        call("sn_hashgrid_bwd_ws_dev", ptr(x), n, None, ptr(e.params), *gargs, None, ptr(genc),
             ptr(gp), None, ptr(ws), nbytes, s)
        torch.cuda.synchronize()
        over.append(int((cursor() - bcap()).clamp_min(0).sum()))
        torch.testing.assert_close(gp, g_ref, rtol=1e-4, atol=1e-5 * scale)
        outs.append(gp)
    assert over[0] > 30000 and over[1] == 0 and over[2] == 0
    assert int(cursor()[0]) > 80000  # the level-0 list
    assert torch.equal(outs[1], outs[2])


def test_encoding_keeps_its_workspace(dev):
    """tcnn.Encoding keeps one bucket workspace between backward calls: the
    same buffer serves every call (fewer points included), the level-0 list
    that overflows the fresh even split no longer overflows from the second
    call on, and from then on the table gradient repeats bit for bit and
    agrees with the atomic path."""
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS["cfg2"]
    e = _enc(cfg, dev, seed=11)
    g = torch.Generator().manual_seed(12)
    n = 12000
    x = (torch.rand(n, 3, generator=g) * 2 - 1).to(dev)
    genc = torch.randn(n, 32, generator=g).to(dev)
    g_ref = torch.zeros_like(e.params)
    call("sn_hashgrid_bwd_dev", ptr(x), n, None, ptr(e.params), 16, 2, 19, 16, 1.38, 16, None,
         ptr(genc), ptr(g_ref), None, stream_of(dev))
    scale = float(g_ref.abs().max())
    outs, bufs = [], []
    for k in range(4):
        m = n if k < 3 else n // 2
        e.params.grad = None
        (e(x[:m]) * genc[:m]).sum().backward()
        bufs.append(e.workspace.buf.data_ptr())
        outs.append(e.params.grad.clone())
    assert len(set(bufs)) == 1
    for gp in outs[:3]:
        torch.testing.assert_close(gp, g_ref, rtol=1e-4, atol=1e-5 * scale)
    assert torch.equal(outs[1], outs[2])
    half = torch.zeros_like(e.params)
    call("sn_hashgrid_bwd_dev", ptr(x), n // 2, None, ptr(e.params), 16, 2, 19, 16, 1.38, 16,
         None, ptr(genc), ptr(half), None, stream_of(dev))
    torch.testing.assert_close(outs[3], half, rtol=1e-4, atol=1e-5 * scale)


def test_double_backward_vs_fp64(dev):
    """L = sum(v * dL1/dx) with dL1/dx = autograd.grad(sum(g * enc), x, create_graph=True):
    d L / d params, d L / d g (through enc's upstream gradient) and d L / d x."""
    cfg = CFGS["cfg1"]
    e = _enc(cfg, dev, seed=3)
    d = _desc(cfg)
    rng = np.random.default_rng(4)
    n = 1500
    x = (rng.random((n, 3)) * 1.8 - 0.9).astype(np.float32)
    g = rng.normal(size=(n, e.n_output_dims)).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)

    def run(xt, pt, gt, f):
        out = f(xt, pt)
        gx, = torch.autograd.grad((out * gt).sum(), xt, create_graph=True)
        L = (gx * torch.as_tensor(v, dtype=gx.dtype, device=gx.device)).sum()
        return torch.autograd.grad(L, (pt, gt, xt))

    xt = torch.from_numpy(x).to(dev).requires_grad_(True)
    gt = torch.from_numpy(g).to(dev).requires_grad_(True)
    dp, dg, dx = run(xt, e.params, gt, lambda a, p: e(a))
    x64 = torch.from_numpy(x).double().requires_grad_(True)
    g64 = torch.from_numpy(g).double().requires_grad_(True)
    p64 = e.params.detach().cpu().double().requires_grad_(True)
    rp, rg, rx = run(x64, p64, g64, lambda a, p: encode(a, p, d))
    torch.testing.assert_close(dp.cpu().double(), rp, rtol=1e-4, atol=1e-4)
    torch.testing.assert_close(dg.cpu().double(), rg, rtol=1e-4, atol=1e-4)
    torch.testing.assert_close(dx.cpu().double(), rx, rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("with_ws", [True, False])
def test_ad_table_gradient_one_pass(dev, with_ws):
    """sn_hashgrid_bwd_ad_ws_dev: the table gradient of
    L = sum(genc * enc) + sum(v * dL2/dx), dL2/dx = autograd.grad(sum(genc2 * enc), x),
    i.e. the first-order term and the double backward's table term, against
    the fp64 oracle's autograd.  With a workspace both terms share each
    corner's bucket record (repeat calls on the same workspace agree bit for
    bit once the layout follows the counts); without one the second term
    takes fp32 atomics."""
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS["cfg2"]
    e = _enc(cfg, dev, seed=21)
    d = _desc(cfg)
    rng = np.random.default_rng(22)
    n = 3000
    x = (rng.random((n, 3)) * 1.8 - 0.9).astype(np.float32)
    g = rng.normal(size=(n, 32)).astype(np.float32)
    g2 = rng.normal(size=(n, 32)).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    x64 = torch.from_numpy(x).double().requires_grad_(True)
    p64 = e.params.detach().cpu().double().requires_grad_(True)
    enc = encode(x64, p64, d)
    gx2, = torch.autograd.grad((enc * torch.from_numpy(g2).double()).sum(), x64,
                               create_graph=True)
    L = (enc * torch.from_numpy(g).double()).sum() + (gx2 * torch.from_numpy(v).double()).sum()
    ref, = torch.autograd.grad(L, p64)
    t = lambda a: torch.from_numpy(a).to(dev)
    xt, gt, g2t, vt = t(x), t(g), t(g2), t(v)
    args = (16, 2, 19, 16, 1.38, 16)
    nbytes = _lib.lib().sn_hashgrid_bwd_workspace_bytes(*args[:5], 16384) if with_ws else 0
    ws = torch.zeros(max(nbytes, 1), dtype=torch.uint8, device=dev)
    outs = []
    for _ in range(3):
        gp = torch.zeros_like(e.params)
        call("sn_hashgrid_bwd_ad_ws_dev", ptr(xt), n, None, ptr(e.params), *args, None, ptr(gt),
             ptr(vt), ptr(g2t), ptr(gp), ptr(ws) if with_ws else None, nbytes, stream_of(dev))
        torch.cuda.synchronize()
        outs.append(gp)
    scale = float(ref.abs().max())
    for gp in outs:
        torch.testing.assert_close(gp.cpu().double(), ref, rtol=1e-4, atol=1e-5 * scale)
    if with_ws:
        assert torch.equal(outs[1], outs[2])


@pytest.mark.parametrize("half", [False, True])
def test_tangent_vs_fp64(dev, half):
    """sn_hashgrid_tangent_dev: the encoding's directional derivative along v
    (d/de enc(x + e v) at e = 0) against fp64 autograd of the oracle; from the
    fp32 table (1e-5 of max|ref|, and equal to sn_hashgrid_bwd_bwd_dev's
    grad_grad_enc within fp32 reassociation) or its fp16 copy (2e-3)."""
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS["cfg2"]
    e = _enc(cfg, dev, seed=31)
    d = _desc(cfg)
    rng = np.random.default_rng(32)
    n = 3000
    x = (rng.random((n, 3)) * 1.8 - 0.9).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    x64 = torch.from_numpy(x).double().requires_grad_(True)
    enc = encode(x64, e.params.detach().cpu().double(), d)
    w = torch.zeros_like(enc, requires_grad=True)
    gx, = torch.autograd.grad(enc, x64, grad_outputs=w, create_graph=True)
    ref, = torch.autograd.grad((gx * torch.from_numpy(v).double()).sum(), w)
    xt, vt = torch.from_numpy(x).to(dev), torch.from_numpy(v).to(dev)
    args = (16, 2, 19, 16, 1.38, 16)
    p16 = e.params.detach().half().contiguous() if half else None
    out = torch.full((n, 32), -9.0, device=dev)
    call("sn_hashgrid_tangent_dev", ptr(xt), n, None, ptr(e.params), ptr(p16), *args, None,
         ptr(vt), ptr(out), stream_of(dev))
    torch.cuda.synchronize()
    scale = float(ref.abs().max())
    err = float((out.cpu().double() - ref).abs().max()) / scale
    print(f"\nhalf={half}: tangent max err {err:.2e} of max|ref|")
    assert err <= (2e-3 if half else 1e-5)
    if not half:
        other = torch.empty_like(out)
        call("sn_hashgrid_bwd_bwd_dev", ptr(xt), n, None, ptr(e.params), *args, None, None,
             ptr(vt), ptr(other), None, None, stream_of(dev))
        torch.testing.assert_close(out, other, rtol=1e-5, atol=1e-6 * scale)


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("n_active", [16, 9])
def test_dx_vs_fp64(dev, half, n_active):
    """sn_hashgrid_dx_dev: grad = dh0[:, :3] + sum_l dh0[:, 3 + l F + f] d enc_lf / dx
    (the analytic gradient's chain through the table) against fp64 autograd of
    the oracle, over the active levels; fp32 table 1e-5 of max|ref|, fp16 copy 2e-3."""
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS["cfg2"]
    e = _enc(cfg, dev, seed=41)
    d = _desc(cfg)
    rng = np.random.default_rng(42)
    n = 3000
    x = (rng.random((n, 3)) * 1.8 - 0.9).astype(np.float32)
    dh0 = rng.normal(size=(n, 35)).astype(np.float32)
    x64 = torch.from_numpy(x).double().requires_grad_(True)
    enc = encode(x64, e.params.detach().cpu().double(), d)
    mask = torch.zeros(32, dtype=torch.float64)
    mask[:2 * n_active] = 1.0
    gx, = torch.autograd.grad((enc * torch.from_numpy(dh0[:, 3:]).double() * mask).sum(), x64)
    ref = gx + torch.from_numpy(dh0[:, :3]).double()
    xt, gt = torch.from_numpy(x).to(dev), torch.from_numpy(dh0).to(dev)
    args = (16, 2, 19, 16, 1.38, n_active)
    p16 = e.params.detach().half().contiguous() if half else None
    out = torch.full((n, 3), -9.0, device=dev)
    call("sn_hashgrid_dx_dev", ptr(xt), n, None, ptr(e.params), ptr(p16), *args, None, ptr(gt),
         35, ptr(out), stream_of(dev))
    torch.cuda.synchronize()
    err = float((out.cpu().double() - ref).abs().max() / ref.abs().max())
    print(f"\nhalf={half} active={n_active}: dx max err {err:.2e} of max|ref|")
    assert err <= (2e-3 if half else 1e-5)


def test_full_size_properties(dev):
    """cfg2 table at the training batch size: partition of unity and linearity in
    the table (size-independent checks)."""
    from supernormal_amd.tcnn import Encoding
    cfg = CFGS["cfg2"]
    e = Encoding(3, {"otype": "HashGrid", **cfg}).to(dev)
    n = 2_359_296
    x = torch.rand(n, 3, device=dev) * 2 - 1
    with torch.no_grad():
        e.params.fill_(0.25)
        out = e(x)
        assert torch.allclose(out, torch.full_like(out, 0.25), atol=1e-6)
        p1 = torch.rand_like(e.params)
        p2 = torch.rand_like(e.params)
        e.params.copy_(p1)
        o1 = e(x)
        e.params.copy_(p2)
        o2 = e(x)
        e.params.copy_(p1 + p2)
        o3 = e(x)
        torch.testing.assert_close(o3, o1 + o2, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("name,n_active", [("cfg2", 16), ("cfg1", 5)])
def test_forward_from_fp16_copy(name, n_active, dev):
    """The fp16-copy forward equals the fp32 forward of the rounded table
    (same gathers, same fp32 blending), bit for bit; and the copy kept by the
    optimizer step (sn_adam_step_copy16) is the rounding of the updated table."""
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS[name]
    e = _enc(cfg, dev, seed=4)
    n = 50_000
    g = torch.Generator(device="cpu").manual_seed(5)
    x = (torch.rand(n, 3, generator=g) * 2 - 1).to(dev)
    args = (cfg["n_levels"], cfg["n_features_per_level"], cfg["log2_hashmap_size"],
            cfg["base_resolution"], cfg["per_level_scale"], n_active)
    t16 = torch.empty(e.params.numel(), dtype=torch.float16, device=dev)
    st = stream_of(dev)
    call("sn_to_half", ptr(e.params), ptr(t16), e.params.numel(), st)
    assert torch.equal(t16, e.params.detach().half())
    rounded = t16.float()
    L, F = cfg["n_levels"], cfg["n_features_per_level"]
    a = torch.empty(n, L * F, device=dev)
    b = torch.empty(n, L * F, device=dev)
    n_dev = torch.tensor([n], dtype=torch.int32, device=dev)
    call("sn_hashgrid_fwd16_dev", ptr(x), n, ptr(n_dev), ptr(t16), *args, None, ptr(a), st)
    call("sn_hashgrid_fwd", ptr(x), n, ptr(rounded), *args, ptr(b), st)
    assert torch.equal(a, b)
    # Adam refreshes the copy of a sub-range of a larger flat buffer
    m = e.params.numel()
    flat = torch.randn(m + 1003, generator=g).to(dev)
    grad = torch.randn(m + 1003, generator=g).to(dev)
    mom1, mom2 = torch.zeros_like(flat), torch.zeros_like(flat)
    call("sn_adam_step_copy16", ptr(flat), ptr(grad), ptr(mom1), ptr(mom2), flat.numel(), 1e-2,
         None, 0.9, 0.999, 1e-8, 1, None, ptr(t16), 1001, m, None, st)
    assert torch.equal(t16, flat[1001:1001 + m].half())


@pytest.mark.parametrize("k", [0, 3, 16])
def test_device_level_count(dev, k):
    """The active level count read from device memory (n_active_dev, bounded
    by the host count 16) gives the same forward and backward, bit for bit,
    as the host count k."""
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of
    cfg = CFGS["cfg2"]
    e = _enc(cfg, dev, seed=9)
    g = torch.Generator().manual_seed(10)
    x = (torch.rand(5000, 3, generator=g) * 2 - 1).to(dev)
    genc = torch.randn(5000, 32, generator=g).to(dev)
    kd = torch.tensor([k], dtype=torch.int32, device=dev)
    s = stream_of(dev)
    outs = []
    for host_k, dev_k in ((k, None), (16, kd)):
        gargs = (16, 2, 19, 16, 1.38, host_k)
        enc = torch.empty(5000, 32, device=dev)
        call("sn_hashgrid_fwd_dev", ptr(x), 5000, None, ptr(e.params), *gargs, ptr(dev_k),
             ptr(enc), s)
        nb = _lib.lib().sn_hashgrid_bwd_workspace_bytes(16, 2, 19, 16, 1.38, 65536)
        # a fresh (zeroed) workspace each: a workspace carries the previous
        # call's list sizes, which could make another input's lists overflow
        ws = torch.zeros(nb, dtype=torch.uint8, device=dev)
        gp = torch.zeros_like(e.params)
        gx = torch.empty_like(x)
        call("sn_hashgrid_bwd_ws_dev", ptr(x), 5000, None, ptr(e.params), *gargs, ptr(dev_k),
             ptr(genc), ptr(gp), ptr(gx), ptr(ws), nb, s)
        torch.cuda.synchronize()
        outs.append((enc, gp, gx))
    for a, b in zip(*outs):
        assert torch.equal(a, b)

