"""GPU parity of the fused per-patch render + loss kernel (sn_render_step)
against the fp64 torch restatement of renderer.py:142-276 / exp_runner.py:169-203
(oracle/render_oracle.py) on identical seeded inputs.

Tolerance: fp32 kernel vs fp64 oracle, |err| <= 2e-3 |ref| + 2e-4 max|ref|
(the NeuS alpha divides differences of two sigmoids, so fp32 rounding is
amplified near the surface; the fp32 run of the same oracle is checked to
sit inside the same bound)."""
import numpy as np
import pytest
import torch

from oracle import render_oracle

pytestmark = pytest.mark.gpu


def _scene(B, Ph, Pw, seed, max_cnt=40):
    g = torch.Generator().manual_seed(seed)
    P = Ph * Pw
    # a camera on a ring looking at the origin, 3x3 pixel patches
    cam = torch.tensor([0.0, -0.4, -2.5], dtype=torch.float64)
    fwd = -cam / cam.norm()
    right = torch.linalg.cross(fwd, torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64))
    right = right / right.norm()
    down = torch.linalg.cross(fwd, right)
    cx = torch.rand(B, generator=g, dtype=torch.float64) * 0.6 - 0.3
    cy = torch.rand(B, generator=g, dtype=torch.float64) * 0.6 - 0.3
    ii, jj = torch.meshgrid(torch.arange(Ph) - Ph // 2, torch.arange(Pw) - Pw // 2, indexing="ij")
    px = cx[:, None, None] + jj[None].double() * 0.002
    py = cy[:, None, None] + ii[None].double() * 0.002
    d = fwd + px[..., None] * right + py[..., None] * down
    d = d / d.norm(dim=-1, keepdim=True)
    o = cam.expand(B, Ph, Pw, 3).clone()
    V = torch.stack([d, right.expand_as(d), down.expand_as(d)], -2)
    vinv = torch.inverse(V)
    n = fwd.expand(B, 3).clone()
    n_true = torch.randn(B, Ph, Pw, 3, generator=g, dtype=torch.float64)
    n_true = n_true / n_true.norm(dim=-1, keepdim=True)
    mask = (torch.rand(B, Ph, Pw, 1, generator=g) > 0.3).double()
    cnt = torch.randint(0, max_cnt + 1, (B,), generator=g)
    cnt[0] = 0
    cnt[1] = 1
    ri, t0, t1 = [], [], []
    for b in range(B):
        t = 1.8 + float(torch.rand(1, generator=g)) * 0.2
        for j in range(int(cnt[b])):
            if torch.rand(1, generator=g) < 0.15:  # skipped empty space: a gap
                t += 0.013
            ri.append(b)
            t0.append(t)
            t += 0.004
            t1.append(t)
    ri = torch.tensor(ri, dtype=torch.long)
    # round t to fp32 so the kernel and the oracle see the same starts/ends
    t0 = torch.tensor(t0, dtype=torch.float32).double()[:, None]
    t1 = torch.tensor(t1, dtype=torch.float32).double()[:, None]
    packed = torch.stack([torch.cumsum(cnt, 0) - cnt, cnt], -1).int()
    return dict(o=o, d=d, n=n, vinv=vinv, n_true=n_true, mask=mask, packed=packed, ri=ri, t0=t0,
                t1=t1)


def _sdf_values(sc, g):
    """Smooth SDF-like values: a plane crossing at t ~ 2.4 seen along each ray."""
    ts = render_oracle.expand_t(sc["t0"], sc["ri"], sc["d"], sc["n"])
    te = render_oracle.expand_t(sc["t1"], sc["ri"], sc["d"], sc["n"])
    dm = render_oracle.diff_mask(sc["ri"], sc["t0"], sc["t1"])
    ps = render_oracle.positions(sc["o"], sc["d"], sc["ri"], ts)
    pe = render_oracle.positions(sc["o"], sc["d"], sc["ri"], te)

    def f(p):
        return (p[..., 2:3] + 0.1 + 0.05 * torch.sin(7 * p[..., 0:1]) * torch.cos(5 * p[..., 1:2]))

    sdf_s = f(ps) + 1e-4 * torch.randn(ps.shape[:-1] + (1,), generator=g, dtype=torch.float64)
    sdf_e = f(pe)[dm]
    return sdf_s.float().double(), sdf_e.float().double(), dm


@pytest.mark.parametrize("impl", [1, 0])
@pytest.mark.parametrize("method,loss_type,Ph,Pw,max_cnt", [
    ("dfd", "l2", 3, 3, 40), ("ad", "l2", 3, 3, 40), ("dfd", "l1", 3, 3, 40),
    ("dfd", "l2", 5, 5, 40), ("dfd", "l2", 2, 4, 40), ("dfd", "l2", 3, 3, 150),
    ("ad", "l1", 3, 3, 150), ("dfd", "l2", 2, 4, 130), ("dfd", "l2", 3, 3, 600),
    ("ad", "l2", 2, 4, 600)])
def test_render_step_vs_fp64(method, loss_type, Ph, Pw, max_cnt, impl, dev):
    """impl 1: sample-parallel kernel (3x3 and 2x4 patches; wavefront scans,
    several 64-sample chunks with carries between them when max_cnt > 64, up
    to 10 at 600); impl 0: ray-parallel kernel."""
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of
    _lib.lib().sn_render_step_set_impl(impl)
    try:
        _run_render_step(method, loss_type, Ph, Pw, max_cnt, dev)
    finally:
        _lib.lib().sn_render_step_set_impl(1)


def _run_render_step(method, loss_type, Ph, Pw, max_cnt, dev):
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of
    B = 300
    sc = _scene(B, Ph, Pw, seed=Ph * 10 + Pw + (1 if method == "ad" else 0) + max_cnt,
                max_cnt=max_cnt)
    g = torch.Generator().manual_seed(99)
    sdf_s, sdf_e, dm = _sdf_values(sc, g)
    S, Nd, P = sc["ri"].shape[0], sdf_e.shape[0], Ph * Pw
    inv_s = torch.tensor(148.4, dtype=torch.float64)
    grad_s = None
    if method == "ad":
        grad_s = torch.randn(S, Ph, Pw, 3, generator=g, dtype=torch.float64).float().double()
    # fp64 oracle
    leaves = [sdf_s.clone().requires_grad_(True), sdf_e.clone().requires_grad_(True),
              inv_s.clone().requires_grad_(True)]
    gl = grad_s.clone().requires_grad_(True) if grad_s is not None else None
    f32 = lambda t: t.float().double()
    loss, parts, cn, ws = render_oracle.render_losses(
        f32(sc["o"]), f32(sc["d"]), f32(sc["n"]), f32(sc["vinv"]), f32(sc["n_true"]), sc["mask"],
        sc["packed"], sc["ri"], sc["t0"], sc["t1"], leaves[0], leaves[1], leaves[2], gl,
        loss_type=loss_type)
    loss.backward()
    # kernel
    D = lambda t, dt=torch.float32: t.to(dt).contiguous().to(dev)
    diff_idx = torch.full((S,), -1, dtype=torch.int32)
    diff_idx[dm] = torch.arange(int(dm.sum()), dtype=torch.int32)
    n_s = torch.tensor([S], dtype=torch.int32, device=dev)
    mask_sum = torch.tensor([float(sc["mask"].sum()) + 1e-5], device=dev)
    inv_d = torch.tensor([148.4], device=dev)
    buf = dict(o=D(sc["o"]), d=D(sc["d"]), n=D(sc["n"]), vinv=D(sc["vinv"]),
               nt=D(sc["n_true"]), m=D(sc["mask"]), pk=D(sc["packed"], torch.int32),
               t0=D(sc["t0"][:, 0]), t1=D(sc["t1"][:, 0]), di=D(diff_idx, torch.int32),
               ss=D(sdf_s), se=D(sdf_e) if Nd else torch.zeros(1, device=dev))
    gs_k = D(grad_s) if grad_s is not None else None
    g_ss = torch.empty(S * P, device=dev)
    g_se = torch.empty(max(Nd, 1) * P, device=dev)
    g_gs = torch.empty(S * P * 3, device=dev) if grad_s is not None else None
    losses = torch.empty(4, device=dev)
    waves = _lib.lib().sn_render_step_waves(B, P)
    partials = torch.empty(waves * 4, device=dev)
    ws_k = torch.empty(B * P, device=dev)
    cn_k = torch.empty(B * P * 3, device=dev)
    call("sn_render_step", B, Ph, Pw, ptr(buf["o"]), ptr(buf["d"]), ptr(buf["n"]),
         ptr(buf["vinv"]), ptr(buf["nt"]), ptr(buf["m"]), ptr(buf["pk"]), ptr(buf["t0"]),
         ptr(buf["t1"]), ptr(buf["di"]), ptr(buf["ss"]), ptr(buf["se"]), ptr(gs_k),
         ptr(inv_d), ptr(mask_sum), ptr(n_s), 1.0, 1.0, 1.0, 1 if loss_type == "l1" else 0,
         ptr(g_ss), ptr(g_se), ptr(g_gs), ptr(losses), ptr(partials), ptr(ws_k), ptr(cn_k),
         stream_of(dev))
    torch.cuda.synchronize()

    def close(got, ref, rel=2e-3, scale=2e-4):
        got = got.detach().double().cpu().reshape(ref.shape)
        ref = ref.detach().double()
        torch.testing.assert_close(got, ref, rtol=rel, atol=scale * float(ref.abs().max()) + 1e-7)

    close(ws_k, ws)
    close(cn_k, cn)
    tot = losses.cpu().double()
    close(tot[0], parts["normal"])
    close(tot[1], parts["mask"])
    close(tot[2], parts["eikonal"])
    close(g_ss, leaves[0].grad)
    if Nd:
        close(g_se[: Nd * P], leaves[1].grad)
    close(tot[3], leaves[2].grad)
    if grad_s is not None:
        close(g_gs, gl.grad)
