"""GPU parity of the fp16/bf16-MFMA MLP kernels (sn_mlp16_*) against fp64
torch autograd (the oracle's field restatement) and against the fp32 MFMA
kernels.

The kernels round the weights once to the MFMA type (fp16 forward, bf16
backward) and carry activations/gradients as hi+lo pairs, so the error is
that of a network whose weights are perturbed by at most 2^-11 (forward) or
2^-8 (backward) relative.  Stated tolerances against fp64 (random weights of
scale 1/sqrt(fan_in), |h0| ~ 1):
    sdf:            |err| <= 2e-3 * max|ref|
    dL/denc, dL/dx: |err| <= 1e-2 * max|ref|
    dL/dweights:    |err| <= 1e-2 * max|ref|  (summed over 4013 points)
The measured maxima are printed (run with -s).
"""
import pytest
import torch

from oracle import field_oracle

pytestmark = pytest.mark.gpu

CFGS = [(32, 1, 16), (64, 2, 32), (64, 1, 28), (64, 2, 16), (32, 2, 32), (64, 2, 60), (64, 2, 27)]


def _layers(H, NH, d, seed):
    g = torch.Generator().manual_seed(seed)
    dims = [d] + [H] * NH + [1]
    out = []
    for i in range(len(dims) - 1):
        W = torch.randn(dims[i + 1], dims[i], generator=g, dtype=torch.float64) / dims[i] ** 0.5
        b = torch.randn(dims[i + 1], generator=g, dtype=torch.float64) * 0.1
        out.append((W, b))
    return out


def _pack(layers):
    return torch.cat([t.reshape(-1) for W, b in layers for t in (W, b)])


def _rel_err(got, ref):
    got = got.detach().double().cpu()
    ref = ref.detach().double().cpu()
    return float((got - ref).abs().max() / ref.abs().max().clamp_min(1e-30))


@pytest.mark.parametrize("H,NH,E", CFGS)
@pytest.mark.parametrize("n", [4013, 37])
def test_mlp16_fwd_bwd_vs_fp64(H, NH, E, n, dev):
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of, workspace
    d = 3 + E
    layers = _layers(H, NH, d, seed=H + NH + E)
    g = torch.Generator().manual_seed(1)
    x = torch.rand(n, 3, generator=g, dtype=torch.float64) * 2 - 1
    enc = torch.randn(n, E, generator=g, dtype=torch.float64) * 0.3
    gs = torch.randn(n, generator=g, dtype=torch.float64) * 1e-6
    h0 = torch.cat([x, enc], 1).requires_grad_(True)
    lay = [(W.clone().requires_grad_(True), b.clone().requires_grad_(True)) for W, b in layers]
    s = field_oracle.mlp(h0, lay)[:, 0]
    (s * gs).sum().backward()
    g_mlp_ref = _pack([(W.grad, b.grad) for W, b in lay])
    xd, encd = x.float().to(dev), enc.float().to(dev).contiguous()
    mlp = _pack(layers).float().to(dev)
    st = stream_of(dev)
    # capacity larger than the count on the device: rows past it stay untouched
    cap = n + 100
    xc = torch.cat([xd, torch.full((100, 3), 7.0, device=dev)])
    ec = torch.cat([encd, torch.full((100, E), 7.0, device=dev)]).contiguous()
    n_dev = torch.tensor([n], dtype=torch.int32, device=dev)
    sdf = torch.full((cap,), -9.0, device=dev)
    call("sn_mlp16_fwd", ptr(xc), ptr(ec), E, cap, ptr(n_dev), ptr(mlp), H, NH, 100.0,
         ptr(sdf), st)
    e_sdf = _rel_err(sdf[:n], s)
    assert (sdf[n:] == -9.0).all()
    gsd = torch.cat([gs.float().to(dev), torch.ones(100, device=dev)])
    g_enc = torch.full((cap, E), -9.0, device=dev)
    g_x = torch.full((cap, 3), -9.0, device=dev)
    g_mlp = torch.empty_like(mlp)
    ws = workspace(_lib.lib().sn_mlp16_bwd_workspace_bytes(cap, E, H, NH), dev)
    call("sn_mlp16_bwd", ptr(xc), ptr(ec), E, cap, ptr(n_dev), ptr(mlp), H, NH, 100.0,
         ptr(gsd), ptr(g_enc), ptr(g_x), ptr(g_mlp), 0, ptr(ws), ws.numel(), st)
    e_enc = _rel_err(g_enc[:n], h0.grad[:, 3:])
    e_x = _rel_err(g_x[:n], h0.grad[:, :3])
    e_w = _rel_err(g_mlp, g_mlp_ref)
    print(f"\nH={H} NH={NH} E={E} n={n}: sdf {e_sdf:.2e} g_enc {e_enc:.2e} g_x {e_x:.2e} "
          f"g_w {e_w:.2e}")
    assert (g_enc[n:] == -9.0).all() and (g_x[n:] == -9.0).all()
    assert e_sdf <= 2e-3
    assert e_enc <= 1e-2 and e_x <= 1e-2
    assert e_w <= 1e-2
    # accumulate mode adds the same (deterministic) sum onto the existing gradient
    g2 = g_mlp.clone()
    call("sn_mlp16_bwd", ptr(xc), ptr(ec), E, cap, ptr(n_dev), ptr(mlp), H, NH, 100.0,
         ptr(gsd), None, None, ptr(g2), 1, ptr(ws), ws.numel(), st)
    assert torch.equal(g2, 2.0 * g_mlp)


@pytest.mark.parametrize("H,NH,E", CFGS)
@pytest.mark.parametrize("n", [4013, 37])
def test_mlp16_fwd_grad_vs_fp64(H, NH, E, n, dev):
    """sn_mlp16_fwd_grad: sdf and d sdf / d h0 (the analytic gradient's MLP
    part, columns [x, enc]) against fp64 autograd, with a capacity above the
    device count (rows past it untouched).  Stated tolerances: sdf 2e-3,
    d sdf / d h0 5e-3 of max|ref|.  The kernel differentiates the network it
    evaluates, whose weights are rounded to fp16 (a fixed perturbation of up
    to 2^-11 relative); the derivative feels it more than the value: the
    softplus derivative s'(z) = sigmoid(100 z) moves by 25 times the shift of
    z (measured ~2e-3 against ~2e-4 for the value).  Cotangents are carried
    as hi + lo fp16 parts, so no further rounding enters."""
    from supernormal_amd._lib import call, ptr, stream_of
    d = 3 + E
    layers = _layers(H, NH, d, seed=7 + H + NH + E)
    g = torch.Generator().manual_seed(3)
    x = torch.rand(n, 3, generator=g, dtype=torch.float64) * 2 - 1
    enc = torch.randn(n, E, generator=g, dtype=torch.float64) * 0.3
    h0 = torch.cat([x, enc], 1).requires_grad_(True)
    s = field_oracle.mlp(h0, layers)[:, 0]
    ref_g, = torch.autograd.grad(s.sum(), h0)
    cap = n + 100
    xc = torch.cat([x.float().to(dev), torch.full((100, 3), 7.0, device=dev)])
    ec = torch.cat([enc.float().to(dev), torch.full((100, E), 7.0, device=dev)]).contiguous()
    n_dev = torch.tensor([n], dtype=torch.int32, device=dev)
    mlp = _pack(layers).float().to(dev)
    sdf = torch.full((cap,), -9.0, device=dev)
    dh0 = torch.full((cap, d), -9.0, device=dev)
    call("sn_mlp16_fwd_grad", ptr(xc), ptr(ec), E, cap, ptr(n_dev), ptr(mlp), H, NH, 100.0,
         ptr(sdf), ptr(dh0), stream_of(dev))
    e_sdf = _rel_err(sdf[:n], s)
    e_g = _rel_err(dh0[:n], ref_g)
    print(f"\nH={H} NH={NH} E={E} n={n}: sdf {e_sdf:.2e} dsdf/dh0 {e_g:.2e}")
    assert (sdf[n:] == -9.0).all() and (dh0[n:] == -9.0).all()
    assert e_sdf <= 2e-3
    assert e_g <= 5e-3


@pytest.mark.parametrize("H,NH,E", CFGS)
@pytest.mark.parametrize("n", [4013, 37])
def test_mlp16_bwd_ad_vs_fp64(H, NH, E, n, dev):
    """sn_mlp16_bwd_ad_enc: the double backward of the analytic gradient,
    J = sum a sdf(h0) + sum hd0 . d sdf / d h0 (hd0 = [v, enc_dot]), against
    fp64 autograd: dJ/denc, dJ/d enc_dot (= d sdf / d enc) and dJ/dweights.
    Stated tolerances (of max|ref|): 1e-2 for dJ/d enc_dot and dJ/dweights
    (bf16 weights and cotangents, as sn_mlp16_bwd); 2e-2 for dJ/denc, whose
    second-order term carries s''(z) = 100 s'(z)(1 - s'(z)) times the bf16
    rounding of the tangent path (measured 4e-3 - 1.2e-2 over the configs)."""
    from supernormal_amd import _lib
    from supernormal_amd._lib import call, ptr, stream_of, workspace
    d = 3 + E
    layers = _layers(H, NH, d, seed=11 + H + NH + E)
    g = torch.Generator().manual_seed(4)
    x = torch.rand(n, 3, generator=g, dtype=torch.float64) * 2 - 1
    enc = torch.randn(n, E, generator=g, dtype=torch.float64) * 0.3
    v = torch.randn(n, 3, generator=g, dtype=torch.float64) * 1e-3
    encd = torch.randn(n, E, generator=g, dtype=torch.float64) * 1e-2
    a = torch.randn(n, generator=g, dtype=torch.float64) * 1e-3
    h0 = torch.cat([x, enc], 1).requires_grad_(True)
    lay = [(W.clone().requires_grad_(True), b.clone().requires_grad_(True)) for W, b in layers]
    s = field_oracle.mlp(h0, lay)[:, 0]
    gs, = torch.autograd.grad(s.sum(), h0, create_graph=True)
    J = (a * s).sum() + (gs * torch.cat([v, encd], 1)).sum()
    J.backward()
    ref_enc = h0.grad[:, 3:]
    ref_encd = gs.detach()[:, 3:]
    ref_w = _pack([(W.grad, b.grad) for W, b in lay])
    cap = n + 100
    pad = lambda t, c: torch.cat([t.float().to(dev), torch.full((100, c), 7.0, device=dev)])
    xc, ec, vc, edc = pad(x, 3), pad(enc, E).contiguous(), pad(v, 3), pad(encd, E).contiguous()
    ac = torch.cat([a.float().to(dev), torch.ones(100, device=dev)])
    n_dev = torch.tensor([n], dtype=torch.int32, device=dev)
    mlp = _pack(layers).float().to(dev)
    g_enc = torch.full((cap, E), -9.0, device=dev)
    g_encd = torch.full((cap, E), -9.0, device=dev)
    g_mlp = torch.empty_like(mlp)
    ws = workspace(_lib.lib().sn_mlp16_bwd_workspace_bytes(cap, E, H, NH), dev)
    call("sn_mlp16_bwd_ad_enc", ptr(xc), ptr(ec), ptr(vc), ptr(edc), E, cap, ptr(n_dev),
         ptr(mlp), H, NH, 100.0, ptr(ac), ptr(g_enc), ptr(g_encd), ptr(g_mlp), 0, ptr(ws),
         ws.numel(), stream_of(dev))
    e_enc = _rel_err(g_enc[:n], ref_enc)
    e_encd = _rel_err(g_encd[:n], ref_encd)
    e_w = _rel_err(g_mlp, ref_w)
    print(f"\nH={H} NH={NH} E={E} n={n}: dJ/denc {e_enc:.2e} dJ/dencd {e_encd:.2e} "
          f"dJ/dW {e_w:.2e}")
    assert (g_enc[n:] == -9.0).all() and (g_encd[n:] == -9.0).all()
    assert e_enc <= 2e-2 and e_encd <= 1e-2 and e_w <= 1e-2


@pytest.mark.parametrize("H,NH,E", [(64, 2, 32), (32, 1, 16)])
def test_mlp16_matches_fp32_kernels_on_trained_scale(H, NH, E, dev):
    """Hash features at their initial scale (U(+-1e-4)): the fp16 forward
    stays within 2e-3 * max|sdf| of the fp32 kernel on |x| <= 1."""
    from supernormal_amd._lib import call, ptr, stream_of
    d = 3 + E
    layers = _layers(H, NH, d, seed=5)
    n = 20000
    g = torch.Generator().manual_seed(2)
    x = (torch.rand(n, 3, generator=g) * 2 - 1).to(dev)
    enc = (torch.rand(n, E, generator=g) * 2e-4 - 1e-4).to(dev).contiguous()
    mlp = _pack(layers).float().to(dev)
    st = stream_of(dev)
    a = torch.empty(n, device=dev)
    b = torch.empty(n, device=dev)
    call("sn_mlp_fwd", ptr(x), ptr(enc), E, n, None, ptr(mlp), H, NH, 100.0, ptr(a), None, st)
    call("sn_mlp16_fwd", ptr(x), ptr(enc), E, n, None, ptr(mlp), H, NH, 100.0, ptr(b), st)
    err = float((a - b).abs().max() / a.abs().max())
    print(f"\nfp16 vs fp32 forward: {err:.2e}")
    assert err < 2e-3
