"""GPU parity of the MFMA MLP kernels and the fused SDF (hash + MLP) against
fp64 torch autograd (the oracle's field restatement).

Tolerances (fp32 kernels vs fp64): values |err| <= 1e-5 + 1e-4 |ref|;
gradients summed over thousands of points |err| <= 2e-4 * max|ref| + 1e-4 |ref|."""
import numpy as np
import pytest
import torch

from oracle import field_oracle
from oracle.hashgrid_oracle import hashgrid_desc

pytestmark = pytest.mark.gpu

CFGS = [(32, 1, 16), (64, 2, 32), (64, 1, 28), (64, 2, 16), (32, 2, 32)]  # (H, NH, E)


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


def _close(got, ref, rel=1e-4, scale_tol=2e-4):
    got = got.detach().double().cpu()
    ref = ref.detach().double().cpu()
    atol = scale_tol * float(ref.abs().max()) + 1e-6
    torch.testing.assert_close(got, ref, rtol=rel, atol=atol)


def _call(name, *a):
    from supernormal_amd._lib import call
    call(name, *a)


@pytest.mark.parametrize("H,NH,E", CFGS)
def test_mlp_fwd_bwd_vs_fp64(H, NH, E, dev):
    from supernormal_amd import _lib
    from supernormal_amd._lib import ptr, stream_of, workspace
    d = 3 + E
    n = 4000 + 13
    layers = _layers(H, NH, d, seed=H + NH + E)
    g = torch.Generator().manual_seed(1)
    x = torch.rand(n, 3, generator=g, dtype=torch.float64) * 2 - 1
    enc = torch.randn(n, E, generator=g, dtype=torch.float64) * 0.3
    gs = torch.randn(n, generator=g, dtype=torch.float64)
    # fp64 reference
    h0 = torch.cat([x, enc], 1).requires_grad_(True)
    lay = [(W.clone().requires_grad_(True), b.clone().requires_grad_(True)) for W, b in layers]
    s = field_oracle.mlp(h0, lay)[:, 0]
    dh0, = torch.autograd.grad(s.sum(), h0, retain_graph=True)
    (s * gs).sum().backward()
    g_mlp_ref = _pack([(W.grad, b.grad) for W, b in lay])
    # kernels
    xd, encd_ = x.float().to(dev), enc.float().to(dev).contiguous()
    mlp = _pack(layers).float().to(dev)
    sdf = torch.empty(n, device=dev)
    dh0_k = torch.empty(n, d, device=dev)
    st = stream_of(dev)
    _call("sn_mlp_fwd", ptr(xd), ptr(encd_), E, n, None, ptr(mlp), H, NH, 100.0, ptr(sdf),
          ptr(dh0_k), st)
    _close(sdf, s, scale_tol=1e-5)
    _close(dh0_k, dh0)
    gsd = gs.float().to(dev)
    g_enc = torch.empty(n, E, device=dev)
    g_x = torch.empty(n, 3, device=dev)
    g_mlp = torch.empty_like(mlp)
    ws = workspace(_lib.lib().sn_mlp_bwd_workspace_bytes(n, E, H, NH, 0), dev)
    _call("sn_mlp_bwd", ptr(xd), ptr(encd_), E, n, None, ptr(mlp), H, NH, 100.0, ptr(gsd),
          ptr(g_enc), ptr(g_x), ptr(g_mlp), 0, ptr(ws), ws.numel(), st)
    _close(g_mlp, g_mlp_ref)
    _close(g_enc, h0.grad[:, 3:])
    _close(g_x, h0.grad[:, :3])
    # device-side count: only the first m rows are used
    m = torch.tensor([1000], dtype=torch.int32, device=dev)
    sdf2 = torch.full((n,), 7.0, device=dev)
    _call("sn_mlp_fwd", ptr(xd), ptr(encd_), E, n, ptr(m), ptr(mlp), H, NH, 100.0, ptr(sdf2),
          None, st)
    assert torch.equal(sdf2[:1000], sdf[:1000]) and (sdf2[1000:] == 7.0).all()


@pytest.mark.parametrize("H,NH,E", CFGS)
def test_mlp_bwd_ad_vs_fp64(H, NH, E, dev):
    from supernormal_amd import _lib
    from supernormal_amd._lib import ptr, stream_of, workspace
    d = 3 + E
    n = 3000 + 7
    layers = _layers(H, NH, d, seed=2 * H + NH + E)
    g = torch.Generator().manual_seed(2)
    x = torch.rand(n, 3, generator=g, dtype=torch.float64) * 2 - 1
    enc = torch.randn(n, E, generator=g, dtype=torch.float64) * 0.3
    v = torch.randn(n, 3, generator=g, dtype=torch.float64)
    encd = torch.randn(n, E, generator=g, dtype=torch.float64) * 0.5
    a = torch.randn(n, generator=g, dtype=torch.float64)
    h0 = torch.cat([x, enc], 1).requires_grad_(True)
    hd0 = torch.cat([v, encd], 1).requires_grad_(True)
    lay = [(W.clone().requires_grad_(True), b.clone().requires_grad_(True)) for W, b in layers]
    s = field_oracle.mlp(h0, lay)[:, 0]
    dh0, = torch.autograd.grad(s.sum(), h0, create_graph=True)
    J = (a * s).sum() + (dh0 * hd0).sum()
    J.backward()
    g_mlp_ref = _pack([(W.grad, b.grad) for W, b in lay])
    dev_ = lambda t: t.float().to(dev).contiguous()
    xd, ed, vd, edd, ad = dev_(x), dev_(enc), dev_(v), dev_(encd), dev_(a)
    mlp = _pack(layers).float().to(dev)
    gh0 = torch.empty(n, d, device=dev)
    gh0d = torch.empty(n, d, device=dev)
    g_mlp = torch.empty_like(mlp)
    ws = workspace(_lib.lib().sn_mlp_bwd_workspace_bytes(n, E, H, NH, 1), dev)
    _call("sn_mlp_bwd_ad", ptr(xd), ptr(ed), ptr(vd), ptr(edd), E, n, None, ptr(mlp), H, NH,
          100.0, ptr(ad), ptr(gh0), ptr(gh0d), ptr(g_mlp), 0, ptr(ws), ws.numel(),
          stream_of(dev))
    _close(g_mlp, g_mlp_ref)
    _close(gh0, h0.grad)
    _close(gh0d, hd0.grad)


def _network(dev, n_layers, d_hidden, grid_cfg, bw):
    from supernormal_amd.fields import SDFNetwork
    torch.manual_seed(0)
    net = SDFNetwork(d_in=3, d_out=1, d_hidden=d_hidden, n_layers=n_layers, skip_in=[-1],
                     bias=0.6, geometric_init=True, weight_norm=True,
                     encoding_config={"otype": "HashGrid", **grid_cfg}, input_concat=True).to(dev)
    with torch.no_grad():
        net.encoding.params.uniform_(-0.05, 0.05)
        # move off the geometric init so every weight matters
        for l in range(net.num_layers - 1):
            lin = getattr(net, f"lin{l}")
            lin.weight_v.add_(torch.randn_like(lin.weight_v) * 0.05)
    net.bindwidth = bw
    return net


GRID1 = dict(n_levels=8, n_features_per_level=2, log2_hashmap_size=15, base_resolution=16,
             per_level_scale=1.5)
GRID2 = dict(n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
             per_level_scale=1.38)


def _oracle_inputs(net):
    desc = hashgrid_desc(net.encoding.grid["n_levels"], net.encoding.grid["log2_hashmap_size"],
                         net.encoding.grid["base_resolution"], net.encoding.grid["per_level_scale"])
    p64 = net.encoding.params.detach().cpu().double().requires_grad_(True)
    vg = []
    layers = []
    for l in range(net.num_layers - 1):
        lin = getattr(net, f"lin{l}")
        v = lin.weight_v.detach().cpu().double().requires_grad_(True)
        gg = lin.weight_g.detach().cpu().double().requires_grad_(True)
        b = lin.bias.detach().cpu().double().requires_grad_(True)
        vg.append((v, gg, b))
        layers.append((field_oracle.weight_norm(v, gg), b))
    return desc, p64, vg, layers


@pytest.mark.parametrize("n_layers,d_hidden,grid,bw", [(1, 32, GRID1, 8), (2, 64, GRID2, 16),
                                                       (2, 64, GRID2, 5)])
def test_fused_sdf_vs_oracle(n_layers, d_hidden, grid, bw, dev):
    net = _network(dev, n_layers, d_hidden, grid, bw)
    n = 3000
    gen = torch.Generator().manual_seed(3)
    x = (torch.rand(n, 3, generator=gen) * 1.6 - 0.8)
    c = torch.randn(n, generator=gen)
    xd = x.to(dev)
    s = net.fused_sdf(xd)
    # the reference-structured path (Encoding kernel + torch MLP) agrees too
    s_ref_path = net.sdf(xd)[:, 0]
    torch.testing.assert_close(s, s_ref_path, rtol=1e-4, atol=1e-5)
    (s * c.to(dev)).sum().backward()
    desc, p64, vg, layers = _oracle_inputs(net)
    so = field_oracle.sdf(x.double(), p64, layers, desc, n_active=net.active_levels())
    (so * c.double()).sum().backward()
    _close(s, so, scale_tol=1e-5)
    _close(net.encoding.params.grad, p64.grad)
    for l, (v, gg, b) in enumerate(vg):
        lin = getattr(net, f"lin{l}")
        _close(lin.weight_v.grad, v.grad)
        _close(lin.weight_g.grad, gg.grad)
        _close(lin.bias.grad, b.grad)


@pytest.mark.parametrize("n_layers,d_hidden,grid,bw", [(1, 32, GRID1, 8), (2, 64, GRID2, 16)])
def test_fused_sdf_gradient_double_backward_vs_oracle(n_layers, d_hidden, grid, bw, dev):
    """L = sum c1*sdf + sum c2 . grad_x sdf: the analytic gradient and the double
    backward the 'ad' gradient method needs (fields.py:107-119, create_graph)."""
    net = _network(dev, n_layers, d_hidden, grid, bw)
    n = 2000
    gen = torch.Generator().manual_seed(4)
    x = (torch.rand(n, 3, generator=gen) * 1.6 - 0.8)
    c1 = torch.randn(n, generator=gen)
    c2 = torch.randn(n, 3, generator=gen)
    s, gx = net.fused_sdf_and_gradient(x.to(dev))
    ((s * c1.to(dev)).sum() + (gx * c2.to(dev)).sum()).backward()
    desc, p64, vg, layers = _oracle_inputs(net)
    so, go = field_oracle.sdf_and_gradient(x.double(), p64, layers, desc,
                                           n_active=net.active_levels())
    ((so * c1.double()).sum() + (go * c2.double()).sum()).backward()
    _close(s, so, scale_tol=1e-5)
    _close(gx, go)
    _close(net.encoding.params.grad, p64.grad)
    for l, (v, gg, b) in enumerate(vg):
        lin = getattr(net, f"lin{l}")
        _close(lin.weight_v.grad, v.grad)
        _close(lin.weight_g.grad, gg.grad)
        _close(lin.bias.grad, b.grad)
    # the reference-structured path's autograd double backward agrees with the fused one
    net.zero_grad()
    gr = net.gradient(x.to(dev))[:, 0]
    torch.testing.assert_close(gr, gx, rtol=1e-3, atol=1e-4)
