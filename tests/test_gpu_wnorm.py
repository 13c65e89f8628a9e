"""Weight-norm packing of the SDF MLP on the device (sn_mlp_pack_fwd/bwd)
against torch autograd in fp64 of the reference's formulation:
torch.nn.utils.weight_norm (W = g v / ||v|| per row, fields.py:69-70),
pack_mlp's concatenation and inv_s = clip(exp(10 variance), 1e-6, 1e6).
Tolerance: rtol 1e-5 (fp32 kernels vs fp64)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _net(dev, weight_norm):
    from supernormal_amd.fields import SDFNetwork, SingleVarianceNetwork
    from supernormal_amd.fused import FlatParams
    enc = dict(otype="HashGrid", n_levels=4, n_features_per_level=2, log2_hashmap_size=12,
               base_resolution=8, per_level_scale=1.5)
    net = SDFNetwork(d_in=3, d_out=1, d_hidden=64, n_layers=2, skip_in=[-1], bias=0.6,
                     geometric_init=True, weight_norm=weight_norm, encoding_config=enc,
                     input_concat=True).to(dev)
    var = SingleVarianceNetwork(0.3).to(dev)
    params = list(net.parameters()) + list(var.parameters())
    flat = FlatParams(params, dev)
    g = torch.Generator().manual_seed(0)
    with torch.no_grad():
        flat.data.copy_(torch.randn(flat.n, generator=g).to(dev) * 0.3)
    return net, var, flat


@pytest.mark.parametrize("weight_norm", [True, False])
def test_mlp_pack_matches_autograd(dev, weight_norm):
    from supernormal_amd._lib import call, ptr, stream_of
    from supernormal_amd.step import flat_offset, mlp_pack_table
    net, var, flat = _net(dev, weight_norm)
    table = mlp_pack_table(net, flat)
    n_mlp = int((table[:, 4] * (table[:, 5] + 1)).sum())
    ov = flat_offset(flat, var.variance)
    packed = torch.empty(n_mlp, device=dev)
    inv_s = torch.empty(1, device=dev)
    s = stream_of(dev)
    call("sn_mlp_pack_fwd", ptr(flat.data), len(table), table.ctypes.data, ptr(packed),
         ptr(flat.data[ov:ov + 1]), ptr(inv_s), s)
    # reference in fp64 over copies of the same parameters
    p64 = [p.detach().double().cpu().requires_grad_(True) for p in flat.params]
    idx = {id(p): k for k, p in enumerate(flat.params)}
    layers = []
    for l in range(net.num_layers - 1):
        lin = getattr(net, "lin" + str(l))
        if weight_norm:
            v, gn = p64[idx[id(lin.weight_v)]], p64[idx[id(lin.weight_g)]]
            W = gn * v / v.norm(dim=1, keepdim=True)
        else:
            W = p64[idx[id(lin.weight)]]
        layers.append((W, p64[idx[id(lin.bias)]]))
    ref = torch.cat([t.reshape(-1) for W, b in layers for t in (W, b)])
    vr = p64[idx[id(var.variance)]]
    inv_ref = torch.exp(vr * 10.0).clip(1e-6, 1e6).reshape(1)
    torch.testing.assert_close(packed.double().cpu(), ref.detach().cpu(), rtol=1e-5, atol=1e-7)
    torch.testing.assert_close(inv_s.double().cpu(), inv_ref.detach().cpu(), rtol=1e-5, atol=0)
    # backward: random dL/dpacked and dL/dinv_s, added onto a non-zero gradient
    g = torch.Generator().manual_seed(1)
    g_packed = torch.randn(n_mlp, generator=g).to(dev)
    g_inv = torch.randn(1, generator=g).to(dev)
    base = torch.randn(flat.n, generator=g).to(dev) * 1e-2
    flat.grad.copy_(base)
    call("sn_mlp_pack_bwd", ptr(flat.data), len(table), table.ctypes.data, ptr(g_packed),
         ptr(flat.grad), ptr(flat.data[ov:ov + 1]), ptr(g_inv), ptr(flat.grad[ov:ov + 1]), s)
    torch.autograd.backward([ref, inv_ref], [g_packed.double().cpu(), g_inv.double().cpu()])
    got = flat.grad.double().cpu() - base.double().cpu()
    for k, (p, o) in enumerate(zip(flat.params, flat.offsets)):
        want = p64[k].grad
        if want is None:
            want = torch.zeros_like(p64[k])
        if p is net.encoding.params:
            continue  # not touched by the packing
        torch.testing.assert_close(got[o:o + p.numel()].reshape(p.shape), want.cpu(),
                                   rtol=1e-5, atol=1e-6)
