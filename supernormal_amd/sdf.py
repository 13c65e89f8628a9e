"""SDF evaluation on the gfx950 kernels: hash-grid encode + MFMA MLP.

``sdf(x)`` and ``sdf_and_gradient(x)`` evaluate SuperNormal's SDFNetwork
(models/fields.py:76-119) with hand-written backward passes:
  * sdf:               x -> hash encode -> MLP -> sdf;  backward: MLP backward
                       (weight gradients summed in fixed order) -> dL/denc ->
                       hash backward (bucket records summed in fixed point, with
                       a workspace kept between calls).
  * sdf_and_gradient:  also the analytic x-gradient
                       grad = dsdf/dx|_direct + sum_l dsdf/denc_l * denc_l/dx;
                       its backward is the double backward the reference gets
                       from autograd.grad(create_graph=True): forward-over-
                       reverse through the MLP and the hash grid's
                       backward-of-backward, recomputing activations.
Positions x are treated as constants (the reference computes them under
no_grad from ray marching); asking for dL/dx raises.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import call, ptr, stream_of, workspace
from .tcnn import hashgrid_bwd


class SDFConfig:
    """Static shape of the field: hash grid + MLP."""

    def __init__(self, grid: dict, d_hidden: int, n_layers: int, beta: float = 100.0,
                 precision: str = "fp32"):
        if precision not in ("fp32", "fp16"):
            raise ValueError("precision must be 'fp32' or 'fp16'")
        self.grid = grid
        self.d_hidden = int(d_hidden)
        self.n_layers = int(n_layers)
        self.beta = float(beta)
        self.E = grid["n_levels"] * grid["n_features"]
        # 'fp16': MLP on the fp16/bf16 matrix cores (sn_mlp16_*) for the sdf
        # value and its backward; the analytic gradient stays on fp32 kernels
        self.precision = precision
        # the hash-grid backward's bucket workspace, kept between calls
        from .tcnn import BucketWorkspace
        self.workspace = BucketWorkspace()

    def ws(self, n_active, n, device):
        return self.workspace.get(self.hash_args(n_active), n, device)

    def hash_args(self, n_active):
        g = self.grid
        return (g["n_levels"], g["n_features"], g["log2_hashmap_size"], g["base_resolution"],
                g["per_level_scale"], int(n_active))


def pack_mlp(layers):
    """[(W [out,in], b [out])] -> flat packed buffer (differentiable)."""
    return torch.cat([t.reshape(-1) for W, b in layers for t in (W, b)])


def _hash_fwd(cfg, x, params, n_active, n_dev=None):
    enc = torch.empty((x.shape[0], cfg.E), dtype=torch.float32, device=x.device)
    call("sn_hashgrid_fwd", ptr(x), x.shape[0], ptr(params), *cfg.hash_args(n_active), ptr(enc),
         stream_of(x.device))
    return enc


def _mlp_fwd(cfg, x, enc, mlp, want_dh0):
    n = x.shape[0]
    sdf = torch.empty(n, dtype=torch.float32, device=x.device)
    if cfg.precision == "fp16" and not want_dh0:
        call("sn_mlp16_fwd", ptr(x), ptr(enc), cfg.E, n, None, ptr(mlp), cfg.d_hidden,
             cfg.n_layers, cfg.beta, ptr(sdf), stream_of(x.device))
        return sdf, None
    dh0 = torch.empty((n, 3 + cfg.E), dtype=torch.float32, device=x.device) if want_dh0 else None
    call("sn_mlp_fwd", ptr(x), ptr(enc), cfg.E, n, None, ptr(mlp), cfg.d_hidden, cfg.n_layers,
         cfg.beta, ptr(sdf), ptr(dh0), stream_of(x.device))
    return sdf, dh0


class _SDFFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params, mlp, cfg, n_active):
        x = x.contiguous()
        enc = _hash_fwd(cfg, x, params, n_active)
        sdf, _ = _mlp_fwd(cfg, x, enc, mlp, False)
        ctx.save_for_backward(x, enc, params, mlp)
        ctx.cfg, ctx.n_active = cfg, n_active
        return sdf

    @staticmethod
    def backward(ctx, g_sdf):
        x, enc, params, mlp = ctx.saved_tensors
        cfg = ctx.cfg
        if ctx.needs_input_grad[0]:
            raise NotImplementedError("the fused SDF path treats positions as constants")
        n = x.shape[0]
        dev = x.device
        g_sdf = g_sdf.contiguous()
        g_enc = torch.empty((n, cfg.E), dtype=torch.float32, device=dev)
        g_mlp = torch.empty_like(mlp)
        lib = _lib.lib()
        s = stream_of(dev)
        if cfg.precision == "fp16":
            ws = workspace(lib.sn_mlp16_bwd_workspace_bytes(n, cfg.E, cfg.d_hidden,
                                                            cfg.n_layers), dev)
            call("sn_mlp16_bwd", ptr(x), ptr(enc), cfg.E, n, None, ptr(mlp), cfg.d_hidden,
                 cfg.n_layers, cfg.beta, ptr(g_sdf), ptr(g_enc), None, ptr(g_mlp), 0, ptr(ws),
                 ws.numel(), s)
        else:
            ws = workspace(lib.sn_mlp_bwd_workspace_bytes(n, cfg.E, cfg.d_hidden, cfg.n_layers,
                                                          0), dev)
            call("sn_mlp_bwd", ptr(x), ptr(enc), cfg.E, n, None, ptr(mlp), cfg.d_hidden,
                 cfg.n_layers, cfg.beta, ptr(g_sdf), ptr(g_enc), None, ptr(g_mlp), 0, ptr(ws),
                 ws.numel(), s)
        g_params = torch.zeros_like(params)
        hashgrid_bwd(x, n, None, params, cfg.hash_args(ctx.n_active), g_enc, g_params, None, s,
                     ws=cfg.ws(ctx.n_active, n, dev))
        return None, g_params, g_mlp, None, None


class _SDFGradFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params, mlp, cfg, n_active):
        x = x.contiguous()
        n = x.shape[0]
        enc = _hash_fwd(cfg, x, params, n_active)
        sdf, dh0 = _mlp_fwd(cfg, x, enc, mlp, True)
        grad = torch.empty((n, 3), dtype=torch.float32, device=x.device)
        g_enc = dh0[:, 3:].contiguous()
        call("sn_hashgrid_bwd", ptr(x), n, ptr(params), *cfg.hash_args(n_active), ptr(g_enc),
             None, ptr(grad), stream_of(x.device))
        grad += dh0[:, :3]
        ctx.save_for_backward(x, enc, params, mlp)
        ctx.cfg, ctx.n_active = cfg, n_active
        return sdf, grad

    @staticmethod
    def backward(ctx, g_sdf, g_grad):
        x, enc, params, mlp = ctx.saved_tensors
        cfg = ctx.cfg
        if ctx.needs_input_grad[0]:
            raise NotImplementedError("the fused SDF path treats positions as constants")
        n = x.shape[0]
        dev = x.device
        s = stream_of(dev)
        lib = _lib.lib()
        if g_sdf is None:
            g_sdf = torch.zeros(n, dtype=torch.float32, device=dev)
        g_sdf = g_sdf.contiguous()
        g_params = torch.zeros_like(params)
        g_mlp = torch.empty_like(mlp)
        if g_grad is None:
            g_enc = torch.empty((n, cfg.E), dtype=torch.float32, device=dev)
            ws = workspace(lib.sn_mlp_bwd_workspace_bytes(n, cfg.E, cfg.d_hidden, cfg.n_layers,
                                                          0), dev)
            call("sn_mlp_bwd", ptr(x), ptr(enc), cfg.E, n, None, ptr(mlp), cfg.d_hidden,
                 cfg.n_layers, cfg.beta, ptr(g_sdf), ptr(g_enc), None, ptr(g_mlp), 0, ptr(ws),
                 ws.numel(), s)
            hashgrid_bwd(x, n, None, params, cfg.hash_args(ctx.n_active), g_enc, g_params,
                         None, s, ws=cfg.ws(ctx.n_active, n, dev))
            return None, g_params, g_mlp, None, None
        v = g_grad.contiguous()
        hargs = cfg.hash_args(ctx.n_active)
        encd = torch.empty((n, cfg.E), dtype=torch.float32, device=dev)
        call("sn_hashgrid_bwd_bwd", ptr(x), n, ptr(params), *hargs, None, ptr(v), ptr(encd),
             None, None, s)
        d = 3 + cfg.E
        g_h0 = torch.empty((n, d), dtype=torch.float32, device=dev)
        g_h0d = torch.empty((n, d), dtype=torch.float32, device=dev)
        ws = workspace(lib.sn_mlp_bwd_workspace_bytes(n, cfg.E, cfg.d_hidden, cfg.n_layers, 1),
                       dev)
        call("sn_mlp_bwd_ad", ptr(x), ptr(enc), ptr(v), ptr(encd), cfg.E, n, None, ptr(mlp),
             cfg.d_hidden, cfg.n_layers, cfg.beta, ptr(g_sdf), ptr(g_h0), ptr(g_h0d),
             ptr(g_mlp), 0, ptr(ws), ws.numel(), s)
        ge = g_h0[:, 3:].contiguous()
        ged = g_h0d[:, 3:].contiguous()
        # the table: first-order term of ge and the double backward's term of
        # (ged, v) in one pass (shared bucket records)
        hws = cfg.ws(ctx.n_active, n, dev)
        call("sn_hashgrid_bwd_ad_ws_dev", ptr(x), n, None, ptr(params), *hargs, None, ptr(ge),
             ptr(v), ptr(ged), ptr(g_params), ptr(hws), 0 if hws is None else hws.numel(), s)
        return None, g_params, g_mlp, None, None


def sdf(x, params, mlp, cfg: SDFConfig, n_active: int):
    """sdf [N] of positions x [N,3] (GPU tensors)."""
    if not x.is_cuda:
        raise NotImplementedError("GPU tensors only")
    return _SDFFn.apply(x.to(torch.float32), params, mlp, cfg, int(n_active))


def sdf_and_gradient(x, params, mlp, cfg: SDFConfig, n_active: int):
    """(sdf [N], d sdf/dx [N,3]); differentiable twice w.r.t. the parameters."""
    if not x.is_cuda:
        raise NotImplementedError("GPU tensors only")
    return _SDFGradFn.apply(x.to(torch.float32), params, mlp, cfg, int(n_active))
