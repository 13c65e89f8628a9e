"""tinycudann-compatible ``Encoding`` (HashGrid) on the gfx950 kernels.

Drop-in for the call sites in models/fields.py:4, 25-31, 37, 78:
``Encoding(n_input_dims, encoding_config, seed=1337, dtype=None)`` with
``.n_output_dims``, a flat fp32 ``.params`` parameter (state-dict key
``params``), and ``forward(x [N,3]) -> [N, n_levels * n_features_per_level]``
differentiable w.r.t. x and params, and twice differentiable through the
x-gradient (tcnn's backward_backward_input: d(dL/dx)/d(dL/dy), /d(params),
/d(x)); like tcnn, the gradient of the parameter-gradient is not provided.

Differences from tcnn (stated, not hidden): parameters are kept and computed in
fp32 (tcnn's default on tensor-core GPUs is an fp16 copy with loss scale 128);
initialisation is U(-1e-4, 1e-4) from torch's generator seeded with ``seed``
(tcnn draws from its own PCG32 stream).  ``n_active_levels`` (default: all)
masks levels >= n without computing them -- SuperNormal's coarse-to-fine
bandwidth (fields.py:80-83) uses it instead of a mask multiply.
"""
from __future__ import annotations

import torch
import torch.nn as nn

from . import _lib
from ._lib import call, ptr, stream_of


def _grid_cfg(n_input_dims, cfg):
    otype = cfg.get("otype", "HashGrid")
    if otype not in ("HashGrid", "Grid"):
        raise NotImplementedError(f"encoding otype {otype!r} is not supported (HashGrid only)")
    if otype == "Grid" and cfg.get("type", "Hash") != "Hash":
        raise NotImplementedError("only the Hash grid type is supported")
    if cfg.get("interpolation", "Linear") != "Linear":
        raise NotImplementedError("only Linear interpolation is supported")
    if n_input_dims != 3:
        raise NotImplementedError("only 3-D inputs are supported")
    F = int(cfg.get("n_features_per_level", 2))
    if F not in (1, 2, 4, 8):
        raise ValueError("n_features_per_level must be 1, 2, 4 or 8")
    return dict(
        n_levels=int(cfg.get("n_levels", 16)),
        n_features=F,
        log2_hashmap_size=int(cfg.get("log2_hashmap_size", 19)),
        base_resolution=int(cfg.get("base_resolution", 16)),
        per_level_scale=float(cfg.get("per_level_scale", 2.0)),
    )


def _gargs(g, n_active):
    return (g["n_levels"], g["n_features"], g["log2_hashmap_size"], g["base_resolution"],
            g["per_level_scale"], int(n_active))


_REC_MIN, _REC_MAX, _WS_MAX = 64, 196608, 4 << 30


def bucket_workspace_bytes(gargs, n_points):
    """Bytes of the bucket workspace of the hash-grid backward over up to
    n_points points (0 when the config has no bucket path).  The pool holds a
    record for every corner of every level of every point (8 per point and
    level), at most ~4 GiB; the kernels share it out over the buckets by the
    counts of the previous call on the same workspace, and records beyond a
    bucket's share take the atomic path.  The pool holds 196608 records per
    bucket (measured on a full training run: the late launches, up to ~180k
    records per fine-level list, 6 % faster than with 131072; no change with
    262144); a list may take up to 262144 of them (the coarsest levels' one
    or two buckets take ~230-270k records per late launch)."""
    n_levels, F, log2T, base, pls = gargs[:5]
    lib = _lib.lib()
    one = lib.sn_hashgrid_bwd_workspace_bytes(n_levels, F, log2T, base, pls, 1)
    if one <= 0:
        return 0
    per = 2 + 4 * F
    more = lib.sn_hashgrid_bwd_workspace_bytes(n_levels, F, log2T, base, pls, 1025)
    n_buckets = max(1, round((more - one) / (1024 * per)))
    total = min(8 * n_levels * max(int(n_points), 1), _WS_MAX // per)
    rpb = min(_REC_MAX, max(_REC_MIN, -(-total // n_buckets)))
    return int(lib.sn_hashgrid_bwd_workspace_bytes(n_levels, F, log2T, base, pls, int(rpb)))


def hashgrid_bwd(x, n, n_dev, params, gargs, genc, gparams, gx, stream, ws=None,
                 n_active_dev=None):
    """sn_hashgrid_bwd_ws_dev over n points (n_dev: device count or None) with a
    bucket workspace (allocated here unless given); n_active_dev: device copy
    of the active level count (or None: gargs' count)."""
    if ws is None and gparams is not None:
        nb = bucket_workspace_bytes(gargs, n)
        ws = torch.empty(nb, dtype=torch.uint8, device=x.device) if nb else None
    call("sn_hashgrid_bwd_ws_dev", ptr(x), n, ptr(n_dev), ptr(params), *gargs,
         ptr(n_active_dev), ptr(genc), ptr(gparams), ptr(gx), ptr(ws),
         0 if ws is None else ws.numel(), stream)


class BucketWorkspace:
    """A bucket workspace kept between backward calls (one per encoder): it
    carries the previous call's list sizes, from which each call shares out
    its record pool, so the layout follows the data and lists seldom
    overflow (records beyond a list take order-dependent fp32 atomics).
    Grown, with a quarter in hand, when a call has more points than it was
    sized for; a call with fewer points uses the whole buffer (the pool size,
    and with it the layout, stays the same).  Calls sharing it must not
    overlap (one stream)."""

    def __init__(self):
        self.buf = None
        self.key = None

    def get(self, gargs, n, device):
        nb = bucket_workspace_bytes(gargs, n)
        if not nb:
            return None
        key = (tuple(gargs[:5]), torch.device(device))
        if self.buf is None or self.key != key or self.buf.numel() < nb:
            self.buf = None
            nb = bucket_workspace_bytes(gargs, n + n // 4)
            self.buf = torch.empty(nb, dtype=torch.uint8, device=device)
            self.key = key
        return self.buf


class _HashBackwardFn(torch.autograd.Function):
    """(dL/denc, x, params) -> (dL/dx, dL/dparams); differentiable in its first
    output (tcnn's bwd_bwd_input)."""

    @staticmethod
    def forward(ctx, genc, x, params, g, n_active, want_x, wsh=None):
        genc = genc.contiguous()
        gparams = torch.zeros_like(params)
        gx = torch.empty_like(x) if want_x else None
        gargs = _gargs(g, n_active)
        ws = wsh.get(gargs, x.shape[0], x.device) if wsh is not None else None
        hashgrid_bwd(x, x.shape[0], None, params, gargs, genc, gparams, gx,
                     stream_of(x.device), ws=ws)
        ctx.save_for_backward(genc, x, params)
        ctx.g, ctx.n_active = g, n_active
        if gx is None:
            gx = torch.zeros_like(x)
        return gx, gparams

    @staticmethod
    def backward(ctx, ggx, ggparams):
        genc, x, params = ctx.saved_tensors
        if ggx is None:
            return None, None, None, None, None, None
        ggx = ggx.contiguous()
        need_genc, need_x, need_p = ctx.needs_input_grad[0], ctx.needs_input_grad[1], \
            ctx.needs_input_grad[2]
        gge = torch.empty_like(genc) if need_genc else None
        gp = torch.zeros_like(params) if need_p else None
        gx2 = torch.empty_like(x) if need_x else None
        if gge is not None or gp is not None or gx2 is not None:
            call("sn_hashgrid_bwd_bwd", ptr(x), x.shape[0], ptr(params),
                 *_gargs(ctx.g, ctx.n_active), ptr(genc), ptr(ggx), ptr(gge), ptr(gp), ptr(gx2),
                 stream_of(x.device))
        return gge, gx2, gp, None, None, None, None


class _HashEncodeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params, g, n_active, wsh=None):
        x = x.contiguous()
        enc = torch.empty((x.shape[0], g["n_levels"] * g["n_features"]), dtype=torch.float32,
                          device=x.device)
        call("sn_hashgrid_fwd", ptr(x), x.shape[0], ptr(params), *_gargs(g, n_active), ptr(enc),
             stream_of(x.device))
        ctx.save_for_backward(x, params)
        ctx.g, ctx.n_active, ctx.wsh = g, n_active, wsh
        return enc

    @staticmethod
    def backward(ctx, genc):
        x, params = ctx.saved_tensors
        gx, gp = _HashBackwardFn.apply(genc, x, params, ctx.g, ctx.n_active,
                                       bool(ctx.needs_input_grad[0]), ctx.wsh)
        return (gx if ctx.needs_input_grad[0] else None), gp, None, None, None


class Encoding(nn.Module):
    """tinycudann.Encoding for ``{"otype": "HashGrid", ...}`` with 3-D inputs."""

    def __init__(self, n_input_dims, encoding_config, seed=1337, dtype=None):
        super().__init__()
        self.n_input_dims = n_input_dims
        self.encoding_config = dict(encoding_config)
        self.seed = seed
        self.dtype = torch.float32 if dtype is None else dtype
        self.grid = _grid_cfg(n_input_dims, self.encoding_config)
        g = self.grid
        n_entries = _lib.lib().sn_hashgrid_layout(g["n_levels"], g["log2_hashmap_size"],
                                                  g["base_resolution"], g["per_level_scale"],
                                                  None)
        if n_entries <= 0:
            raise ValueError(f"invalid hash-grid config {encoding_config}")
        self.n_entries = int(n_entries)
        self.n_output_dims = g["n_levels"] * g["n_features"]
        gen = torch.Generator().manual_seed(seed)
        init = (torch.rand(self.n_entries * g["n_features"], generator=gen) * 2 - 1) * 1e-4
        self.params = nn.Parameter(init.to(torch.float32), requires_grad=True)
        self.n_active_levels = g["n_levels"]
        # the backward's bucket workspace, kept between calls
        self.workspace = BucketWorkspace()

    @property
    def n_params(self):
        return self.params.numel()

    def level_offsets(self):
        import ctypes
        g = self.grid
        arr = (ctypes.c_uint32 * (g["n_levels"] + 1))()
        _lib.lib().sn_hashgrid_layout(g["n_levels"], g["log2_hashmap_size"],
                                      g["base_resolution"], g["per_level_scale"], arr)
        return list(arr)

    def forward(self, x):
        if not x.is_cuda:
            raise NotImplementedError("Encoding: only GPU tensors are supported")
        assert x.dim() == 2 and x.shape[1] == self.n_input_dims, \
            f"expected input of shape [N, {self.n_input_dims}], got {tuple(x.shape)}"
        x = x.to(torch.float32)
        n_active = max(0, min(int(self.n_active_levels), self.grid["n_levels"]))
        out = _HashEncodeFn.apply(x, self.params, self.grid, n_active, self.workspace)
        return out if self.dtype == torch.float32 else out.to(self.dtype)

    def extra_repr(self):
        return f"n_input_dims={self.n_input_dims}, n_output_dims={self.n_output_dims}, " \
               f"config={self.encoding_config}"
