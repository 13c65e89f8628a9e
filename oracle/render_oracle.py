"""Patch rendering and losses in torch (fp64 by default).  Test infrastructure only.

Restates, from the visibility-filtered patch-centre samples onwards,
models/renderer.py:142-276 (patch expansion, NeuS alpha, patch weights,
'dfd' / 'ad' gradients, accumulation) and exp_runner.py:169-203 (normal,
eikonal and mask losses).  Differentiable (torch autograd) in the SDF values
at sample starts / ends, the analytic gradients and inv_s, which gives the
reference's backward.

The one intended difference from renderer.py:152-154: a sample's end is
looked up in the next sample's start only when the next sample belongs to
the same patch ray (the reference compares t values across rays, which can
pair two different rays when their t happen to coincide).
"""
from __future__ import annotations

import torch


def diff_mask(ri, t0, t1):
    """True where the sample's end is not the next sample's start (renderer.py:150-154)."""
    S = ri.shape[0]
    d = torch.ones(S, dtype=torch.bool)
    if S > 1:
        same = ri[1:] == ri[:-1]
        d[:-1] = ~(same & (t1[:-1, 0] == t0[1:, 0]))
    return d


def expand_t(t_c, ri, rays_d_all, plane_n):
    """t on every patch ray: t_c * (d_c . n) / (d_k . n) (renderer.py:146-149). -> [S, Ph, Pw, 1]"""
    Ph, Pw = rays_d_all.shape[1:3]
    dc = rays_d_all[:, Ph // 2, Pw // 2]
    dcn = (dc * plane_n).sum(-1, keepdim=True)[ri][:, None, None, :]
    dkn = (rays_d_all * plane_n[:, None, None, :]).sum(-1, keepdim=True)[ri]
    return t_c[:, None, None, :] * dcn / dkn


def positions(rays_o_all, rays_d_all, ri, t_all):
    return rays_o_all[ri] + rays_d_all[ri] * t_all


def patch_weights(alpha, packed):
    """w = alpha * exclusive cumprod(1 - alpha) along each patch ray. alpha [S, P]."""
    S, P = alpha.shape
    B = packed.shape[0]
    maxc = int(packed[:, 1].max()) if B else 0
    pad = torch.zeros(B, max(maxc, 1), P, dtype=alpha.dtype)
    rows = []
    for b in range(B):
        base, cnt = int(packed[b, 0]), int(packed[b, 1])
        rows.append(torch.arange(base, base + cnt))
    idx_b = torch.cat([torch.full((int(packed[b, 1]),), b) for b in range(B)]) if S else \
        torch.zeros(0, dtype=torch.long)
    idx_j = torch.cat([torch.arange(int(packed[b, 1])) for b in range(B)]) if S else \
        torch.zeros(0, dtype=torch.long)
    src = torch.cat(rows) if S else torch.zeros(0, dtype=torch.long)
    pad = pad.index_put((idx_b, idx_j), alpha[src])
    one_minus = 1.0 - pad
    T = torch.cumprod(torch.cat([torch.ones(B, 1, P, dtype=alpha.dtype), one_minus[:, :-1]], 1), 1)
    w_pad = pad * T
    w = torch.zeros_like(alpha).index_put((src,), w_pad[idx_b, idx_j])
    return w


def neus_alpha(sdf_s, sdf_end, inv_s):
    """NeuS discrete opacity (renderer.py:171-179):
    clip((sigmoid(s f_i) - sigmoid(s f_i+1) + 1e-5) / (sigmoid(s f_i) + 1e-5), 0, 1)."""
    prev_cdf = torch.sigmoid(sdf_s * inv_s)
    next_cdf = torch.sigmoid(sdf_end * inv_s)
    return ((prev_cdf - next_cdf + 1e-5) / (prev_cdf + 1e-5)).clip(0.0, 1.0)


def dfd_gradients(rays_o_all, rays_d_all, ri, ts, te, sdf_s, sdf_end, v_inv):
    """'dfd' SDF gradients (renderer.py:187-223): forward difference along each
    patch ray, central differences across the patch on the marching plane
    (one-sided at the border, distances between neighbouring start points),
    then V^-1 [df/dt, df/dx, df/dy].  ts/te [S, Ph, Pw, 1]; -> [S, Ph, Pw, 3]."""
    pos = positions(rays_o_all, rays_d_all, ri, ts).detach()
    dist_x = torch.norm(pos[:, :, 1:] - pos[:, :, :-1], dim=-1, keepdim=True)
    dist_y = torch.norm(pos[:, 1:] - pos[:, :-1], dim=-1, keepdim=True)
    df_dt = (sdf_end - sdf_s) / (te - ts)
    f = sdf_s
    mid_x = (f[:, :, 2:] - f[:, :, :-2]) / (dist_x[:, :, :-1] + dist_x[:, :, 1:])
    mid_y = (f[:, 2:] - f[:, :-2]) / (dist_y[:, 1:] + dist_y[:, :-1])
    left = (f[:, :, 1:2] - f[:, :, 0:1]) / dist_x[:, :, 0:1]
    right = (f[:, :, -1:] - f[:, :, -2:-1]) / dist_x[:, :, -1:]
    top = (f[:, 1:2] - f[:, 0:1]) / dist_y[:, 0:1]
    bottom = (f[:, -1:] - f[:, -2:-1]) / dist_y[:, -1:]
    df_dx = torch.cat([left, mid_x, right], 2)
    df_dy = torch.cat([top, mid_y, bottom], 1)
    proj = torch.cat([df_dt, df_dx, df_dy], -1)
    return (v_inv[ri] @ proj[..., None])[..., 0]


def patch_losses(cn, ws, grads, n_true, mask, normal_w=1.0, eik_w=1.0, mask_w=1.0,
                 loss_type="l2"):
    """exp_runner.py:169-203: masked normal loss (sum of squared / absolute
    errors over the mask's sum + 1e-5), eikonal mean((|grad| - 1)^2) over every
    sample, BCE of the clipped weight sum against the mask.
    Returns (loss, parts)."""
    mask_sum = mask.sum() + 1e-5
    err = (cn - n_true) * mask
    if loss_type == "l1":
        l_n = err.abs().sum() / mask_sum
    else:
        l_n = (err ** 2).sum() / mask_sum
    gn = torch.linalg.norm(grads, ord=2, dim=-1)
    l_e = ((gn - 1.0) ** 2).mean()
    l_m = torch.nn.functional.binary_cross_entropy(ws.clip(1e-5, 1.0 - 1e-5), mask)
    loss = normal_w * l_n + mask_w * l_m + eik_w * l_e
    return loss, dict(normal=normal_w * l_n, mask=mask_w * l_m, eikonal=eik_w * l_e)


def render_losses(rays_o_all, rays_d_all, plane_n, v_inv, n_true, mask, packed, ri, t0, t1,
                  sdf_s, sdf_e, inv_s, grad_s=None, normal_w=1.0, eik_w=1.0, mask_w=1.0,
                  loss_type="l2", out=None):
    """All inputs torch tensors.  sdf_s [S, Ph, Pw, 1], sdf_e [Nd, Ph, Pw, 1] (ends of the
    samples where diff_mask is true, in order), grad_s [S, Ph, Pw, 3] for 'ad' (None = dfd),
    mask [B, Ph, Pw, 1] already binarised.  Returns (loss, parts, comp_normal, weight_sum);
    `out` (a dict, optional) receives alpha, weights and gradients."""
    B, Ph, Pw, _ = rays_o_all.shape
    P = Ph * Pw
    S = ri.shape[0]
    dm = diff_mask(ri, t0, t1)
    ts = expand_t(t0, ri, rays_d_all, plane_n)
    te = expand_t(t1, ri, rays_d_all, plane_n)
    # ends: shifted starts, replaced by their own values where dm
    sdf_end = torch.cat([sdf_s[1:], sdf_s[-1:]], 0).clone()
    sdf_end[dm] = sdf_e
    alpha = neus_alpha(sdf_s, sdf_end, inv_s)
    w = patch_weights(alpha.reshape(S, P), packed).reshape(S, Ph, Pw, 1)
    if grad_s is None:
        grads = dfd_gradients(rays_o_all, rays_d_all, ri, ts, te, sdf_s, sdf_end, v_inv)
    else:
        grads = grad_s
    ws = torch.zeros(B, Ph, Pw, 1, dtype=w.dtype).index_add(0, ri, w)
    cn = torch.zeros(B, Ph, Pw, 3, dtype=w.dtype).index_add(0, ri, w * grads)
    loss, parts = patch_losses(cn, ws, grads, n_true, mask, normal_w, eik_w, mask_w, loss_type)
    if out is not None:
        out.update(alpha=alpha, weights=w, gradients=grads)
    return loss, parts, cn, ws
