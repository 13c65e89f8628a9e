"""One SuperNormal training iteration on the gfx950 kernels.

The iteration of exp_runner.py:147-210 (occupancy update, patch sampling,
rendering, losses, backward, Adam, learning-rate schedule) with the
reference's per-op torch graph replaced by:
  * sn_ray_march_count / _fill (occupancy-grid marching, bit-exact),
  * the visibility pre-pass: fused SDF forward at the centre samples, NeuS
    alpha, sn_visibility (sequential transmittance), compaction,
  * the patch expansion onto the marching planes,
  * the fused SDF (hash encode + MFMA MLP, optionally the analytic gradient),
  * sn_render_step: alpha, weights, gradients, accumulation and the three
    losses with their whole backward in one kernel,
  * the SDF backward kernels (MLP weight gradients, hash-table atomics; the
    analytic-gradient double backward), and
  * sn_adam_step over all parameters held in one flat buffer.

Sizes that depend on the data (samples per iteration) are read back three
times per iteration (marched, visible, non-contiguous ends), as the
reference does (ray_marching.cu:261, ray_marching.py:216-220, renderer.py:157).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import torch

from . import _lib
from ._lib import call, ptr, stream_of, workspace
from .nerfacc import march_raw, pack_info, render_visibility


class FlatParams:
    """All trainable tensors as views of one flat fp32 buffer (and their
    gradients as views of one flat gradient buffer), so Adam is one launch."""

    def __init__(self, params, device):
        self.params = [p for p in params]
        n = sum(p.numel() for p in self.params)
        pad = lambda k: (k + 3) // 4 * 4
        self.offsets = []
        off = 0
        for p in self.params:
            self.offsets.append(off)
            off += pad(p.numel())
        self.n = off
        self.data = torch.zeros(off, dtype=torch.float32, device=device)
        self.grad = torch.zeros(off, dtype=torch.float32, device=device)
        self.exp_avg = torch.zeros_like(self.data)
        self.exp_avg_sq = torch.zeros_like(self.data)
        for p, o in zip(self.params, self.offsets):
            self.data[o:o + p.numel()].copy_(p.detach().reshape(-1))
            p.data = self.data[o:o + p.numel()].view(p.shape)
        self.step_count = 0
        self.attach_grads()

    def attach_grads(self):
        for p, o in zip(self.params, self.offsets):
            p.grad = self.grad[o:o + p.numel()].view(p.shape)

    def zero_grad(self):
        self.grad.zero_()
        self.attach_grads()

    def collect(self):
        """Make sure every gradient landed in the flat buffer."""
        for p, o in zip(self.params, self.offsets):
            g = p.grad
            if g is None:
                continue
            if g.data_ptr() != self.grad[o:o + p.numel()].data_ptr():
                self.grad[o:o + p.numel()].copy_(g.reshape(-1))

    def adam(self, lr, betas=(0.9, 0.999), eps=1e-8):
        self.step_count += 1
        call("sn_adam_step", ptr(self.data), ptr(self.grad), ptr(self.exp_avg),
             ptr(self.exp_avg_sq), self.n, float(lr), None, float(betas[0]), float(betas[1]),
             float(eps), self.step_count, None, stream_of(self.data.device))


class _RenderStepFn(torch.autograd.Function):
    """Loss of the rendered patches; the kernel computes the gradients in the
    same pass, backward only scales them."""

    @staticmethod
    def forward(ctx, sdf_s, sdf_e, grad_s, inv_s, st):
        dev = sdf_s.device
        S, P, Nd = st["S"], st["P"], st["Nd"]
        g_ss = torch.empty(S * P, device=dev)
        g_se = torch.empty(max(Nd, 1) * P, device=dev)
        g_gs = torch.empty(S * P, 3, device=dev) if st["ad"] else None
        losses = torch.empty(4, device=dev)
        waves = _lib.lib().sn_render_step_waves(st["B"], P)
        partials = torch.empty(4 * waves, device=dev)
        se = sdf_e if Nd else torch.zeros(P, device=dev)
        call("sn_render_step", st["B"], st["Ph"], st["Pw"], ptr(st["rays_o"]), ptr(st["rays_d"]),
             ptr(st["plane_n"]), ptr(st["v_inv"]), ptr(st["n_true"]), ptr(st["mask"]),
             ptr(st["packed"]), ptr(st["t0"]), ptr(st["t1"]), ptr(st["diff_idx"]),
             ptr(sdf_s), ptr(se), ptr(grad_s if st["ad"] else None), ptr(inv_s),
             ptr(st["mask_sum"]), ptr(st["n_s"]), st["normal_w"], st["eik_w"], st["mask_w"],
             1 if st["l1"] else 0, ptr(g_ss), ptr(g_se), ptr(g_gs), ptr(losses), ptr(partials),
             ptr(st.get("weight_sum")), ptr(st.get("comp_normal")), stream_of(dev))
        st["losses"] = losses
        ctx.save_for_backward(g_ss, g_se, g_gs if g_gs is not None else g_ss, losses)
        ctx.ad, ctx.Nd, ctx.P = st["ad"], Nd, P
        return losses[0] + losses[1] + losses[2]

    @staticmethod
    def backward(ctx, g):
        g_ss, g_se, g_gs, losses = ctx.saved_tensors
        g_e = (g_se[: ctx.Nd * ctx.P] * g) if ctx.Nd else None
        return g_ss * g, g_e, (g_gs * g) if ctx.ad else None, losses[3:4] * g, None


@dataclass
class StepResult:
    loss: Optional[torch.Tensor]
    parts: Optional[torch.Tensor]
    n_marched: int
    n_samples: int
    n_eval: int


class FusedStep:
    """The per-iteration work of the fused path for an SDFNetwork +
    SingleVarianceNetwork pair and an OccupancyGrid."""

    def __init__(self, sdf_network, deviation_network, occupancy_grid, gradient_method="dfd",
                 max_samples_per_ray=0, normal_weight=1.0, eikonal_weight=1.0, mask_weight=1.0,
                 loss_type="l2", early_stop_eps=1e-8):
        if gradient_method not in ("dfd", "ad"):
            raise ValueError("fused path supports gradient_method 'dfd' or 'ad'")
        self.net = sdf_network
        self.dev_net = deviation_network
        self.grid = occupancy_grid
        self.method = gradient_method
        self.cap = int(max_samples_per_ray)
        self.normal_w, self.eik_w, self.mask_w = float(normal_weight), float(eikonal_weight), \
            float(mask_weight)
        self.l1 = loss_type == "l1"
        self.early_stop_eps = early_stop_eps
        # the visibility decision of the last call's marched samples (bool [S0])
        self.last_keep = None

    def inv_s(self):
        return torch.exp(self.dev_net.variance * 10.0).clip(1e-6, 1e6).reshape(1)

    @torch.no_grad()
    def occ_eval_fn(self, x):
        """renderer.py:56-60 on the fused SDF."""
        return torch.sigmoid(-self.net.fused_sdf(x) * 80.0)[:, None]

    def forward_backward(self, rays_o, rays_d, plane_n, v_inv, n_true, mask, near, far,
                         step_size, stratified=True) -> StepResult:
        """Render the patches, compute the losses and accumulate gradients into
        the parameters' .grad.  Returns the loss (device tensor) or None when no
        sample survives (the reference would raise KeyError, exp_runner.py:182)."""
        net = self.net
        dev = rays_o.device
        B, Ph, Pw, _ = rays_o.shape
        P = Ph * Pw
        kc = (Ph // 2, Pw // 2)
        o_c = rays_o[:, kc[0], kc[1]].contiguous()
        d_c = rays_d[:, kc[0], kc[1]].contiguous()
        t_min = near
        if stratified:
            t_min = t_min + torch.rand_like(t_min) * step_size
        roi = self.grid.roi_aabb.contiguous()
        packed0, ri0, t00, t10 = march_raw(o_c, d_c, t_min.contiguous(), far.contiguous(), roi,
                                           self.grid.binary, step_size, 0.0, self.cap)
        S0 = ri0.numel()
        if S0 == 0:
            return StepResult(None, None, 0, 0, 0)
        inv_s = self.inv_s()
        # ---- visibility pre-pass on the centre rays (renderer.py:80-122, 124-135)
        with torch.no_grad():
            dm0 = _diff_mask(ri0, t00, t10)
            xs = o_c[ri0] + d_c[ri0] * t00
            xe = (o_c[ri0] + d_c[ri0] * t10)[dm0]
            f_all = net.fused_sdf(torch.cat([xs, xe], 0))
            fs = f_all[:S0]
            fe = fs.roll(-1).clone()
            fe[dm0] = f_all[S0:]
            c = torch.sigmoid(fs * inv_s)
            e = torch.sigmoid(fe * inv_s)
            alpha = ((c - e + 1e-5) / (c + 1e-5)).clip(0.0, 1.0)
            keep = render_visibility(alpha[:, None], packed_info=packed0,
                                     early_stop_eps=self.early_stop_eps, alpha_thre=0.0)
            self.last_keep = keep
            ri, t0, t1 = ri0[keep], t00[keep], t10[keep]
        S = ri.numel()
        if S == 0:
            return StepResult(None, None, S0, 0, 0)
        # ---- patch expansion onto the marching planes (renderer.py:142-159)
        with torch.no_grad():
            packed = pack_info(ri, n_rays=B)
            dm = _diff_mask(ri, t0, t1)
            Nd = int(dm.sum())
            diff_idx = torch.full((S,), -1, dtype=torch.int32, device=dev)
            diff_idx[dm] = torch.arange(Nd, dtype=torch.int32, device=dev)
            dcn = (d_c * plane_n).sum(-1, keepdim=True)              # [B,1]
            dkn = (rays_d * plane_n[:, None, None, :]).sum(-1)        # [B,Ph,Pw]
            ts = t0[:, :, None] * dcn[ri][:, :, None] / dkn[ri]       # [S,Ph,Pw]
            te = t1[:, :, None] * dcn[ri][:, :, None] / dkn[ri]
            ps = rays_o[ri] + rays_d[ri] * ts[..., None]
            pe = (rays_o[ri] + rays_d[ri] * te[..., None])[dm]
            pts = torch.cat([ps.reshape(-1, 3), pe.reshape(-1, 3)], 0)
        # ---- SDF on every patch ray (renderer.py:161) and the gradients
        if self.method == "ad":
            sdf_s, grad_s = net.fused_sdf_and_gradient(ps.reshape(-1, 3))
            sdf_e = net.fused_sdf(pe.reshape(-1, 3)) if Nd else torch.zeros(0, device=dev)
        else:
            f = net.fused_sdf(pts)
            sdf_s, sdf_e = f[: S * P], f[S * P:]
            grad_s = None
        mask_b = (mask > 0.5).float() if self.mask_w > 0 else torch.ones_like(mask)
        st = dict(B=B, Ph=Ph, Pw=Pw, P=P, S=S, Nd=Nd, ad=self.method == "ad",
                  rays_o=rays_o.contiguous(), rays_d=rays_d.contiguous(),
                  plane_n=plane_n.contiguous(), v_inv=v_inv.contiguous(),
                  n_true=n_true.contiguous(), mask=mask_b.reshape(B, P).contiguous(),
                  packed=packed, t0=t0[:, 0].contiguous(), t1=t1[:, 0].contiguous(),
                  diff_idx=diff_idx, mask_sum=(mask_b.sum() + 1e-5).reshape(1),
                  n_s=torch.tensor([S], dtype=torch.int32, device=dev),
                  normal_w=self.normal_w, eik_w=self.eik_w, mask_w=self.mask_w, l1=self.l1)
        loss = _RenderStepFn.apply(sdf_s.contiguous(), sdf_e.contiguous(),
                                   grad_s.contiguous() if grad_s is not None else None, inv_s, st)
        loss.backward()
        return StepResult(loss.detach(), st["losses"], S0, S, pts.shape[0])


def _diff_mask(ri, t0, t1):
    """A sample's end is its own point unless the next sample (same ray)
    starts exactly there (renderer.py:150-154)."""
    S = ri.shape[0]
    dm = torch.ones(S, dtype=torch.bool, device=ri.device)
    if S > 1:
        dm[:-1] = ~((ri[1:] == ri[:-1]) & (t1[:-1, 0] == t0[1:, 0]))
    return dm
