"""NeuSRenderer drop-in (models/renderer.py) on the supernormal_amd kernels.

``NeuSRenderer(sdf_network, deviation_network, gradient_method)`` keeps the
reference's constructor, attributes (scene_aabb, contraction_type,
occupancy_grid, sampling_step_size, gradient_method), ``occ_eval_fn``,
``render(...)`` (same arguments and returned dict) and
``extract_geometry``.  ``render`` is the reference's structure -- torch
autograd over the drop-in nerfacc functions, the hash-grid kernel and torch
Linear layers -- i.e. the plain path; the fused path is
supernormal_amd.fused.FusedStep.
"""
from __future__ import annotations

import numpy as np
import torch

from .nerfacc import (ContractionType, OccupancyGrid, accumulate_along_rays,
                      accumulate_along_rays_patch_based, ray_marching, render_weight_from_alpha,
                      render_weight_from_alpha_patch_based)


def _neus_alpha(sdf_start, sdf_end, inv_s):
    """NeuS discrete opacity (renderer.py:108-118, 171-179)."""
    prev_cdf = torch.sigmoid(sdf_start * inv_s)
    next_cdf = torch.sigmoid(sdf_end * inv_s)
    return ((prev_cdf - next_cdf + 1e-5) / (prev_cdf + 1e-5)).clip(0.0, 1.0)


def _end_values(sdf_start, sdf_end_diff, diff_mask):
    """SDF at every sample's end: the next start, or the own end where they
    differ (renderer.py:164-169)."""
    end = torch.cat([sdf_start[1:], sdf_start[-1:]], 0)
    return end.masked_scatter(diff_mask.view(-1, *([1] * (end.dim() - 1))).expand_as(end),
                              sdf_end_diff) if diff_mask.any() else end


def _diff_mask(ri, t_starts, t_ends):
    S = ri.shape[0]
    dm = torch.ones(S, dtype=torch.bool, device=ri.device)
    if S > 1:
        dm[:-1] = ~((ri[1:] == ri[:-1]) & (t_ends[:-1, 0] == t_starts[1:, 0]))
    return dm


class NeuSRenderer:
    def __init__(self, sdf_network, deviation_network, gradient_method="dfd",
                 occ_resolution=128, device="cuda"):
        self.sdf_network = sdf_network
        self.deviation_network = deviation_network
        self.scene_aabb = torch.as_tensor([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0], dtype=torch.float32)
        self.contraction_type = ContractionType.AABB
        self.occupancy_grid = OccupancyGrid(roi_aabb=self.scene_aabb, resolution=occ_resolution,
                                            contraction_type=self.contraction_type).to(device)
        self.sampling_step_size = 0.01
        self.gradient_method = gradient_method

    def occ_eval_fn(self, x):
        sdf = self.sdf_network(x)[..., :1]
        return torch.sigmoid(-sdf * 80)

    def _inv_s(self):
        return self.deviation_network(torch.zeros([1, 3]))[:, :1].clip(1e-6, 1e6)

    def render(self, rays_o_patch_all, rays_d_patch_all, marching_plane_normal, near, far,
               V_inverse_patch_all, val_gradient_method="dfd", mode="train",
               max_samples_per_ray=0):
        Ph, Pw = rays_o_patch_all.shape[1], rays_o_patch_all.shape[2]
        B = rays_o_patch_all.shape[0]
        o_c = rays_o_patch_all[:, Ph // 2, Pw // 2]
        d_c = rays_d_patch_all[:, Ph // 2, Pw // 2]

        def alpha_fn(t_starts, t_ends, ray_indices):
            ri = ray_indices.long()
            xs = o_c[ri] + d_c[ri] * t_starts
            xe = o_c[ri] + d_c[ri] * t_ends
            dm = _diff_mask(ri, t_starts, t_ends)
            f = self.sdf_network(torch.cat([xs, xe[dm].reshape(-1, 3)], 0))
            fs = f[: xs.shape[0]]
            fe = _end_values(fs, f[xs.shape[0]:], dm)
            return _neus_alpha(fs, fe, self._inv_s().expand(fs.shape[0], 1)).reshape(-1, 1)

        with torch.no_grad():
            ri, t0, t1 = ray_marching(o_c, d_c, t_min=near, t_max=far, grid=self.occupancy_grid,
                                      render_step_size=self.sampling_step_size, stratified=True,
                                      cone_angle=0.0, early_stop_eps=1e-8, alpha_fn=alpha_fn,
                                      max_samples_per_ray=max_samples_per_ray)
            samples_per_ray = ri.shape[0] / B
            if ri.shape[0] == 0:
                return {"comp_normal": torch.zeros([B, Ph, Pw, 3], device=o_c.device),
                        "gradients": None, "samples_per_ray": 0.0}
            S = ri.shape[0]
            ri = ri.long()
            dcn = (d_c * marching_plane_normal).sum(-1, keepdim=True)[ri][:, None, None, :]
            dkn = (rays_d_patch_all * marching_plane_normal[:, None, None, :]).sum(-1, keepdim=True)[ri]
            ts = t0[:, None, None, :] * dcn / dkn
            te = t1[:, None, None, :] * dcn / dkn
            dm = _diff_mask(ri, t0, t1)
            ps = rays_o_patch_all[ri] + rays_d_patch_all[ri] * ts
            pe = rays_o_patch_all[ri] + rays_d_patch_all[ri] * te
            pts = torch.cat([ps, pe[dm]], 0)
        sdf_all = self.sdf_network(pts.reshape(-1, 3)).reshape(*pts.shape[:-1], 1)
        sdf_s = sdf_all[:S]
        sdf_e = _end_values(sdf_s, sdf_all[S:], dm)
        inv_s = self._inv_s()
        alpha = _neus_alpha(sdf_s, sdf_e, inv_s)
        weights = render_weight_from_alpha_patch_based(alpha.reshape(S, Ph * Pw, 1), ri)
        method = self.gradient_method if mode == "train" else val_gradient_method
        if method == "dfd":
            with torch.no_grad():
                dist_x = torch.norm(ps[:, :, 1:] - ps[:, :, :-1], dim=-1, keepdim=True)
                dist_y = torch.norm(ps[:, 1:] - ps[:, :-1], dim=-1, keepdim=True)
            df_dt = (sdf_e - sdf_s) / (te - ts)
            f = sdf_s
            dx = torch.cat([(f[:, :, 1:2] - f[:, :, 0:1]) / dist_x[:, :, 0:1],
                            (f[:, :, 2:] - f[:, :, :-2]) / (dist_x[:, :, :-1] + dist_x[:, :, 1:]),
                            (f[:, :, -1:] - f[:, :, -2:-1]) / dist_x[:, :, -1:]], 2)
            dy = torch.cat([(f[:, 1:2] - f[:, 0:1]) / dist_y[:, 0:1],
                            (f[:, 2:] - f[:, :-2]) / (dist_y[:, 1:] + dist_y[:, :-1]),
                            (f[:, -1:] - f[:, -2:-1]) / dist_y[:, -1:]], 1)
            proj = torch.cat([df_dt, dx, dy], -1)
            gradients = (V_inverse_patch_all[ri] @ proj[..., None])[..., 0]
        elif method == "ad":
            gradients = self.sdf_network.gradient(ps.reshape(-1, 3)).reshape(S, Ph, Pw, 3)
        elif method == "fd":
            eps = 1e-3
            offs = torch.eye(3, device=ps.device) * eps
            xs = torch.cat([ps + s * offs[i] for i in range(3) for s in (-1.0, 1.0)], 0)
            f6 = self.sdf_network(xs.reshape(-1, 3)).reshape(6, S, Ph, Pw)
            gradients = torch.stack([(f6[2 * i + 1] - f6[2 * i]) / (2 * eps) for i in range(3)], -1)
        else:
            raise ValueError(f"unknown gradient method {method!r}")
        ws = accumulate_along_rays_patch_based(weights, ri, n_patches=B).reshape(B, Ph, Pw, 1)
        cn = accumulate_along_rays_patch_based(weights, ri, values=gradients.reshape(S, Ph * Pw, 3),
                                               n_patches=B).reshape(B, Ph, Pw, 3)
        return {"s_val": 1 / inv_s, "weight_sum": ws, "gradients": gradients,
                "comp_normal": cn, "samples_per_ray": samples_per_ray}

    def extract_geometry(self, bound_min, bound_max, resolution, threshold=0.0):
        from .mesh import extract_geometry
        return extract_geometry(bound_min, bound_max, resolution, threshold,
                                lambda pts: -self.sdf_network.sdf(pts))
