"""nerfacc-compatible functions used by SuperNormal, on the gfx950 kernels.

Drop-in for the vendored NerfAcc 0.3.5 (third_parties/nerfacc-0.3.5/nerfacc-0.3.5/nerfacc,
"NA/" below): same names, argument order, defaults, return shapes and dtypes
(packed_info int32 [n,2], ray_indices int64, data float32) and the same
exception types (NotImplementedError for CPU tensors, AssertionError for bad
ranks, RuntimeError from every_n_step outside training).

Semantics that deliberately differ from the reference (each also listed in the
design notes):
  * render_visibility / ray_marching's visibility filter use the sequential
    per-ray product (NA csrc/render_transmittance.cu:85-112) instead of CUB's
    scan order, so the keep mask is reproducible bit for bit;
  * patch-weight backward is the division-free recurrence (finite at alpha=1);
  * the transmittance-only patch functions read alphas[ray] (the reference
    kernel reads alphas[j], csrc/render_transmittance.cu:173) and have a real
    backward (the reference's passes wrong arguments, vol_rendering.py:936);
  * OccupancyGrid's EMA is a deterministic scatter-max and the mean is an fp64
    sum in fixed order.
"""
from __future__ import annotations

from enum import Enum
from typing import Callable, List, Optional, Tuple, Union

import torch
import torch.nn as nn
from torch import Tensor

from . import _lib
from ._lib import call, ptr, require_cuda, stream_of, workspace


# --------------------------------------------------------------------------- #
# contraction (NA/nerfacc/contraction.py)
# --------------------------------------------------------------------------- #
class ContractionType(Enum):
    """Space contraction options (NA/nerfacc/contraction.py:12-62).  Only AABB is
    on SuperNormal's path (renderer.py:46); the others raise."""

    AABB = 0
    UN_BOUNDED_TANH = 1
    UN_BOUNDED_SPHERE = 2

    def to_cpp_version(self):
        return self.value


@torch.no_grad()
def contract_inv(x: Tensor, roi: Tensor, type: ContractionType = ContractionType.AABB) -> Tensor:
    """[0,1]^3 -> roi (NA/nerfacc/contraction.py:85-102, AABB only)."""
    if type != ContractionType.AABB:
        raise NotImplementedError("only ContractionType.AABB is supported")
    require_cuda(x, roi, name="contract_inv")
    roi = roi.to(x.device, torch.float32)
    return x * (roi[3:] - roi[:3]) + roi[:3]


# --------------------------------------------------------------------------- #
# packing (NA/nerfacc/pack.py:46-77)
# --------------------------------------------------------------------------- #
@torch.no_grad()
def pack_info(ray_indices: Tensor, n_rays: Optional[int] = None) -> Tensor:
    """Pack ray_indices to packed_info int32 [n_rays, 2] = (base, count)."""
    assert ray_indices.dim() == 1, "ray_indices must be a 1D tensor with shape (n_samples)."
    if not ray_indices.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    ri = ray_indices.contiguous().to(torch.int64)
    if n_rays is None:
        n_rays = int(ri.max()) + 1 if ri.numel() else 0
    packed = torch.empty((n_rays, 2), dtype=torch.int32, device=ri.device)
    ws = workspace(_lib.lib().sn_pack_info_workspace_bytes(n_rays), ri.device)
    call("sn_pack_info", ptr(ri), ri.numel(), n_rays, ptr(packed), None, ptr(ws), ws.numel(),
         stream_of(ri.device))
    return packed


def _packed_info_from(ray_indices, packed_info, n_rays):
    assert ray_indices is not None or packed_info is not None, \
        "Either ray_indices or packed_info should be provided."
    if packed_info is None:
        packed_info = pack_info(ray_indices, n_rays=n_rays)
    return packed_info.contiguous().to(torch.int32)


# --------------------------------------------------------------------------- #
# patch weights (NA/nerfacc/vol_rendering.py:533-622, 990-1033)
# --------------------------------------------------------------------------- #
def _as3(alphas):
    assert alphas.dim() == 3 and alphas.shape[2] == 1, \
        f"alphas must have shape (n_samples, patch_size, 1), got {tuple(alphas.shape)}"
    return alphas


class _PatchWeightFn(torch.autograd.Function):
    """w (and T) from alpha per patch ray; backward = division-free recurrence."""

    @staticmethod
    def forward(ctx, packed_info, alphas, want_T: bool):
        a = alphas.contiguous()
        w = torch.empty_like(a)
        T = torch.empty_like(a) if want_T else None
        # samples outside every patch stay defined (the reference leaves them
        # uninitialised; zero is the value a caller summing them expects)
        w.zero_()
        if T is not None:
            T.zero_()
        call("sn_patch_weight_fwd", ptr(packed_info), packed_info.shape[0], a.shape[1],
             a.shape[0], ptr(a), ptr(w), ptr(T), stream_of(a.device))
        ctx.save_for_backward(packed_info, a)
        ctx.want_T = want_T
        return (w, T) if want_T else w

    @staticmethod
    def backward(ctx, gw, gT=None):
        packed_info, a = ctx.saved_tensors
        gw = gw.contiguous() if gw is not None else None
        gT = gT.contiguous() if gT is not None else None
        ga = torch.zeros_like(a)
        call("sn_patch_weight_bwd", ptr(packed_info), packed_info.shape[0], a.shape[1],
             a.shape[0], ptr(a), ptr(gw), ptr(gT), ptr(ga), stream_of(a.device))
        return None, ga, None


def render_weight_from_alpha_patch_based(
        alphas: Tensor, ray_indices: Tensor, *, n_rays: Optional[int] = None) -> Tensor:
    """w_i = T_i alpha_i per ray of each patch; alphas (n_samples, patch_size, 1)."""
    _as3(alphas)
    if not alphas.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    packed_info = pack_info(ray_indices, n_rays=n_rays)
    return _PatchWeightFn.apply(packed_info, alphas, False)


def render_weight_and_transmittance_from_alpha_patch_based(
        alphas: Tensor, ray_indices: Tensor, *, n_rays: Optional[int] = None):
    """(weights, transmittance), both (n_samples, patch_size, 1), both differentiable."""
    _as3(alphas)
    if not alphas.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    packed_info = pack_info(ray_indices, n_rays=n_rays)
    return _PatchWeightFn.apply(packed_info, alphas, True)


def render_transmittance_from_alpha_patch_based(
        alphas: Tensor, ray_indices: Optional[Tensor] = None, *,
        packed_info: Optional[Tensor] = None, n_patches: Optional[int] = None) -> Tensor:
    """Exclusive transmittance per patch ray (corrected form of
    _RenderingTransmittanceFromAlphaPatchBasedNaive, vol_rendering.py:920-939)."""
    _as3(alphas)
    if not alphas.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    packed_info = _packed_info_from(ray_indices, packed_info, n_patches)
    return _PatchWeightFn.apply(packed_info, alphas, True)[1]


def render_weight_from_alpha(alphas: Tensor, *, packed_info: Optional[Tensor] = None,
                             ray_indices: Optional[Tensor] = None,
                             n_rays: Optional[int] = None) -> Tensor:
    """Non-patch weights (vol_rendering.py:624-677): the patch kernel with one ray."""
    assert alphas.dim() == 2 and alphas.shape[1] == 1
    if not alphas.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    packed_info = _packed_info_from(ray_indices, packed_info, n_rays)
    return _PatchWeightFn.apply(packed_info, alphas[:, :, None], False)[:, :, 0]


def render_transmittance_from_alpha(alphas: Tensor, *, packed_info: Optional[Tensor] = None,
                                    ray_indices: Optional[Tensor] = None,
                                    n_rays: Optional[int] = None) -> Tensor:
    """Non-patch exclusive transmittance (vol_rendering.py:406-460)."""
    assert alphas.dim() == 2 and alphas.shape[1] == 1
    if not alphas.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    packed_info = _packed_info_from(ray_indices, packed_info, n_rays)
    return _PatchWeightFn.apply(packed_info, alphas[:, :, None], True)[1][:, :, 0]


# --------------------------------------------------------------------------- #
# visibility (NA/nerfacc/vol_rendering.py:680-819)
# --------------------------------------------------------------------------- #
@torch.no_grad()
def render_visibility(alphas: Tensor, *, ray_indices: Optional[Tensor] = None,
                      packed_info: Optional[Tensor] = None, n_rays: Optional[int] = None,
                      early_stop_eps: float = 1e-4, alpha_thre: float = 0.0) -> Tensor:
    """Visible samples: T >= early_stop_eps and (alpha >= alpha_thre if alpha_thre > 0)."""
    if not alphas.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    packed_info = _packed_info_from(ray_indices, packed_info, n_rays)
    a = alphas.contiguous().view(-1)
    keep = torch.zeros(a.shape[0], dtype=torch.bool, device=a.device)
    if a.numel():
        call("sn_visibility", ptr(packed_info), packed_info.shape[0], a.shape[0], ptr(a),
             float(early_stop_eps), float(alpha_thre), ptr(keep), stream_of(a.device))
    return keep


@torch.no_grad()
def render_visibility_patch_based(alphas: Tensor, *, ray_indices: Optional[Tensor] = None,
                                  packed_info: Optional[Tensor] = None,
                                  n_patches: Optional[int] = None,
                                  early_stop_eps: float = 1e-4,
                                  alpha_thre: float = 0.0) -> Tensor:
    """A patch sample is visible if any of its rays has T >= eps (vol_rendering.py:815)."""
    _as3(alphas)
    if not alphas.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    T = render_transmittance_from_alpha_patch_based(alphas, ray_indices, packed_info=packed_info,
                                                    n_patches=n_patches)
    vis = torch.any(T >= early_stop_eps, dim=1, keepdim=True)
    if alpha_thre > 0:
        vis = vis & (alphas >= alpha_thre)
    return vis.squeeze()


# --------------------------------------------------------------------------- #
# accumulation (NA/nerfacc/vol_rendering.py:132-197, 269-335)
# --------------------------------------------------------------------------- #
class _ScatterAddRowsFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, src, ray_indices, n_out):
        src = src.contiguous()
        ri = ray_indices.contiguous()
        out = torch.zeros((n_out,) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
        row = 1
        for s in src.shape[1:]:
            row *= s
        call("sn_scatter_add_rows", ptr(ri), ri.numel(), row, ptr(src), ptr(out), n_out,
             stream_of(src.device))
        ctx.save_for_backward(ri)
        return out

    @staticmethod
    def backward(ctx, g):
        (ri,) = ctx.saved_tensors
        return g.contiguous()[ri], None, None


def accumulate_along_rays_patch_based(weights: Tensor, ray_indices: Tensor,
                                      values: Optional[Tensor] = None,
                                      n_patches: Optional[int] = None) -> Tensor:
    """out[p, k, :] = sum over the samples s of patch p of weights[s,k] * values[s,k,:]."""
    assert ray_indices.dim() == 1 and weights.dim() == 3
    if not weights.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    if values is not None:
        assert values.dim() == 3 and values.shape[0] == weights.shape[0], \
            "Invalid shapes: {} vs {}".format(values.shape, weights.shape)
        src = weights * values
    else:
        src = weights
    if ray_indices.numel() == 0:
        assert n_patches is not None
        return torch.zeros((n_patches, src.shape[1], src.shape[-1]), device=weights.device)
    if n_patches is None:
        n_patches = int(ray_indices.max()) + 1
    return _ScatterAddRowsFn.apply(src, ray_indices.to(torch.int64), n_patches)


def accumulate_along_rays(weights: Tensor, ray_indices: Tensor,
                          values: Optional[Tensor] = None,
                          n_rays: Optional[int] = None) -> Tensor:
    """Non-patch accumulation (vol_rendering.py:132-197)."""
    assert ray_indices.dim() == 1 and weights.dim() == 2
    if not weights.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    if values is not None:
        assert values.dim() == 2 and values.shape[0] == weights.shape[0], \
            "Invalid shapes: {} vs {}".format(values.shape, weights.shape)
        src = weights * values
    else:
        src = weights
    if ray_indices.numel() == 0:
        assert n_rays is not None
        return torch.zeros((n_rays, src.shape[-1]), device=weights.device)
    if n_rays is None:
        n_rays = int(ray_indices.max()) + 1
    return _ScatterAddRowsFn.apply(src, ray_indices.to(torch.int64), n_rays)


# --------------------------------------------------------------------------- #
# occupancy grid (NA/nerfacc/grid.py:113-295)
# --------------------------------------------------------------------------- #
class OccupancyGrid(nn.Module):
    """Occupancy grid with EMA update (NA/nerfacc/grid.py:113-277).

    Buffers and state-dict keys match the reference: _roi_aabb, _binary (bool
    [rx,ry,rz]), resolution (int32 [3]), occs (float [rx*ry*rz])."""

    NUM_DIM: int = 3

    def __init__(self, roi_aabb: Union[List[float], Tensor],
                 resolution: Union[int, List[int], Tensor] = 128,
                 contraction_type: ContractionType = ContractionType.AABB) -> None:
        super().__init__()
        if isinstance(resolution, int):
            resolution = [resolution] * self.NUM_DIM
        if isinstance(resolution, (list, tuple)):
            resolution = torch.tensor(resolution, dtype=torch.int32)
        assert isinstance(resolution, Tensor), f"Invalid type: {type(resolution)}"
        assert resolution.shape == (self.NUM_DIM,), f"Invalid shape: {resolution.shape}"
        if isinstance(roi_aabb, (list, tuple)):
            roi_aabb = torch.tensor(roi_aabb, dtype=torch.float32)
        assert isinstance(roi_aabb, Tensor), f"Invalid type: {type(roi_aabb)}"
        assert roi_aabb.shape == torch.Size([self.NUM_DIM * 2]), f"Invalid shape: {roi_aabb.shape}"
        if contraction_type != ContractionType.AABB:
            raise NotImplementedError("only ContractionType.AABB is supported")
        self.num_cells = int(resolution.prod().item())
        self.register_buffer("_roi_aabb", roi_aabb.to(torch.float32))
        self.register_buffer("_binary", torch.zeros(resolution.tolist(), dtype=torch.bool))
        self._contraction_type = contraction_type
        self.register_buffer("resolution", resolution.to(torch.int32))
        self.register_buffer("occs", torch.zeros(self.num_cells))
        self._res = tuple(int(r) for r in resolution.tolist())
        self._all = None
        # extension: a torch.Generator for the update's cell sampling and jitter
        # (None = torch's default generator, the reference's behaviour)
        self.generator = None
        # extension: cells evaluated by the updates so far (a host count)
        self.points_evaluated = 0

    @property
    def roi_aabb(self) -> Tensor:
        return self._roi_aabb

    @property
    def binary(self) -> Tensor:
        return self._binary

    @property
    def contraction_type(self) -> ContractionType:
        return self._contraction_type

    @property
    def device(self):
        return self.occs.device

    @torch.no_grad()
    def _get_all_cells(self) -> Tensor:
        if self._all is None or self._all.device != self.device:
            self._all = torch.arange(self.num_cells, device=self.device)
        return self._all

    @torch.no_grad()
    def _occupied_cells(self) -> Tensor:
        dev = self.device
        out = torch.empty(self.num_cells, dtype=torch.int64, device=dev)
        cnt = torch.empty(1, dtype=torch.int32, device=dev)
        ws = workspace(_lib.lib().sn_compact_workspace_bytes(self.num_cells), dev)
        call("sn_occ_nonzero", ptr(self._binary), self.num_cells, ptr(out), ptr(cnt), ptr(ws),
             ws.numel(), stream_of(dev))
        return out[: int(cnt.item())]

    @torch.no_grad()
    def _sample_uniform_and_occupied_cells(self, n: int) -> Tensor:
        uniform = torch.randint(self.num_cells, (n,), device=self.device, generator=self.generator)
        occupied = self._occupied_cells()
        if n < len(occupied):
            sel = torch.randint(len(occupied), (n,), device=self.device, generator=self.generator)
            occupied = occupied[sel]
        return torch.cat([uniform, occupied], dim=0)

    @torch.no_grad()
    def _update(self, step: int, occ_eval_fn: Callable, occ_thre: float = 0.01,
                ema_decay: float = 0.95, warmup_steps: int = 256) -> None:
        require_cuda(self.occs, name="OccupancyGrid")
        dev = self.device
        if step < warmup_steps:
            indices = self._get_all_cells()
        else:
            indices = self._sample_uniform_and_occupied_cells(self.num_cells // 4)
        n = indices.numel()
        jitter = torch.rand((n, 3), device=dev, dtype=torch.float32, generator=self.generator)
        self.update_cells(indices, jitter, occ_eval_fn, occ_thre, ema_decay)

    @torch.no_grad()
    def update_cells(self, indices: Tensor, jitter: Tensor, occ_eval_fn: Callable,
                     occ_thre: float = 0.01, ema_decay: float = 0.95) -> None:
        """The update of the given cells with the given jitter in [0,1)^3
        (grid.py:214-239): points in the cells, occ_eval_fn, EMA by scatter-max,
        binarisation by min(mean, occ_thre).  Extension: _update draws the cells
        and the jitter, then calls this."""
        dev = self.device
        indices = indices.to(device=dev, dtype=torch.int64).contiguous()
        jitter = jitter.to(device=dev, dtype=torch.float32).contiguous()
        n = indices.numel()
        self.points_evaluated += n
        x = torch.empty((n, 3), device=dev, dtype=torch.float32)
        rx, ry, rz = self._res
        s = stream_of(dev)
        call("sn_occ_cell_points", ptr(indices), n, ptr(jitter), rx, ry, rz, ptr(self._roi_aabb),
             ptr(x), s)
        occ = occ_eval_fn(x).squeeze(-1).contiguous().to(torch.float32)
        ws = workspace(_lib.lib().sn_occ_workspace_bytes(self.num_cells), dev)
        call("sn_occ_ema_max", ptr(self.occs), self.num_cells, ptr(indices), ptr(occ), n,
             float(ema_decay), ptr(ws), ws.numel(), s)
        call("sn_occ_binarize", ptr(self.occs), self.num_cells, float(occ_thre),
             ptr(self._binary), None, ptr(ws), ws.numel(), s)

    @torch.no_grad()
    def every_n_step(self, step: int, occ_eval_fn: Callable, occ_thre: float = 1e-2,
                     ema_decay: float = 0.95, warmup_steps: int = 256, n: int = 16) -> None:
        if not self.training:
            raise RuntimeError(
                "You should only call this function only during training. "
                "Please call _update() directly if you want to update the "
                "field during inference.")
        if step % n == 0 and self.training:
            self._update(step=step, occ_eval_fn=occ_eval_fn, occ_thre=occ_thre,
                         ema_decay=ema_decay, warmup_steps=warmup_steps)


# --------------------------------------------------------------------------- #
# ray marching (NA/nerfacc/ray_marching.py:13-222)
# --------------------------------------------------------------------------- #
@torch.no_grad()
def ray_aabb_intersect(rays_o: Tensor, rays_d: Tensor, aabb: Tensor) -> Tuple[Tensor, Tensor]:
    """Slab test (NA/nerfacc/intersection.py); t_min=t_max=1e10 for misses."""
    inv = 1.0 / rays_d
    t0 = (aabb[:3] - rays_o) * inv
    t1 = (aabb[3:] - rays_o) * inv
    tmin = torch.minimum(t0, t1).amax(-1).clamp(min=0.0)
    tmax = torch.maximum(t0, t1).amin(-1)
    miss = tmax <= tmin
    tmin = torch.where(miss, torch.full_like(tmin, 1e10), tmin)
    tmax = torch.where(miss, torch.full_like(tmax, 1e10), tmax)
    return tmin, tmax


def march_raw(rays_o: Tensor, rays_d: Tensor, t_min: Tensor, t_max: Tensor, roi: Tensor,
              binary: Tensor, render_step_size: float, cone_angle: float = 0.0,
              max_samples_per_ray: int = 0):
    """_C.ray_marching (csrc/ray_marching.cu:194-289): (packed_info, ray_indices,
    t_starts [n,1], t_ends [n,1]).  Reads the sample count back to size the
    outputs, as the reference does (its .item() at ray_marching.cu:261)."""
    dev = rays_o.device
    n_rays = rays_o.shape[0]
    rx, ry, rz = (int(s) for s in binary.shape)
    s = stream_of(dev)
    lib = _lib.lib()
    counts = torch.empty(n_rays, dtype=torch.int32, device=dev)
    packed = torch.empty((n_rays, 2), dtype=torch.int32, device=dev)
    total = torch.zeros(1, dtype=torch.int32, device=dev)
    args = (ptr(rays_o), ptr(rays_d), ptr(t_min), ptr(t_max), n_rays, ptr(roi), ptr(binary),
            rx, ry, rz, float(render_step_size), None, float(cone_angle),
            int(max_samples_per_ray))
    if n_rays:
        call("sn_ray_march_count", *args, ptr(counts), s)
        ws = workspace(lib.sn_scan_workspace_bytes(n_rays), dev)
        call("sn_counts_to_packed_info", ptr(counts), n_rays, ptr(packed), ptr(total), ptr(ws),
             ws.numel(), s)
    n = int(total.item())
    ray_indices = torch.empty(n, dtype=torch.int64, device=dev)
    t_starts = torch.empty((n, 1), dtype=torch.float32, device=dev)
    t_ends = torch.empty((n, 1), dtype=torch.float32, device=dev)
    if n:
        call("sn_ray_march_fill", *args, ptr(packed), ptr(ray_indices), ptr(t_starts),
             ptr(t_ends), s)
    return packed, ray_indices, t_starts, t_ends


@torch.no_grad()
def ray_marching(rays_o: Tensor, rays_d: Tensor, t_min: Optional[Tensor] = None,
                 t_max: Optional[Tensor] = None, scene_aabb: Optional[Tensor] = None,
                 grid: Optional[OccupancyGrid] = None, sigma_fn: Optional[Callable] = None,
                 alpha_fn: Optional[Callable] = None, early_stop_eps: float = 1e-4,
                 alpha_thre: float = 0.0, near_plane: Optional[float] = None,
                 far_plane: Optional[float] = None, render_step_size: float = 1e-3,
                 stratified: bool = False, cone_angle: float = 0.0,
                 max_samples_per_ray: int = 0) -> Tuple[Tensor, Tensor, Tensor]:
    """Ray marching with occupancy-grid skipping and the visibility filter.

    Returns (ray_indices int64 [n], t_starts [n,1], t_ends [n,1]).
    `max_samples_per_ray` (extension, default 0 = unbounded like the reference)
    truncates each ray before the visibility filter."""
    if not rays_o.is_cuda:
        raise NotImplementedError("Only support cuda inputs.")
    if alpha_fn is not None and sigma_fn is not None:
        raise ValueError("Only one of `alpha_fn` and `sigma_fn` should be provided.")
    rays_o = rays_o.contiguous().to(torch.float32)
    rays_d = rays_d.contiguous().to(torch.float32)
    if t_min is None or t_max is None:
        if scene_aabb is not None:
            t_min, t_max = ray_aabb_intersect(rays_o, rays_d, scene_aabb.to(rays_o))
        else:
            t_min = torch.zeros_like(rays_o[..., 0])
            t_max = torch.ones_like(rays_o[..., 0]) * 1e10
    if near_plane is not None:
        t_min = torch.clamp(t_min, min=near_plane)
    if far_plane is not None:
        t_max = torch.clamp(t_max, max=far_plane)
    if stratified:
        t_min = t_min + torch.rand_like(t_min) * render_step_size
    if grid is not None:
        roi = grid.roi_aabb.contiguous().to(torch.float32)
        binary = grid.binary.contiguous()
        if grid.contraction_type != ContractionType.AABB:
            raise NotImplementedError("only ContractionType.AABB is supported")
    else:
        roi = torch.tensor([-1e10, -1e10, -1e10, 1e10, 1e10, 1e10], dtype=torch.float32,
                           device=rays_o.device)
        binary = torch.ones([1, 1, 1], dtype=torch.bool, device=rays_o.device)
    packed, ray_indices, t_starts, t_ends = march_raw(
        rays_o, rays_d, t_min.contiguous().to(torch.float32), t_max.contiguous().to(torch.float32),
        roi, binary, render_step_size, cone_angle, max_samples_per_ray)
    if sigma_fn is not None or alpha_fn is not None:
        if sigma_fn is not None:
            sigmas = sigma_fn(t_starts, t_ends, ray_indices)
            assert sigmas.shape == t_starts.shape, \
                "sigmas must have shape of (N, 1)! Got {}".format(sigmas.shape)
            alphas = 1.0 - torch.exp(-sigmas * (t_ends - t_starts))
        else:
            alphas = alpha_fn(t_starts, t_ends, ray_indices)
            assert alphas.shape == t_starts.shape, \
                "alphas must have shape of (N, 1)! Got {}".format(alphas.shape)
        masks = render_visibility(alphas, ray_indices=ray_indices, packed_info=packed,
                                  early_stop_eps=early_stop_eps, alpha_thre=alpha_thre,
                                  n_rays=rays_o.shape[0])
        ray_indices, t_starts, t_ends = ray_indices[masks], t_starts[masks], t_ends[masks]
    return ray_indices, t_starts, t_ends
