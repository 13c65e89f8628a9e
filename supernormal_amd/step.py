"""The training iteration with no host round trip, optionally as one HIP graph.

FusedStep (fused.py) reads three data-dependent sizes back to the host per
iteration, as the reference does (ray_marching.cu:261, ray_marching.py:216-220,
renderer.py:157).  Here every such size stays on the device: buffers are
allocated once at capacity (rays x max_samples_per_ray centre samples, and
x patch size for the patch rays), each kernel launch is sized by the capacity
and reads the valid count from device memory.  The iteration is then a fixed
sequence of launches on one stream:

  patch sampling (dataset_loader.py:223-277, device RNG)
  -> stratified jitter, sn_ray_march_padded / scan / sn_march_gather (one
     march per ray instead of count + fill, renderer.py:66-79)
  -> sn_centre_points, SDF forward, sn_alpha_visibility, sn_compact_samples_dev
     (with the kept samples' packed_info; the visibility pre-pass,
     renderer.py:80-135)
  -> sn_patch_points (patch expansion, renderer.py:142-159)
  -> hash-grid + MLP forward on every patch sample (renderer.py:161)
  -> sn_render_step (weights, normals, losses and their backward)
  -> MLP backward, hash-grid backward into the flat gradient buffer, weight-
     norm and variance chain (autograd over the small parameter tensors)
  -> sn_adam_step with the learning rate and step count read from the device

and is captured once into four HIP graphs (torch.cuda.CUDAGraph): sampling +
march, the rest of the forward and backward, the hash-grid backward, Adam.
The march reads no parameter, so the next iteration's sampling + march can be
replayed on a side stream as soon as this iteration's second graph is done,
overlapping the hash-grid backward and Adam (which fill the device while the
march, one lane per centre ray, occupies only a few compute units).  The
occupancy-grid update (every n iterations) runs between replays (no march
runs ahead across it); the active hash levels are a device scalar, so a
bandwidth increase changes no launch.  While a march is pending, the buffers it writes
(inputs, counts, packed0, n0, ri0/t00/t10) already hold the next iteration's.

Both gradient methods run here: 'dfd' (finite differences of the SDF at the
patch points) and 'ad' (the analytic x-gradient and its double backward,
forward-over-reverse, on the fp16/bf16 MLP kernels when the precision is
fp16, on the fp32 ones otherwise).  max_samples_per_ray must be set (it
bounds the capacity).
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import call, ptr, stream_of
from .tcnn import bucket_workspace_bytes, hashgrid_bwd


def flat_offset(flat, t):
    """Offset of parameter tensor t in the flat parameter buffer."""
    k = [i for i, q in enumerate(flat.params) if q is t]
    if not k:
        raise ValueError("a network tensor is not among the optimised parameters")
    return flat.offsets[k[0]]


def mlp_pack_table(net, flat):
    """Host table (int64 [layers, 6]) of the SDF MLP's tensors in the flat
    parameter buffer for sn_mlp_pack_fwd/bwd: offsets of v, g (-1 without
    weight norm) and b, the layer's offset in the packed buffer, out, in."""
    import numpy as np
    rows, off = [], 0
    for l in range(net.num_layers - 1):
        lin = getattr(net, "lin" + str(l))
        if net.weight_norm:
            v, g = lin.weight_v, lin.weight_g
        else:
            v, g = lin.weight, None
        out, inn = v.shape
        rows.append([flat_offset(flat, v), -1 if g is None else flat_offset(flat, g),
                     flat_offset(flat, lin.bias), off, out, inn])
        off += out * inn + out
    return np.ascontiguousarray(np.array(rows, dtype=np.int64))


class SyncFreeStep:
    def __init__(self, sdf_network, deviation_network, occupancy_grid, flat, dataset,
                 batch_size, patch_size=3, max_samples_per_ray=128, normal_weight=1.0,
                 eikonal_weight=1.0, mask_weight=1.0, loss_type="l2", early_stop_eps=1e-8,
                 precision="fp16", device="cuda", seed=None, half_table=None, method="dfd"):
        if max_samples_per_ray <= 0:
            raise ValueError("SyncFreeStep needs max_samples_per_ray > 0 (it bounds the buffers)")
        if precision not in ("fp32", "fp16"):
            raise ValueError("precision must be 'fp32' or 'fp16'")
        self.net, self.dev_net, self.grid, self.flat = sdf_network, deviation_network, \
            occupancy_grid, flat
        self.dataset = dataset
        self.dev = torch.device(device)
        self.B, self.Ph, self.Pw = int(batch_size), int(patch_size), int(patch_size)
        self.P = self.Ph * self.Pw
        self.msr = int(max_samples_per_ray)
        self.normal_w, self.eik_w, self.mask_w = float(normal_weight), float(eikonal_weight), \
            float(mask_weight)
        self.l1 = loss_type == "l1"
        self.eps = float(early_stop_eps)
        self.precision = precision
        if method not in ("dfd", "ad"):
            raise ValueError("gradient method must be 'dfd' or 'ad'")
        # 'ad': the rendered normal is the analytic x-gradient of the SDF
        # (renderer.py:225-226), its loss differentiated twice; with precision
        # fp16 the patch points' forward, gradient and forward-over-reverse
        # backward run on the fp16/bf16 matrix cores, else on the fp32 kernels
        self.ad = method == "ad"
        # forward gathers from an fp16 copy of the hash table, refreshed by each
        # Adam step (fp32 master table and gradients); default with precision fp16
        self.half_table = (precision == "fp16") if half_table is None else bool(half_table)
        cfg = sdf_network.sdf_config()
        self.cfg = cfg
        self.E = cfg.E
        self._alloc()
        self.adam_t.fill_(int(flat.step_count))
        # host copies of what adam_t / levels_dev hold, so that unchanged values
        # are not written again (a launch each, ahead of every iteration)
        self._adam_t_host = int(flat.step_count)
        self._levels_host = None
        self.graph = None
        self.graph_given = None  # the march alone, for iterations on given inputs
        # the next iteration's march when it ran ahead: (event, step size)
        self._pending = None
        self._rng_pending = None
        self.side = None
        # optional timing of the hash-grid backward (events between replays)
        self.time_hash_bwd = False
        self.hash_times = []
        self.hash_counts = torch.zeros(100000, dtype=torch.int32, device=self.dev)
        # a generator of its own (seeded) makes the sampled patches independent
        # of other work on the device, e.g. other jobs' steps in between
        self.gen = None
        if seed is not None:
            self.gen = torch.Generator(device=self.dev)
            self.gen.manual_seed(int(seed))

    # ------------------------------------------------------------------ buffers
    def _alloc(self):
        dev, B, P, E = self.dev, self.B, self.P, self.E
        lib = _lib.lib()
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        c0 = B * self.msr                       # centre samples
        c1 = 2 * c0 * P                         # patch points (starts + non-contiguous ends)
        self.c0, self.c1 = c0, c1
        Ph, Pw = self.Ph, self.Pw
        # inputs of one iteration (written by the sampling step)
        self.rays_o = torch.zeros(B, Ph, Pw, 3, **f32)
        self.rays_d = torch.zeros(B, Ph, Pw, 3, **f32)
        self.plane_n = torch.zeros(B, 3, **f32)
        self.v_inv = torch.zeros(B, Ph, Pw, 3, 3, **f32)
        self.n_true = torch.zeros(B, Ph, Pw, 3, **f32)
        self.mask = torch.zeros(B, P, **f32)
        self.o_c = torch.zeros(B, 3, **f32)
        self.d_c = torch.zeros(B, 3, **f32)
        self.t_min = torch.zeros(B, **f32)
        self.t_max = torch.zeros(B, **f32)
        self.mask_sum = torch.zeros(1, **f32)
        # schedule scalars read by the kernels
        self.step_size = torch.zeros(1, **f32)
        self.lr = torch.zeros(1, **f32)
        self.adam_t = torch.zeros(1, **i32)
        self.levels_dev = torch.zeros(1, **i32)  # active hash levels (bandwidth)
        self.inv_s = torch.zeros(1, **f32)
        # device counts of one iteration, views of one buffer: marched centre
        # samples, visible ones, centre points and patch points evaluated; then
        # two running counters: rays that filled their sample capacity and
        # iterations skipped for want of samples
        self.stat = torch.zeros(8, **i32)
        self.n0, self.ns, self.ne0, self.ne = (self.stat[k:k + 1] for k in range(4))
        self.n_capped, self.n_skipped = self.stat[4:5], self.stat[5:6]
        # whether this iteration's optimiser step runs (samples, finite losses
        # and gradients), and a running count of iterations skipped for a
        # non-finite loss or gradient
        self.run, self.n_nonfinite = self.stat[6:7], self.stat[7:8]
        # the four per-iteration counts summed over iterations (int64)
        self.stat_acc = torch.zeros(4, dtype=torch.int64, device=dev)
        # marching
        self.counts = torch.zeros(B, **i32)
        self.packed0 = torch.zeros(B, 2, **i32)
        self.ri0 = torch.zeros(c0, dtype=torch.int64, device=dev)
        self.t00 = torch.zeros(c0, **f32)
        self.t10 = torch.zeros(c0, **f32)
        self.t0_pad = torch.zeros(c0, **f32)   # ray i's samples at i * max_samples_per_ray
        self.t1_pad = torch.zeros(c0, **f32)
        # the dataset's tensors as the sampling kernel reads them (contiguous,
        # fixed dtypes: torch.inverse returns a column-major K^-1)
        ds = self.dataset
        self._ds_tensors = None
        if all(hasattr(ds, k) for k in ("normals", "V_inverse_all", "masks",
                                        "intrinsics_all_inv", "pose_all", "_train_idx")):
            self._ds_tensors = dict(
                normals=ds.normals.float().contiguous(), v_inv=ds.V_inverse_all.float().contiguous(),
                masks=ds.masks.float().contiguous(),
                k_inv=ds.intrinsics_all_inv.float().contiguous(),
                pose=ds.pose_all.float().contiguous(),
                train_idx=ds._train_idx.to(torch.int64).contiguous())
        # visibility pre-pass
        self.didx0 = torch.zeros(c0, **i32)
        self.nd0 = torch.zeros(1, **i32)
        self.pts0 = torch.zeros(2 * c0, 3, **f32)
        self.enc0 = torch.zeros(2 * c0, E, **f32)
        self.sdf0 = torch.zeros(2 * c0, **f32)
        self.keep = torch.zeros(c0, dtype=torch.uint8, device=dev)
        # visible samples, patch expansion
        self.ri = torch.zeros(c0, dtype=torch.int64, device=dev)
        self.t0 = torch.zeros(c0, **f32)
        self.t1 = torch.zeros(c0, **f32)
        self.packed = torch.zeros(B, 2, **i32)
        self.didx = torch.zeros(c0, **i32)
        self.nd = torch.zeros(1, **i32)
        self.pts = torch.zeros(self.c1, 3, **f32)
        self.enc = torch.zeros(self.c1, E, **f32)
        self.sdf = torch.zeros(self.c1, **f32)
        self.g_sdf = torch.zeros(self.c1, **f32)
        self.g_enc = torch.zeros(self.c1, E, **f32)
        if self.ad:
            # d sdf / d h0 rows, the analytic gradient of every patch point, its
            # cotangent from the render step, and the tangents of the backward
            self.dh0 = torch.zeros(self.c1, 3 + E, **f32)
            self.grad_pts = torch.zeros(self.c1, 3, **f32)
            self.g_grad = torch.zeros(self.c1, 3, **f32)
            self.enc_dot = torch.zeros(self.c1, E, **f32)
            self.g_enc_dot = torch.zeros(self.c1, E, **f32)
        # render step
        self.losses = torch.zeros(4, **f32)
        waves = lib.sn_render_step_waves(B, P)
        self.partials = torch.zeros(4 * waves, **f32)
        # workspaces
        nb = lambda n: torch.zeros(max(int(n), 1), dtype=torch.uint8, device=dev)
        self.ws_scan = nb(lib.sn_scan_workspace_bytes(B))
        self.ws_step = nb(lib.sn_step_workspace_bytes(c0))
        self.ws_compact = nb(lib.sn_compact_workspace_bytes(c0))
        hb = bucket_workspace_bytes(self.cfg.hash_args(self.cfg.grid["n_levels"]), self.c1)
        self.ws_hash = nb(hb) if hb else None
        cfg = self.cfg
        if self.ad and self.precision != "fp16":
            wsb = lib.sn_mlp_bwd_workspace_bytes(self.c1, E, cfg.d_hidden, cfg.n_layers, 1)
        elif self.precision == "fp16":
            wsb = lib.sn_mlp16_bwd_workspace_bytes(self.c1, E, cfg.d_hidden, cfg.n_layers)
        else:
            wsb = lib.sn_mlp_bwd_workspace_bytes(self.c1, E, cfg.d_hidden, cfg.n_layers, 0)
        self.ws_mlp = nb(wsb)
        n_mlp = lib.sn_mlp_n_params(3 + E, cfg.d_hidden, cfg.n_layers)
        self.g_mlp = torch.zeros(n_mlp, **f32)
        self.mlp_packed = torch.zeros(n_mlp, **f32)
        self._pack_table(n_mlp)
        self.table16 = None
        if self.half_table:
            table = self.net.encoding.params
            self.table16 = torch.empty(table.numel(), dtype=torch.float16, device=dev)
            k = [i for i, q in enumerate(self.flat.params) if q is table]
            if not k:
                raise ValueError("the hash table must be one of the optimised parameters")
            self.table_off = self.flat.offsets[k[0]]
            self.refresh_copies()

    def _pack_table(self, n_mlp):
        self.pack_layers = mlp_pack_table(self.net, self.flat)
        if int((self.pack_layers[:, 4] * (self.pack_layers[:, 5] + 1)).sum()) != n_mlp:
            raise ValueError("packed MLP size differs from the MLP kernels' layout")
        self.n_pack = len(self.pack_layers)
        ov = flat_offset(self.flat, self.dev_net.variance)
        self.var_data = self.flat.data[ov:ov + 1]
        self.var_grad = self.flat.grad[ov:ov + 1]

    def refresh_copies(self):
        """Re-derive the fp16 table copy from the parameters (after they were
        changed outside the step, e.g. a checkpoint load)."""
        if self.table16 is not None:
            t = self.net.encoding.params
            call("sn_to_half", ptr(t), ptr(self.table16), t.numel(), stream_of(self.dev))

    # ------------------------------------------------------------- one iteration
    def _sample(self):
        """Patch sampling into the static input buffers (device RNG only): the
        random numbers of gen_random_patches and the stratified start, drawn in
        the same order, then one sampling kernel (sn_patch_sample)."""
        ds = self.dataset
        t = self._ds_tensors
        if t is None:
            o, d, pn, vi, nt, m = ds.gen_random_patches(self.B, patch_H=self.Ph,
                                                         patch_W=self.Pw, generator=self.gen)
            self._load(o, d, pn, vi, nt, m, stratified=True)
            return
        B, Ph, Pw, dev, g = self.B, self.Ph, self.Pw, self.dev, self.gen
        cx = torch.randint(Pw // 2, ds.W - 1 - Pw // 2, (B,), device=dev, generator=g)
        cy = torch.randint(Ph // 2, ds.H - 1 - Ph // 2, (B,), device=dev, generator=g)
        sel = torch.randint(0, len(ds.train_images), (B,), device=dev, generator=g)
        u = torch.rand((B,), generator=g, device=dev, dtype=torch.float32)
        call("sn_patch_sample", ptr(cx), ptr(cy), ptr(sel), ptr(u), ptr(self.step_size),
             ptr(t["train_idx"]), ptr(t["normals"]), ptr(t["v_inv"]), ptr(t["masks"]),
             ptr(t["k_inv"]), int(t["k_inv"].shape[-1]), ptr(t["pose"]),
             int(t["normals"].shape[0]), int(t["train_idx"].numel()), ds.H, ds.W, B, Ph, Pw,
             1 if self.mask_w > 0 else 0,
             ptr(self.rays_o), ptr(self.rays_d), ptr(self.plane_n), ptr(self.v_inv),
             ptr(self.n_true), ptr(self.mask), ptr(self.mask_sum), ptr(self.o_c),
             ptr(self.d_c), ptr(self.t_min), ptr(self.t_max), stream_of(dev))

    def _load(self, o, d, pn, vi, nt, m, stratified, t_min=None, t_max=None):
        """Given patches into the static input buffers; t_min / t_max of the
        centre rays when given (e.g. an already jittered start), else near / far
        on the unit sphere (+ the stratified jitter)."""
        ds = self.dataset
        self.rays_o.copy_(o)
        self.rays_d.copy_(d)
        self.plane_n.copy_(pn)
        self.v_inv.copy_(vi)
        self.n_true.copy_(nt)
        mb = (m > 0.5).float() if self.mask_w > 0 else torch.ones_like(m)
        self.mask.copy_(mb.reshape(self.B, self.P))
        self.mask_sum.copy_(self.mask.sum().reshape(1) + 1e-5)
        kc = (self.Ph // 2, self.Pw // 2)
        self.o_c.copy_(self.rays_o[:, kc[0], kc[1]])
        self.d_c.copy_(self.rays_d[:, kc[0], kc[1]])
        if t_min is not None:
            self.t_min.copy_(t_min)
            self.t_max.copy_(t_max)
            return
        near, far = ds.near_far_from_sphere(self.o_c, self.d_c)
        # stratified start (renderer.py:69-71)
        if stratified:
            u = torch.rand(near.shape, generator=self.gen, device=near.device, dtype=near.dtype)
            near = near + u * self.step_size
        self.t_min.copy_(near)
        self.t_max.copy_(far)

    def _load_given(self, inputs):
        """Given inputs (rays_o, rays_d, plane_n, v_inv, n_true, mask[, t_min,
        t_max]) into the static buffers; without t_min / t_max the centre rays
        start at near on the unit sphere, unjittered."""
        if len(inputs) not in (6, 8):
            raise RuntimeError(f"given inputs: 6 or 8 tensors, got {len(inputs)}")
        t_min, t_max = (inputs[6], inputs[7]) if len(inputs) == 8 else (None, None)
        self._load(*inputs[:6], stratified=False, t_min=t_min, t_max=t_max)

    def _hash_args(self):
        """Hash-grid arguments with every level as the bound; the active count
        is read from self.levels_dev (set by step()), so a bandwidth increase
        changes no launch and needs no re-capture."""
        return self.cfg.hash_args(self.cfg.grid["n_levels"])

    def _sdf_fwd(self, pts, enc, sdf, cap, n_dev, mlp, s):
        cfg = self.cfg
        if self.table16 is not None:
            call("sn_hashgrid_fwd16_dev", ptr(pts), cap, ptr(n_dev), ptr(self.table16),
                 *self._hash_args(), ptr(self.levels_dev), ptr(enc), s)
        else:
            call("sn_hashgrid_fwd_dev", ptr(pts), cap, ptr(n_dev),
                 ptr(self.net.encoding.params), *self._hash_args(), ptr(self.levels_dev),
                 ptr(enc), s)
        if self.precision == "fp16":
            call("sn_mlp16_fwd", ptr(pts), ptr(enc), cfg.E, cap, ptr(n_dev), ptr(mlp),
                 cfg.d_hidden, cfg.n_layers, cfg.beta, ptr(sdf), s)
        else:
            call("sn_mlp_fwd", ptr(pts), ptr(enc), cfg.E, cap, ptr(n_dev), ptr(mlp),
                 cfg.d_hidden, cfg.n_layers, cfg.beta, ptr(sdf), None, s)

    def _iteration(self, inputs=None):
        self._part_march(inputs)
        self._part_rest()
        self._part_hash()
        self._part_adam()

    def _part_a(self, inputs=None):
        self._part_march(inputs)
        self._part_rest()

    def _part_march(self, inputs=None):
        """Patch sampling (or the given inputs) and the march of the centre rays.
        Reads only the inputs, the step size and the occupancy grid, so it can run
        ahead of the previous iteration's hash-grid backward and Adam step."""
        s = stream_of(self.dev)
        B = self.B
        grid = self.grid
        rx, ry, rz = (int(v) for v in grid.binary.shape)
        if inputs is None:
            self._sample()
        elif not isinstance(inputs, str):  # "keep": the buffers hold the inputs already
            self._load_given(inputs)
        # one march into per-ray slots, then packed (= count + fill of the reference);
        # (the byte grid: brick masks, sn_ray_march_padded_bricks, measured no faster
        # with the wavefront-per-ray march)
        call("sn_ray_march_padded", ptr(self.o_c), ptr(self.d_c), ptr(self.t_min),
             ptr(self.t_max), B, ptr(grid.roi_aabb), ptr(grid.binary), rx, ry, rz, 1.0,
             ptr(self.step_size), 0.0, self.msr, ptr(self.counts), ptr(self.t0_pad),
             ptr(self.t1_pad), s)
        call("sn_counts_to_packed_info", ptr(self.counts), B, ptr(self.packed0), ptr(self.n0),
             ptr(self.ws_scan), self.ws_scan.numel(), s)
        call("sn_march_gather", ptr(self.packed0), B, self.msr, ptr(self.t0_pad),
             ptr(self.t1_pad), ptr(self.ri0), ptr(self.t00), ptr(self.t10),
             ptr(self.n_capped), s)

    def _part_rest(self):
        """From the marched samples up to the hash-grid backward: the visibility
        pre-pass, SDF, render + losses, MLP backward and the weight-norm / variance
        chain."""
        s = stream_of(self.dev)
        B, c0, c1 = self.B, self.c0, self.c1
        cfg = self.cfg
        f = self.flat
        f.grad.zero_()
        # weight-normed MLP weights packed for the kernels, and inv_s
        call("sn_mlp_pack_fwd", ptr(f.data), self.n_pack, self.pack_layers.ctypes.data,
             ptr(self.mlp_packed), ptr(self.var_data), ptr(self.inv_s), s)
        mlp_d = self.mlp_packed
        # visibility pre-pass on the centre rays
        call("sn_centre_points", ptr(self.ri0), ptr(self.t00), ptr(self.t10), c0, ptr(self.n0),
             ptr(self.o_c), ptr(self.d_c), 3, ptr(self.didx0), ptr(self.nd0), ptr(self.pts0),
             ptr(self.ne0), ptr(self.ws_step), self.ws_step.numel(), s)
        self._sdf_fwd(self.pts0, self.enc0, self.sdf0, 2 * c0, self.ne0, mlp_d, s)
        call("sn_alpha_visibility", ptr(self.packed0), B, ptr(self.sdf0), ptr(self.didx0), c0,
             ptr(self.n0), ptr(self.inv_s), self.eps, ptr(self.keep), s)
        # compaction also derives the packed_info of the kept samples
        call("sn_compact_samples_dev", ptr(self.keep), c0, ptr(self.n0), ptr(self.ri0),
             ptr(self.t00), ptr(self.t10), ptr(self.ri), ptr(self.t0), ptr(self.t1),
             ptr(self.ns), ptr(self.packed0), B, ptr(self.packed), ptr(self.ws_compact),
             self.ws_compact.numel(), s)
        # patch expansion and the SDF on every patch sample
        call("sn_patch_points", ptr(self.ri), ptr(self.t0), ptr(self.t1), c0, ptr(self.ns),
             ptr(self.rays_o), ptr(self.rays_d), ptr(self.plane_n), self.Ph, self.Pw,
             ptr(self.didx), ptr(self.nd), ptr(self.pts), ptr(self.ne), ptr(self.ws_step),
             self.ws_step.numel(), s)
        if self.ad:
            self._sdf_and_gradient(c1, mlp_d, s)
        else:
            self._sdf_fwd(self.pts, self.enc, self.sdf, c1, self.ne, mlp_d, s)
        ad = self.ad
        call("sn_render_step", B, self.Ph, self.Pw, ptr(self.rays_o), ptr(self.rays_d),
             ptr(self.plane_n), ptr(self.v_inv), ptr(self.n_true), ptr(self.mask),
             ptr(self.packed), ptr(self.t0), ptr(self.t1), ptr(self.didx), ptr(self.sdf), None,
             ptr(self.grad_pts) if ad else None, ptr(self.inv_s), ptr(self.mask_sum),
             ptr(self.ns), self.normal_w, self.eik_w, self.mask_w, 1 if self.l1 else 0,
             ptr(self.g_sdf), None, ptr(self.g_grad) if ad else None, ptr(self.losses),
             ptr(self.partials), None, None, s)
        # backward: the MLP here; the hash grid in its own part (_part_hash)
        if ad:
            self._ad_backward(c1, mlp_d, s)
        elif self.precision == "fp16":
            call("sn_mlp16_bwd", ptr(self.pts), ptr(self.enc), cfg.E, c1, ptr(self.ne),
                 ptr(mlp_d), cfg.d_hidden, cfg.n_layers, cfg.beta, ptr(self.g_sdf),
                 ptr(self.g_enc), None, ptr(self.g_mlp), 0, ptr(self.ws_mlp),
                 self.ws_mlp.numel(), s)
        else:
            call("sn_mlp_bwd", ptr(self.pts), ptr(self.enc), cfg.E, c1, ptr(self.ne),
                 ptr(mlp_d), cfg.d_hidden, cfg.n_layers, cfg.beta, ptr(self.g_sdf),
                 ptr(self.g_enc), None, ptr(self.g_mlp), 0, ptr(self.ws_mlp),
                 self.ws_mlp.numel(), s)
        # weight norm (packed weights -> weight_v, weight_g, bias) and inv_s -> variance
        call("sn_mlp_pack_bwd", ptr(f.data), self.n_pack, self.pack_layers.ctypes.data,
             ptr(self.g_mlp), ptr(f.grad), ptr(self.var_data), ptr(self.losses[3:4]),
             ptr(self.var_grad), s)
        # this iteration's counts (the march of the next one may run ahead once
        # this part is done, and it rewrites n0)
        self.stat_acc.add_(self.stat[:4])

    def _sdf_and_gradient(self, c1, mlp_d, s):
        """'ad': sdf and the analytic x-gradient of every patch point
        (fields.py:107-119): MLP forward with d sdf / d h0 (on the fp16 matrix
        cores with precision fp16), then the table's part of the gradient by a
        gather."""
        cfg = self.cfg
        params = self.net.encoding.params
        if self.table16 is not None:
            call("sn_hashgrid_fwd16_dev", ptr(self.pts), c1, ptr(self.ne), ptr(self.table16),
                 *self._hash_args(), ptr(self.levels_dev), ptr(self.enc), s)
        else:
            call("sn_hashgrid_fwd_dev", ptr(self.pts), c1, ptr(self.ne), ptr(params),
                 *self._hash_args(), ptr(self.levels_dev), ptr(self.enc), s)
        fwd = "sn_mlp16_fwd_grad" if self.precision == "fp16" else "sn_mlp_fwd"
        call(fwd, ptr(self.pts), ptr(self.enc), cfg.E, c1, ptr(self.ne), ptr(mlp_d),
             cfg.d_hidden, cfg.n_layers, cfg.beta, ptr(self.sdf), ptr(self.dh0), s)
        call("sn_hashgrid_dx_dev", ptr(self.pts), c1, ptr(self.ne), ptr(params),
             ptr(self.table16), *self._hash_args(), ptr(self.levels_dev), ptr(self.dh0),
             3 + cfg.E, ptr(self.grad_pts), s)

    def _ad_backward(self, c1, mlp_d, s):
        """'ad': the double backward through the analytic gradient (the
        reference's autograd.grad(create_graph=True)): the gradient's cotangent
        v (zero on the samples' end points, which carry no normal), the
        encoding's tangent along v, the forward-over-reverse MLP backward."""
        cfg = self.cfg
        params = self.net.encoding.params
        call("sn_fill_span_dev", ptr(self.g_grad), c1 * 3, ptr(self.ns), self.P * 3,
             ptr(self.ne), 3, 0.0, s)
        call("sn_hashgrid_tangent_dev", ptr(self.pts), c1, ptr(self.ne), ptr(params),
             ptr(self.table16), *self._hash_args(), ptr(self.levels_dev), ptr(self.g_grad),
             ptr(self.enc_dot), s)
        bwd = "sn_mlp16_bwd_ad_enc" if self.precision == "fp16" else "sn_mlp_bwd_ad_enc"
        call(bwd, ptr(self.pts), ptr(self.enc), ptr(self.g_grad),
             ptr(self.enc_dot), cfg.E, c1, ptr(self.ne), ptr(mlp_d), cfg.d_hidden,
             cfg.n_layers, cfg.beta, ptr(self.g_sdf), ptr(self.g_enc), ptr(self.g_enc_dot),
             ptr(self.g_mlp), 0, ptr(self.ws_mlp), self.ws_mlp.numel(), s)

    def _part_hash(self):
        """Hash-grid backward into the flat gradient buffer (bucket records,
        then per-bucket sums); with 'ad' the records also carry the table's
        part of the double backward (one pass, no fp32 atomics)."""
        params = self.net.encoding.params
        s = stream_of(self.dev)
        if self.ad:
            ws = self.ws_hash
            call("sn_hashgrid_bwd_ad_ws_dev", ptr(self.pts), self.c1, ptr(self.ne), ptr(params),
                 *self._hash_args(), ptr(self.levels_dev), ptr(self.g_enc), ptr(self.g_grad),
                 ptr(self.g_enc_dot), ptr(params.grad), ptr(ws), 0 if ws is None else ws.numel(),
                 s)
            return
        hashgrid_bwd(self.pts, self.c1, self.ne, params, self._hash_args(), self.g_enc,
                     params.grad, None, s, ws=self.ws_hash, n_active_dev=self.levels_dev)

    def _part_guard(self):
        """Whether this iteration's optimiser step runs, and the device step
        count.  An iteration with no visible sample is skipped (parameters,
        moments and step count unchanged; the reference raises KeyError there,
        exp_runner.py:182) and counted in n_skipped, and so is one whose losses
        or gradients are not finite (also counted in n_nonfinite)."""
        s = stream_of(self.dev)
        f = self.flat
        call("sn_step_guard", ptr(self.ns), ptr(self.losses), 3, ptr(f.grad), f.n,
             ptr(self.run), ptr(self.n_nonfinite), s)
        call("sn_step_advance", ptr(self.adam_t), ptr(self.run), ptr(self.n_skipped), s)

    def _part_adam(self):
        """Adam with the learning rate and step count on the device."""
        s = stream_of(self.dev)
        self._part_guard()
        f = self.flat
        t16 = self.table16
        call("sn_adam_step_copy16", ptr(f.data), ptr(f.grad), ptr(f.exp_avg),
             ptr(f.exp_avg_sq), f.n, 0.0, ptr(self.lr), 0.9, 0.999, 1e-8, 0,
             ptr(self.adam_t), ptr(t16), self.table_off if t16 is not None else 0,
             t16.numel() if t16 is not None else 0, ptr(self.run), s)

    # ------------------------------------------------------------------- driver
    def step(self, lr, step_size, use_graph=True, inputs=None, next_step_size=None):
        """One training iteration; lr / step_size for this iteration.  Returns
        the device tensor of losses (normal, mask, eikonal, dL/dinv_s).
        `inputs` = (rays_o, rays_d, plane_n, V_inv, normals, mask[, t_min,
        t_max]) of given patches instead of sampling (without t_min / t_max:
        unjittered starts on the unit sphere), or "keep" to rerun on the inputs
        of the previous iteration (eager only).

        Graph mode with `next_step_size`: the next iteration's sampling and march
        are launched on a side stream once this iteration no longer reads their
        buffers, so they overlap this iteration's hash-grid backward and Adam
        step.  The caller must not change the occupancy grid before the next
        step (the march has read it) and passes the same step size to it."""
        if inputs is not None and not isinstance(inputs, str) and use_graph:
            return self._step_given(lr, step_size, inputs)
        self.lr.fill_(float(lr))
        levels = int(self.net.active_levels())
        if levels != self._levels_host:
            self.levels_dev.fill_(levels)
            self._levels_host = levels
        # the optimizer's step count is the host's (other drivers may have stepped
        # it); the iteration's Adam step increments the device copy itself
        t = int(self.flat.step_count)
        if t != self._adam_t_host:
            self.adam_t.fill_(t)
        self._adam_t_host = None  # unknown until the iteration is submitted
        timing = self.time_hash_bwd
        if not use_graph or inputs is not None:
            if self._pending is not None:
                raise RuntimeError("an iteration's march is pending; run it in graph mode")
            self.step_size.fill_(float(step_size))
            self._part_a(inputs)
            e = self._hash_events() if timing else None
            self._part_hash()
            if e is not None:
                e[1].record()
            self._part_adam()
            self.flat.step_count += 1
            self._adam_t_host = t + 1
            return self.losses
        if self._pending is None:
            self.step_size.fill_(float(step_size))
        if self.graph is None:
            self._capture()
        cur = torch.cuda.current_stream(self.dev)
        if self._pending is None:
            self.graph[0].replay()
        else:
            ev, pending_step = self._pending
            if float(step_size) != pending_step:
                raise ValueError(f"step size {step_size} differs from the one the march "
                                 f"ran ahead with ({pending_step})")
            cur.wait_event(ev)
            self._pending = None
        self.graph[1].replay()
        done = None
        if next_step_size is not None:
            # the last reader of the march's buffers was the graph just replayed
            done = torch.cuda.Event()
            done.record(cur)
        # the hash-grid backward alone in its graph, so that events between the
        # replays time it
        e = self._hash_events() if timing else None
        self.graph[2].replay()
        if e is not None:
            e[1].record()
        self.graph[3].replay()
        # submitted after this iteration's last graphs, so that their launch does
        # not wait for the host to submit the (many-node) sampling graph
        if done is not None:
            self._march_ahead(done, float(next_step_size))
        self.flat.step_count += 1
        self._adam_t_host = t + 1
        return self.losses

    # ------------------------------------------- driven by another graph owner
    def prepare(self, lr, step_size):
        """Writes this iteration's device scalars (learning rate, step size,
        active levels, the optimiser's step count if the host changed it) for
        an iteration that another driver replays (the serving loop's joint
        graph, which holds capture_parts() and an Adam launch); returns the
        host step count to hand to finished()."""
        if self._pending is not None:
            raise RuntimeError("a march is pending; drain() it first")
        self.lr.fill_(float(lr))
        levels = int(self.net.active_levels())
        if levels != self._levels_host:
            self.levels_dev.fill_(levels)
            self._levels_host = levels
        t = int(self.flat.step_count)
        if t != self._adam_t_host:
            self.adam_t.fill_(t)
        self._adam_t_host = None
        self.step_size.fill_(float(step_size))
        return t

    def finished(self, t):
        """After the replay of an iteration prepared with prepare()."""
        self.flat.step_count = t + 1
        self._adam_t_host = t + 1

    def capture_parts(self):
        """The iteration up to the optimiser step (sampling, march, rendering,
        backward, step guard), enqueued on the current stream, for a graph that
        another driver captures; the Adam step is the caller's
        (adam_segment())."""
        self._part_march()
        self._part_rest()
        self._part_hash()
        self._part_guard()

    def adam_segment(self):
        """This optimiser's row of sn_adam_multi's segment table, without the
        block range: (params, grads, exp_avg, exp_avg_sq, n, lr, step, run,
        fp16 copy, copy offset, copy length)."""
        f, t16 = self.flat, self.table16
        return (ptr(f.data), ptr(f.grad), ptr(f.exp_avg), ptr(f.exp_avg_sq), f.n, ptr(self.lr),
                ptr(self.adam_t), ptr(self.run), ptr(t16),
                self.table_off if t16 is not None else 0, t16.numel() if t16 is not None else 0)

    def _step_given(self, lr, step_size, inputs):
        """Graph mode on given inputs: they are copied into the static buffers,
        then a graph holding the march alone (no sampling) and the iteration's
        other three graphs are replayed."""
        if self._pending is not None:
            raise RuntimeError("an iteration's march is pending; run it in graph mode")
        self.lr.fill_(float(lr))
        levels = int(self.net.active_levels())
        if levels != self._levels_host:
            self.levels_dev.fill_(levels)
            self._levels_host = levels
        t = int(self.flat.step_count)
        if t != self._adam_t_host:
            self.adam_t.fill_(t)
        self._adam_t_host = None
        if self.graph is None:
            self._capture()
        if self.graph_given is None:
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize(self.dev)
            with torch.cuda.graph(g):
                self._part_march("keep")
            self.graph_given = g
        self.step_size.fill_(float(step_size))
        self._load_given(inputs)
        self.graph_given.replay()
        self.graph[1].replay()
        e = self._hash_events() if self.time_hash_bwd else None
        self.graph[2].replay()
        if e is not None:
            e[1].record()
        self.graph[3].replay()
        self.flat.step_count += 1
        self._adam_t_host = t + 1
        return self.losses

    def _march_ahead(self, done, step_size):
        """Next iteration's sampling + march on the side stream, once `done`
        (recorded after this iteration's last reader of those buffers) has
        completed."""
        if self.side is None:
            self.side = torch.cuda.Stream(device=self.dev)
        self.side.wait_event(done)
        # the sampler's generator as of this sampling, for a checkpoint taken
        # while the march is pending (it resumes by sampling these patches)
        self._rng_pending = self._rng_state()
        with torch.cuda.stream(self.side):
            self.step_size.fill_(step_size)
            self.graph[0].replay()
            ev = torch.cuda.Event()
            ev.record(self.side)
        self._pending = (ev, step_size)

    def _rng_state(self):
        return self.gen.get_state() if self.gen is not None else \
            torch.cuda.get_rng_state(self.dev)

    def sampler_rng_state(self):
        """State of the patch sampler's generator (its own, or torch's default
        one) as of the next iteration's sampling, also when that iteration's
        march already ran ahead (then the state before it sampled)."""
        if self._pending is not None:
            return self._rng_pending
        return self._rng_state()

    def set_sampler_rng_state(self, state):
        if self._pending is not None:
            raise RuntimeError("a march is pending; drain() it first")
        if self.gen is not None:
            self.gen.set_state(state)
        else:
            torch.cuda.set_rng_state(state, self.dev)

    def sync_step_count(self):
        """Sets the host's optimiser step count (flat.step_count) to the device's,
        which does not advance on skipped iterations; reads device memory."""
        t = int(self.adam_t)
        self.flat.step_count = t
        self._adam_t_host = t
        return t

    @property
    def skipped_iterations(self):
        """Iterations skipped so far for want of visible samples (reads the device)."""
        return int(self.n_skipped)

    @property
    def nonfinite_iterations(self):
        """Iterations so far skipped for a non-finite loss or gradient (reads
        the device; also counted in skipped_iterations)."""
        return int(self.n_nonfinite)

    @property
    def capped_rays(self):
        """Centre rays so far whose samples filled max_samples_per_ray (the cap may
        have truncated them; reads the device)."""
        return int(self.n_capped)

    def counts_total(self):
        """Sums over the iterations so far of (marched centre samples, visible
        centre samples, centre points, patch points evaluated)."""
        return [int(v) for v in self.stat_acc.tolist()]

    @property
    def march_pending(self):
        return self._pending is not None

    def drain(self):
        """Discards a march that ran ahead once it has finished (the current
        stream waits for it); call before freeing the step or changing what the
        march read."""
        if self._pending is not None:
            torch.cuda.current_stream(self.dev).wait_event(self._pending[0])
            self._pending = None

    def _hash_events(self):
        """Records the start event of a timed hash-grid backward and keeps the
        device count and level count it ran with."""
        e = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        e[0].record()
        k = len(self.hash_times)
        if k < self.hash_counts.numel():
            self.hash_counts[k:k + 1].copy_(self.ne)
        self.hash_times.append((e[0], e[1], self.net.active_levels()))
        return e

    def hash_bwd_timings(self):
        """[(ms, n points, active levels)] of the timed hash-grid backwards."""
        torch.cuda.synchronize(self.dev)
        counts = self.hash_counts[:len(self.hash_times)].tolist()
        return [(a.elapsed_time(b), int(n), la)
                for (a, b, la), n in zip(self.hash_times, counts)]

    def _capture(self):
        # one eager warm-up on a side stream (allocator pools, kernel attributes),
        # then capture; the warm-up's update (parameters, moments, step count,
        # fp16 table copy, counters) is undone from saved state.  A march
        # that ran ahead keeps its inputs: the warm-up re-marches them ("keep").
        f = self.flat
        torch.cuda.synchronize(self.dev)
        kept = (f.data, f.exp_avg, f.exp_avg_sq, self.adam_t, self.stat, self.stat_acc)
        saved = [t.clone() for t in kept]
        gen_state = self._rng_state()
        side = torch.cuda.Stream(device=self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            self._iteration("keep" if self._pending is not None else None)
        torch.cuda.current_stream(self.dev).wait_stream(side)
        torch.cuda.synchronize(self.dev)
        for t, v in zip(kept, saved):
            t.copy_(v)
        # the warm-up's Adam step also rewrote the fp16 table copy
        self.refresh_copies()
        if self.gen is not None:
            self.gen.set_state(gen_state)
        else:
            torch.cuda.set_rng_state(gen_state, self.dev)
        graphs = [torch.cuda.CUDAGraph() for _ in range(4)]
        if self.gen is not None:
            graphs[0].register_generator_state(self.gen)
        with torch.cuda.graph(graphs[0]):
            self._part_march()
        with torch.cuda.graph(graphs[1]):
            self._part_rest()
        with torch.cuda.graph(graphs[2]):
            self._part_hash()
        with torch.cuda.graph(graphs[3]):
            self._part_adam()
        self.graph = graphs
