"""One training iteration of the reference restated in torch on the CPU.

TEST INFRASTRUCTURE ONLY (the "plain method on CPU" of the CPU baseline and
the end-to-end checker).  Follows exp_runner.py:147-210 with
models/renderer.py:63-276 and models/fields.py:76-99:
  occupancy update           NA/nerfacc/grid.py:196-239   (c_oracle + torch)
  ray marching               csrc/ray_marching.cu          (c_oracle)
  visibility pre-pass        renderer.py:80-135, vol_rendering.py:680-748
  patch expansion, alpha, weights, dfd/ad gradients, accumulation, losses
                             render_oracle.render_losses
  backward                   torch autograd
  Adam                       torch.optim.Adam
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F

from . import c_oracle, render_oracle
from .field_oracle import weight_norm
from .hashgrid_oracle import encode, hashgrid_desc


class OracleField(torch.nn.Module):
    """SDFNetwork (fields.py) with the oracle hash encoding; parameters copied
    from a supernormal_amd SDFNetwork so both start from the same state."""

    def __init__(self, net, dtype=torch.float32):
        super().__init__()
        g = net.encoding.grid
        self.desc = hashgrid_desc(g["n_levels"], g["log2_hashmap_size"], g["base_resolution"],
                                  g["per_level_scale"])
        self.F = g["n_features"]
        self.params = torch.nn.Parameter(net.encoding.params.detach().cpu().to(dtype).clone())
        self.v, self.g, self.b = torch.nn.ParameterList(), torch.nn.ParameterList(), \
            torch.nn.ParameterList()
        for l in range(net.num_layers - 1):
            lin = getattr(net, f"lin{l}")
            self.v.append(torch.nn.Parameter(lin.weight_v.detach().cpu().to(dtype).clone()))
            self.g.append(torch.nn.Parameter(lin.weight_g.detach().cpu().to(dtype).clone()))
            self.b.append(torch.nn.Parameter(lin.bias.detach().cpu().to(dtype).clone()))
        self.bindwidth = net.bindwidth
        self.dtype = dtype

    def n_active(self):
        L = self.desc["n_levels"]
        return min(self.bindwidth * 2, L * self.F) // self.F

    def forward(self, x):
        enc = encode(x.to(self.dtype), self.params, self.desc, self.F, self.n_active())
        h = torch.cat([x.to(self.dtype), enc], 1)
        n = len(self.v)
        for l in range(n):
            W = weight_norm(self.v[l], self.g[l])
            h = h @ W.T + self.b[l]
            if l < n - 1:
                h = F.softplus(h, beta=100, threshold=20)
        return h[:, 0]


class OracleTrainer:
    def __init__(self, net, variance, occ_res, step_cfg, dtype=torch.float32):
        self.field = OracleField(net, dtype)
        self.variance = torch.nn.Parameter(torch.tensor(float(variance), dtype=dtype))
        self.opt = torch.optim.Adam(list(self.field.parameters()) + [self.variance],
                                    lr=step_cfg["lr"])
        self.occ_res = occ_res
        self.occs = np.zeros(occ_res ** 3, np.float32)
        self.binary = np.zeros((occ_res,) * 3, np.uint8)
        self.cfg = step_cfg
        self.dtype = dtype
        self.rng = np.random.default_rng(step_cfg.get("seed", 0))

    def inv_s(self):
        return torch.exp(self.variance * 10.0).clip(1e-6, 1e6)

    def occupancy_cells(self, step, warmup=256):
        """Cells and jitter of an occupancy update (grid.py:196-223): every cell
        during the warm-up, else N/4 uniform + up to N/4 occupied cells."""
        N = self.occ_res ** 3
        if step < warmup:
            idx = np.arange(N)
        else:
            uni = self.rng.integers(0, N, N // 4)
            occ = np.nonzero(self.binary.reshape(-1))[0]
            if N // 4 < occ.size:
                occ = occ[self.rng.integers(0, occ.size, N // 4)]
            idx = np.concatenate([uni, occ])
        jit = self.rng.random((idx.size, 3), dtype=np.float32)
        return idx, jit

    @torch.no_grad()
    def occupancy_update(self, step, n=8, thre=0.1, warmup=256, cells=None):
        """grid.py:196-239 with occ = sigmoid(-80 sdf) (renderer.py:56-60);
        `cells` = (indices, jitter) to use instead of drawing them."""
        if step % n:
            return
        idx, jit = self.occupancy_cells(step, warmup) if cells is None else cells
        x = c_oracle.occ_cell_points(idx, jit, (self.occ_res,) * 3, [-1, -1, -1, 1, 1, 1])
        occ_v = torch.sigmoid(-self.field(torch.from_numpy(x)) * 80.0).numpy().astype(np.float32)
        self.occs = c_oracle.occ_ema_max(self.occs, idx, occ_v, 0.95)
        b, _ = c_oracle.occ_binarize(self.occs, thre)
        self.binary = b.reshape((self.occ_res,) * 3)

    def forward_backward(self, o, d, plane_n, v_inv, n_true, mask, t_min, t_max, step_size,
                         cap=0, method="dfd", keep=None):
        """The iteration up to the gradients, on CPU tensors ([B,Ph,Pw,*] patches)
        with the centre rays' jittered start t_min and end t_max given.  Leaves
        the gradients in the parameters' .grad and returns a dict of the
        intermediate results: march (packed0, ri0, t00, t10), the visibility
        keep mask and the transmittance in front of each marched sample (T0),
        the visible samples (packed, ri, t0, t1), alpha / weights / gradients
        of every patch sample, losses.  `keep` (bool [n0]) replaces the
        visibility decision (None = the oracle's own); None is returned when no
        sample survives (the reference raises KeyError, exp_runner.py:182)."""
        B, Ph, Pw, _ = o.shape
        o_c, d_c = o[:, Ph // 2, Pw // 2], d[:, Ph // 2, Pw // 2]
        res = {}
        packed0, ri0, t00, t10 = c_oracle.ray_marching(
            o_c.numpy(), d_c.numpy(), np.ascontiguousarray(t_min.numpy(), np.float32),
            np.ascontiguousarray(t_max.numpy(), np.float32),
            np.array([-1, -1, -1, 1, 1, 1], np.float32), self.binary, step_size, 0.0, cap)
        res.update(packed0=packed0, ri0=ri0, t00=t00, t10=t10)
        if ri0.size == 0:
            return None
        ri0t = torch.from_numpy(ri0)
        t00t, t10t = torch.from_numpy(t00), torch.from_numpy(t10)
        with torch.no_grad():
            dm0 = render_oracle.diff_mask(ri0t, t00t, t10t)
            xs = o_c[ri0t] + d_c[ri0t] * t00t
            xe = (o_c[ri0t] + d_c[ri0t] * t10t)[dm0]
            f = self.field(torch.cat([xs, xe]))
            fs = f[: xs.shape[0]]
            fe = torch.cat([fs[1:], fs[-1:]]).clone()
            fe[dm0] = f[xs.shape[0]:]
            inv_s = self.inv_s()
            c, e = torch.sigmoid(fs * inv_s), torch.sigmoid(fe * inv_s)
            alpha0 = ((c - e + 1e-5) / (c + 1e-5)).clip(0, 1).numpy()
        own_keep = c_oracle.visibility(packed0, alpha0.astype(np.float32), 1e-8, 0.0)
        # transmittance in front of each sample (sequential product in fp64)
        T0 = np.ones(ri0.size)
        for r in range(B):
            b, n = int(packed0[r, 0]), int(packed0[r, 1])
            if n:
                T0[b:b + n] = np.concatenate([[1.0], np.cumprod(1.0 - alpha0[b:b + n - 1])])
        res.update(keep_oracle=own_keep, T0=T0)
        keep = own_keep if keep is None else np.asarray(keep, bool)
        res["keep"] = keep
        kt = torch.from_numpy(keep)
        ri = ri0t[kt]
        t0, t1 = t00t[kt], t10t[kt]
        if ri.numel() == 0:
            return None
        packed = torch.from_numpy(c_oracle.pack_info(ri.numpy(), B))
        res.update(packed=packed.numpy(), ri=ri.numpy(), t0=t0.numpy(), t1=t1.numpy())
        dm = render_oracle.diff_mask(ri, t0, t1)
        with torch.no_grad():
            ts = render_oracle.expand_t(t0, ri, d, plane_n)
            te = render_oracle.expand_t(t1, ri, d, plane_n)
            ps = render_oracle.positions(o, d, ri, ts)
            pe = render_oracle.positions(o, d, ri, te)[dm]
        S, P = ri.numel(), Ph * Pw
        self.opt.zero_grad()
        grad_s = None
        if method == "ad":
            x = ps.reshape(-1, 3).to(self.dtype).clone().requires_grad_(True)
            fs_all = self.field(x)
            grad_s, = torch.autograd.grad(fs_all.sum(), x, create_graph=True)
            sdf_s = fs_all.reshape(S, Ph, Pw, 1)
            sdf_e = self.field(pe.reshape(-1, 3)).reshape(-1, Ph, Pw, 1)
            grad_s = grad_s.reshape(S, Ph, Pw, 3)
        else:
            f = self.field(torch.cat([ps.reshape(-1, 3), pe.reshape(-1, 3)]))
            sdf_s = f[: S * P].reshape(S, Ph, Pw, 1)
            sdf_e = f[S * P:].reshape(-1, Ph, Pw, 1)
        m = (mask > 0.5).to(self.dtype)
        out = {}
        loss, parts, cn, ws = render_oracle.render_losses(
            o.to(self.dtype), d.to(self.dtype), plane_n.to(self.dtype), v_inv.to(self.dtype),
            n_true.to(self.dtype), m, packed, ri, t0.to(self.dtype), t1.to(self.dtype), sdf_s,
            sdf_e, self.inv_s(), grad_s, loss_type=self.cfg.get("loss_type", "l2"), out=out)
        loss.backward()
        res.update(loss=float(loss), parts={k: float(v) for k, v in parts.items()},
                   comp_normal=cn.detach(), weight_sum=ws.detach(),
                   n_points=S * P + int(pe.shape[0]) * P, n_centre_points=int(xs.shape[0] + xe.shape[0]),
                   **{k: v.detach() for k, v in out.items()})
        return res

    def gradients(self):
        """{name: gradient} of the parameters after forward_backward: the hash
        table, each layer's weight_v / weight_g / bias, the variance."""
        g = {"params": self.field.params.grad}
        for l in range(len(self.field.v)):
            g[f"lin{l}.weight_v"] = self.field.v[l].grad
            g[f"lin{l}.weight_g"] = self.field.g[l].grad
            g[f"lin{l}.bias"] = self.field.b[l].grad
        g["variance"] = self.variance.grad
        return g

    def step(self, o, d, plane_n, v_inv, n_true, mask, near, far, step_size, lr, cap=0,
             method="dfd", t_min=None):
        """One iteration on CPU tensors ([B,Ph,Pw,*] patches): stratified start
        (near + U[0,1) step, renderer.py:69-71, unless t_min is given), then
        forward_backward and the Adam step; returns the loss (None = skipped)."""
        B = o.shape[0]
        if t_min is None:
            t_min = near + torch.from_numpy(self.rng.random(B, dtype=np.float32)) * step_size
        res = self.forward_backward(o, d, plane_n, v_inv, n_true, mask, t_min, far, step_size,
                                    cap, method)
        if res is None:
            return None
        for gp in self.opt.param_groups:
            gp["lr"] = lr
        self.opt.step()
        return res["loss"]
