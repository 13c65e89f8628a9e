"""SDF field: drop-in for models/fields.py (SDFNetwork, SingleVarianceNetwork).

Same constructor signature, parameter names and state-dict keys
(``encoding.params``, ``lin{l}.weight_g``, ``lin{l}.weight_v``, ``lin{l}.bias``,
``variance``), same geometric initialisation (fields.py:47-65), weight norm
(66-67), Softplus(beta=100) (70), coarse-to-fine bandwidth (73-83) and
analytic gradient (107-119).

Two evaluation paths share the parameters:
  * ``forward`` -- the reference's structure: hash encoding on the HIP kernel
    (supernormal_amd.tcnn.Encoding) followed by torch nn.Linear layers;
  * ``fused_sdf`` / ``fused_sdf_and_gradient`` -- one HIP kernel for encode +
    MLP (+ analytic x-gradient), with hand-written backward and double
    backward (supernormal_amd.sdf_fused).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn as nn

from .tcnn import Encoding


class SDFNetwork(nn.Module):
    def __init__(self, d_in, d_out, d_hidden, n_layers, skip_in=(4,), bias=0.5,
                 geometric_init=True, weight_norm=True, inside_outside=False,
                 encoding_config=None, input_concat=False):
        super().__init__()
        self.input_concat = input_concat
        dims = [d_in] + [d_hidden for _ in range(n_layers)] + [d_out]
        if encoding_config is not None:
            self.encoding = Encoding(d_in, encoding_config).to(torch.float32)
            dims[0] = self.encoding.n_output_dims
            if input_concat:
                dims[0] += d_in
        else:
            self.encoding = None
        self.num_layers = len(dims)
        self.skip_in = skip_in
        self.dims = dims
        self.d_in = d_in
        self.bindwidth = 0
        self.enc_dim = self.encoding.n_output_dims if self.encoding is not None else 0
        self.weight_norm = weight_norm
        for l in range(0, self.num_layers - 1):
            out_dim = dims[l + 1] - dims[0] if (l + 1) in self.skip_in else dims[l + 1]
            lin = nn.Linear(dims[l], out_dim)
            if geometric_init:
                if l == self.num_layers - 2:
                    mean = np.sqrt(np.pi) / np.sqrt(dims[l])
                    if not inside_outside:
                        torch.nn.init.normal_(lin.weight, mean=mean, std=0.0001)
                        torch.nn.init.constant_(lin.bias, -bias)
                    else:
                        torch.nn.init.normal_(lin.weight, mean=-mean, std=0.0001)
                        torch.nn.init.constant_(lin.bias, bias)
                elif self.encoding is not None and l == 0:
                    torch.nn.init.constant_(lin.bias, 0.0)
                    torch.nn.init.constant_(lin.weight[:, 3:], 0.0)
                    torch.nn.init.normal_(lin.weight[:, :3], 0.0, np.sqrt(2) / np.sqrt(out_dim))
                elif self.encoding is not None and l in self.skip_in:
                    torch.nn.init.constant_(lin.bias, 0.0)
                    torch.nn.init.normal_(lin.weight, 0.0, np.sqrt(2) / np.sqrt(out_dim))
                    torch.nn.init.constant_(lin.weight[:, -(dims[0] - 3):], 0.0)
                else:
                    torch.nn.init.constant_(lin.bias, 0.0)
                    torch.nn.init.normal_(lin.weight, 0.0, np.sqrt(2) / np.sqrt(out_dim))
            if weight_norm:
                lin = nn.utils.weight_norm(lin)
            setattr(self, "lin" + str(l), lin)
        self.activation = nn.Softplus(beta=100)

    # -- bandwidth (fields.py:73-83) ---------------------------------------
    def increase_bandwidth(self):
        self.bindwidth += 1

    def active_levels(self) -> int:
        """Levels left on by the mask enc_mask[bindwidth*2:] = 0 (2 features/level)."""
        if self.encoding is None:
            return 0
        F = self.encoding.grid["n_features"]
        L = self.encoding.grid["n_levels"]
        feats = min(self.bindwidth * 2, L * F)
        return (feats + F - 1) // F

    def _enc_mask(self, device):
        m = torch.ones(self.enc_dim, dtype=torch.bool, device=device)
        m[self.bindwidth * 2:] = 0
        return m

    # -- reference-structured forward (fields.py:76-99) ----------------------
    def forward(self, inputs):
        if self.encoding is not None:
            self.encoding.n_active_levels = self.active_levels()
            encoded = self.encoding(inputs).to(torch.float32)
            if (self.bindwidth * 2) % self.encoding.grid["n_features"] != 0:
                encoded = encoded * self._enc_mask(encoded.device)
        if self.input_concat:
            inputs = torch.cat([inputs, encoded], dim=1)
        x = inputs
        for l in range(0, self.num_layers - 1):
            lin = getattr(self, "lin" + str(l))
            if l in self.skip_in:
                x = torch.cat([x, inputs], 1) / np.sqrt(2)
            x = lin(x)
            if l < self.num_layers - 2:
                x = self.activation(x)
        return x

    def sdf(self, x):
        return self.forward(x)[:, :1]

    def sdf_hidden_appearance(self, x):
        return self.forward(x)

    @torch.enable_grad()
    def gradient(self, x):
        x.requires_grad_(True)
        y = self.sdf(x)
        d_output = torch.ones_like(y, requires_grad=False, device=y.device)
        gradients = torch.autograd.grad(outputs=y, inputs=x, grad_outputs=d_output,
                                        create_graph=True, retain_graph=True,
                                        only_inputs=True)[0]
        return gradients.unsqueeze(1)

    # -- fused HIP evaluation ------------------------------------------------
    def sdf_config(self):
        """Shape of this field for the fused kernels (raises if unsupported)."""
        from .sdf import SDFConfig
        n_hidden = self.num_layers - 2
        ok = (self.encoding is not None and self.input_concat and self.d_in == 3
              and self.dims[-1] == 1 and all(l not in self.skip_in for l in range(self.num_layers))
              and n_hidden in (1, 2) and all(h == self.dims[1] for h in self.dims[1:-1])
              and self.dims[1] in (32, 64)
              and (self.bindwidth * 2) % self.encoding.grid["n_features"] == 0)
        if not ok:
            raise ValueError("fused SDF kernels need input_concat, no skips, d_out 1, 1-2 hidden "
                             "layers of width 32 or 64 and 2*bandwidth divisible by the "
                             "features per level")
        return SDFConfig(self.encoding.grid, self.dims[1], n_hidden, beta=100.0,
                         precision=getattr(self, "mlp_precision", "fp32"))

    def packed_mlp(self):
        from .sdf import pack_mlp
        return pack_mlp(self.mlp_layers())

    def fused_sdf(self, x):
        """sdf [N] via hash encode + MFMA MLP kernels (same values as sdf(x)[:, 0])."""
        from .sdf import sdf
        return sdf(x, self.encoding.params, self.packed_mlp(), self.sdf_config(),
                   self.active_levels())

    def fused_sdf_and_gradient(self, x):
        """(sdf [N], d sdf / dx [N,3]), twice differentiable w.r.t. the parameters."""
        from .sdf import sdf_and_gradient
        return sdf_and_gradient(x, self.encoding.params, self.packed_mlp(), self.sdf_config(),
                                self.active_levels())

    # -- folded MLP weights (used by the fused kernels) ----------------------
    def mlp_layers(self):
        """[(W [out,in], b [out])] with weight norm applied (differentiable)."""
        out = []
        for l in range(0, self.num_layers - 1):
            lin = getattr(self, "lin" + str(l))
            if self.weight_norm:
                v, g = lin.weight_v, lin.weight_g
                W = g * v / v.norm(dim=1, keepdim=True)
            else:
                W = lin.weight
            out.append((W, lin.bias))
        return out


class SingleVarianceNetwork(nn.Module):
    """inv_s = exp(10 * variance) (fields.py:133-139)."""

    def __init__(self, init_val):
        super().__init__()
        self.register_parameter("variance", nn.Parameter(torch.tensor(init_val)))

    def forward(self, x):
        return torch.ones([len(x), 1], device=self.variance.device) * torch.exp(self.variance * 10.0)
