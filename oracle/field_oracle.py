"""SDF network in fp64 torch.  Test infrastructure only.

Restates models/fields.py:76-119: h0 = [x, enc(x)] (input_concat), hidden
layers Linear + Softplus(beta=100, threshold=20), linear output; weights are
the weight-normed matrices g * v / ||v||_row (fields.py:66-67).  The hash
encoding is oracle.hashgrid_oracle.encode.  torch autograd (a library, not
reference code) supplies first and second derivatives.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

from .hashgrid_oracle import encode


def mlp(h0, layers, beta=100.0):
    """layers: [(W, b)] hidden ... output; h0 [n, d_in]."""
    h = h0
    for W, b in layers[:-1]:
        h = F.softplus(h @ W.T + b, beta=beta, threshold=20)
    W, b = layers[-1]
    return h @ W.T + b


def sdf(x, params, layers, desc, n_features=2, n_active=None, beta=100.0):
    enc = encode(x, params, desc, n_features=n_features, n_active=n_active)
    return mlp(torch.cat([x, enc], dim=1), layers, beta)[:, 0]


def sdf_and_gradient(x, params, layers, desc, n_features=2, n_active=None, beta=100.0):
    """(sdf, d sdf/dx) with create_graph, as fields.py:107-119."""
    x = x.detach().requires_grad_(True)
    s = sdf(x, params, layers, desc, n_features, n_active, beta)
    g, = torch.autograd.grad(s.sum(), x, create_graph=True)
    return s, g


def weight_norm(v, g):
    return g * v / v.norm(dim=1, keepdim=True)
