"""Hash-grid encoding in fp64 torch.  Test infrastructure only.

Restates tiny-cuda-nn's GridEncoding (HashGrid, 3-D, linear interpolation) as
used by the reference at models/fields.py:25-31, 78 (tcnn commit 2ec562e8 is
pinned at create_env.sh:11 but is not vendored in the reference, so the
method is restated from the Instant-NGP paper and tcnn's published source):
per level l, scale = base*b^l - 1, res = ceil(scale)+1, table size
min(next_multiple(res^3, 8), T), pos = scale*x + 0.5, trilinear blend of the
8 corners, corner index dense or xor-prime hashed modulo the table size.

The integer corner selection (incl. the fp32 fmaf that decides the cell) comes
from the C oracle; the interpolation is fp64 and differentiable, so torch
autograd supplies the backward and double backward checked against the HIP
kernels.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import c_oracle


def hashgrid_desc(n_levels, log2_hashmap_size, base_resolution, per_level_scale):
    """Per-level layout; offsets in table entries (not floats)."""
    T = 1 << log2_hashmap_size
    l2s = math.log2(float(np.float32(per_level_scale)))
    scale, res, hsize, dense, offset = [], [], [], [], [0]
    for l in range(n_levels):
        sc = float(np.float32(2.0 ** (l * l2s) * base_resolution - 1.0))
        r = int(math.ceil(sc)) + 1
        p = min(r ** 3, 0xFFFFFFFF // 2)
        p = (p + 7) // 8 * 8
        p = min(p, T)
        scale.append(sc)
        res.append(r)
        hsize.append(p)
        dense.append(1 if r ** 3 <= p else 0)
        offset.append(offset[-1] + p)
    return {
        "n_levels": n_levels,
        "scale": np.array(scale, np.float32),
        "res": np.array(res, np.uint32),
        "hsize": np.array(hsize, np.uint32),
        "dense": np.array(dense, np.uint8),
        "offset": np.array(offset, np.int64),
        "n_entries": offset[-1],
    }


def encode(x, params, desc, n_features=2, n_active=None, fp32_pos=True):
    """enc[i, l*F+f] for x [n,3] and flat params (fp64 tensors; autograd-able).

    fp32_pos=True (the method as specified): fractional positions are the fp32
    values of fmaf(scale, x, 0.5) - cell.  fp32_pos=False: the same corners with
    fp64 positions, a smooth model for finite-difference checks."""
    L = desc["n_levels"]
    F = n_features
    if n_active is None:
        n_active = L
    n = x.shape[0]
    xnp = x.detach().to(torch.float32).cpu().numpy()
    idx, frac32, cell32 = c_oracle.hashgrid_corners(xnp, desc)
    idx_t = torch.from_numpy(idx.astype(np.int64))
    frac32_t = torch.from_numpy(frac32).to(x.dtype)
    table = params.view(-1, F)
    outs = []
    for l in range(L):
        if l >= n_active:
            outs.append(torch.zeros(n, F, dtype=x.dtype))
            continue
        sc = float(desc["scale"][l])
        # The position is part of the method's definition in fp32 (fmaf(scale, x, 0.5)
        # rounded to fp32, ulp ~3e-5 at scale ~270): the value is the fp32 fraction,
        # the derivative d frac / dx = scale is exact.
        if fp32_pos:
            frac = frac32_t[:, l] + (sc * x - (sc * x).detach())  # [n,3]
        else:
            frac = sc * x + 0.5 - torch.from_numpy(cell32[:, l]).to(x.dtype)
        acc = 0.0
        for c in range(8):
            w = torch.ones(n, dtype=x.dtype)
            for d in range(3):
                w = w * (frac[:, d] if (c >> d) & 1 else 1.0 - frac[:, d])
            rows = idx_t[:, l, c] + int(desc["offset"][l])
            acc = acc + w[:, None] * table[rows]
        outs.append(acc)
    return torch.cat(outs, dim=1)
