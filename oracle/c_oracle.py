"""numpy wrappers of the C oracle (oracle/sn_oracle.c).  Test infrastructure only.

Each wrapper names the reference routine it restates; see sn_oracle.c for the
file:line of each.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_SO = _DIR / "_build" / "liboracle_sn.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_DIR)], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not _SO.exists():
            build()
        _lib = C.CDLL(str(_SO))
        _lib.or_occ_binarize.restype = C.c_float
        _lib.or_pack_info.restype = C.c_int
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def pack_info(ray_indices, n_rays):
    """nerfacc.pack_info (NA/nerfacc/pack.py:46-77) -> int32 [n_rays, 2]."""
    ri = _c(ray_indices, np.int64)
    out = np.zeros((n_rays, 2), np.int32)
    lib().or_pack_info(_p(ri), C.c_int64(ri.size), C.c_int32(n_rays), _p(out))
    return out


def patch_weight_fwd(packed_info, alphas):
    """weight_and_transmittance_from_alpha_patch_based forward (csrc/render_weight.cu:120-152).

    alphas: float32 [n, P] (or [n, P, 1]).  Returns (weights, transmittance) with the
    same shape; samples outside every patch are 0."""
    a = _c(alphas, np.float32)
    shp = a.shape
    a2 = a.reshape(shp[0], -1)
    pi = _c(packed_info, np.int32)
    w = np.zeros_like(a2)
    t = np.zeros_like(a2)
    lib().or_patch_weight_fwd(_p(pi), C.c_int32(pi.shape[0]), C.c_int32(a2.shape[1]), _p(a2),
                              _p(w), _p(t))
    return w.reshape(shp), t.reshape(shp)


def patch_weight_bwd_ref(packed_info, alphas, weights, grad_weights):
    """The reference's backward formula (csrc/render_weight.cu:298-340), fp32."""
    a = _c(alphas, np.float32)
    shp = a.shape
    a2 = a.reshape(shp[0], -1)
    w = _c(weights, np.float32).reshape(a2.shape)
    gw = _c(grad_weights, np.float32).reshape(a2.shape)
    pi = _c(packed_info, np.int32)
    ga = np.zeros_like(a2)
    lib().or_patch_weight_bwd_ref(_p(pi), C.c_int32(pi.shape[0]), C.c_int32(a2.shape[1]),
                                  _p(a2), _p(w), _p(gw), _p(ga))
    return ga.reshape(shp)


def patch_weight_bwd_f64(packed_info, alphas, grad_weights=None, grad_trans=None):
    """d/d alpha of sum(gw*w) + sum(gT*T), division-free recurrence, fp64."""
    a = _c(alphas, np.float64)
    shp = a.shape
    a2 = a.reshape(shp[0], -1)
    gw = None if grad_weights is None else _c(grad_weights, np.float64).reshape(a2.shape)
    gT = None if grad_trans is None else _c(grad_trans, np.float64).reshape(a2.shape)
    pi = _c(packed_info, np.int32)
    ga = np.zeros_like(a2)
    lib().or_patch_weight_bwd_f64(_p(pi), C.c_int32(pi.shape[0]), C.c_int32(a2.shape[1]),
                                  _p(a2), _p(gw), _p(gT), _p(ga))
    return ga.reshape(shp)


def visibility(packed_info, alphas, early_stop_eps, alpha_thre):
    """render_visibility with the sequential product (vol_rendering.py:680-748,
    csrc/render_transmittance.cu:85-112) -> bool [n]."""
    a = _c(alphas, np.float32).reshape(-1)
    pi = _c(packed_info, np.int32)
    keep = np.zeros(a.size, np.uint8)
    lib().or_visibility(_p(pi), C.c_int32(pi.shape[0]), _p(a), C.c_float(early_stop_eps),
                        C.c_float(alpha_thre), _p(keep))
    return keep.astype(bool)


def ray_marching(rays_o, rays_d, t_min, t_max, roi, grid, step, cone=0.0, max_per_ray=0):
    """_C.ray_marching (csrc/ray_marching.cu:194-289) -> packed_info, ray_indices,
    t_starts [n,1], t_ends [n,1]."""
    o = _c(rays_o, np.float32)
    d = _c(rays_d, np.float32)
    tn = _c(t_min, np.float32)
    tf = _c(t_max, np.float32)
    roi = _c(roi, np.float32)
    g = _c(grid, np.uint8)
    rx, ry, rz = g.shape
    n = o.shape[0]
    counts = np.zeros(n, np.int32)
    L = lib()
    args = (_p(o), _p(d), _p(tn), _p(tf), C.c_int32(n), _p(roi), _p(g), C.c_int32(rx),
            C.c_int32(ry), C.c_int32(rz), C.c_float(step), C.c_float(cone),
            C.c_int32(max_per_ray))
    L.or_ray_march_count(*args, _p(counts))
    cum = np.cumsum(counts, dtype=np.int64)
    packed = np.stack([cum - counts, counts], -1).astype(np.int32)
    tot = int(cum[-1]) if n else 0
    ri = np.zeros(tot, np.int64)
    t0 = np.zeros(tot, np.float32)
    t1 = np.zeros(tot, np.float32)
    L.or_ray_march_fill(*args, _p(packed), _p(ri), _p(t0), _p(t1))
    return packed, ri, t0[:, None], t1[:, None]


def occ_cell_points(cell_idx, jitter, res, roi):
    """grid.py:214-228 + AABB contract_inv: cell -> world sample point."""
    idx = _c(cell_idx, np.int64)
    jit = _c(jitter, np.float32)
    roi = _c(roi, np.float32)
    x = np.zeros((idx.size, 3), np.float32)
    lib().or_occ_cell_points(_p(idx), C.c_int32(idx.size), _p(jit), C.c_int32(res[0]),
                             C.c_int32(res[1]), C.c_int32(res[2]), _p(roi), _p(x))
    return x


def occ_ema_max(occs, cell_idx, occ, decay):
    """grid.py:232 as a deterministic scatter-max; returns the new occs."""
    o = _c(occs, np.float32).copy()
    idx = _c(cell_idx, np.int64)
    oc = _c(occ, np.float32)
    lib().or_occ_ema_max(_p(o), C.c_int32(o.size), _p(idx), _p(oc), C.c_int32(idx.size),
                         C.c_float(decay))
    return o


def occ_binarize(occs, thre):
    """grid.py:237-239 -> (binary uint8, threshold used)."""
    o = _c(occs, np.float32)
    b = np.zeros(o.size, np.uint8)
    t = lib().or_occ_binarize(_p(o), C.c_int32(o.size), C.c_float(thre), _p(b))
    return b, float(t)


def hashgrid_corners(x, desc):
    """Integer part of the hash-grid encode: corner indices [n,L,8] (level-local)
    and fp32 fractional positions [n,L,3], cells [n,L,3] (as float)."""
    xx = _c(x, np.float32)
    n = xx.shape[0]
    L = desc["n_levels"]
    scale = _c(desc["scale"], np.float32)
    res = _c(desc["res"], np.uint32)
    hs = _c(desc["hsize"], np.uint32)
    dense = _c(desc["dense"], np.uint8)
    idx = np.zeros((n, L, 8), np.uint32)
    frac = np.zeros((n, L, 3), np.float32)
    cell = np.zeros((n, L, 3), np.float32)
    lib().or_hashgrid_corners(_p(xx), C.c_int64(n), C.c_int32(L), _p(scale), _p(res), _p(hs),
                              _p(dense), _p(idx), _p(frac), _p(cell))
    return idx, frac, cell
