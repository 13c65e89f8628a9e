"""Closed-form pins of the render oracle (oracle/render_oracle.py), which
restates models/renderer.py:142-276 and exp_runner.py:169-203.  CPU only.

Each check is chosen so that a plausible slip in the restatement fails it:
  * an affine SDF seen through a real camera's 3x3 patch (anisotropic focal
    lengths, rotated pose): forward and central differences are exact on affine
    fields, so the 'dfd' gradient must equal the field's gradient at every pixel
    and sample -- a swapped dist_x / dist_y, a wrong V^-1 row order, a sign
    error in df/dt or a wrong border stencil all break it;
  * NeuS alpha and the patch weights against the logistic-CDF reading of NeuS:
    along a run of contiguous samples 1 - alpha_i = Phi(f_i+1) / Phi(f_i) up to
    the 1e-5 regularisers, so T telescopes to Phi(f_i) / Phi(f_first) and
    w_i = (Phi(f_i) - Phi(f_i+1)) / Phi(f_first) -- a swapped prev / next CDF,
    a missing inv_s or a shifted transmittance breaks it;
  * the three losses on literal inputs whose values are summed by hand;
  * the patch expansion puts every patch ray's point of a sample on the same
    marching plane (renderer.py:146-149);
  * the non-contiguous-end mask on hand-made sample lists.
"""
import math

import numpy as np
import torch

from oracle import render_oracle as ro

D = torch.float64


def _rot(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _patch(K, R, C, px, py):
    """3x3 patch centred at pixel (px, py) as gen_random_patches builds it
    (dataset_loader.py:240-270): rays [3,3,3] indexed [row (y), col (x)],
    V^-1 from V = [d, camera x, camera y] (dataset_loader.py:114-137)."""
    Kinv = np.linalg.inv(K)
    o = np.zeros((3, 3, 3))
    d = np.zeros((3, 3, 3))
    vinv = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            p = Kinv @ np.array([px + j - 1, py + i - 1, 1.0])
            dd = R @ (p / np.linalg.norm(p))
            o[i, j] = C
            d[i, j] = dd
            vinv[i, j] = np.linalg.inv(np.stack([dd, R[:, 0], R[:, 1]]))
    return o, d, R[:, 2].copy(), vinv


def _cameras():
    K = np.array([[520.0, 0, 306.0], [0, 410.0, 256.0], [0, 0, 1]])  # fx != fy
    cams = []
    for k, (C, px, py) in enumerate([((0.3, -2.2, 1.1), 300, 250), ((-1.7, 0.4, 1.9), 120, 401)]):
        R = _rot(k + 7)
        C = np.array(C)
        # look roughly at the origin: rotate so that camera z points there
        z = -C / np.linalg.norm(C)
        x = np.cross(R[:, 1], z)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        cams.append(_patch(K, np.stack([x, y, z], 1), C, px, py))
    o, d, pn, vinv = (torch.tensor(np.stack(a), dtype=D) for a in zip(*cams))
    return o, d, pn, vinv


def _samples():
    """Patch 0: a contiguous run of 3 samples and one isolated sample; patch 1:
    two contiguous samples."""
    ri = torch.tensor([0, 0, 0, 0, 1, 1])
    t0 = torch.tensor([[2.0], [2.05], [2.1], [2.3], [1.5], [1.52]], dtype=D)
    t1 = torch.tensor([[2.05], [2.1], [2.15], [2.36], [1.52], [1.54]], dtype=D)
    return ri, t0, t1


def _affine_sdf_at(o, d, pn, ri, t0, t1, n, c):
    ts = ro.expand_t(t0, ri, d, pn)
    te = ro.expand_t(t1, ri, d, pn)
    f = lambda t: (ro.positions(o, d, ri, t) @ n)[..., None] + c
    dm = ro.diff_mask(ri, t0, t1)
    sdf_s = f(ts)
    sdf_end = torch.cat([sdf_s[1:], sdf_s[-1:]], 0).clone()
    sdf_end[dm] = f(te)[dm]
    return ts, te, sdf_s, sdf_end, f(te)[dm]


def test_dfd_gradient_of_affine_sdf_is_exact():
    o, d, pn, vinv = _cameras()
    ri, t0, t1 = _samples()
    n = torch.tensor([0.3, -0.8, 0.5], dtype=D) * 1.7
    ts, te, sdf_s, sdf_end, _ = _affine_sdf_at(o, d, pn, ri, t0, t1, n, 0.1)
    g = ro.dfd_gradients(o, d, ri, ts, te, sdf_s, sdf_end, vinv)
    assert g.shape == (6, 3, 3, 3)
    torch.testing.assert_close(g, n.expand_as(g), rtol=0, atol=1e-9)
    # the dfd stencil is not symmetric in x / y: the focal lengths differ, so the
    # distances between neighbours differ along the two image axes
    pos = ro.positions(o, d, ri, ts)
    dx = (pos[:, :, 1:] - pos[:, :, :-1]).norm(dim=-1)
    dy = (pos[:, 1:] - pos[:, :-1]).norm(dim=-1)
    assert (dx.mean() / dy.mean() - 410.0 / 520.0).abs() < 0.01


def test_render_losses_affine_sdf_normals():
    """Through the whole restatement: comp_normal = weight_sum * n for an affine
    SDF (every sample's gradient is n), and the eikonal term is (|n| - 1)^2."""
    o, d, pn, vinv = _cameras()
    ri, t0, t1 = _samples()
    n = torch.tensor([0.6, -0.48, 0.64], dtype=D)  # |n| = 1
    _, _, sdf_s, _, sdf_e = _affine_sdf_at(o, d, pn, ri, t0, t1, n, -0.7)
    B = 2
    mask = torch.ones(B, 3, 3, 1, dtype=D)
    packed = torch.tensor([[0, 4], [4, 2]], dtype=torch.int32)
    out = {}
    loss, parts, cn, ws = ro.render_losses(o, d, pn, vinv, n.expand(B, 3, 3, 3), mask, packed,
                                           ri, t0, t1, sdf_s, sdf_e, torch.tensor(40.0, dtype=D),
                                           out=out)
    torch.testing.assert_close(cn, ws * n, rtol=0, atol=1e-9)
    assert float(parts["eikonal"]) < 1e-18
    assert float(ws.max()) > 0.05  # the samples do carry weight


def test_neus_alpha_and_weights_follow_the_logistic_cdf():
    s = 30.0
    f = torch.tensor([0.3, 0.2, 0.12, 0.05, 0.0, -0.03, -0.05], dtype=D)
    sdf_s, sdf_end = f[:-1, None], f[1:, None]
    alpha = ro.neus_alpha(sdf_s, sdf_end, torch.tensor(s, dtype=D))
    Phi = lambda v: 1.0 / (1.0 + math.exp(-s * v))
    # 1 - alpha = Phi(f_i+1) / (Phi(f_i) + 1e-5)
    for i in range(6):
        assert abs((1 - float(alpha[i])) - Phi(float(f[i + 1])) / (Phi(float(f[i])) + 1e-5)) < 1e-12
    packed = torch.tensor([[0, 6]], dtype=torch.int32)
    w = ro.patch_weights(alpha.reshape(6, 1), packed)[:, 0]
    f0 = Phi(float(f[0]))
    for i in range(6):
        expect = (Phi(float(f[i])) - Phi(float(f[i + 1]))) / f0
        assert abs(float(w[i]) - expect) < 2e-4, (i, float(w[i]), expect)
    assert abs(float(w.sum()) - (1 - Phi(float(f[-1])) / f0)) < 5e-4
    # moving away from the surface: alpha clips to 0; equal values: the regulariser
    a = ro.neus_alpha(torch.tensor([[0.1], [0.2]], dtype=D), torch.tensor([[0.2], [0.2]], dtype=D),
                      torch.tensor(s, dtype=D))
    assert float(a[0]) == 0.0
    assert abs(float(a[1]) - 1e-5 / (Phi(0.2) + 1e-5)) < 1e-15
    # a sharp crossing (s large): the first crossing sample takes all the weight
    a = ro.neus_alpha(torch.tensor([[0.5]], dtype=D), torch.tensor([[-0.5]], dtype=D),
                      torch.tensor(400.0, dtype=D))
    assert abs(float(a[0]) - 1.0) < 1e-9


def test_patch_losses_by_hand():
    """Two patches of 1x2 pixels (values summed by hand)."""
    cn = torch.tensor([[[[1.0, 2.0, 2.0], [0.0, 0.0, 0.0]]],
                       [[[0.5, 0.0, 0.0], [0.0, -1.0, 0.0]]]], dtype=D)  # [2,1,2,3]
    n_true = torch.zeros(2, 1, 2, 3, dtype=D)
    n_true[1, 0, 1] = torch.tensor([0.0, 0.0, 1.0])
    mask = torch.tensor([[[[1.0], [0.0]]], [[[1.0], [1.0]]]], dtype=D)
    ws = torch.tensor([[[[0.5], [0.25]]], [[[1.0], [0.0]]]], dtype=D)
    grads = torch.tensor([[3.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.5]], dtype=D)
    loss, parts = ro.patch_losses(cn, ws, grads, n_true, mask)
    # normal: masked squared errors (1+4+4) + 0.25 + (1+1) over 3 + 1e-5
    assert abs(float(parts["normal"]) - 11.25 / (3 + 1e-5)) < 1e-12
    # eikonal: mean of (3-1)^2, 0, (0.5-1)^2
    assert abs(float(parts["eikonal"]) - (4.0 + 0.25) / 3) < 1e-12
    # mask BCE, mean over 4 pixels, weight sums clipped to [1e-5, 1 - 1e-5]
    bce = -(math.log(0.5) + math.log(1 - 0.25) + math.log(1 - 1e-5) + math.log(1e-5)) / 4
    assert abs(float(parts["mask"]) - bce) < 1e-12
    assert abs(float(loss) - (11.25 / (3 + 1e-5) + 4.25 / 3 + bce)) < 1e-12
    _, p1 = ro.patch_losses(cn, ws, grads, n_true, mask, loss_type="l1")
    assert abs(float(p1["normal"]) - (5.0 + 0.5 + 2.0) / (3 + 1e-5)) < 1e-12


def test_expand_t_puts_patch_points_on_the_marching_plane():
    o, d, pn, _ = _cameras()
    ri, t0, _ = _samples()
    pos = ro.positions(o, d, ri, ro.expand_t(t0, ri, d, pn))  # [S,3,3,3]
    depth = ((pos - o[ri]) * pn[ri][:, None, None, :]).sum(-1)  # [S,3,3]
    centre = t0[:, 0] * (d[ri, 1, 1] * pn[ri]).sum(-1)
    torch.testing.assert_close(depth, centre[:, None, None].expand_as(depth), rtol=0, atol=1e-12)
    # the centre ray keeps its own t
    torch.testing.assert_close(pos[:, 1, 1], o[ri, 1, 1] + d[ri, 1, 1] * t0, rtol=0, atol=1e-12)


def test_diff_mask_cases():
    ri = torch.tensor([0, 0, 0, 1, 1, 2])
    t0 = torch.tensor([[1.0], [1.1], [1.5], [1.6], [1.7], [0.3]], dtype=D)
    t1 = torch.tensor([[1.1], [1.2], [1.6], [1.7], [1.8], [0.4]], dtype=D)
    dm = ro.diff_mask(ri, t0, t1)
    # contiguous / gap / end of ray 0 whose end equals ray 1's start (another ray:
    # not paired) / contiguous / end of ray 1 / last sample
    assert dm.tolist() == [False, True, True, False, True, True]
