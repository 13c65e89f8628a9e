"""Evaluation path (visible surface points, Chamfer / F-score) on the CPU.

The ground-truth SDF of a synthetic scene stands in for the learned field:
the first hits along the masked pixel rays must then land on the same
surface as the scene's own visible points, so the Chamfer distance is at the
bisection resolution and the F-score is 1.  Chamfer / F-score themselves are
checked against brute-force nearest neighbours in numpy.
"""
import numpy as np
import torch

from supernormal_amd import mesh, scene


def test_chamfer_matches_bruteforce():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(300, 3))
    b = rng.normal(size=(200, 3)) * 1.1
    d = np.sqrt(((a[:, None, :] - b[None, :, :]) ** 2).sum(-1))
    e2r, r2e = d.min(1), d.min(0)
    cd_ref = (e2r.mean() + r2e.mean()) / 2
    p, r = (e2r < 0.3).mean(), (r2e < 0.3).mean()
    f_ref = 2 * p * r / (p + r)
    # eval_points = a, ref_points = b
    cd, f = mesh.chamfer_distance_and_f1_score(b, a, 0.3)
    assert abs(cd - cd_ref) < 1e-12
    assert abs(f - f_ref) < 1e-12


def test_visible_points_of_true_sdf_match_ground_truth():
    ds = scene.config_dataset("sphere128", torch.device("cpu"))
    gt = ds.surface_points(n=2000)
    assert np.isfinite(gt).all()
    # On the sphere of radius 0.5 every ground-truth hit lies on the surface.
    assert np.abs(np.linalg.norm(gt, axis=1) - 0.5).max() < 1e-4
    ev = mesh.visible_surface_points(lambda x: ds.scene(x.double()).float(), ds,
                                     rays_per_view=250)
    assert ev.shape[0] > 1500
    cd, f = mesh.chamfer_distance_and_f1_score(gt, ev, 0.005)
    assert cd < 5e-3 and f > 0.99


def test_first_hits_skip_rays_that_miss():
    sph = lambda x: x.norm(dim=-1) - 0.3
    o = torch.tensor([[0.0, 0.0, -0.9], [0.0, 0.8, -0.9]])
    d = torch.tensor([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    near = torch.tensor([0.0, 0.0])
    far = torch.tensor([1.8, 1.8])
    pts, hit = mesh.first_hits(sph, o, d, near, far, step=1e-2)
    assert hit.tolist() == [True, False]
    assert abs(float(pts[0, 2]) + 0.3) < 1e-4
