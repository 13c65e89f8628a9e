"""Pins the CPU oracle against the reference's own golden vectors, closed forms,
invariants and brute force.  CPU only."""
import math
import numpy as np
import pytest
import torch

from oracle import c_oracle
from oracle.hashgrid_oracle import encode, hashgrid_desc


# ---------------------------------------------------------------- pack_info
def test_pack_info_golden(golden):
    g = golden["pack"]
    pi = c_oracle.pack_info(np.array(g["ray_indices"]), len(g["packed_info"]))
    assert pi.tolist() == g["packed_info"]


def test_pack_info_invariants():
    rng = np.random.default_rng(0)
    ri = np.sort(rng.integers(0, 50, 1000))
    pi = c_oracle.pack_info(ri, 60)
    assert pi[:, 1].sum() == 1000
    assert (pi[1:, 0] == np.cumsum(pi[:, 1])[:-1]).all() and pi[0, 0] == 0
    np.testing.assert_array_equal(np.repeat(np.arange(60), pi[:, 1]), ri)


# ---------------------------------------------------------------- weights
def _ri_to_packed(ri, n):
    return c_oracle.pack_info(np.array(ri), n)


def test_weights_golden_test_rendering(golden):
    g = golden["weights"]
    pi = _ri_to_packed(g["ray_indices"], g["n_rays"])
    w, T = c_oracle.patch_weight_fwd(pi, np.array(g["alphas"], np.float32)[:, None])
    np.testing.assert_allclose(w[:, 0], g["weights"], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(T[:, 0], golden["visibility"]["transmittance"], rtol=1e-6)


def test_weights_golden_docstring(golden):
    g = golden["docstring"]
    pi = _ri_to_packed(g["ray_indices"], 3)
    w, T = c_oracle.patch_weight_fwd(pi, np.array(g["alphas"], np.float32)[:, None])
    np.testing.assert_allclose(w[:, 0], g["weights"], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(T[:, 0], g["transmittance"], rtol=1e-6)


def test_visibility_golden(golden):
    g = golden["visibility"]
    pi = _ri_to_packed(g["ray_indices"], g["n_rays"])
    for case in g["cases"]:
        vis = c_oracle.visibility(pi, np.array(g["alphas"], np.float32), case["early_stop_eps"],
                                  case["alpha_thre"])
        assert vis.tolist() == case["visibility"]
    d = golden["docstring"]
    vis = c_oracle.visibility(_ri_to_packed(d["ray_indices"], 3), np.array(d["alphas"]),
                              d["visibility"]["early_stop_eps"], d["visibility"]["alpha_thre"])
    assert vis.tolist() == d["visibility"]["expected"]


def test_weight_grads_golden(golden):
    """d(sum w)/d sigma through alpha = 1 - exp(-sigma dt) (test_rendering.py:137-214)."""
    g = golden["grads"]
    sig = np.array(g["sigmas"], np.float64)
    a = 1.0 - np.exp(-sig * g["dt"])
    pi = np.array(g["packed_info"], np.int32)
    w, _ = c_oracle.patch_weight_fwd(pi, a.astype(np.float32)[:, None])
    np.testing.assert_allclose(w[:, 0], g["weights"], atol=g["atol"])
    ga = c_oracle.patch_weight_bwd_f64(pi, a[:, None], np.ones((5, 1)))[:, 0]
    np.testing.assert_allclose(ga * np.exp(-sig * g["dt"]), g["sigmas_grad"], atol=g["atol"])
    ga_ref = c_oracle.patch_weight_bwd_ref(pi, a.astype(np.float32)[:, None], w,
                                           np.ones((5, 1), np.float32))[:, 0]
    np.testing.assert_allclose(ga_ref * np.exp(-sig), g["sigmas_grad"], atol=g["atol"])


def _brute_weights(a):
    """w_j = a_j prod_{m<j}(1-a_m), T_j = prod_{m<j}(1-a_m), fp64, one ray."""
    T = np.cumprod(np.concatenate([[1.0], 1.0 - a]))[:-1]
    return a * T, T


def _random_patches(rng, n_patches, P, max_steps, alpha_hi=1.0):
    steps = rng.integers(0, max_steps + 1, n_patches)
    pi = np.stack([np.concatenate([[0], np.cumsum(steps)[:-1]]), steps], -1).astype(np.int32)
    a = rng.uniform(0, alpha_hi, (int(steps.sum()), P))
    return pi, a


def test_patch_weights_reduce_to_per_ray():
    rng = np.random.default_rng(1)
    pi, a = _random_patches(rng, 40, 9, 12)
    w, T = c_oracle.patch_weight_fwd(pi, a.astype(np.float32))
    for p in range(pi.shape[0]):
        b, s = pi[p]
        for k in range(9):
            wb, Tb = _brute_weights(a[b:b + s, k].astype(np.float32).astype(np.float64))
            np.testing.assert_allclose(w[b:b + s, k], wb, rtol=2e-6, atol=1e-7)
            np.testing.assert_allclose(T[b:b + s, k], Tb, rtol=2e-6, atol=1e-7)


def test_patch_weight_bwd_fd_and_reference_formula():
    rng = np.random.default_rng(2)
    pi, a = _random_patches(rng, 12, 9, 8, alpha_hi=0.95)
    gw = rng.normal(size=a.shape)
    gT = rng.normal(size=a.shape)
    ga = c_oracle.patch_weight_bwd_f64(pi, a, gw, gT)

    def loss(aa):
        tot = 0.0
        for p in range(pi.shape[0]):
            b, s = pi[p]
            for k in range(9):
                w, T = _brute_weights(aa[b:b + s, k])
                tot += (gw[b:b + s, k] * w).sum() + (gT[b:b + s, k] * T).sum()
        return tot

    eps = 1e-6
    for idx in [(0, 0), (3, 4), (a.shape[0] - 1, 8), (a.shape[0] // 2, 2)]:
        ap, am = a.copy(), a.copy()
        ap[idx] += eps
        am[idx] -= eps
        fd = (loss(ap) - loss(am)) / (2 * eps)
        assert abs(fd - ga[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, ga[idx])
    # the reference's divide-by-(1-a) formula agrees where it is stable
    a32 = a.astype(np.float32)
    w32, _ = c_oracle.patch_weight_fwd(pi, a32)
    ga_ref = c_oracle.patch_weight_bwd_ref(pi, a32, w32, gw.astype(np.float32))
    ga_nT = c_oracle.patch_weight_bwd_f64(pi, a32.astype(np.float64), gw)
    np.testing.assert_allclose(ga_ref, ga_nT, rtol=1e-4, atol=1e-5)


def test_patch_weight_bwd_alpha_one_is_finite():
    pi = np.array([[0, 3]], np.int32)
    a = np.array([[0.3], [1.0], [0.5]])
    ga = c_oracle.patch_weight_bwd_f64(pi, a, np.ones_like(a))
    assert np.isfinite(ga).all()
    # closed form: w = [a0, (1-a0) a1, (1-a0)(1-a1) a2]; d sum w / d a1 = (1-a0)(1 - a2)
    np.testing.assert_allclose(ga[1, 0], (1 - 0.3) * (1 - 0.5))


# ---------------------------------------------------------------- marching
def _rays(rng, n):
    o = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


ROI = np.array([-1, -1, -1, 1, 1, 1], np.float32)


def test_march_all_occupied_is_progression():
    rng = np.random.default_rng(3)
    o, d = _rays(rng, 64)
    tmin = np.zeros(64, np.float32)
    tmax = np.ones(64, np.float32) * 0.5
    grid = np.ones((16, 16, 16), np.uint8)
    pi, ri, t0, t1 = c_oracle.ray_marching(o, d, tmin, tmax, ROI, grid, 1e-2)
    for r in range(64):
        b, s = pi[r]
        mids = o[r] + d[r] * ((t0[b:b + s] + t1[b:b + s]) / 2)
        inside = (np.abs(mids) <= 1).all(1)
        # while inside the roi every step is emitted, contiguous from t_min
        if s:
            assert t0[b, 0] == tmin[r]
            np.testing.assert_array_equal(t0[b + 1:b + s, 0], t1[b:b + s - 1, 0])
            assert inside.all()
        assert (ri[b:b + s] == r).all()
    assert pi[:, 1].sum() == ri.size


def test_march_empty_grid_gives_nothing():
    rng = np.random.default_rng(4)
    o, d = _rays(rng, 32)
    pi, ri, _, _ = c_oracle.ray_marching(o, d, np.zeros(32), np.full(32, 3.0), ROI,
                                         np.zeros((8, 8, 8), np.uint8), 1e-2)
    assert ri.size == 0 and (pi[:, 1] == 0).all()


def test_march_samples_in_occupied_cells_and_cap_is_prefix():
    rng = np.random.default_rng(5)
    o = rng.uniform(-0.5, 0.5, (128, 3)).astype(np.float32) * 0 + np.array([0, 0, -2.5], np.float32)
    d = rng.normal(size=(128, 3)) * 0.15 + np.array([0, 0, 1.0])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    res = 32
    c = (np.arange(res) + 0.5) / res * 2 - 1
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    grid = (np.abs(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.5) < 0.1).astype(np.uint8)
    tmin = np.full(128, 1.0, np.float32)
    tmax = np.full(128, 4.0, np.float32)
    pi, ri, t0, t1 = c_oracle.ray_marching(o, d, tmin, tmax, ROI, grid, 5e-3)
    assert ri.size > 0
    mids = o[ri] + d[ri] * ((t0 + t1) / 2)
    u = (mids - ROI[:3]) / 2
    ix = np.clip((u * res).astype(np.int64), 0, res - 1)
    assert grid[ix[:, 0], ix[:, 1], ix[:, 2]].all()
    assert ((mids >= -1) & (mids <= 1)).all()
    pic, ric, t0c, _ = c_oracle.ray_marching(o, d, tmin, tmax, ROI, grid, 5e-3, max_per_ray=7)
    assert (pic[:, 1] == np.minimum(pi[:, 1], 7)).all()
    for r in range(128):
        np.testing.assert_array_equal(t0c[pic[r, 0]:pic[r, 0] + pic[r, 1]],
                                      t0[pi[r, 0]:pi[r, 0] + pic[r, 1]])


def _slab_march(near, step, far=2.0):
    """Ray along +x from x = -1 (t = x + 1) through an 8^3 grid over the roi
    [-1, 1]^3 whose only occupied cells are the slab ix = 5 (x in [0.25, 0.5),
    t in [1.25, 1.5)); y = z = 0.15 so the DDA's y / z terms are +inf."""
    o = np.array([[-1.0, 0.15, 0.15]], np.float32)
    d = np.array([[1.0, 0.0, 0.0]], np.float32)
    grid = np.zeros((8, 8, 8), np.uint8)
    grid[5] = 1
    return c_oracle.ray_marching(o, d, np.array([near], np.float32),
                                 np.array([far], np.float32), ROI, grid, step)


def test_march_skip_then_slab_aligned():
    """Closed form of the empty-space skip (ray_marching.cu:47-75, 165-171) with
    dyadic numbers: every skip moves t_mid by a whole cell (4 steps of 1/16), so
    t_mid = 1/32 + k/4 lands at 1.28125 in the slab, the interval is re-centred
    on it (t0 = t_mid - dt/2 = 1.25, the slab's boundary) and the slab's four
    steps are emitted contiguously; the skips after the slab end the ray."""
    pi, ri, t0, t1 = _slab_march(0.0, 0.0625)
    assert pi.tolist() == [[0, 4]] and ri.tolist() == [0, 0, 0, 0]
    np.testing.assert_array_equal(t0[:, 0], [1.25, 1.3125, 1.375, 1.4375])
    np.testing.assert_array_equal(t1[:, 0], [1.3125, 1.375, 1.4375, 1.5])


def test_march_skip_then_slab_jittered():
    """With a stratified start the skip does not re-align to the cell grid: the
    first sample's midpoint is near + dt/2 + k*dt with the smallest k that puts
    it at or past the slab boundary t = 1.25 (advance_to_next_voxel steps by
    dt_min from t_mid, ray_marching.cu:62-75); then the slab is sampled every dt
    while the midpoint stays inside, t0 / t1 = midpoint -/+ dt/2."""
    step = 0.0625
    for near in (0.01, 0.02, 0.037, 0.05):
        pi, ri, t0, t1 = _slab_march(near, step)
        m0 = near + step / 2
        k = math.ceil((1.25 - m0) / step)
        mids = [m0 + (k + j) * step for j in range(8)]
        mids = [m for m in mids if m < 1.5]
        assert int(pi[0, 1]) == len(mids) == ri.size
        np.testing.assert_allclose((t0[:, 0] + t1[:, 0]) / 2, mids, atol=2e-6)
        np.testing.assert_allclose(t1[:, 0] - t0[:, 0], step, atol=2e-6)
        # contiguous inside the slab
        np.testing.assert_array_equal(t0[1:, 0], t1[:-1, 0])


# ---------------------------------------------------------------- occupancy
def test_contract_inv_golden(golden):
    g = golden["contraction_aabb"]
    unit = np.array(g["unit"], np.float32)
    x = c_oracle.occ_cell_points(np.zeros(len(unit), np.int64), unit, (1, 1, 1), g["roi"])
    np.testing.assert_allclose(x, g["world"], atol=1e-6)


def test_occ_update_sphere():
    """test_grid.py:9-21 occ_eval_fn (ball of radius 0.5 in the [0,1]^3 roi)."""
    res = 16
    n = res ** 3
    rng = np.random.default_rng(6)
    idx = np.arange(n)
    jit = rng.random((n, 3), dtype=np.float32)
    x = c_oracle.occ_cell_points(idx, jit, (res, res, res), [0, 0, 0, 1, 1, 1])
    coord = np.stack(np.meshgrid(*[np.arange(res)] * 3, indexing="ij"), -1).reshape(-1, 3)
    np.testing.assert_allclose(x, (coord + jit) / res, rtol=0, atol=1e-6)
    occ = (np.linalg.norm(x - 0.5, axis=1) < 0.5).astype(np.float32)
    occs = c_oracle.occ_ema_max(np.zeros(n, np.float32), idx, occ, 0.95)
    np.testing.assert_array_equal(occs, occ)
    b, t = c_oracle.occ_binarize(occs, 0.1)
    assert t == pytest.approx(0.1)  # mean ~ pi/6 > 0.1
    np.testing.assert_array_equal(b.astype(bool), occ > 0.1)


def test_occ_scatter_max_duplicates_bruteforce():
    rng = np.random.default_rng(7)
    n_cells = 50
    occs = rng.random(n_cells, dtype=np.float32)
    idx = rng.integers(0, n_cells, 400)
    occ = rng.random(400, dtype=np.float32)
    got = c_oracle.occ_ema_max(occs, idx, occ, 0.95)
    want = occs.copy()
    for c in range(n_cells):
        hit = occ[idx == c]
        if hit.size:
            want[c] = max(np.float32(occs[c] * np.float32(0.95)), hit.max())
    np.testing.assert_array_equal(got, want)
    # the threshold is min(mean, thre)
    b, t = c_oracle.occ_binarize(got, 0.99)
    assert t == pytest.approx(float(np.mean(got.astype(np.float64))), rel=1e-6)


# ---------------------------------------------------------------- hash grid
def test_hashgrid_layout_golden(golden):
    for c in golden["hashgrid_layout"]["cases"]:
        d = hashgrid_desc(c["n_levels"], c["log2_hashmap_size"], c["base_resolution"],
                          c["per_level_scale"])
        assert d["n_entries"] == c["entries"]
        assert int(d["dense"].sum()) == c["dense_levels"]


def _small_desc():
    return hashgrid_desc(4, 10, 4, 1.7)


def test_hashgrid_partition_of_unity():
    d = _small_desc()
    x = torch.rand(200, 3, dtype=torch.float64) * 2 - 1
    params = torch.full((d["n_entries"] * 2,), 0.37, dtype=torch.float64)
    enc = encode(x, params, d)
    torch.testing.assert_close(enc, torch.full_like(enc, 0.37), rtol=1e-12, atol=1e-12)


def test_hashgrid_dense_level_reproduces_affine():
    """On a dense level with table = affine(corner coordinate), trilinear
    interpolation reproduces the affine function of pos = scale*x + 0.5."""
    d = _small_desc()
    assert d["dense"][0] == 1
    res, sc = int(d["res"][0]), float(d["scale"][0])
    e = np.arange(d["hsize"][0])
    cx, cy, cz = e % res, (e // res) % res, e // (res * res)
    A = np.array([0.3, -0.2, 0.5])
    val = A[0] * cx + A[1] * cy + A[2] * cz + 0.1
    params = torch.zeros(d["n_entries"] * 2, dtype=torch.float64)
    params.view(-1, 2)[: len(e), 0] = torch.from_numpy(val)
    # inputs inside [0, 1) so no corner wraps: cells stay within [0, res-1)
    x = torch.rand(300, 3, dtype=torch.float64) * (res - 1.5) / sc
    enc = encode(x, params, d, n_active=1)
    # pos = fmaf(scale, x, 0.5) in fp32 (= cell + frac from the corner selection)
    _, frac, cell = c_oracle.hashgrid_corners(x.numpy().astype(np.float32), d)
    pos = torch.from_numpy(cell[:, 0].astype(np.float64) + frac[:, 0].astype(np.float64))
    torch.testing.assert_close(pos, sc * x.float().double() + 0.5, rtol=0, atol=1e-5)
    want = pos @ torch.tensor(A) + 0.1
    torch.testing.assert_close(enc[:, 0], want, rtol=1e-9, atol=1e-9)


def test_hashgrid_wrap_and_hash_indices_bruteforce():
    d = hashgrid_desc(6, 12, 4, 2.0)
    x = (np.random.default_rng(8).random((50, 3)) * 2 - 1).astype(np.float32)
    idx, frac, cell = c_oracle.hashgrid_corners(x, d)
    for i in range(50):
        for l in range(6):
            sc = np.float32(d["scale"][l])
            for k in range(3):
                p = np.float64(sc) * np.float64(x[i, k]) + 0.5
                assert cell[i, l, k] == np.floor(np.float32(p)) or abs(p - round(p)) < 1e-6
            g = cell[i, l].astype(np.int64).astype(np.uint32)
            for c in range(8):
                gg = [(int(g[k]) + ((c >> k) & 1)) & 0xFFFFFFFF for k in range(3)]
                if d["dense"][l]:
                    r = int(d["res"][l])
                    want = (gg[0] + gg[1] * r + gg[2] * r * r) & 0xFFFFFFFF
                else:
                    want = gg[0] ^ ((gg[1] * 2654435761) & 0xFFFFFFFF) ^ ((gg[2] * 805459861) & 0xFFFFFFFF)
                assert idx[i, l, c] == want % int(d["hsize"][l])


def test_hashgrid_grad_fd():
    d = _small_desc()
    g = torch.Generator().manual_seed(0)
    params = (torch.rand(d["n_entries"] * 2, generator=g, dtype=torch.float64) - 0.5)
    x = (torch.rand(6, 3, generator=g, dtype=torch.float64) * 1.6 - 0.8).requires_grad_(True)
    p = params.clone().requires_grad_(True)

    def f(xx, pp):
        return encode(xx, pp, d, fp32_pos=False)

    # finite differences stay inside one cell for tiny eps
    assert torch.autograd.gradcheck(f, (x, p), eps=1e-7, atol=1e-5, rtol=1e-4)


# ---------------------------------------------------------------- sanitizers
def test_oracle_under_address_and_ub_sanitizers():
    """The C oracle built with -fsanitize=address,undefined runs every
    function on exact-size heap buffers (empty rays, the sample cap,
    out-of-range indices, NaN occupancies, axis-aligned rays) without a
    report; any report aborts the driver."""
    import shutil
    import subprocess
    from pathlib import Path
    if not shutil.which("make") or not shutil.which("gcc"):
        pytest.skip("make / gcc not present")
    root = Path(__file__).resolve().parents[1] / "oracle"
    r = subprocess.run(["make", "-s", "-C", str(root), "sanitize"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "all functions ran clean" in r.stdout
