"""Patch sampling kernel (sn_patch_sample) against the torch restatement of
the reference's gen_random_patches + near_far_from_sphere + stratified start
(Dataset.gen_random_patches / near_far_from_sphere, dataset_loader.py:223-297,
renderer.py:69-71) on the same random numbers: gathered supervision
(normals, V^-1, mask) bit-exact; rays, near/far within fp32 rounding
(rtol 1e-5); the mask sum in a fixed order within 1e-6 relative."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ph,pw,mask_mode", [(3, 3, 1), (2, 4, 0)])
def test_patch_sample_matches_torch(dev, ph, pw, mask_mode):
    from supernormal_amd._lib import call, ptr, stream_of
    from supernormal_amd.scene import config_dataset
    ds = config_dataset("sphere128", dev, 0)
    B = 700
    step = torch.tensor([7e-3], device=dev)
    g1 = torch.Generator(device=dev).manual_seed(5)
    o, d, pn, vi, nt, m = ds.gen_random_patches(B, patch_H=ph, patch_W=pw, generator=g1)
    kc = (ph // 2, pw // 2)
    near, far = ds.near_far_from_sphere(o[:, kc[0], kc[1]], d[:, kc[0], kc[1]])
    u_ref = torch.rand(near.shape, generator=g1, device=dev, dtype=near.dtype)
    near = near + u_ref * step
    mref = (m > 0.5).float() if mask_mode else torch.ones_like(m)
    # the kernel on the same random numbers
    g2 = torch.Generator(device=dev).manual_seed(5)
    cx = torch.randint(pw // 2, ds.W - 1 - pw // 2, (B,), device=dev, generator=g2)
    cy = torch.randint(ph // 2, ds.H - 1 - ph // 2, (B,), device=dev, generator=g2)
    sel = torch.randint(0, len(ds.train_images), (B,), device=dev, generator=g2)
    u = torch.rand((B,), generator=g2, device=dev, dtype=torch.float32)
    P = ph * pw
    f = dict(dtype=torch.float32, device=dev)
    out = dict(ro=torch.empty(B, ph, pw, 3, **f), rd=torch.empty(B, ph, pw, 3, **f),
               pn=torch.empty(B, 3, **f), vi=torch.empty(B, ph, pw, 3, 3, **f),
               nt=torch.empty(B, ph, pw, 3, **f), m=torch.empty(B, P, **f),
               ms=torch.empty(1, **f), oc=torch.empty(B, 3, **f), dc=torch.empty(B, 3, **f),
               t0=torch.empty(B, **f), t1=torch.empty(B, **f))
    k_inv = ds.intrinsics_all_inv.contiguous()  # torch.inverse returns column-major
    pose = ds.pose_all.contiguous()
    call("sn_patch_sample", ptr(cx), ptr(cy), ptr(sel), ptr(u), ptr(step), ptr(ds._train_idx),
         ptr(ds.normals), ptr(ds.V_inverse_all), ptr(ds.masks), ptr(k_inv), int(k_inv.shape[-1]),
         ptr(pose), int(ds.normals.shape[0]), int(ds._train_idx.numel()), ds.H, ds.W, B,
         ph, pw, mask_mode, *(ptr(out[k]) for k in ("ro", "rd", "pn", "vi", "nt", "m", "ms",
                                                    "oc", "dc", "t0", "t1")), stream_of(dev))
    torch.cuda.synchronize()
    assert torch.equal(out["vi"], vi) and torch.equal(out["nt"], nt)
    assert torch.equal(out["m"], mref.reshape(B, P))
    torch.testing.assert_close(out["ro"], o, rtol=0, atol=0)
    torch.testing.assert_close(out["pn"], pn, rtol=0, atol=0)
    torch.testing.assert_close(out["rd"], d, rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(out["oc"], o[:, kc[0], kc[1]], rtol=0, atol=0)
    torch.testing.assert_close(out["dc"], d[:, kc[0], kc[1]], rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(out["t0"], near, rtol=1e-5, atol=1e-5)
    torch.testing.assert_close(out["t1"], far, rtol=1e-5, atol=1e-5)
    ms = float(mref.sum()) + 1e-5
    assert abs(float(out["ms"]) - ms) <= 1e-6 * ms
