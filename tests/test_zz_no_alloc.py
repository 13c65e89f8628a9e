"""sn_pack_info_noalloc (scratch in a fixed device array, no allocation)
against sn_pack_info (workspace from the caller) through the C ABI: two calls
in a row, the second with fewer rays and samples than the first, so stale
counts left in the shared scratch by the first call would show; plus the
size limit."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _with_workspace(lib, ri, n_rays, dev, s):
    out = torch.full((n_rays, 2), -7, dtype=torch.int32, device=dev)
    bad = torch.full((1,), -7, dtype=torch.int32, device=dev)
    ws = torch.empty(max(int(lib.sn_pack_info_workspace_bytes(n_rays)), 1), dtype=torch.uint8,
                     device=dev)
    rc = lib.sn_pack_info(ri.data_ptr(), ri.numel(), n_rays, out.data_ptr(), bad.data_ptr(),
                          ws.data_ptr(), ws.numel(), s)
    assert rc == 0, lib.sn_last_error()
    return out, bad


def _no_alloc(lib, ri, n_rays, dev, s):
    out = torch.full((n_rays, 2), -7, dtype=torch.int32, device=dev)
    bad = torch.full((1,), -7, dtype=torch.int32, device=dev)
    rc = lib.sn_pack_info_noalloc(ri.data_ptr(), ri.numel(), n_rays, out.data_ptr(),
                                  bad.data_ptr(), s)
    assert rc == 0, lib.sn_last_error()
    return out, bad


def test_noalloc_matches_workspace_entry_twice(dev):
    from supernormal_amd import _lib
    lib = _lib.load()
    s = torch.cuda.current_stream(dev).cuda_stream
    rng = np.random.default_rng(5)
    for n_rays, n in ((300000, 2000000), (1000, 7000)):
        ri_np = np.sort(rng.integers(-3, n_rays + 3, n))  # a few out-of-range indices
        ri = torch.from_numpy(ri_np).to(dev)
        a, abad = _with_workspace(lib, ri, n_rays, dev, s)
        b, bbad = _no_alloc(lib, ri, n_rays, dev, s)
        torch.cuda.synchronize()
        print(f"n_rays {n_rays}, samples {n}: bad {int(bbad)}", flush=True)
        assert torch.equal(a, b)
        assert int(abad) == int(bbad) == int(((ri_np < 0) | (ri_np >= n_rays)).sum())
        # and the counts themselves (bincount of the valid indices)
        valid = ri_np[(ri_np >= 0) & (ri_np < n_rays)]
        cnt = np.bincount(valid, minlength=n_rays)
        np.testing.assert_array_equal(b[:, 1].cpu().numpy(), cnt)
        np.testing.assert_array_equal(b[:, 0].cpu().numpy(),
                                      np.concatenate([[0], np.cumsum(cnt)[:-1]]))


def test_noalloc_rejects_beyond_limit(dev):
    from supernormal_amd import _lib
    lib = _lib.load()
    lim = int(lib.sn_pack_info_noalloc_max_rays())
    assert lim == 1 << 22
    s = torch.cuda.current_stream(dev).cuda_stream
    ri = torch.zeros(4, dtype=torch.int64, device=dev)
    out = torch.zeros((8, 2), dtype=torch.int32, device=dev)
    rc = lib.sn_pack_info_noalloc(ri.data_ptr(), 4, lim + 1, out.data_ptr(), None, s)
    assert rc == -1
    assert b"scratch limit" in lib.sn_last_error()
