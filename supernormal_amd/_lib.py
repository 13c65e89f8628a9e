"""ctypes binding of libsupernormal_amd.so (the C ABI declared in include/sn_api.h).

The library is the only compute path of this package: if it is missing the
import of any op fails with a RuntimeError telling how to build it.  There is
no CPU or pure-PyTorch fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("SN_LIB_PATH", _PKG / "_lib" / "libsupernormal_amd.so"))

_i32, _i64, _f32, _sz = C.c_int32, C.c_int64, C.c_float, C.c_size_t
_p = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "sn_last_error": (C.c_char_p, []),
    "sn_abi_version": (C.c_int, []),
    "sn_pack_info_workspace_bytes": (_sz, [_i32]),
    "sn_pack_info": (C.c_int, [_p, _i64, _i32, _p, _p, _p, _sz, _p]),
    "sn_pack_info_noalloc_max_rays": (_i64, []),
    "sn_pack_info_noalloc": (C.c_int, [_p, _i64, _i32, _p, _p, _p]),
    "sn_scan_workspace_bytes": (_sz, [_i64]),
    "sn_fill_span_dev": (C.c_int, [_p, _i64, _p, _i32, _p, _i32, _f32, _p]),
    "sn_counts_to_packed_info": (C.c_int, [_p, _i32, _p, _p, _p, _sz, _p]),
    "sn_patch_weight_fwd": (C.c_int, [_p, _i32, _i32, _i64, _p, _p, _p, _p]),
    "sn_patch_weight_bwd": (C.c_int, [_p, _i32, _i32, _i64, _p, _p, _p, _p, _p]),
    "sn_patch_transmittance_fwd": (C.c_int, [_p, _i32, _i32, _i64, _p, _p, _p]),
    "sn_visibility": (C.c_int, [_p, _i32, _i64, _p, _f32, _f32, _p, _p]),
    "sn_compact_workspace_bytes": (_sz, [_i64]),
    "sn_compact_samples": (C.c_int, [_p, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "sn_ray_march_count": (C.c_int, [_p, _p, _p, _p, _i32, _p, _p, _i32, _i32, _i32, _f32,
                                     _p, _f32, _i32, _p, _p]),
    "sn_ray_march_fill": (C.c_int, [_p, _p, _p, _p, _i32, _p, _p, _i32, _i32, _i32, _f32,
                                    _p, _f32, _i32, _p, _p, _p, _p, _p]),
    "sn_scatter_add_rows": (C.c_int, [_p, _i64, _i32, _p, _p, _i32, _p]),
    "sn_occ_cell_points": (C.c_int, [_p, _i32, _p, _i32, _i32, _i32, _p, _p, _p]),
    "sn_occ_workspace_bytes": (_sz, [_i32]),
    "sn_occ_ema_max": (C.c_int, [_p, _i32, _p, _p, _i32, _f32, _p, _sz, _p]),
    "sn_occ_binarize": (C.c_int, [_p, _i32, _f32, _p, _p, _p, _sz, _p]),
    "sn_occ_nonzero": (C.c_int, [_p, _i32, _p, _p, _p, _sz, _p]),
    "sn_hashgrid_layout": (_i64, [_i32, _i32, _i32, _f32, _p]),
    "sn_hashgrid_set_aggregation": (None, [_i32]),
    "sn_hashgrid_fwd": (C.c_int, [_p, _i64, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p]),
    "sn_hashgrid_bwd": (C.c_int, [_p, _i64, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p,
                                  _p, _p]),
    "sn_hashgrid_bwd_bwd": (C.c_int, [_p, _i64, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p,
                                      _p, _p, _p, _p, _p]),
    "sn_mlp_n_params": (C.c_int, [_i32, _i32, _i32]),
    "sn_mlp_bwd_workspace_bytes": (_sz, [_i64, _i32, _i32, _i32, _i32]),
    "sn_mlp_fwd": (C.c_int, [_p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p, _p, _p]),
    "sn_mlp_bwd": (C.c_int, [_p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p, _p, _p, _p,
                             _i32, _p, _sz, _p]),
    "sn_mlp_bwd_ad": (C.c_int, [_p, _p, _p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p, _p,
                                _p, _p, _i32, _p, _sz, _p]),
    "sn_mlp_bwd_ad_enc": (C.c_int, [_p, _p, _p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p,
                                    _p, _p, _p, _i32, _p, _sz, _p]),
    "sn_mlp16_bwd_ad_enc": (C.c_int, [_p, _p, _p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p,
                                      _p, _p, _p, _i32, _p, _sz, _p]),
    "sn_render_step_waves": (C.c_int, [_i32, _i32]),
    "sn_render_step": (C.c_int, [_i32, _i32, _i32] + [_p] * 16 + [_f32, _f32, _f32, _i32]
                       + [_p] * 8),
    "sn_adam_step": (C.c_int, [_p, _p, _p, _p, _i64, _f32, _p, C.c_double, C.c_double,
                               C.c_double, _i32, _p, _p]),
    "sn_mlp16_fwd": (C.c_int, [_p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p, _p]),
    "sn_mlp16_fwd_grad": (C.c_int, [_p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p, _p, _p]),
    "sn_mlp16_bwd_workspace_bytes": (_sz, [_i64, _i32, _i32, _i32]),
    "sn_mlp16_bwd": (C.c_int, [_p, _p, _i32, _i64, _p, _p, _i32, _i32, _f32, _p, _p, _p, _p,
                               _i32, _p, _sz, _p]),
    "sn_render_step_set_impl": (None, [_i32]),
    "sn_ray_march_padded": (C.c_int, [_p, _p, _p, _p, _i32, _p, _p, _i32, _i32, _i32, _f32,
                                      _p, _f32, _i32, _p, _p, _p, _p]),
    "sn_ray_march_padded_bricks": (C.c_int, [_p, _p, _p, _p, _i32, _p, _p, _p, _i32, _i32,
                                             _i32, _f32, _p, _f32, _i32, _p, _p, _p, _p]),
    "sn_occ_pack_bricks": (C.c_int, [_p, _i32, _i32, _i32, _p, _p]),
    "sn_mlp_pack_fwd": (C.c_int, [_p, _i32, _p, _p, _p, _p, _p]),
    "sn_patch_sample": (C.c_int, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i32, _p, _i32, _i32, _i32,
                                  _i32, _i32, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _p,
                                  _p, _p, _p, _p]),
    "sn_mlp_pack_bwd": (C.c_int, [_p, _i32, _p, _p, _p, _p, _p, _p, _p]),
    "sn_march_gather": (C.c_int, [_p, _i32, _i32, _p, _p, _p, _p, _p, _p, _p]),
    "sn_mt_count": (C.c_int, [_p, _i32, _i32, _i32, _f32, _p, _p]),
    "sn_mt_emit": (C.c_int, [_p, _i32, _i32, _i32, _f32, _p, _p, _p, _p, _p]),
    "sn_mc_count": (C.c_int, [_p, _i32, _i32, _i32, _f32, _p, _p]),
    "sn_mc_emit": (C.c_int, [_p, _i32, _i32, _i32, _f32, _p, _p, _p, _p, _p, _p]),
    "sn_mesh_first_hit": (C.c_int, [_p, _p, _p, _p, _i32, _i32, _i32, _i32, _p, _p, _p, _p,
                                    _p, _p, _p]),
    "sn_adam_step_copy16": (C.c_int, [_p, _p, _p, _p, _i64, _f32, _p, C.c_double, C.c_double,
                                      C.c_double, _i32, _p, _p, _i64, _i64, _p, _p]),
    "sn_step_advance": (C.c_int, [_p, _p, _p, _p]),
    "sn_to_half": (C.c_int, [_p, _p, _i64, _p]),
    "sn_step_guard": (C.c_int, [_p, _p, _i32, _p, _i64, _p, _p, _p]),
    "sn_adam_multi_blocks": (_i32, [_i64]),
    "sn_adam_multi": (C.c_int, [_p, _i32, _i32, C.c_double, C.c_double, C.c_double, _p]),
    "sn_hashgrid_fwd16_dev": (C.c_int, [_p, _i64, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p, _p]),
    "sn_pack_info_dev": (C.c_int, [_p, _i64, _p, _i32, _p, _p, _p, _sz, _p]),
    "sn_compact_samples_dev": (C.c_int, [_p, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i32,
                                         _p, _p, _sz, _p]),
    "sn_step_workspace_bytes": (_sz, [_i64]),
    "sn_centre_points": (C.c_int, [_p, _p, _p, _i64, _p, _p, _p, _i32, _p, _p, _p, _p, _p,
                                   _sz, _p]),
    "sn_alpha_visibility": (C.c_int, [_p, _i32, _p, _p, _i64, _p, _p, _f32, _p, _p]),
    "sn_patch_points": (C.c_int, [_p, _p, _p, _i64, _p, _p, _p, _p, _i32, _i32, _p, _p, _p,
                                  _p, _p, _sz, _p]),
    "sn_hashgrid_fwd_dev": (C.c_int, [_p, _i64, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p, _p]),
    "sn_hashgrid_bwd_dev": (C.c_int, [_p, _i64, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p, _p, _p, _p]),
    "sn_hashgrid_dx_dev": (C.c_int, [_p, _i64, _p, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32,
                                     _p, _p, _i32, _p, _p]),
    "sn_hashgrid_bwd_workspace_bytes": (_i64, [_i32, _i32, _i32, _i32, _f32, _i32]),
    "sn_hashgrid_bwd_ws_dev": (C.c_int, [_p, _i64, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p, _p, _p, _p, _sz, _p]),
    "sn_hashgrid_bwd_bwd_dev": (C.c_int, [_p, _i64, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p, _p, _p, _p, _p, _p]),
    "sn_hashgrid_tangent_dev": (C.c_int, [_p, _i64, _p, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p, _p, _p]),
    "sn_hashgrid_bwd_ad_ws_dev": (C.c_int, [_p, _i64, _p, _p, _i32, _i32, _i32, _i32, _f32, _i32, _p, _p, _p, _p, _p, _p, _sz, _p]),
    "sn_occ_sample_workspace_bytes": (_sz, [_i32]),
    "sn_occ_sample_cells": (C.c_int, [_p, _i32, _p, _p, _p, _p, _p, _sz, _p]),
}

class SnAdamSegment(C.Structure):
    """One row of sn_adam_multi's segment table (include/sn_api.h)."""
    _fields_ = [("params", _p), ("grads", _p), ("exp_avg", _p), ("exp_avg_sq", _p),
                ("n", _i64), ("lr_dev", _p), ("step_dev", _p), ("run_dev", _p),
                ("copy16", _p), ("copy_off", _i64), ("copy_n", _i64), ("block0", _i32),
                ("n_blocks", _i32)]


_lib = None


class SNError(RuntimeError):
    """A failed call into libsupernormal_amd (message from sn_last_error).

    A RuntimeError, as the reference's TORCH_CHECK / CHECK_INPUT failures in
    its _C extension are (helpers_cuda.h:21-26): a bad argument (SN_EINVAL)
    and a failed HIP call (SN_EHIP) raise the same type; `code` tells them
    apart."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"libsupernormal_amd.so not found at {LIB_PATH}; build it with "
            "`python -m supernormal_amd.build` (requires hipcc / ROCm). "
            "There is no fallback implementation.")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib() -> C.CDLL:
    return load()


def check(rc: int, name: str) -> None:
    if rc != 0:
        msg = load().sn_last_error().decode(errors="replace")
        if rc == -1:
            raise SNError(f"{name}: {msg}", rc)
        raise SNError(f"{name} failed ({rc}): {msg}", rc)


class LaunchTimer:
    """Brackets every call of one entry point with HIP events on the current
    stream (the stream the kernels are enqueued on) and records the argument
    tuple, so a caller can compute per-launch bytes / duration."""

    def __init__(self, name: str):
        self.name = name
        self.records = []  # (start_event, end_event, args)

    def durations_ms(self):
        torch.cuda.synchronize()
        return [a.elapsed_time(b) for a, b, _ in self.records]


_timers = {}


def time_calls(name: str) -> LaunchTimer:
    t = LaunchTimer(name)
    _timers[name] = t
    return t


def stop_timing(name: str) -> None:
    _timers.pop(name, None)


def call(name: str, *args) -> None:
    t = _timers.get(name)
    if t is None:
        check(getattr(load(), name)(*args), name)
        return
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    check(getattr(load(), name)(*args), name)
    e1.record()
    t.records.append((e0, e1, args))


def ptr(t) -> int | None:
    """Device address of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_of(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_cuda(*tensors, name: str = "op") -> None:
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise NotImplementedError(f"{name}: only GPU tensors are supported")


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
