"""The C ABI library loads without a GPU and exports every symbol the headers declare."""
import ctypes
import re

import pytest

from conftest import ROOT


def _declared_functions():
    names = []
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(sn_[a-z0-9_]+)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_functions():
    names = _declared_functions()
    assert "sn_patch_weight_fwd" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from supernormal_amd import _lib
    lib = _lib.load()
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.sn_abi_version() >= 1


def test_python_signatures_cover_header():
    from supernormal_amd import _lib
    missing = [n for n in _declared_functions() if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_error_reporting_without_gpu():
    """Argument validation happens before any device work, so it runs on CPU."""
    from supernormal_amd import _lib
    lib = _lib.load()
    rc = lib.sn_patch_weight_fwd(None, -1, 9, 0, None, None, None, None)
    assert rc == -1
    assert b"bad sizes" in lib.sn_last_error()
    rc = lib.sn_ray_march_count(None, None, None, None, 4, None, None, 8, 8, 8,
                                ctypes.c_float(0.0), None, ctypes.c_float(0.0), 0, None, None)
    assert rc == -1 and b"render_step_size" in lib.sn_last_error()
    # a bad argument raises RuntimeError, as the reference's TORCH_CHECK does
    with pytest.raises(RuntimeError) as e:
        _lib.check(lib.sn_hashgrid_fwd(None, 10, None, 0, 2, 19, 16, ctypes.c_float(1.5), 0,
                                       None, None), "sn_hashgrid_fwd")
    assert e.value.code == -1
