"""The C-ABI library loads and exports every entry point declared in
include/pyflow_hip.h (no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'pyflow_hip.h')


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(pf_[a-z0-9_]+)\s*\(', txt)))


def test_header_declares_entry_points():
    names = declared()
    for n in ('pf_box_create', 'pf_rk4_step', 'pf_bicgstab_solve', 'pf_vel_corr', 'pf_stats'):
        assert n in names


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (load torch's HIP runtime first, as the package does)
    from pyflow_amd import _build, _native
    if not os.path.exists(_build.LIB_PATH):
        pytest.skip('library not built')
    lib = ctypes.CDLL(_build.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers all of them
    bound = {n for n, _, _ in _native.SIGNATURES}
    assert set(declared()) <= bound, set(declared()) - bound


def test_abi_version_and_error_path_without_gpu():
    import torch  # noqa: F401
    from pyflow_amd import _native
    L = _native.lib()
    assert L.pf_abi_version() == 1
    # invalid arguments are rejected before touching the device
    h = ctypes.c_void_p()
    rc = L.pf_box_create(ctypes.byref(h), 2, 8, 8, 1.0, 1.0, 1.0, 0)
    assert rc < 0 and 'dimension' in _native.last_error()
    assert L.pf_fill_halo(None, None, None) < 0


def test_noalloc_rk4_rejects_oversized_box_before_any_gpu_work():
    """pf_rk4_step_noalloc checks its scratch limit before its first HIP call
    (the pointers below are never dereferenced)."""
    import torch  # noqa: F401
    from pyflow_amd import _native
    L = _native.lib()
    cap = int(L.pf_rk4_step_noalloc_max_bytes())
    assert cap == 64 << 20
    fake = [ctypes.c_void_p(4096 * (q + 1)) for q in range(6)]
    # 128^3 FP64 needs 3*(128^3 + 132^3)*8 bytes, more than the scratch holds
    rc = L.pf_rk4_step_noalloc(128, 128, 128, 0.1, 0.1, 0.1, _native.PF_F64, *fake,
                               1e-3, 0.01, 1.0, None, None)
    assert rc == -1 and 'bytes of scratch' in _native.last_error()
    rc = L.pf_rk4_step_noalloc(3, 8, 8, 0.1, 0.1, 0.1, _native.PF_F64, *fake,
                               1e-3, 0.01, 1.0, None, None)
    assert rc == -1 and 'dimension' in _native.last_error()
    rc = L.pf_rk4_step_noalloc(8, 8, 8, 0.1, 0.1, 0.1, _native.PF_F64, *fake[:5], None,
                               1e-3, 0.01, 1.0, None, None)
    assert rc == -1 and 'NULL' in _native.last_error()
