"""The C ABI library loads and exports every entry point include/lbm_hip.h declares
(no compute calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes as ct
import os
import re

from lbm3d_bb_amd import _ffi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in os.listdir(os.path.join(REPO, "include")):
        if fn.endswith(".h"):
            with open(os.path.join(REPO, "include", fn), encoding="utf-8") as fh:
                names += re.findall(r"^int\s+(lbm_\w+)\s*\(", fh.read(), flags=re.M)
    return names


def test_header_declares_entry_points():
    from lbm3d_bb_amd.warmstart import UNET_SYMBOLS

    names = _declared()
    assert "lbm_step" in names and "lbm_create" in names and "lbm_unet_train_step" in names
    assert sorted(names) == sorted(_ffi.exported_symbols() + UNET_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(_ffi.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_struct_layout_matches_header_sizes():
    # lbm_face: 8 + 24; lbm_config: 8 ints (32) + 2 doubles + g[3] + rho0 + u0[3] (72) + 6 faces (192)
    assert ct.sizeof(_ffi.LbmFace) == 32
    assert ct.sizeof(_ffi.LbmConfig) == 32 + 72 + 6 * 32
    # lbm_body: 8 + 3f + 3f + 9f + 1f (64) + 9 doubles (72)
    assert ct.sizeof(_ffi.LbmBody) == 8 + 64 + 72
    assert ct.sizeof(_ffi.LbmMotion) == 8 + 8 * (3 + 3 + 9 + 3 + 3)


def test_missing_library_fails_loudly(tmp_path):
    import pytest

    with pytest.raises(ImportError):
        _ffi.load_library(str(tmp_path / "nope.so"))
