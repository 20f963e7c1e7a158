"""The C ABI library loads and exports every entry point include/lbm_hip.h declares
(no compute calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes as ct
import os
import re

from lbm3d_bb_amd import _ffi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(REPO, "include")


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


def test_struct_layout_matches_header(tmp_path):
    """Every ctypes field offset and struct size equals the C compiler's for
    include/lbm_hip.h (gcc on a generated probe)."""
    import shutil
    import subprocess

    import pytest

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"lbm_face": _ffi.LbmFace, "lbm_config": _ffi.LbmConfig, "lbm_body": _ffi.LbmBody,
               "lbm_motion": _ffi.LbmMotion}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "lbm_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line.strip()}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ct.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_missing_library_fails_loudly(tmp_path):
    import pytest

    with pytest.raises(ImportError):
        _ffi.load_library(str(tmp_path / "nope.so"))


def test_lattice_other_than_19_or_27_rejected():
    """lbm_create refuses lattices other than 19 / 27 before touching a device; the
    host layer names the lattice as the reference does."""
    import pytest

    from lbm3d_bb_amd.solver import Config, periodic

    c = Config(8, 8, 8, collision="BGK", storage="fp32", omega_plus=1.2, omega_minus=1.2, faces=[periodic()] * 6)
    c.q = 23
    with pytest.raises(ValueError):
        c.to_c()
    c.q = 27
    raw = c.to_c()
    raw.q = 23
    ctx = ct.c_void_p()
    assert _ffi.load_library().lbm_create(ct.byref(raw), ct.byref(ctx)) == -1  # LBM_ERR_ARG
