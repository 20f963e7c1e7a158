"""Build recipe for the gfx950 shared library ``liblbm_hip.so`` (in-tree).

``python -m lbm3d_bb_amd.build`` compiles ``csrc/lbm_capi.hip`` with hipcc for
gfx950 only.  Packed FP32 instructions are disabled
(``-target-feature -packed-fp32-ops``): hipcc would otherwise emit them from
ordinary float code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "liblbm_hip.so")
REPO = os.path.dirname(PKG_DIR)

SOURCES = ["lbm_capi.hip", "lbm_geom.hip", "lbm_store_f64.hip", "lbm_store_f32.hip", "lbm_store_f16s.hip",
           "lbm_store_bf16s.hip", "lbm_store_f64_ibb.hip", "lbm_store_f32_ibb.hip", "lbm_store_f16s_ibb.hip",
           "lbm_store_bf16s_ibb.hip", "lbm_q27.hip", "lbm_q27_ibb.hip", "lbm_q27_16.hip",
           "lbm_q27_16_ibb.hip", "unet_capi.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-function",
    "-I", os.path.join(REPO, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _headers_mtime() -> float:
    heads = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(REPO, "include")
    heads += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return max(os.path.getmtime(h) for h in heads)


def build_lib(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles each translation unit in parallel, then links liblbm_hip.so.

    incremental: skip a unit whose object is newer than its source and every
    header (a development shortcut; the default rebuilds everything)."""
    import concurrent.futures as cf

    obj_dir = os.path.join(PKG_DIR, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    cc = hipcc()
    hdr_t = _headers_mtime()

    def compile_one(src: str) -> str:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        if incremental and os.path.exists(obj):
            t = os.path.getmtime(obj)
            if t > hdr_t and t > os.path.getmtime(os.path.join(CSRC, src)):
                return obj
        cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src} ({res.returncode}):\n{res.stdout}\n{res.stderr}")
        return obj

    jobs = min(len(SOURCES), max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_lib(verbose=True, incremental="--incremental" in sys.argv[1:]))
