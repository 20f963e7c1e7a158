"""The reference's HDF5 outputs, written through the HDF5 C library (ctypes).

* ``final_fields.h5`` (write_fields_h5_time!, src/I_O.jl:296-405): datasets
  ``rho``, ``u``, ``v``, ``w`` (float32) of Julia shape (Nx, Ny, Nz, num_snaps)
  and ``steps`` (int32, num_snaps); snapshot s (1-based) holds step s * n_save.
* ``aero_coeffs.h5`` (save_coefficients!, src/I_O.jl:587-662): datasets ``CL``,
  ``CD``, ``CZ`` (float32, n_max), written block by block.

HDF5.jl stores a Julia array of size (Nx, Ny, Nz, S) as a dataset of
dimensions (S, Nz, Ny, Nx) with x fastest, so the bytes are those of the
column-major Julia array; this module writes the same dimensions and byte
order.  Unwritten elements read as the HDF5 fill value 0, the zeros the
reference creates its datasets with.

The library is located through LBM_HDF5_LIB, the dynamic loader, or the usual
install paths; without it the writers raise ImportError.
"""
from __future__ import annotations

import ctypes as ct
import ctypes.util
import glob
import os

import numpy as np

_hid = ct.c_int64
_herr = ct.c_int
_hsize = ct.c_ulonglong
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SELECT_SET = 0
H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0x0000, 0x0001, 0x0002
H5T_INTEGER, H5T_FLOAT = 0, 1

_LIB = None


def _find() -> str:
    env = os.environ.get("LBM_HDF5_LIB")
    if env and os.path.exists(env):
        return env
    name = ctypes.util.find_library("hdf5")
    if name:
        return name
    for pat in ("/opt/conda/lib/libhdf5.so*", "/usr/lib/x86_64-linux-gnu/libhdf5*.so*", "/usr/lib64/libhdf5.so*",
                "/usr/local/lib/libhdf5.so*"):
        hits = sorted(p for p in glob.glob(pat) if "_hl" not in p and "fortran" not in p and "cpp" not in p)
        if hits:
            return hits[0]
    raise ImportError("the HDF5 C library was not found (set LBM_HDF5_LIB)")


def lib() -> ct.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    h = ct.CDLL(_find())
    sig = {
        "H5open": ([], _herr),
        "H5Fcreate": ([ct.c_char_p, ct.c_uint, _hid, _hid], _hid),
        "H5Fopen": ([ct.c_char_p, ct.c_uint, _hid], _hid),
        "H5Fclose": ([_hid], _herr),
        "H5Screate_simple": ([ct.c_int, ct.POINTER(_hsize), ct.POINTER(_hsize)], _hid),
        "H5Sclose": ([_hid], _herr),
        "H5Sselect_hyperslab": ([_hid, ct.c_int, ct.POINTER(_hsize), ct.POINTER(_hsize), ct.POINTER(_hsize),
                                 ct.POINTER(_hsize)], _herr),
        "H5Sget_simple_extent_ndims": ([_hid], ct.c_int),
        "H5Sget_simple_extent_dims": ([_hid, ct.POINTER(_hsize), ct.POINTER(_hsize)], ct.c_int),
        "H5Dcreate2": ([_hid, ct.c_char_p, _hid, _hid, _hid, _hid, _hid], _hid),
        "H5Dopen2": ([_hid, ct.c_char_p, _hid], _hid),
        "H5Dclose": ([_hid], _herr),
        "H5Dwrite": ([_hid, _hid, _hid, _hid, _hid, ct.c_void_p], _herr),
        "H5Dread": ([_hid, _hid, _hid, _hid, _hid, ct.c_void_p], _herr),
        "H5Dget_space": ([_hid], _hid),
        "H5Dget_type": ([_hid], _hid),
        "H5Tget_class": ([_hid], ct.c_int),
        "H5Tget_size": ([_hid], ct.c_size_t),
        "H5Tclose": ([_hid], _herr),
    }
    for name, (args, res) in sig.items():
        fn = getattr(h, name)
        fn.argtypes = args
        fn.restype = res
    if h.H5open() < 0:
        raise ImportError("H5open failed")
    _LIB = h
    return h


def _native(name: str) -> int:
    return _hid.in_dll(lib(), name).value


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise OSError(f"HDF5: {what} failed ({rc})")
    return rc


def _dims(shape) -> ct.Array:
    return (_hsize * len(shape))(*[int(v) for v in shape])


class H5File:
    """A file opened for writing ("w": truncate) or reading and writing ("r+", "r")."""

    def __init__(self, path: str, mode: str = "r"):
        h = lib()
        p = os.fsencode(path)
        if mode == "w":
            self.fid = _check(h.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"create {path}")
        else:
            flags = H5F_ACC_RDWR if mode == "r+" else H5F_ACC_RDONLY
            self.fid = _check(h.H5Fopen(p, flags, H5P_DEFAULT), f"open {path}")

    def close(self) -> None:
        if self.fid is not None:
            lib().H5Fclose(self.fid)
            self.fid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def create(self, name: str, dtype, shape) -> None:
        """A dataset of the given C-order shape (fill value 0)."""
        h = lib()
        t = _native("H5T_NATIVE_FLOAT_g") if np.dtype(dtype) == np.float32 else _native("H5T_NATIVE_INT_g")
        if np.dtype(dtype) not in (np.dtype(np.float32), np.dtype(np.int32)):
            raise TypeError("only float32 and int32 datasets")
        sp = _check(h.H5Screate_simple(len(shape), _dims(shape), None), "dataspace")
        try:
            d = _check(h.H5Dcreate2(self.fid, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                       f"create dataset {name}")
            h.H5Dclose(d)
        finally:
            h.H5Sclose(sp)

    def _slab(self, name: str, start, count, buf: np.ndarray, write: bool) -> None:
        h = lib()
        d = _check(h.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT), f"open dataset {name}")
        fsp = h.H5Dget_space(d)
        msp = _check(h.H5Screate_simple(len(count), _dims(count), None), "memory space")
        try:
            _check(h.H5Sselect_hyperslab(fsp, H5S_SELECT_SET, _dims(start), None, _dims(count), None), "hyperslab")
            t = _native("H5T_NATIVE_FLOAT_g") if buf.dtype == np.float32 else _native("H5T_NATIVE_INT_g")
            ptr = buf.ctypes.data_as(ct.c_void_p)
            if write:
                _check(h.H5Dwrite(d, t, msp, fsp, H5P_DEFAULT, ptr), f"write {name}")
            else:
                _check(h.H5Dread(d, t, msp, fsp, H5P_DEFAULT, ptr), f"read {name}")
        finally:
            h.H5Sclose(msp)
            h.H5Sclose(fsp)
            h.H5Dclose(d)

    def write_slab(self, name: str, start, data: np.ndarray) -> None:
        a = np.ascontiguousarray(data)
        self._slab(name, start, a.shape, a, True)

    def info(self, name: str) -> tuple[tuple[int, ...], str, int]:
        """(C-order dimensions, "float" | "integer" | other, element size) of a dataset."""
        h = lib()
        d = _check(h.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT), f"open dataset {name}")
        sp = h.H5Dget_space(d)
        t = h.H5Dget_type(d)
        try:
            n = h.H5Sget_simple_extent_ndims(sp)
            dims = (_hsize * max(n, 1))()
            h.H5Sget_simple_extent_dims(sp, dims, None)
            cls = h.H5Tget_class(t)
            kind = {H5T_FLOAT: "float", H5T_INTEGER: "integer"}.get(cls, str(cls))
            return tuple(int(dims[i]) for i in range(n)), kind, int(h.H5Tget_size(t))
        finally:
            h.H5Tclose(t)
            h.H5Sclose(sp)
            h.H5Dclose(d)

    def read(self, name: str) -> np.ndarray:
        dims, kind, size = self.info(name)
        out = np.empty(dims, dtype=np.float32 if kind == "float" else np.int32)
        self._slab(name, (0,) * len(dims), dims, out, False)
        return out


class FieldSnapshots:
    """final_fields.h5: snapshot s (1-based) of rho, u, v, w as float32 and its step."""

    def __init__(self, path: str, nx: int, ny: int, nz: int, num_snaps: int):
        self.path, self.shape, self.num_snaps = path, (nz, ny, nx), num_snaps
        with H5File(path, "w") as f:
            for name in ("rho", "u", "v", "w"):
                f.create(name, np.float32, (num_snaps, nz, ny, nx))
            f.create("steps", np.int32, (num_snaps,))

    def write(self, snap: int, step: int, rho: np.ndarray, u: np.ndarray) -> None:
        """rho (N), u (3, N) in cell order x fastest; snap is 1-based."""
        nz, ny, nx = self.shape
        with H5File(self.path, "r+") as f:
            f.write_slab("rho", (snap - 1, 0, 0, 0), np.asarray(rho, np.float32).reshape(1, nz, ny, nx))
            for k, name in enumerate(("u", "v", "w")):
                f.write_slab(name, (snap - 1, 0, 0, 0), np.asarray(u[k], np.float32).reshape(1, nz, ny, nx))
            f.write_slab("steps", (snap - 1,), np.array([step], np.int32))


class AeroCoefficients:
    """aero_coeffs.h5: CL, CD, CZ per step (float32, n_max), written in blocks."""

    def __init__(self, path: str, n_max: int):
        self.path, self.n_max = path, n_max
        with H5File(path, "w") as f:
            for name in ("CL", "CD", "CZ"):
                f.create(name, np.float32, (n_max,))

    def write(self, first_step: int, cl, cd, cz) -> None:
        """Coefficients of steps first_step .. first_step + len - 1 (1-based)."""
        with H5File(self.path, "r+") as f:
            for name, v in (("CL", cl), ("CD", cd), ("CZ", cz)):
                f.write_slab(name, (first_step - 1,), np.asarray(v, np.float32))
