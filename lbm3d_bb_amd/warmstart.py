"""Learned warm-start: the U-Net of include/lbm_unet.h and its use as a start state.

``UNet`` wraps the C ABI (training and inference run on this library's BF16
MFMA kernels).  ``warm_start(solver, net, mask, re, ...)`` predicts (rho, u) and
hands it to ``lbm_init_equilibrium``; the solver's own drag residual
(``DragMonitor``) still decides convergence.  The reference has no learned
component: it always starts from a uniform equilibrium
(src/Initialization.jl:57-83).
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np

from . import _ffi


class UNetConfig(ct.Structure):
    _fields_ = [("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32), ("batch", ct.c_int32),
                ("cin", ct.c_int32), ("cout", ct.c_int32), ("base", ct.c_int32), ("levels", ct.c_int32),
                ("seed", ct.c_uint64), ("lr", ct.c_float), ("beta1", ct.c_float), ("beta2", ct.c_float),
                ("eps", ct.c_float)]


_F = ct.POINTER(ct.c_float)
_SIGS = {
    "lbm_unet_create": [ct.POINTER(UNetConfig), ct.POINTER(ct.c_void_p)],
    "lbm_unet_destroy": [ct.c_void_p],
    "lbm_unet_last_error": [ct.c_void_p, ct.c_char_p, ct.c_size_t],
    "lbm_unet_num_params": [ct.c_void_p, ct.POINTER(ct.c_int64)],
    "lbm_unet_get_params": [ct.c_void_p, _F],
    "lbm_unet_set_params": [ct.c_void_p, _F],
    "lbm_unet_param_layout": [ct.c_void_p, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32),
                              ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)],
    "lbm_unet_forward": [ct.c_void_p, _F, ct.c_int32, _F],
    "lbm_unet_train_step": [ct.c_void_p, _F, _F, ct.c_int32, _F],
    "lbm_unet_gradient": [ct.c_void_p, _F, _F, ct.c_int32, _F, _F],
    "lbm_unet_last_timing": [ct.c_void_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)],
}
UNET_SYMBOLS = list(_SIGS)


def _lib():
    lib = _ffi.load_library()
    if not getattr(lib, "_unet_bound", False):
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ct.c_int
        lib._unet_bound = True
    return lib


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_F)


@dataclass
class Layer:
    kind: int  # 0 conv3, 1 up2, 2 head
    cin: int
    cout: int
    w_off: int
    b_off: int


class UNet:
    def __init__(self, shape, batch=1, cin=2, cout=4, base=16, levels=4, seed=0, lr=1e-3):
        nz, ny, nx = shape
        self.shape = (nz, ny, nx)
        self.cin, self.cout, self.batch = cin, cout, batch
        cfg = UNetConfig(nx, ny, nz, batch, cin, cout, base, levels, seed, lr, 0.9, 0.999, 1e-8)
        self.lib = _lib()
        self._h = ct.c_void_p()
        rc = self.lib.lbm_unet_create(ct.byref(cfg), ct.byref(self._h))
        if rc != 0:
            raise _ffi.LbmError(rc, "lbm_unet_create failed")
        n = ct.c_int64(0)
        self.lib.lbm_unet_num_params(self._h, ct.byref(n))
        self.nparams = n.value

    def _chk(self, rc):
        if rc != 0:
            buf = ct.create_string_buffer(512)
            self.lib.lbm_unet_last_error(self._h, buf, 512)
            raise _ffi.LbmError(rc, buf.value.decode(errors="replace"))

    def close(self):
        if self._h:
            self.lib.lbm_unet_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layers(self) -> list[Layer]:
        nl = ct.c_int32(0)
        self._chk(self.lib.lbm_unet_param_layout(self._h, ct.byref(nl), None, None, None, None, None))
        k = (ct.c_int32 * nl.value)()
        ci = (ct.c_int32 * nl.value)()
        co = (ct.c_int32 * nl.value)()
        wo = (ct.c_int64 * nl.value)()
        bo = (ct.c_int64 * nl.value)()
        self._chk(self.lib.lbm_unet_param_layout(self._h, ct.byref(nl), k, ci, co, wo, bo))
        return [Layer(k[i], ci[i], co[i], wo[i], bo[i]) for i in range(nl.value)]

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, dtype=np.float32)
        self._chk(self.lib.lbm_unet_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(self.lib.lbm_unet_set_params(self._h, _fp(a)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        b = x.shape[0]
        y = np.empty((b, self.cout) + self.shape, dtype=np.float32)
        self._chk(self.lib.lbm_unet_forward(self._h, _fp(x), b, _fp(y)))
        return y

    def train_step(self, x: np.ndarray, target: np.ndarray) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = ct.c_float(0.0)
        self._chk(self.lib.lbm_unet_train_step(self._h, _fp(x), _fp(t), x.shape[0], ct.byref(loss)))
        return loss.value

    def gradient(self, x: np.ndarray, target: np.ndarray) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = ct.c_float(0.0)
        g = np.empty(self.nparams, dtype=np.float32)
        self._chk(self.lib.lbm_unet_gradient(self._h, _fp(x), _fp(t), x.shape[0], ct.byref(loss), _fp(g)))
        return loss.value, g

    def last_timing(self) -> tuple[float, float]:
        ms, fl = ct.c_double(0.0), ct.c_double(0.0)
        self._chk(self.lib.lbm_unet_last_timing(self._h, ct.byref(ms), ct.byref(fl)))
        return ms.value, fl.value


# --------------------------------------------------------------------------
# Network inputs/outputs for the flow problem
# --------------------------------------------------------------------------
RE_SCALE = 100.0


def encode_input(mask: np.ndarray, re: float) -> np.ndarray:
    """(mask, Re/100) channels, shape (2, nz, ny, nx)."""
    m = (np.asarray(mask) != 0).astype(np.float32)
    return np.stack([m, np.full_like(m, re / RE_SCALE)])


def encode_target(rho: np.ndarray, u: np.ndarray, u_ref: float, shape) -> np.ndarray:
    """((rho - 1) / u_ref, u / u_ref): O(1) channels, shape (4, nz, ny, nx)."""
    out = np.empty((4,) + tuple(shape), dtype=np.float32)
    out[0] = ((rho - 1.0) / u_ref).reshape(shape)
    for k in range(3):
        out[1 + k] = (u[k] / u_ref).reshape(shape)
    return out


def decode_output(y: np.ndarray, u_ref: float) -> tuple[np.ndarray, np.ndarray]:
    rho = 1.0 + u_ref * y[0].reshape(-1).astype(np.float64)
    u = u_ref * y[1:4].reshape(3, -1).astype(np.float64)
    return rho, u
