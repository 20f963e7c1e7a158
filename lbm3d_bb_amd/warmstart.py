"""Learned warm-start: the U-Net of include/lbm_unet.h and its use as a start state.

``UNet`` wraps the C ABI (training and inference run on this library's BF16
MFMA kernels).  ``warm_start(solver, net, mask, re, ...)`` predicts (rho, u) and
hands it to ``lbm_init_equilibrium``; the solver's own drag residual
(``DragMonitor``) still decides convergence.  ``warm_start_tiled`` does the
same on lattices larger than the network by running it on overlapping tiles
and keeping each tile's core.  The reference has no learned
component: it always starts from a uniform equilibrium
(src/Initialization.jl:57-83).
"""
from __future__ import annotations

import ctypes as ct
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _ffi


class UNetConfig(ct.Structure):
    _fields_ = [("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32), ("batch", ct.c_int32),
                ("cin", ct.c_int32), ("cout", ct.c_int32), ("base", ct.c_int32), ("levels", ct.c_int32),
                ("seed", ct.c_uint64), ("lr", ct.c_float), ("beta1", ct.c_float), ("beta2", ct.c_float),
                ("eps", ct.c_float), ("periodic", ct.c_int32)]


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
    "lbm_unet_set_lr": [ct.c_void_p, ct.c_float],
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
    def __init__(self, shape, batch=1, cin=2, cout=4, base=16, levels=4, seed=0, lr=1e-3, periodic=0):
        """periodic: bit 0 x, bit 1 y, bit 2 z -- axes along which the 3x3x3
        convolutions wrap around instead of padding with zeros."""
        nz, ny, nx = shape
        self.shape = (nz, ny, nx)
        self.cin, self.cout, self.batch, self.periodic = cin, cout, batch, periodic
        cfg = UNetConfig(nx, ny, nz, batch, cin, cout, base, levels, seed, lr, 0.9, 0.999, 1e-8, periodic)
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
        if x.ndim != 5 or x.shape[1:] != (self.cin,) + self.shape or x.shape[0] > self.batch:
            raise ValueError(f"input of shape {x.shape} for a network of {self.cin} channels on {self.shape}, "
                             f"batch <= {self.batch}")
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


    def set_lr(self, lr: float) -> None:
        self._chk(self.lib.lbm_unet_set_lr(self._h, float(lr)))

    def save(self, path: str) -> None:
        np.save(path, self.get_params())

    def load(self, path: str) -> None:
        self.set_params(np.load(path, allow_pickle=False))


# --------------------------------------------------------------------------
# Network inputs/outputs for the flow problem
# --------------------------------------------------------------------------
LOG_RE_MAX = float(np.log(100.0))

# The pack network: 4 levels of 16/32/64/128 channels; inputs solid mask,
# ln(Re)/ln(100) and the position along the channel x/(nx - 1) (the inlet is
# at x = 0, the outlet at nx - 1: the pressure falls along x, and a cell far
# from both ends could not tell its distance to either from the mask alone);
# 3x3x3 convolutions wrap around in y and z like the packs.
PACK_NET = dict(cin=3, cout=4, base=16, levels=4, periodic=6)


# Parameters of the pack network trained at 128^3 (1042 packs, seeds 0-1041,
# 24 epochs; stored in FP16, the kernels round them to BF16 anyway).
PACK_PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_pack128.npy")


def pack_net(shape, batch: int = 1, seed: int = 0, lr: float = 1e-3, trained: bool = False) -> "UNet":
    """The pack network; with trained=True it holds the committed 128^3
    parameters (any shape the 4-level network accepts)."""
    net = UNet(shape, batch=batch, seed=seed, lr=lr, **PACK_NET)
    if trained:
        net.set_params(np.load(PACK_PARAMS, allow_pickle=False).astype(np.float32))
    return net


def encode_input(mask: np.ndarray, re: float, channels: int = 2) -> np.ndarray:
    """Network input, shape (channels, nz, ny, nx): the solid mask,
    ln(Re)/ln(100) and, for channels = 3, x/(nx - 1)."""
    m = (np.asarray(mask) != 0).astype(np.float32)
    ch = [m, np.full_like(m, np.log(re) / LOG_RE_MAX)]
    if channels >= 3:
        nx = m.shape[2]
        ch.append(np.broadcast_to(np.linspace(0.0, 1.0, nx, dtype=np.float32), m.shape).copy())
    if channels not in (2, 3):
        raise ValueError("the network input has 2 or 3 channels")
    return np.stack(ch)


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


def _grad(a: np.ndarray, axis: int, periodic: bool) -> np.ndarray:
    """Central difference along an axis of a (nz, ny, nx) array: wraps on a
    periodic axis, one-sided at the two ends otherwise."""
    if periodic:
        return 0.5 * (np.roll(a, -1, axis=axis) - np.roll(a, 1, axis=axis))
    return np.gradient(a, axis=axis)


def populations_from_fields(rho: np.ndarray, u: np.ndarray, shape, tau_plus: float,
                            periodic=(False, False, False)) -> np.ndarray:
    """f = f_eq(rho, u) + f_neq with the first-order Chapman-Enskog
    non-equilibrium part f_neq_i = -3 tau+ w_i rho Q_i : grad(u),
    Q_i = c_i c_i - I/3, gradients by central differences.  Starting from
    f_eq alone leaves the viscous stress of the field to build up again (a
    transient that the drag residual has to wait out); f_neq carries it.
    shape = (nz, ny, nx); rho (N,), u (3, N); returns (19, N)."""
    from .lattice import CX, CY, CZ, W, equilibrium

    nz, ny, nx = shape
    rho = np.asarray(rho, dtype=np.float64).reshape(-1)
    u = np.asarray(u, dtype=np.float64).reshape(3, -1)
    f = equilibrium(rho, u[0], u[1], u[2])
    U = u.reshape(3, nz, ny, nx)
    # grad[a][b] = d u_b / d x_a ; array axes: 0 = z, 1 = y, 2 = x
    ax_of = {0: 2, 1: 1, 2: 0}
    grad = [[_grad(U[b], ax_of[a], periodic[a]).reshape(-1) for b in range(3)] for a in range(3)]
    C = np.stack([CX, CY, CZ], axis=1).astype(np.float64)
    for d in range(19):
        qs = np.zeros_like(rho)
        for a in range(3):
            for b in range(3):
                q = C[d, a] * C[d, b] - (1.0 / 3.0 if a == b else 0.0)
                if q != 0.0:
                    qs += q * grad[a][b]
        f[d] += -3.0 * tau_plus * W[d] * rho * qs
    return f


# --------------------------------------------------------------------------
# Warm start with fallback, steady solves, data, training, evaluation
# --------------------------------------------------------------------------
def plausible(rho: np.ndarray, u: np.ndarray, u_ref: float) -> tuple[bool, str]:
    """A predicted field is used only if it is finite, its density stays in
    (0.5, 1.5) and its speed below max(0.3, 6 u_ref): anything else is a
    garbage prediction and the uniform start is taken instead."""
    if not (np.all(np.isfinite(rho)) and np.all(np.isfinite(u))):
        return False, "non-finite prediction"
    if rho.min() <= 0.5 or rho.max() >= 1.5:
        return False, f"density out of range [{rho.min():.3g}, {rho.max():.3g}]"
    vmax = float(np.sqrt((u * u).sum(axis=0)).max())
    if vmax >= max(0.3, 6.0 * u_ref):
        return False, f"speed {vmax:.3g} too large"
    return True, "ok"


def flux_correct(rho: np.ndarray, u: np.ndarray, shape, u_in: float, solid: np.ndarray) -> np.ndarray:
    """u with its x-component rescaled slice by slice so that every x-slice
    carries the inlet's mass flux u_in * sum_{x=0 slice} rho, as a steady flow
    through the channel must (continuity).  A predicted field whose slices
    carry too much or too little flux would otherwise start a channel-long
    pressure wave that the drag residual has to wait out.  rho (N,), u (3, N),
    solid (N,) bool; shape = (nz, ny, nx)."""
    nz, ny, nx = shape
    R = rho.reshape(nz, ny, nx)
    U = u.copy().reshape(3, nz, ny, nx)
    fl = ~solid.reshape(nz, ny, nx)
    target = u_in * R[:, :, 0].sum()
    flux = (R * U[0] * fl).sum(axis=(0, 1))
    scale = np.where(flux > 0.0, target / np.where(flux > 0.0, flux, 1.0), 1.0)
    U[0] *= scale[None, None, :]
    return U.reshape(3, -1)


def predict(net: UNet, case) -> tuple[np.ndarray, np.ndarray]:
    """(rho, u) predicted for a pack case; u is (3, N), solid cells at rest."""
    L = case.L
    mask = case.mask()
    y = net.forward(encode_input(mask, case.re, net.cin)[None])[0]
    rho, u = decode_output(y, case.u_in)
    solid = mask.reshape(-1)
    u[:, solid] = 0.0
    return rho, u


def tile_spans(n: int, t: int, halo: int) -> list[tuple[int, int, int]]:
    """Tiles of t cells covering an axis of n cells, as (start, lo, hi): the
    tile reads [start, start + t) and keeps its core [lo, hi).  The cores
    partition [0, n) in order; a core stays `halo` cells away from its tile's
    ends except where that end is the end of the lattice, and the tiles at the
    two ends are moved inward so that every tile lies inside the lattice."""
    if t > n:
        raise ValueError(f"tile of {t} cells is longer than the lattice axis of {n}")
    if halo < 0 or 2 * halo >= t:
        raise ValueError(f"halo {halo} leaves no core in a tile of {t} cells")
    out = []
    lo = 0
    while lo < n:
        start = min(max(lo - halo, 0), n - t)
        hi = n if start + t >= n else start + t - halo
        out.append((start, lo, hi))
        lo = hi
    return out


def predict_tiled(net, mask: np.ndarray, re: float, u_ref: float, halo: int = 16) -> tuple[np.ndarray, np.ndarray]:
    """(rho, u) of a lattice larger than the network: the network (fixed
    input shape ``net.shape`` = (tz, ty, tx), at most ``net.batch`` tiles per
    call) runs on overlapping tiles placed by tile_spans along each axis, and
    each tile contributes only its core, so that every kept cell sees at least
    `halo` cells of context on each side inside the lattice.  mask is the
    (nz, ny, nx) solid indicator; u is (3, N) with solid cells at rest."""
    mask = np.asarray(mask) != 0
    nz, ny, nx = mask.shape
    tz, ty, tx = net.shape
    spans = [tile_spans(n, t, halo) for n, t in zip((nz, ny, nx), (tz, ty, tx))]
    tiles = [(a, b, c) for a in spans[0] for b in spans[1] for c in spans[2]]
    rho = np.empty((nz, ny, nx), dtype=np.float64)
    u = np.empty((3, nz, ny, nx), dtype=np.float64)
    for i in range(0, len(tiles), net.batch):
        group = tiles[i:i + net.batch]
        x = np.stack([encode_input(mask[sz:sz + tz, sy:sy + ty, sx:sx + tx], re, net.cin)
                      for (sz, _, _), (sy, _, _), (sx, _, _) in group])
        y = net.forward(x)
        for b, ((sz, lz, hz), (sy, ly, hy), (sx, lx, hx)) in enumerate(group):
            core = y[b][:, lz - sz:hz - sz, ly - sy:hy - sy, lx - sx:hx - sx].astype(np.float64)
            rho[lz:hz, ly:hy, lx:hx] = 1.0 + u_ref * core[0]
            u[:, lz:hz, ly:hy, lx:hx] = u_ref * core[1:4]
    rho = rho.reshape(-1)
    u = u.reshape(3, -1)
    u[:, mask.reshape(-1)] = 0.0
    return rho, u


def solid_mask(solver) -> np.ndarray:
    """(nz, ny, nx) solid indicator of a solver's current flags (0 = fluid)."""
    c = solver.config
    return (solver.flags() != 0).reshape(c.nz, c.ny, c.nx)


def warm_start_tiled(solver, net, re: float, u_ref: float, halo: int = 16) -> dict:
    """Initialises ``solver`` (any lattice, D3Q19 or D3Q27) at the equilibrium
    of the tile-wise prediction of the network from the solver's own solid
    flags, or at the uniform state when the prediction is not plausible.  The
    bodies must be set first.  Returns which start was taken, why, and the
    inference time."""
    if net is None:
        solver.init_equilibrium()
        return {"start": "uniform", "reason": "no network", "tiles": 0, "seconds": 0.0}
    t0 = time.perf_counter()
    mask = solid_mask(solver)
    rho, u = predict_tiled(net, mask, re, u_ref, halo)
    seconds = time.perf_counter() - t0
    ntiles = int(np.prod([len(tile_spans(n, t, halo)) for n, t in zip(mask.shape, net.shape)]))
    ok, why = plausible(rho, u, u_ref)
    if ok:
        solver.init_equilibrium(rho, u)
        return {"start": "net", "reason": why, "tiles": ntiles, "seconds": seconds}
    solver.init_equilibrium()
    return {"start": "uniform", "reason": why, "tiles": ntiles, "seconds": seconds}


def warm_start(solver, net: UNet | None, case) -> dict:
    """Initialises ``solver`` from the network's (rho, u): equilibrium plus
    the first-order non-equilibrium part of that field
    (populations_from_fields), or at the uniform state when there is no
    network or its prediction is not plausible.  Returns which start was
    taken and why."""
    if net is not None:
        rho, u = predict(net, case)
        ok, why = plausible(rho, u, case.u_in)
        if ok:
            f = populations_from_fields(rho, u, (case.L,) * 3, case.tau_plus, periodic=(False, True, True))
            solver.set_populations(f)
            return {"start": "net", "reason": why}
    else:
        why = "no network"
    solver.init_equilibrium()
    return {"start": "uniform", "reason": why}


def solve_steady(case, net: UNet | None = None, use_net: bool = False, tol: float = 1e-4, every: int = 100,
                 patience: int = 3, max_steps: int = 100000, storage: str = "fp32", fields: bool = False) -> dict:
    """Runs one pack case to a converged drag (the solver's own residual rule,
    lbm3d_bb_amd.convergence) from the uniform or the network start."""
    from .convergence import run_to_convergence
    from .solver import Solver

    s = Solver(case.config(storage))
    try:
        s.set_bodies(case.bodies)
        start = warm_start(s, net if use_net else None, case)
        res = run_to_convergence(s, tol=tol, every=every, patience=patience, max_steps=max_steps)
        if use_net and start["start"] == "net" and not res["converged"]:
            # a prediction that does not lead to convergence: restart uniformly
            s.init_equilibrium()
            start = {"start": "uniform", "reason": "net start did not converge"}
            r2 = run_to_convergence(s, tol=tol, every=every, patience=patience, max_steps=max_steps)
            r2["steps"] += res["steps"]
            res = r2
        res.update(start)
        res.pop("history", None)
        if fields:
            res["rho"], res["u"] = s.macros()
    finally:
        s.close()
    return res


def generate_dataset(seeds, L: int, log=None, budget_s: float | None = None, cin: int = PACK_NET["cin"],
                     **kw) -> dict:
    """Steady FP32 fields of the packs ``seeds`` from uniform starts: network
    inputs X (n, cin, L, L, L), targets Y (n, 4, L, L, L) (both held in FP16 on
    the host) and the plain steps-to-converged-drag of every sample.  Packs
    whose run does not reach a converged, finite drag (an unsteady wake) are
    left out of X and Y.  With budget_s, generation stops once that many
    seconds have passed."""
    from .packs import make_pack

    seeds = list(seeds)
    n = len(seeds)
    X = np.empty((n, cin, L, L, L), dtype=np.float16)
    Y = np.empty((n, 4, L, L, L), dtype=np.float16)
    steps = np.zeros(n, dtype=np.int64)
    conv = np.zeros(n, dtype=bool)
    re = np.zeros(n)
    k = 0
    t0 = time.perf_counter()
    done = 0
    for i, sd in enumerate(seeds):
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
        done = i + 1
        case = make_pack(int(sd), L)
        r = solve_steady(case, fields=True, **kw)
        ok = bool(r["converged"]) and bool(np.all(np.isfinite(r["rho"]))) and bool(np.all(np.isfinite(r["u"])))
        steps[i], conv[i], re[i] = r["steps"], ok, case.re
        if ok:
            X[k] = encode_input(case.mask(), case.re, cin)
            Y[k] = encode_target(r["rho"], r["u"], case.u_in, (L, L, L))
            k += 1
        if log:
            log(f"sample {i + 1}/{n} seed {sd} Re {case.re:.2f} spheres {len(case.bodies)} "
                f"steps {r['steps']} converged {r['converged']}{'' if ok else ' (left out)'}")
    return {"X": X[:k], "Y": Y[:k], "steps": steps[:done], "converged": conv[:done], "re": re[:done],
            "seeds": np.asarray(seeds[:done])}


def augment(x: np.ndarray, y: np.ndarray, rng) -> tuple[np.ndarray, np.ndarray]:
    """An exact symmetry of the pack flow applied to one sample (x (cin, z, y, x),
    y (4, z, y, x)): a periodic shift in y and z, reflections of y and z (the
    velocity component flips sign) and the y <-> z exchange."""
    sy, sz = (int(v) for v in rng.integers(0, x.shape[2], size=2))
    x = np.roll(x, (sz, sy), axis=(1, 2))
    y = np.roll(y, (sz, sy), axis=(1, 2))
    if rng.random() < 0.5:  # y -> -y
        x = x[:, :, ::-1, :]
        y = y[:, :, ::-1, :].copy()
        y[2] = -y[2]
    if rng.random() < 0.5:  # z -> -z
        x = x[:, ::-1, :, :]
        y = y[:, ::-1, :, :].copy()
        y[3] = -y[3]
    if rng.random() < 0.5:  # y <-> z
        x = np.swapaxes(x, 1, 2)
        y = np.swapaxes(y, 1, 2)[[0, 1, 3, 2]]
    return np.ascontiguousarray(x), np.ascontiguousarray(y)


def train(net: UNet, X: np.ndarray, Y: np.ndarray, epochs: int, batch: int, lr: float = 1e-3, seed: int = 0,
          log=None, augment_data: bool = True, warmup: int = 200, on_epoch=None) -> list[float]:
    """Adam on the mean-squared error with a linear warm-up over `warmup`
    steps and a cosine decay after it; the samples are reshuffled every epoch
    and, by default, each draw is a random symmetry image of its sample
    (augment).  The next batch is assembled on a host thread while the device
    runs the current step.  on_epoch(epoch, loss), if given, runs after every
    epoch (a checkpoint).  Returns the mean loss per epoch."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    n = X.shape[0]
    per_epoch = max(n // batch, 1)
    total = epochs * per_epoch
    order = np.concatenate([rng.permutation(n)[:per_epoch * batch] for _ in range(epochs)]) if n else np.zeros(0, int)

    def make(t: int):
        idx = np.sort(order[t * batch:(t + 1) * batch])
        if augment_data:
            pairs = [augment(X[i], Y[i], rng) for i in idx]
            return (np.stack([p[0] for p in pairs]).astype(np.float32),
                    np.stack([p[1] for p in pairs]).astype(np.float32))
        return X[idx].astype(np.float32), Y[idx].astype(np.float32)

    hist = []
    losses = []
    with ThreadPoolExecutor(max_workers=1) as pool:
        nxt = pool.submit(make, 0) if total else None
        for t in range(total):
            xb, yb = nxt.result()
            if t + 1 < total:
                nxt = pool.submit(make, t + 1)
            f = (t + 1) / warmup if t < warmup else 0.5 * (1.0 + np.cos(np.pi * (t - warmup) / max(total - warmup, 1)))
            net.set_lr(lr * f + 1e-6)
            losses.append(net.train_step(xb, yb))
            if (t + 1) % per_epoch == 0:
                hist.append(float(np.mean(losses)))
                losses = []
                if log:
                    log(f"epoch {len(hist)}/{epochs} loss {hist[-1]:.5f}")
                if on_epoch:
                    on_epoch(len(hist), hist[-1])
    return hist


def prediction_error(net: UNet, case, rho: np.ndarray, u: np.ndarray) -> dict:
    """RMS error over the fluid cells of the network's prediction against a
    steady field: velocity in units of the inlet speed (per component), and
    rho - 1 relative to the rms of the true rho - 1."""
    rp, up = predict(net, case)
    fluid = ~case.mask().reshape(-1)
    du = (up - u)[:, fluid] / case.u_in
    dr = (rp - rho)[fluid]
    return {"u_rms": float(np.sqrt((du * du).mean())),
            "rho_rel_rms": float(np.sqrt((dr * dr).mean()) / max(np.sqrt(((rho[fluid] - 1.0) ** 2).mean()), 1e-300))}


def evaluate(net: UNet, seeds, L: int, log=None, **kw) -> dict:
    """Steps to a converged drag from the uniform and from the network start on
    held-out packs, the relative difference of the two converged drags, and the
    network's prediction error against the converged field of the uniform run."""
    from .packs import make_pack

    rows = []
    for i, sd in enumerate(seeds):
        case = make_pack(int(sd), L)
        ru = solve_steady(case, fields=True, **kw)
        rn = solve_steady(case, net=net, use_net=True, **kw)
        pe = prediction_error(net, case, ru.pop("rho"), ru.pop("u"))
        d_rel = abs(rn["drag"] - ru["drag"]) / max(abs(ru["drag"]), 1e-300)
        rows.append({"seed": int(sd), "re": case.re, "spheres": len(case.bodies), "steps_uniform": ru["steps"],
                     "steps_net": rn["steps"], "converged_uniform": ru["converged"],
                     "converged_net": rn["converged"], "start": rn["start"], "drag_uniform": ru["drag"],
                     "drag_net": rn["drag"], "drag_rel_diff": d_rel, "pred_u_rms": pe["u_rms"],
                     "pred_rho_rel_rms": pe["rho_rel_rms"]})
        if log:
            log(f"eval {i + 1}/{len(seeds)} seed {sd} Re {case.re:.2f} steps uniform {ru['steps']} net {rn['steps']} "
                f"({rn['start']}) drag diff {d_rel:.2e} pred err u {pe['u_rms']:.3f} rho {pe['rho_rel_rms']:.3f}")
    su = np.array([r["steps_uniform"] for r in rows], dtype=np.float64)
    sn = np.array([r["steps_net"] for r in rows], dtype=np.float64)
    return {"rows": rows, "mean_steps_uniform": float(su.mean()), "mean_steps_net": float(sn.mean()),
            "speedup_of_means": float(su.mean() / sn.mean()), "median_ratio": float(np.median(su / sn)),
            "median_pred_u_rms": float(np.median([r["pred_u_rms"] for r in rows])),
            "max_drag_rel_diff": float(max(r["drag_rel_diff"] for r in rows)),
            "net_starts": int(sum(r["start"] == "net" for r in rows))}
