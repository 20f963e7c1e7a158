"""Warm-start host logic (CPU): the non-equilibrium start, the symmetry
augmentation and the drag-residual rule."""
from __future__ import annotations

import numpy as np

from lbm3d_bb_amd.convergence import DragMonitor
from lbm3d_bb_amd.lattice import CX, CY, CZ, equilibrium
from lbm3d_bb_amd.warmstart import augment, populations_from_fields


def _shear_field(n=12):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ux = 0.01 * np.sin(2 * np.pi * y / n)
    uy = 0.005 * np.cos(2 * np.pi * x / n)
    u = np.stack([ux.reshape(-1), uy.reshape(-1), np.zeros(n ** 3)])
    rho = 1.0 + 0.001 * np.cos(2 * np.pi * z / n).reshape(-1)
    return n, rho, u


def test_nonequilibrium_part_keeps_mass_and_momentum():
    n, rho, u = _shear_field()
    f = populations_from_fields(rho, u, (n, n, n), 0.8, (True, True, True))
    assert np.abs(f.sum(0) - rho).max() < 1e-15
    j = np.stack([(f * c[:, None]).sum(0) for c in (CX, CY, CZ)])
    assert np.abs(j - rho * u).max() < 1e-15


def test_nonequilibrium_stress_is_chapman_enskog():
    # sum_i f_neq c_x c_y = -(tau rho / 3)(d_x u_y + d_y u_x) for D3Q19
    # (fourth-order isotropy sum w c c c c = (dd + dd + dd) / 9)
    n, rho, u = _shear_field()
    tau = 0.8
    f = populations_from_fields(rho, u, (n, n, n), tau, (True, True, True))
    fneq = f - equilibrium(rho, *u)
    pxy = (fneq * (CX * CY)[:, None]).sum(0)
    U = u.reshape(3, n, n, n)
    dyux = (0.5 * (np.roll(U[0], -1, axis=1) - np.roll(U[0], 1, axis=1))).reshape(-1)
    dxuy = (0.5 * (np.roll(U[1], -1, axis=2) - np.roll(U[1], 1, axis=2))).reshape(-1)
    expect = -(tau * rho / 3.0) * (dyux + dxuy)
    assert np.abs(pxy - expect).max() < 1e-14 * max(1.0, np.abs(expect).max()) + 1e-17
    # and no normal-stress imbalance for a divergence-free field: trace part zero
    trace = (fneq * (CX * CX + CY * CY + CZ * CZ)[:, None]).sum(0)
    assert np.abs(trace).max() < 1e-15


def test_augmentation_is_a_symmetry_of_the_encoding():
    rng = np.random.default_rng(1)
    n = 8
    mask = rng.random((n, n, n)) < 0.2
    x = np.stack([mask.astype(np.float32), np.full((n, n, n), 0.4, np.float32)])
    # a target tied to the mask so the same transform must apply to both
    y = np.stack([mask * 1.0, mask * 2.0, mask * 3.0 + 0.5, mask * 4.0 - 0.5]).astype(np.float32)
    for _ in range(20):
        xa, ya = augment(x, y, rng)
        m = xa[0] > 0.5
        assert np.allclose(ya[0], m * 1.0) and np.allclose(ya[1], m * 2.0)
        # the y and z components keep their magnitude pattern, possibly swapped and sign-flipped
        vy, vz = np.abs(ya[2]), np.abs(ya[3])
        assert (np.allclose(vy, np.abs(m * 3.0 + 0.5)) and np.allclose(vz, np.abs(m * 4.0 - 0.5))) or \
               (np.allclose(vy, np.abs(m * 4.0 - 0.5)) and np.allclose(vz, np.abs(m * 3.0 + 0.5)))
        assert np.isclose(xa[1].mean(), 0.4)


def test_drag_monitor_needs_consecutive_small_residuals():
    mon = DragMonitor(tol=1e-4, every=100, patience=3)
    seq = [1.0, 1.5, 1.2, 1.20001, 1.200011, 1.3, 1.30001, 1.300011, 1.300012]
    hits = [mon.update(100 * (i + 1), v) for i, v in enumerate(seq)]
    assert hits[-1] and not any(hits[:-1])


def test_offline_rule_matches_the_drag_monitor():
    """warmstart_ceiling evaluates the residual rule on recorded per-step drags:
    it must stop where DragMonitor, fed the same window means, stops."""
    from lbm3d_bb_amd.convergence import DragMonitor
    from lbm3d_bb_amd.warmstart_ceiling import smooth_noise, steps_to_converge

    t = np.arange(40000)
    rng = np.random.default_rng(1)
    d = 1.0 + 0.3 * np.exp(-t / 3000.0) + 0.01 * np.exp(-t / 900.0) * np.sin(2 * np.pi * t / 443.0)
    d += 1e-7 * rng.standard_normal(t.size)
    for every in (100, 500):
        m = DragMonitor(every=every)
        stop = None
        for i in range(1, t.size // every + 1):
            if m.update(i * every, d[(i - 1) * every:i * every].mean()):
                stop = i * every
                break
        assert stop is not None and steps_to_converge(d, every) == stop
    assert steps_to_converge(np.ones(150), 100) is None  # one window: nothing to compare
    a = smooth_noise((16, 16, 16), rng, 3.0)
    assert abs(a.mean()) < 1e-12 and abs(a.std() - 1.0) < 1e-12
