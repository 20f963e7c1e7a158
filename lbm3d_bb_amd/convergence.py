"""Drag-residual convergence monitor.

The run is converged at step n (a multiple of ``every``) when the relative
change of the window-mean drag, r(n) = |Fm(n) - Fm(n - every)| / |Fm(n)|, is
below ``tol`` for ``patience`` consecutive checks, where Fm(n) is the mean
over the ``every`` steps ending at n of F, the x-force summed over all bodies
(the reference's C_D numerator, src/MEA.jl:57-59).  The window mean filters
the pressure waves that every start state launches (they bounce between the
faces with a period of the order of the lattice crossing time), which would
otherwise reset the count without bearing on the steady drag.  The same rule
is applied to every start state, so a warm start can shorten a run but not
change the converged answer.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class DragMonitor:
    tol: float = 1e-4
    every: int = 100
    patience: int = 3
    axis: int = 0
    history: list = field(default_factory=list)  # (step, drag)
    _ok: int = 0

    def update(self, step: int, drag: float) -> bool:
        self.history.append((step, float(drag)))
        if len(self.history) < 2:
            return False
        prev = self.history[-2][1]
        r = abs(drag - prev) / max(abs(drag), 1e-300)
        self._ok = self._ok + 1 if r < self.tol else 0
        return self._ok >= self.patience

    def residuals(self) -> np.ndarray:
        d = np.array([h[1] for h in self.history])
        return np.abs(np.diff(d)) / np.maximum(np.abs(d[1:]), 1e-300)


def run_to_convergence(solver, tol: float = 1e-4, every: int = 100, patience: int = 3, max_steps: int = 200000,
                       axis: int = 0) -> dict:
    """Steps a solver until the drag residual criterion holds (or max_steps)."""
    mon = DragMonitor(tol=tol, every=every, patience=patience, axis=axis)
    steps = 0
    converged = False
    while steps < max_steps:
        F = solver.step(every, forces=True)
        steps += every
        drag = float(F[:, :, axis].sum(axis=1).mean())
        if not np.isfinite(drag):
            break
        if mon.update(steps, drag):
            converged = True
            break
    return {"steps": steps, "converged": converged, "drag": mon.history[-1][1] if mon.history else float("nan"),
            "history": mon.history}
