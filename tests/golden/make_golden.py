"""Writes the oracle's golden fixtures (tests/golden/*.npz).

  python tests/golden/make_golden.py [name ...]    (default: every case)

Each fixture holds the inputs of a small case (as arrays the loader turns
back into the same Config and bodies) and the oracle's outputs after a few
steps: solid flags, boundary links, per-step force/torque, macroscopic fields
and populations.  The CPU tests check that the oracle still reproduces them;
the GPU tests check the HIP path against them.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(HERE))

from cases import ConstantMotion, channel, periodic_box, sphere  # noqa: E402
from oracle.oracle import OracleSolver  # noqa: E402

CASES = {
    # periodic box, BGK tau 0.8, body force, one fixed sphere (the 64^3 case's shape at 12^3)
    "box12_bgk_sphere": dict(kind="box", n=12, collision="BGK", tau=0.8, g=1e-5, steps=6,
                             bodies=[(6.0, 6.0, 6.0, 2.5)]),
    # channel: velocity inlet, zero-gradient outlet, no-slip walls, translating sphere, TRT 3/16
    "channel16_trt_moving": dict(kind="channel", nx=16, ny=10, nz=10, tau=0.6, steps=6,
                                 motion=(5.3, 5.1, 4.9, 2.5, 0.05)),
    # the same two shapes on the D3Q27 lattice, the channel with the interpolated bounce-back
    "box10_q27_trt_sphere": dict(kind="box", n=10, collision="TRT", tau=0.8, g=1e-5, steps=6, q=27,
                                 bodies=[(5.0, 5.2, 4.8, 2.3)]),
    "channel14_q27_ibb_moving": dict(kind="channel", nx=14, ny=10, nz=10, tau=0.6, steps=6, q=27, bc="IBB",
                                     motion=(4.8, 5.1, 4.9, 2.4, 0.05)),
}


def build(name: str):
    c = CASES[name]
    if c["kind"] == "box":
        cfg = periodic_box(n=c["n"], storage="fp64", collision=c["collision"], tau=c["tau"], g=(c["g"], 0.0, 0.0))
        cfg.q = c.get("q", 19)
        cfg.solid_bc = c.get("bc", "SBB")
        o = OracleSolver(cfg, "fp64")
        o.set_bodies([sphere((b[0], b[1], b[2]), b[3]) for b in c["bodies"]])
    else:
        cfg = channel(nx=c["nx"], ny=c["ny"], nz=c["nz"], storage="fp64", tau=c["tau"])
        cfg.q = c.get("q", 19)
        cfg.solid_bc = c.get("bc", "SBB")
        o = OracleSolver(cfg, "fp64")
        m = c["motion"]
        o.set_motion([ConstantMotion("sphere", (m[0], m[1], m[2]), (m[3],), vel=(m[4], 0.0, 0.0))])
    return cfg, o


def setup(name: str, solver) -> None:
    c = CASES[name]
    if c["kind"] == "box":
        solver.set_bodies([sphere((b[0], b[1], b[2]), b[3]) for b in c["bodies"]])
    else:
        m = c["motion"]
        solver.set_motion([ConstantMotion("sphere", (m[0], m[1], m[2]), (m[3],), vel=(m[4], 0.0, 0.0))])


def main(argv=None) -> int:
    names = (argv if argv is not None else sys.argv[1:]) or list(CASES)
    for name in names:
        c = CASES[name]
        cfg, o = build(name)
        o.init_equilibrium()
        F = o.step(c["steps"])
        rho, u = o.macros()
        f = o.populations()
        cells, dirs = o.links()
        flags = o.flags()
        o.close()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, forces=F, rho=rho, u=u, f=f, link_cell=cells, link_dir=dirs, flags=flags)
        print(path, os.path.getsize(path))
    return 0


if __name__ == "__main__":
    sys.exit(main())
