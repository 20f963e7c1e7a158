"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must keep reproducing them (CPU), and the HIP
path must match them (GPU) to the FP64 parity bar.  Flags and links are
integer data and must match exactly."""
from __future__ import annotations

import importlib.util
import os

import numpy as np
import pytest

from cases import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _maker():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


MG = _maker()
NAMES = sorted(MG.CASES)


def _load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _scale(gold):
    from lbm3d_bb_amd.lattice import W, W27

    w = W27 if gold["f"].shape[0] == 27 else W
    return float(2.0 * w[gold["link_dir"].astype(np.int64)].sum())


def _check(s, gold, steps, tol):
    F = s.step(steps)
    assert np.array_equal(s.flags(), gold["flags"])
    cells, dirs = s.links()
    assert np.array_equal(cells, gold["link_cell"]) and np.array_equal(dirs, gold["link_dir"])
    assert np.abs(F - gold["forces"]).max() / _scale(gold) < tol
    rho, u = s.macros()
    assert rel_err(rho, gold["rho"]) < tol
    assert np.abs(u - gold["u"]).max() < tol
    assert rel_err(s.populations(), gold["f"]) < tol


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    cfg, o = MG.build(name)
    o.init_equilibrium()
    _check(o, _load(name), MG.CASES[name]["steps"], 1e-14)
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_golden(name):
    from lbm3d_bb_amd.solver import Solver

    cfg, o = MG.build(name)
    o.close()
    s = Solver(cfg)
    MG.setup(name, s)
    s.init_equilibrium()
    _check(s, _load(name), MG.CASES[name]["steps"], 1e-12)
    s.close()
