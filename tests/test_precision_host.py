"""Host logic of the precision summaries and the bench line (CPU only): the
per-body drag and torque errors against an FP64 run, the start-pair
histories of precision_run combine, and bench.py's drag error and 16-bit
target."""
from __future__ import annotations

import json

import numpy as np
import pytest


def _run(storage, start, drag, torque, hist):
    return {"runs": {storage: {"storage": storage, "steps": 4000, "seconds": 1.0, "mlups": 1.0, "drag": drag,
                               "force": [[d, 0.0, 0.0] for d in drag], "torque": torque, "rho_min": 1.0,
                               "rho_max": 1.0, "drag_total_history": [], "chunk_mean_history": hist,
                               "start": {"start": start}, "workload": "w"}}}


def _hist(drags):
    # chunk means every 1000 steps; body force (Fx, 0, 0, 0, 0, 0)
    return [[1000 * (i + 1), [[d, 0, 0, 0, 0, 0] for d in row]] for i, row in enumerate(drags)]


def test_combine_errors_and_start_pairs(tmp_path):
    from lbm3d_bb_amd.precision_run import combine

    t64 = [[0.0, 0.0, 2.0], [0.0, 1.0, 0.0]]
    runs = {
        "fp64": _run("fp64", "uniform", [1.0, 2.0], t64, _hist([[1.0, 2.0]] * 4)),
        "uniform": _run("fp16s", "uniform", [1.001, 1.996], [[0.0, 0.0, 2.002], [0.0, 1.0, 0.0]],
                        _hist([[1.0, 2.0], [1.0, 2.0], [1.0, 2.0], [1.0, 2.0]])),
        # the network start: 17 % off after 1000 steps, 0.5 % after 2000, 0.25 % from 3000 on
        "net": _run("fp16s", "net", [1.0, 2.0], t64, _hist([[1.0, 2.4], [1.0, 2.01], [1.0, 2.005], [1.0, 2.005]])),
    }
    named = []
    for k, r in runs.items():
        p = tmp_path / f"{k}.json"
        p.write_text(json.dumps(r))
        named.append(f"{k}={p}")
    out = tmp_path / "summary.json"
    s = combine(named, str(out))
    e = s["errors_vs_fp64"]["uniform"]
    assert e["drag_err_per_body"] == pytest.approx([1e-3, 2e-3])
    assert e["drag_err_max"] == pytest.approx(2e-3)
    assert e["drag_err_total"] == pytest.approx(abs(2.997 - 3.0) / 3.0)
    # torque error relative to the largest FP64 torque (2.0)
    assert e["torque_err_per_body"] == pytest.approx([0.001, 0.0])
    assert "net" not in s["errors_vs_fp64"]  # another start: not an error against the uniform FP64 run
    pair = s["start_pairs"]["uniform_vs_net"]
    rel = pair["max_body_drag_rel_diff"]
    assert rel == pytest.approx([0.4 / 2.4, 0.01 / 2.01, 0.005 / 2.005, 0.005 / 2.005])
    assert pair["steps_until_within"] == {"0.1": 2000, "0.01": 2000, "0.003": 3000, "0.002": None}
    assert json.loads(out.read_text())["start_pairs"]["uniform_vs_net"]["final_rel_diff"] == pytest.approx(rel[-1])


def test_bench_drag_error_and_target():
    import bench

    F = np.zeros((300, 2, 6))
    F64 = np.zeros((300, 2, 6))
    F[:, 0, 0], F[:, 1, 0] = 1.01, 2.0
    F64[:, 0, 0], F64[:, 1, 0] = 1.0, 2.0
    F[:200, :, 0] = 50.0  # outside the last-100-step window
    e = bench.drag_err(F, F64)
    assert e["window"] == 100 and e["steps"] == 300
    assert e["max"] == pytest.approx(0.01)
    assert e["median"] == pytest.approx(0.005)
    assert e["total"] == pytest.approx(0.01 / 3.0)
    assert bench.drag_err(F, None)["max"] == 0.0
    # the D3Q19 16-bit target: 1.7 x (60 % of 8 TB/s over 152 bytes per cell) in MLUPS
    assert 1.7 * 0.6 * 8000e9 / bench.bytes_per_cell("fp32", 19) / 1e6 == pytest.approx(53684.2, abs=0.1)
    same = bench._case_key(bench.WORKLOADS["channel256"]) == bench._case_key(bench.WORKLOADS["channel256_fp16s"])
    other = bench._case_key(bench.WORKLOADS["channel256"]) == bench._case_key(bench.WORKLOADS["channel256_q27"])
    assert same and not other
