"""Writes the body lists of the two benchmark configurations that draw them from
seed 0 (bench.make_case): the 64 spheres of the 512x256x256 box and the 8
ellipsoids of the 1024x384x384 channel.  Run from the repository root:
``python tests/golden/make_config_bodies.py``."""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import bench  # noqa: E402


def bodies_of(name: str) -> dict:
    _, laws, bodies, desc = bench.make_case(bench.WORKLOADS[name])
    if bodies is not None:
        return {"workload": desc, "spheres": [[*map(float, b.center), float(b.semi[0])] for b in bodies]}
    return {"workload": desc,
            "ellipsoids": [{"center0": list(map(float, l.center0)), "semi": list(map(float, l.semi)),
                            "omega": list(map(float, l.omega)), "rot0": [list(map(float, r)) for r in l.rot0]}
                           for l in laws]}


def main() -> None:
    for name, fname in (("box512_spheres64", "box512_spheres64_seed0.json"),
                        ("channel1024_ellipsoids8", "channel1024_ellipsoids8_seed0.json")):
        with open(os.path.join(HERE, fname), "w", encoding="utf-8") as fh:
            json.dump(bodies_of(name), fh, indent=1)
            fh.write("\n")


if __name__ == "__main__":
    main()
