"""Measure the algorithmic work of one walk step on the reference BVH.

For each benchmark workload, walks from a fixed sample of inside grid points
are run in the oracle with its query tally on; the mean node records read
(V_node) and primitive tests (V_prim) per walk-step query are written to
profiles/algorithmic_counts.json, together with the bytes per step on the
reference layout at FP32 and FP64 sizes:

    B_step = V_node * S_node + V_prim * S_prim
    S_node: 2D 40 B (FP64) / 24 B (FP32); 3D 56 B / 32 B  (BVHNode, bvh.hpp:10-15)
    S_prim: 2D 48 B / 32 B (2 vertices + indices);  3D 116 B / 60 B (EberlyCache + index)

TEST INFRASTRUCTURE (run offline here; bench.py only reads the JSON).
Usage: python oracle/algorithmic_counts.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "walk-on-spheres_amd"))

import pyoracle  # noqa: E402
from wos_amd import meshgen, problems  # noqa: E402

S_NODE = {2: {"f64": 40, "f32": 24}, 3: {"f64": 56, "f32": 32}}
S_PRIM = {2: {"f64": 48, "f32": 32}, 3: {"f64": 116, "f32": 60}}

WORKLOADS = {
    "annulus_laplace_64x64": dict(mesh=lambda: meshgen.annulus(128), eq=problems.annulus_laplace,
                                  grid=(64, 64, 1), eps=1e-3, seed=12345),
    "icosphere4_screened_32cubed": dict(mesh=lambda: meshgen.icosphere(4),
                                        eq=problems.sphere_screened_constant,
                                        grid=(32, 32, 32), eps=1e-4, seed=7),
}


def measure(name, spec, sample_points=400, walks=16):
    orc = pyoracle.Oracle()
    V, F, B = spec["mesh"]()
    om = orc.mesh(V, F, B)
    eq = spec["eq"]()
    nx, ny, nz = spec["grid"]
    dim = F.shape[1]
    loc = om.solve_grid(eq, nx, ny, nz, 1, spec["eps"], seed=spec["seed"], rng_mode=1,
                        n_threads=8)["location"]
    pts = pyoracle.grid_points(om.bbox(), nx, ny, nz, dim)
    inside = np.nonzero(loc == 1)[0]
    pick = inside[np.linspace(0, len(inside) - 1, min(sample_points, len(inside))).astype(int)]
    orc.count_reset(True)
    om.solve_points(eq, pts[pick], walks, spec["eps"], seed=spec["seed"], rng_mode=1,
                    n_threads=1, seed_index=pick)
    c = orc.count_get()
    orc.count_reset(False)
    vn = c["node_reads"] / c["queries"]
    vp = c["prim_tests"] / c["queries"]
    out = dict(dim=dim, n_prims=len(F), V_node=vn, V_prim=vp, queries=c["queries"],
               sample_points=len(pick), walks_per_point=walks)
    for prec in ("f32", "f64"):
        out["B_step_" + prec] = vn * S_NODE[dim][prec] + vp * S_PRIM[dim][prec]
    return out


def main():
    res = {k: measure(k, v) for k, v in WORKLOADS.items()}
    res["_command"] = "python oracle/algorithmic_counts.py"
    path = os.path.join(ROOT, "profiles", "algorithmic_counts.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
