"""Specimens of the linear state-based (ElasticState) tests.

qs2d / qs1d rebuild the meshes of the reference's quasi-static examples
(examples/qsModel/{1D,2D}), whose output_res.vtu are the goldens of
quasistatic.{1D,2D}.elastic (test/CMakeLists.txt:12-54).  The reference reads
its mesh from a .vtu file whose Points are float32, so the 2D node
coordinates are the mesh tool's grid (fe2D.cpp:148-158) rounded to float32.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from nlmech_amd import problems as pr

GOLDEN = Path(__file__).resolve().parent / "golden"


def qs2d():
    """17x17 nodes on [0, 1.6]^2, h = 0.1, horizon 0.5; K = 4000, G = 1500,
    plane stress (examples/qsModel/2D/input.yaml, input_mesh.yaml)."""
    g = np.load(GOLDEN / "qs2d_output_res.npz")
    x, _ = pr.grid2d(17, 17, 0.1, jitter=0.0)
    x = x.astype(np.float32).astype(np.float64)
    p = pr.Problem(name="qs2d", dim=2, x=x, vol=g["Node_Volume"].copy(),
                   horizon=0.5, h=0.1, dt=1e-3, nsteps=1,
                   material=pr.es_material(dim=2, K=4000.0, G=1500.0,
                                           density=1200.0, sc=-1.0))
    return p, g


def qs1d():
    """33 nodes on [0, 16], h = 0.5, horizon 1; E = 4000
    (examples/qsModel/1D/input.yaml, mesh.vtu)."""
    g = np.load(GOLDEN / "qs1d_output_res.npz")
    m = np.load(GOLDEN / "qs1d_mesh.npz")
    x = m["Points"].astype(np.float64)
    p = pr.Problem(name="qs1d", dim=1, x=x, vol=m["Node_Volume"].copy(),
                   horizon=1.0, h=0.5, dt=1e-3, nsteps=1,
                   material=pr.es_material(dim=1, K=-1.0, G=-1.0, E=4000.0,
                                           density=1.0, sc=-1.0))
    return p, g


def small_cube(n=8, ratio=3, sc=1e-3, pull=1.0, sigma_v=1e-3, dt=2e-8,
               h=6.25e-4, **kw):
    return pr.cube3d(n, h, ratio, dt, 10, pull=pull, sigma_v=sigma_v, sc=sc,
                     **kw)
