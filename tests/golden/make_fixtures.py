"""Writes the committed fixtures under tests/golden/ (run in the build
container, where the reference tree is mounted):

  python tests/golden/make_fixtures.py [/root/reference]

1. Reference-held data (inputs and expected outputs of the reference's own
   tests), decoded to .npz so the tests need no reference tree:
   - qs2d_output_res.npz  <- examples/qsModel/2D/output_res.vtu
       (289 nodes, horizon 0.5 = 5h; the CTest golden of
        quasistatic.2D.elastic, test/CMakeLists.txt:36-54)
   - qs1d_output_res.npz  <- examples/qsModel/1D/output_res.vtu
       (33 nodes; golden of quasistatic.1D.elastic, test/CMakeLists.txt:12-32)
   - qs1d_mesh.npz        <- examples/qsModel/1D/mesh.vtu (the input mesh of
       that test: Points as stored, float32, and Node_Volume)
   - dc_fd_simple.npz     <- examples/tools/dc/fd_simple/output_{2,4}.vtu and
       compare.txt (the dc tool's known answer, test/CMakeLists.txt:61-64)
   - testFracture_nodes.csv <- data/test/geometry/testFracture_nodes.csv
       (the node set of TestGeom, src/test/geometry/testGeomLib.cpp:42-86)
   - qs1d/input_legacy_vtk.yaml <- examples/qsModel/1D/input_legacy_vtk.yaml
       (the deck of quasistatic.1D.elastic_legacy_vtk, test/CMakeLists.txt:
       88-91)
   - gmsh/ <- examples/io/{input_2d_gmsh_quad,input_2d_gmsh_triangle}.yaml
       and their meshes mesh_2d_{quad,triangle}.msh (gmsh format 4.1; the
       inputs of mesh.2d.read.{quad,triangle}.gmsh, test/CMakeLists.txt:
       72-82), examples/fdModel/logo/input.yaml and mesh.msh (format 2.2)
       as logo_input.yaml / logo_mesh.msh, copied unchanged
2. Oracle-generated trajectories (oracle_*.npz), pinning the CPU restatement
   against later edits; written by make_oracle_fixtures().
"""
from __future__ import annotations

import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent.parent))

from vtu_read import read_vtu  # noqa: E402


def reference_fixtures(ref: Path) -> None:
    for dim in (1, 2):
        d = read_vtu(str(ref / f"examples/qsModel/{dim}D/output_res.vtu"))
        keep = {k: v for k, v in d.items()
                if k in ("Points", "Displacement", "Force", "Strain_Energy",
                         "Fixity", "Node_Volume", "Neighbors",
                         "Strain_Tensor", "Stress_Tensor")}
        np.savez_compressed(HERE / f"qs{dim}d_output_res.npz", **keep)
    m = read_vtu(str(ref / "examples/qsModel/1D/mesh.vtu"))
    np.savez_compressed(HERE / "qs1d_mesh.npz", Points=m["Points"],
                        Node_Volume=m["Node_Volume"])
    a = read_vtu(str(ref / "examples/tools/dc/fd_simple/output_2.vtu"))
    b = read_vtu(str(ref / "examples/tools/dc/fd_simple/output_4.vtu"))
    txt = (ref / "examples/tools/dc/fd_simple/compare.txt").read_text()
    vals = [float(t) for t in txt.strip().splitlines()[1].split(",")]
    np.savez_compressed(HERE / "dc_fd_simple.npz",
                        points_2=a["Points"], velocity_2=a["Velocity"],
                        points_4=b["Points"], velocity_4=b["Velocity"],
                        expected_l2_sup=np.array(vals))
    shutil.copyfile(ref / "data/test/geometry/testFracture_nodes.csv",
                    HERE / "testFracture_nodes.csv")
    shutil.copyfile(ref / "examples/qsModel/1D/input_legacy_vtk.yaml",
                    HERE / "qs1d/input_legacy_vtk.yaml")
    g = HERE / "gmsh"
    g.mkdir(exist_ok=True)
    for f in ("input_2d_gmsh_quad.yaml", "input_2d_gmsh_triangle.yaml",
              "mesh_2d_quad.msh", "mesh_2d_triangle.msh"):
        shutil.copyfile(ref / "examples/io" / f, g / f)
    shutil.copyfile(ref / "examples/fdModel/logo/input.yaml",
                    g / "logo_input.yaml")
    shutil.copyfile(ref / "examples/fdModel/logo/mesh.msh", g / "logo_mesh.msh")


def make_oracle_fixtures() -> None:
    from nlmech_amd import problems as pr
    from oracle import pyoracle as po

    cases = {
        # 24x24 plate, notch, pull 0.5 m/s, 40 steps
        "oracle_plate24_pull05": (pr.plate2d(24, 24, 1e-3, 3, 1e-7, 40,
                                             notch_len=0.008), 40),
        # 24x24 plate, pull 30 m/s: bonds break near the loaded layers
        "oracle_plate24_pull30": (pr.plate2d(24, 24, 1e-3, 3, 1e-7, 60,
                                             notch_len=0.008, pull=30.0), 60),
    }
    for name, (p, steps) in cases.items():
        o = po.OracleFD(p)
        o.init()
        o.step(steps)
        u, v, f = o.get_state()
        np.savez_compressed(HERE / f"{name}.npz", u=u, v=v, f=f,
                            bonds=o.get_bonds(), row_ptr=o.row_ptr,
                            col=o.col, steps=np.array(steps))


if __name__ == "__main__":
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    if ref.exists():
        reference_fixtures(ref)
    make_oracle_fixtures()
    print("fixtures written to", HERE)
