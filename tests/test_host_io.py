"""CPU checks of the C++ host layer (nlmech_amd/host): the YAML deck reader,
the .vtu writer / reader, the mesh tool and the dc tool.

- dc: the reference's own known answer (examples/tools/dc/fd_simple:
  compare.txt "1.2062, 0.261072" from output_2.vtu / output_4.vtu, both
  reference-held files), read through the C++ .vtu reader.
- mesh tool: the 1D mesh of examples/qsModel/1D equals the reference-held
  mesh.vtu array by array (Float32 points, Float64 volumes, cells); the 2D
  generators follow fe2D.cpp's node numbering, and their volumes sum to the
  domain area.
- deck reader: values of the reference's example decks (fixtures copied from
  examples/) as inp::Input reads them: dt = Final_Time / Time_Steps, mesh
  size from Horizon / Horizon_h_Ratio, Rectangle [x1, y1, x2, y2], PDBond
  renamed RNPBond, a repeated key keeping its first value; error messages
  and exit status 1 where the reference exits.
- gmsh meshes (formats 4.1 and 2.2, the reference's examples/io and
  examples/fdModel/logo meshes): node counts, element counts, and nodal
  volumes pinned by closed forms -- the unit quadrangles' corner / edge /
  interior volumes 1/4, 1/2, 1, every mesh's volumes summing to its domain's
  area (5 x 5, and 2.25 for the logo polygon of logo.geo) -- and by the
  element areas split in thirds (triangles); binary files and 3D decks are
  refused.
No GPU work is issued here.
"""
import json
import re
import shutil
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, run_tool
from vtu_read import read_vtu

DECKS = GOLDEN / "decks"


def dump(host_bin, deck: Path):
    rc, out = run_tool(host_bin / "nlm_fd", "--dump-deck", "-i", deck)
    line = [l for l in out.splitlines() if l.startswith("{")][-1]
    return json.loads(line)


def test_dc_known_answer(host_bin, tmp_path):
    src = GOLDEN / "dc_fd_simple"
    (tmp_path / "input.yaml").write_text(
        "Dimension: 2\n"
        f"Filename_1: {src / 'output_2.vtu'}\n"
        f"Filename_2: {src / 'output_4.vtu'}\n"
        "Out_Filename: compare.txt\nPrint_Screen: false\n"
        "Compare_Tags:\n    - Velocity\n")
    run_tool(host_bin / "nlm_dc", "-i", "input.yaml", "-k", "fd_simple",
             cwd=tmp_path)
    lines = (tmp_path / "compare.txt").read_text().splitlines()
    assert lines == ["Velocity_L2_Error, Velocity_Sup_Error",
                     "1.2062, 0.261072"]


def test_dc_tolerance_exit_status(host_bin, tmp_path):
    src = GOLDEN / "dc_fd_simple"
    for tol, rc_expect in (("1.0", 1), ("2.0", 0)):
        (tmp_path / "input.yaml").write_text(
            "Dimension: 2\n"
            f"Filename_1: {src / 'output_2.vtu'}\n"
            f"Filename_2: {src / 'output_4.vtu'}\n"
            f"Out_Filename: compare.txt\nPrint_Screen: false\n"
            f"Tolerance: {tol}\nCompare_Tags:\n    - Velocity\n")
        rc, _ = run_tool(host_bin / "nlm_dc", "-i", "input.yaml", "-k",
                         "fd_simple", cwd=tmp_path, check=False)
        assert rc == rc_expect


def test_dc_corrupt_vtu_headers(host_bin, tmp_path):
    """Sizes read from a .vtu file's compressed-block header are checked
    before they are used: a block count or a block size far beyond the file
    ends in a reader error (non-zero exit, "vtu:" message), not a crash."""
    import base64
    import struct
    src = GOLDEN / "dc_fd_simple"
    data = (src / "output_4.vtu").read_bytes()
    us = data.index(b"_", data.index(b"<AppendedData"))
    m = re.search(rb'Name="Velocity"[^>]*offset="(\d+)"', data)
    at = us + 1 + int(m.group(1))
    hdr = base64.b64decode(data[at:at + 24])     # UInt32 [nb, bs, last, cs]
    assert struct.unpack("<I", hdr[:4])[0] == 1
    for word, value in ((0, 0xFFFFFFF0), (1, 0x90000000)):
        h = bytearray(hdr)
        h[4 * word:4 * word + 4] = struct.pack("<I", value)
        bad = data[:at] + base64.b64encode(bytes(h)) + data[at + 24:]
        (tmp_path / "bad.vtu").write_bytes(bad)
        (tmp_path / "input.yaml").write_text(
            "Dimension: 2\n"
            f"Filename_1: {src / 'output_2.vtu'}\n"
            f"Filename_2: {tmp_path / 'bad.vtu'}\n"
            "Out_Filename: compare.txt\nPrint_Screen: false\n"
            "Compare_Tags:\n    - Velocity\n")
        rc, out = run_tool(host_bin / "nlm_dc", "-i", "input.yaml", "-k",
                           "fd_simple", cwd=tmp_path, check=False)
        assert rc > 0 and "vtu:" in out, (word, rc, out[-300:])


def test_mesh_tool_1d_equals_reference_mesh(host_bin, tmp_path):
    shutil.copy(DECKS / "qs1d_input_mesh.yaml", tmp_path / "input_mesh.yaml")
    run_tool(host_bin / "nlm_mesh", "-i", "input_mesh.yaml", "-d", "1",
             cwd=tmp_path)
    ours = read_vtu(str(tmp_path / "mesh.vtu"))
    ref = read_vtu(str(GOLDEN / "qs1d_mesh.vtu"))
    assert set(ref) <= set(ours)
    for k, v in ref.items():
        assert ours[k].dtype == v.dtype, k
        assert np.array_equal(ours[k], v), k


@pytest.mark.parametrize("kind", ["uniform_square", "uniform_tri",
                                  "uniform_tri_sym"])
@pytest.mark.parametrize("compress", ["zlib", ""])
def test_mesh_tool_2d(host_bin, tmp_path, kind, compress):
    L, h = 0.01, 1e-3
    (tmp_path / "m.yaml").write_text(
        "Output:\n  Path: ./\n  Mesh: mesh\n"
        f"Domain: [0.0, 0.0, {L}, {L}]\nHorizon: 0.003\nHorizon_h_Ratio: 3\n"
        f"Mesh_Type: {kind}\n" + (f"Compress_Type: {compress}\n"
                                  if compress else ""))
    run_tool(host_bin / "nlm_mesh", "-i", "m.yaml", "-d", "2", cwd=tmp_path)
    m = read_vtu(str(tmp_path / "mesh.vtu"))
    x, vol = m["Points"], m["Node_Volume"]
    assert x.dtype == np.float32 and vol.dtype == np.float64
    assert abs(vol.sum() - L * L) <= 1e-12 * L * L
    if kind != "uniform_tri_sym":
        nx = 10
        j, i = np.divmod(np.arange((nx + 1) ** 2), nx + 1)
        assert np.array_equal(x[:, 0], (i * h).astype(np.float32))
        assert np.array_equal(x[:, 1], (j * h).astype(np.float32))
        nv = 4 if kind == "uniform_square" else 3
        assert m["connectivity"].size == nv * (nx * nx if nv == 4 else
                                               2 * nx * nx)
        assert np.all(m["types"] == (9 if nv == 4 else 5))
    else:
        assert x.shape[0] == 6 * 6 + 5 * 5
        assert m["connectivity"].size == 3 * 4 * 25


def test_deck_values_crack_propagation(host_bin):
    d = dump(host_bin, DECKS / "crack_propagation.yaml")
    assert d["model"]["dim"] == 2
    assert d["model"]["time"] == "central_difference"
    assert d["model"]["dt"] == 0.000140 / 35000
    assert d["model"]["h"] == 0.002 / 4
    assert d["material"]["type"] == "RNPBond"
    assert d["material"]["check_sc"] == 10.0
    assert d["material"]["E"] == 37500000000.0
    assert d["material"]["K"] == -1.0
    c = d["cracks"][0]
    assert c["o"] == -1 and c["pb"] == [5.0005e-2, 0.0]
    assert c["pt"] == [5.0005e-2, 2.0e-2] and c["activation"] == -1.0
    u = d["ubc"]
    assert len(u) == 3
    assert u[0]["region"] == "rectangle"
    assert (u[0]["x1"], u[0]["y1"], u[0]["x2"], u[0]["y2"]) == (0.0, 9.8e-2,
                                                                0.1, 0.1)
    assert u[0]["direction"] == [1, 2]
    assert u[1]["time_fn"] == "linear" and u[1]["time_params"] == [-1.0]
    assert u[2]["space_fn"] == "constant" and u[2]["space_params"] == []
    o = d["output"]
    assert o["tags"] == ["Displacement", "Velocity", "Force", "Damage_Z"]
    assert o["interval"] == 500 and o["compress"] == "zlib"
    assert o["perform_fe_out"] is False
    assert d["postprocessing"] is False


def test_deck_values_restart(host_bin):
    d = dump(host_bin, DECKS / "restart_input.yaml")
    assert d["model"]["time"] == "velocity_verlet"
    assert d["mesh"]["keep_element_conn"] is True
    f = d["fbc"][0]
    # Time_Function lists Parameters twice; the first one is kept
    assert f["time_params"] == [-5000000000.0]
    assert f["space_params"] == [1.0]
    assert d["model"]["restart"] is False
    r = dump(host_bin, DECKS / "restart_restart.yaml")
    assert r["model"]["restart"] is True
    assert r["restart"] == {"file": "./output_4.vtu", "step": 4}


def test_deck_values_quasi_static(host_bin):
    d = dump(host_bin, GOLDEN / "qs2d" / "input.yaml")
    assert d["model"]["time"] == "quasi_static"
    assert d["material"]["type"] == "ElasticState"
    assert (d["material"]["K"], d["material"]["G"]) == (4000.0, 1500.0)
    assert d["solver"] == {"type": "BiCGSTAB", "max_iter": 200, "tol": 1e-6,
                           "perturbation": 1e-7}
    assert d["ubc"][0]["region"] == "line"
    assert (d["ubc"][0]["x1"], d["ubc"][0]["x2"]) == (1.55, 1.65)
    assert d["fbc"][0]["space_params"] == [-250.0]


@pytest.mark.parametrize("edit,message", [
    (("  Density: 1200.000000 \n", ""),
     "Error: Please specify the density of the material."),
    (("  Time_Steps: 35000 \n", ""),
     "Error: Check Final_Time and Time_Steps data."),
    (("  File: mesh.vtu \n", ""), "Error: Please specify mesh filename."),
    (("      Line: [5.000500e-02, 0.000000e+00, 5.000500e-02, 2.000000e-02]\n",
      "      Line: [5.000500e-02, 0.000000e+00, 5.000500e-02]\n"),
     "Error: Input data in tag Line does not have enough elements."),
])
def test_deck_errors_exit_like_reference(host_bin, tmp_path, edit, message):
    text = (DECKS / "crack_propagation.yaml").read_text()
    assert edit[0] in text
    (tmp_path / "d.yaml").write_text(text.replace(edit[0], edit[1]))
    rc, out = run_tool(host_bin / "nlm_fd", "--dump-deck", "-i",
                       tmp_path / "d.yaml", check=False)
    assert rc == 1
    assert message in out


def test_yaml_subset_forms(host_bin, tmp_path):
    # block sequences at the key's indentation, quoted scalars, comments,
    # flow sequences split over lines
    (tmp_path / "d.yaml").write_text(
        "# comment\nModel:\n  Dimension: 2  # trailing comment\n"
        "  Discretization_Type: {Spatial: finite_difference, "
        "Time: central_difference}\n"
        "  Final_Time: 1.0e-6\n  Time_Steps: 10\n  Horizon: 0.003\n"
        "  Horizon_h_Ratio: 3\nMesh:\n  File: 'a b.vtu'\n"
        "Material:\n  Type: \"RNPBond\"\n  Density: 1200\n  E: 3.24e9\n"
        "  Gc: 500\n  Influence_Function:\n    Type: 0\n    Parameters:\n"
        "    - 1\nFracture:\n  Cracks:\n    Sets: 1\n    Set_1:\n"
        "      Orientation: 1\n      Line: [0.0, 0.005,\n"
        "             0.002, 0.005]\n"
        "Output:\n  Tags:\n  - Displacement\n  - Damage_Z\n"
        "  Output_Interval: 5\n")
    d = dump(host_bin, tmp_path / "d.yaml")
    assert d["mesh"]["file"] == "a b.vtu"
    assert d["material"]["influence_params"] == [1.0]
    assert d["cracks"][0]["pt"] == [0.002, 0.005]
    assert d["output"]["tags"] == ["Displacement", "Damage_Z"]
    assert d["model"]["time"] == "central_difference"


REF_EXAMPLES = Path("/root/reference/examples")


@pytest.mark.skipif(not REF_EXAMPLES.exists(),
                    reason="reference example decks not present")
def test_every_reference_example_deck_parses(host_bin):
    decks = [p for p in REF_EXAMPLES.rglob("*.yaml")
             if "Model" in p.read_text()]
    assert len(decks) >= 8
    for p in decks:
        d = dump(host_bin, p)
        assert d["model"]["dim"] in (1, 2, 3), p


# ------------------------------------------------------------ gmsh meshes
GMSH = GOLDEN / "gmsh"


def dump_mesh(host_bin, deck: Path):
    rc, out = run_tool(host_bin / "nlm_fd", "--dump-mesh", "-i", deck.name,
                       cwd=deck.parent)
    line = [l for l in out.splitlines() if l.startswith("{")][-1]
    d = json.loads(line)
    d["nodes"] = np.asarray(d["nodes"]).reshape(-1, 3)
    d["volumes"] = np.asarray(d["volumes"])
    d["elements"] = np.asarray(d["elements"], np.int64)
    return d


def gmsh_case(tmp_path, case):
    """The deck of case (its mesh copied beside it) and the mesh's domain
    area: examples/io covers [0, 5]^2; the logo is the polygon of logo.geo,
    whose shoelace area is 2.25."""
    if case == "logo":
        shutil.copy(GMSH / "logo_input.yaml", tmp_path / "input.yaml")
        shutil.copy(GMSH / "logo_mesh.msh", tmp_path / "mesh.msh")
        return tmp_path / "input.yaml", 2.25
    deck = f"input_2d_gmsh_{case}.yaml"
    shutil.copy(GMSH / deck, tmp_path / deck)
    shutil.copy(GMSH / f"mesh_2d_{case}.msh", tmp_path / f"mesh_2d_{case}.msh")
    return tmp_path / deck, 25.0


@pytest.mark.parametrize("case, nodes, elems, etype",
                         [("quad", 36, 25, 9), ("triangle", 45, 68, 5),
                          ("logo", 5107, 9486, 5)])
def test_gmsh_mesh_volumes(host_bin, tmp_path, case, nodes, elems, etype):
    deck, area = gmsh_case(tmp_path, case)
    d = dump_mesh(host_bin, deck)
    x, v, e = d["nodes"], d["volumes"], d["elements"]
    assert x.shape == (nodes, 3) and v.shape == (nodes,)
    assert d["element_type"] == etype
    nv = 3 if etype == 5 else 4
    assert e.size == nv * elems and e.min() >= 0 and e.max() < nodes
    assert abs(v.sum() - area) <= 1e-12 * area
    assert (v > 0).all()
    c = e.reshape(-1, nv)
    if etype == 5:
        a, b, q = x[c[:, 0], :2], x[c[:, 1], :2], x[c[:, 2], :2]
        area_e = 0.5 * np.abs((b[:, 0] - a[:, 0]) * (q[:, 1] - a[:, 1]) -
                              (q[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))
        ref = np.zeros(nodes)
        for k in range(3):
            np.add.at(ref, c[:, k], area_e / 3.0)
        assert np.allclose(v, ref, rtol=1e-12, atol=0)
    else:
        # unit squares: corner nodes 1/4, edge nodes 1/2, interior nodes 1
        on_x = np.isclose(x[:, 0], 0) | np.isclose(x[:, 0], 5)
        on_y = np.isclose(x[:, 1], 0) | np.isclose(x[:, 1], 5)
        ref = np.where(on_x & on_y, 0.25, np.where(on_x | on_y, 0.5, 1.0))
        assert np.allclose(v, ref, rtol=1e-10, atol=0)


def test_gmsh_refusals(host_bin, tmp_path):
    deck, _ = gmsh_case(tmp_path, "triangle")
    msh = tmp_path / "mesh_2d_triangle.msh"
    text = msh.read_text()
    msh.write_text(text.replace("4.1 0 8", "4.1 1 8", 1))
    rc, out = run_tool(host_bin / "nlm_fd", "--dump-mesh", "-i", deck.name,
                       cwd=tmp_path, check=False)
    assert rc != 0 and "binary" in out
    msh.write_text(text)
    deck.write_text(deck.read_text().replace("Dimension: 2", "Dimension: 3"))
    rc, out = run_tool(host_bin / "nlm_fd", "--dump-mesh", "-i", deck.name,
                       cwd=tmp_path, check=False)
    assert rc != 0 and "dimension 2" in out


def test_gmsh_malformed_counts(host_bin, tmp_path):
    """Counts read from a damaged gmsh file are checked before they size
    anything: a negative node-block size (format 4.1) and an absurd element
    tag count (format 2.2) end in a deck error, not an allocation failure or
    a run-away read loop."""
    deck, _ = gmsh_case(tmp_path, "triangle")
    msh = tmp_path / "mesh_2d_triangle.msh"
    lines = msh.read_text().split("\n")
    k = lines.index("$Nodes") + 2          # first node block's header
    head = lines[k].split()
    lines[k] = " ".join(head[:3] + ["-5"])
    msh.write_text("\n".join(lines))
    rc, out = run_tool(host_bin / "nlm_fd", "--dump-mesh", "-i", deck.name,
                       cwd=tmp_path, check=False)
    assert rc != 0 and "bad node block size" in out
    deck, _ = gmsh_case(tmp_path, "logo")
    msh = tmp_path / "mesh.msh"
    lines = msh.read_text().split("\n")
    k = lines.index("$Elements") + 2       # first element line
    e = lines[k].split()
    lines[k] = " ".join(e[:2] + ["2000000000"] + e[3:])
    msh.write_text("\n".join(lines))
    rc, out = run_tool(host_bin / "nlm_fd", "--dump-mesh", "-i", deck.name,
                       cwd=tmp_path, check=False)
    assert rc != 0 and "bad element tag count" in out

