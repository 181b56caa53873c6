"""The C++ host layer on the GPU: nlm_fd runs input decks end to end (mesh
file, deck, FDModel / QuasiStaticModel, .vtu output) as the reference's
NLMech executable does.

Bars:
  - a bond-based plate deck (pre-crack, pulled layers, central difference)
    matches the oracle given the same mesh, loading and crack: relative L2
    displacement and force errors <= 1e-9 (FP64 exact mode), broken bonds in
    the output's Bond_State identical to the oracle's, bit for bit;
  - the same deck with velocity_verlet matches the oracle's Verlet run;
  - restarting from an output file written with Bond_State continues
    exactly: the restarted run's final output equals the uninterrupted
    run's, bit for bit (displacement, velocity, force, bond bits);
  - stepping through a captured HIP graph (GPU: Graph_Chunk) after the
    joint autotuner (GPU: Autotune) writes the same
    output bit for bit;
  - a 3D state-based deck (ElasticState with Critical_Stretch, .csv mesh,
    cuboid displacement sets, explicit dynamics) matches the oracle: relative
    L2 displacement error <= 1e-9 and identical bond bits;
  - the post-processing tags (work done, strain / kinetic / fracture
    energies and totals, reaction force and its integer-accumulated total)
    equal the oracle's on the GPU's own state at each output;
  - the reference's quasi-static example deck (examples/qsModel/2D, state-
    based material, its mesh made by nlm_mesh) reproduces the reference-held
    result output_res.vtu within the reference test's own dc tolerance
    (displacement L2 < 1e-7, test/CMakeLists.txt:39-52); the internal force
    too;
  - gmsh meshes: the reference's mesh.2d.read.{quad,triangle}.gmsh CTests
    (examples/io decks, format 4.1, unchanged) run and write the nodal
    volumes of their meshes; the logo deck (examples/fdModel/logo, format
    2.2, shortened) matches the oracle on the same mesh;
  - the legacy VTK output (File_Format: legacy_vtk) of the reference's 1D
    quasi-static CTest equals the reference-held result to its printed
    precision.
"""
import shutil
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, progress, run_tool
from nlmech_amd import problems as pr
from vtu_read import read_vtu

pytestmark = pytest.mark.gpu

PLATE_MESH = """Output:
  Path: ./
  Mesh: mesh
Domain: [0.0, 0.0, 0.031, 0.031]
Horizon: 0.003
Horizon_h_Ratio: 3
Mesh_Type: uniform_square
Is_FD: true
Compress_Type: zlib
"""

PLATE_DECK = """Model:
  Dimension: 2
  Discretization_Type:
    Spatial: finite_difference
    Time: {time}
  Final_Time: 4.0e-6
  Time_Steps: 40
  Horizon: 0.003
  Horizon_h_Ratio: 3
Policy:
  Enable_PostProcessing: true
Mesh:
  File: {mesh}
Material:
  Type: RNPBond
  Density: 1200.0
  Compute_From_Classical: true
  E: 3.24e9
  Gc: 500.0
  Bond_Potential:
    Type: 1
    Check_Sc_Factor: 1.0
    Irreversible_Bond_Fracture: true
  Influence_Function:
    Type: 0
    Parameters: [1.0]
Fracture:
  Cracks:
    Sets: 1
    Set_1:
      Orientation: 1
      Line: [-0.0005, 0.0155, 0.0195, 0.0155]
Displacement_BC:
  Sets: 2
  Set_1:
    Location:
      Rectangle: [-0.001, 0.0285, 0.032, 0.0315]
    Direction: [2]
    Time_Function:
      Type: linear
      Parameters:
        - 20.0
    Spatial_Function:
      Type: constant
      Parameters: [1.0]
  Set_2:
    Location:
      Rectangle: [-0.001, -0.0005, 0.032, 0.0025]
    Direction: [2]
    Time_Function:
      Type: linear
      Parameters:
        - -20.0
    Spatial_Function:
      Type: constant
      Parameters: [1.0]
Output:
  Path: ./
  Tags:
    - Displacement
    - Velocity
    - Force
    - Damage_Z
    - Fixity
    - Node_Volume
    - Neighbors
  Output_Interval: 20
  Compress_Type: zlib
{extra}"""


def plate_dir(host_bin, root: Path, name: str, time="central_difference",
              extra="") -> Path:
    d = root / name
    d.mkdir()
    mesh = root / "mesh.vtu"
    if not mesh.exists():
        (root / "m.yaml").write_text(PLATE_MESH)
        run_tool(host_bin / "nlm_mesh", "-i", "m.yaml", "-d", "2", cwd=root)
    (d / "input.yaml").write_text(PLATE_DECK.format(time=time, mesh=mesh,
                                                    extra=extra))
    return d


def run_deck(host_bin, d: Path):
    progress(f"[gpu] nlm_fd {d.name}")
    rc, out = run_tool(host_bin / "nlm_fd", "-i", "input.yaml", cwd=d)
    return out


def plate_problem(root: Path) -> pr.Problem:
    """The deck's model as the oracle takes it (same mesh file, same BCs,
    crack and material)."""
    m = read_vtu(str(root / "mesh.vtu"))
    x = m["Points"].astype(np.float64)
    vol = m["Node_Volume"].astype(np.float64)
    ubc = [pr.BC("rectangle", [2], "linear", [20.0], "constant", [1.0],
                 x1=-0.001, x2=0.032, y1=0.0285, y2=0.0315),
           pr.BC("rectangle", [2], "linear", [-20.0], "constant", [1.0],
                 x1=-0.001, x2=0.032, y1=-0.0005, y2=0.0025)]
    cracks = [pr.CrackSpec(1, (-0.0005, 0.0155), (0.0195, 0.0155))]
    mat = pr.rnp_material(dim=2, E=3.24e9, Gc=500.0, density=1200.0,
                          check_sc=1.0, inf_type=0, inf_params=(1.0,))
    return pr.Problem(name="deck_plate", dim=2, x=x, vol=vol, horizon=0.003,
                      h=0.003 / 3.0, dt=4.0e-6 / 40, nsteps=40, material=mat,
                      ubc=ubc, cracks=cracks)


def pack_bits(row_ptr, bits):
    """Per-bond bits (CSR order) in the reference packing: ceil(n_i / 8)
    bytes per node (Fracture::d_fracture)."""
    out = []
    for i in range(len(row_ptr) - 1):
        b = bits[row_ptr[i]:row_ptr[i + 1]]
        nb = (len(b) + 7) // 8
        by = np.zeros(nb, np.uint8)
        for j, v in enumerate(b):
            if v:
                by[j // 8] |= np.uint8(1 << (j % 8))
        out.append(by)
    return np.concatenate(out) if out else np.zeros(0, np.uint8)


@pytest.mark.parametrize("scheme", ["central_difference", "velocity_verlet"])
def test_plate_deck_matches_oracle(host_bin, oracle, tmp_path, scheme):
    d = plate_dir(host_bin, tmp_path, "run", time=scheme)
    run_deck(host_bin, d)
    out = read_vtu(str(d / "output_2.vtu"))
    p = plate_problem(tmp_path)
    o = oracle.OracleFD(p)
    o.init()
    o.step(40, scheme="verlet" if scheme == "velocity_verlet" else "cd")
    u_ref, v_ref, f_ref = o.get_state()
    bits = o.get_bonds()
    bytes_ref = o.get_bond_bytes()
    rp = o.row_ptr
    # damage of the GPU state (postproc parity itself is covered elsewhere)
    o.set_state(u=out["Displacement"], v=out["Velocity"],
                f=out["Force"] / p.vol[:, None])
    Zr, _, _ = o.postproc()
    o.close()
    u = out["Displacement"]
    assert np.linalg.norm(u - u_ref) <= 1e-9 * np.linalg.norm(u_ref)
    F_ref = f_ref * p.vol[:, None]
    assert np.linalg.norm(out["Force"] - F_ref) <= 1e-9 * np.linalg.norm(F_ref)
    assert bits.sum() > 0
    assert np.array_equal(out["Bond_State"], bytes_ref)
    assert np.array_equal(bytes_ref, pack_bits(rp, bits))
    Z = out["Damage_Z"]
    assert np.max(np.abs(Z - Zr) / np.maximum(Zr, 1e-30)) <= 1e-6
    assert np.array_equal(out["Node_Volume"], p.vol)
    assert np.array_equal(out["Neighbors"], np.diff(rp).astype(np.float64))
    assert out["TIME"][0] == pytest.approx(40 * p.dt, rel=1e-12)
    assert (d / "output_0.vtu").exists() and (d / "output_1.vtu").exists()


def test_restart_with_bond_state_is_exact(host_bin, tmp_path):
    a = plate_dir(host_bin, tmp_path, "a")
    run_deck(host_bin, a)
    b = plate_dir(host_bin, tmp_path, "b",
                  extra=f"Restart:\n  File: {a / 'output_1.vtu'}\n  Step: 20\n")
    run_deck(host_bin, b)
    assert not (b / "output_1.vtu").exists()
    ra = read_vtu(str(a / "output_2.vtu"))
    rb = read_vtu(str(b / "output_2.vtu"))
    assert ra["Bond_State"].sum() > 0
    for k in ("Displacement", "Velocity", "Force", "Bond_State", "Points",
              "TIME"):
        assert np.array_equal(ra[k], rb[k]), k


def test_graph_stepping_deck_is_bitwise_equal(host_bin, tmp_path):
    a = plate_dir(host_bin, tmp_path, "a")
    run_deck(host_bin, a)
    # graph replays after the joint tile x chunk autotuner (the deck's
    # precision, FP64 exact, gives the same bits at every tile)
    g = plate_dir(host_bin, tmp_path, "g",
                  extra="GPU:\n  Graph_Chunk: 10\n  Autotune: true\n")
    out = run_deck(host_bin, g)
    assert "autotuner chose tiles of" in out
    ra = read_vtu(str(a / "output_2.vtu"))
    rg = read_vtu(str(g / "output_2.vtu"))
    for k in ("Displacement", "Velocity", "Force", "Bond_State"):
        assert np.array_equal(ra[k], rg[k]), k


def test_reference_quasi_static_deck(host_bin, tmp_path):
    src = GOLDEN / "qs2d"
    shutil.copy(src / "input_mesh.yaml", tmp_path / "input_mesh.yaml")
    shutil.copy(src / "input.yaml", tmp_path / "input.yaml")
    run_tool(host_bin / "nlm_mesh", "-i", "input_mesh.yaml", "-d", "2",
             cwd=tmp_path)
    out = run_deck(host_bin, tmp_path)
    progress("[gpu]   " + " | ".join(l for l in out.splitlines()
                                     if l.startswith("Load step")))
    ours = read_vtu(str(tmp_path / "output_1.vtu"))
    ref = read_vtu(str(src / "output_res.vtu"))
    du = np.linalg.norm(ours["Displacement"] - ref["Displacement"])
    dF = np.linalg.norm(ours["Force"] - ref["Force"])
    progress(f"[gpu]   displacement L2 {du:.3e}, force L2 {dF:.3e}")
    assert du < 1e-7
    assert dF < 1e-4 * np.linalg.norm(ref["Force"])
    (tmp_path / "compare.yaml").write_text(
        "Dimension: 2\nFilename_1: output_1.vtu\n"
        f"Filename_2: {src / 'output_res.vtu'}\n"
        "Out_Filename: compare.txt\nPrint_Screen: false\nTolerance: 1e-7\n"
        "Compare_Tags:\n    - Displacement\n")
    run_tool(host_bin / "nlm_dc", "-i", "compare.yaml", "-k", "fd_simple",
             cwd=tmp_path)


ES_DECK = """Model:
  Dimension: 3
  Discretization_Type:
    Spatial: finite_difference
    Time: central_difference
  Final_Time: 2.0e-4
  Time_Steps: 20
  Horizon: 0.1875
  Horizon_h_Ratio: 3
Mesh:
  File: mesh.csv
Material:
  Type: ElasticState
  Density: 1200.0
  Compute_From_Classical: true
  K: 2.16e9
  G: 1.30e9
  Critical_Stretch: 1.0e-3
Displacement_BC:
  Sets: 2
  Set_1:
    Location:
      Cuboid: [-0.0625, -0.0625, {ztop1}, {xmax}, {xmax}, {ztop2}]
    Direction: [3]
    Time_Function:
      Type: linear
      Parameters: [1.0]
    Spatial_Function:
      Type: constant
      Parameters: [1.0]
  Set_2:
    Location:
      Cuboid: [-0.0625, -0.0625, -0.03125, {xmax}, {xmax}, {zbot2}]
    Direction: [3]
    Time_Function:
      Type: linear
      Parameters: [-1.0]
    Spatial_Function:
      Type: constant
      Parameters: [1.0]
Output:
  Path: ./
  Tags:
    - Displacement
    - Velocity
    - Force
  Output_Interval: 20
"""


def test_state_based_3d_deck_matches_oracle(host_bin, oracle, tmp_path):
    """FDModel<ElasticState> (explicit, with the critical-stretch breaking)
    from a deck with a 3D .csv mesh, against the oracle."""
    nx, h = 10, 0.0625
    x, vol = pr.grid3d(nx, nx, nx, h, jitter=0.0)
    with open(tmp_path / "mesh.csv", "w") as f:
        f.write("id,x,y,z,volume\n")
        for i in range(x.shape[0]):
            f.write(f"{i},{x[i, 0]:.17g},{x[i, 1]:.17g},{x[i, 2]:.17g},"
                    f"{vol[i]:.17g}\n")
    zmax = (nx - 1) * h
    xmax = zmax + h
    ztop1, ztop2 = zmax - 2 * h - 0.5 * h, zmax + 0.5 * h
    zbot2 = 2 * h + 0.5 * h
    (tmp_path / "input.yaml").write_text(ES_DECK.format(
        ztop1=f"{ztop1:.17g}", ztop2=f"{ztop2:.17g}", zbot2=f"{zbot2:.17g}",
        xmax=f"{xmax:.17g}"))
    run_deck(host_bin, tmp_path)
    out = read_vtu(str(tmp_path / "output_1.vtu"))
    box = dict(x1=-0.0625, x2=xmax, y1=-0.0625, y2=xmax)
    ubc = [pr.BC("cuboid", [3], "linear", [1.0], "constant", [1.0],
                 z1=ztop1, z2=ztop2, **box),
           pr.BC("cuboid", [3], "linear", [-1.0], "constant", [1.0],
                 z1=-0.03125, z2=zbot2, **box)]
    p = pr.Problem(name="deck_cube", dim=3, x=x, vol=vol, horizon=0.1875,
                   h=0.1875 / 3.0, dt=2.0e-4 / 20, nsteps=20,
                   material=pr.es_material(dim=3, K=2.16e9, G=1.30e9,
                                           density=1200.0, sc=1.0e-3),
                   ubc=ubc)
    o = oracle.OracleES(p)
    o.init()
    o.step(20)
    u_ref, _, _ = o.get_state()
    bits = o.get_bonds()
    rp = o.row_ptr
    o.close()
    u = out["Displacement"]
    progress(f"[gpu]   broken bonds {int(bits.sum())}")
    assert np.abs(u_ref[:, 2]).max() > 0
    assert np.linalg.norm(u - u_ref) <= 1e-9 * np.linalg.norm(u_ref)
    assert np.array_equal(out["Bond_State"], pack_bits(rp, bits))


ENERGY_TAGS = """    - Neighbors
    - Strain_Energy
    - Work_Done
    - Reaction_Force
    - Total_Reaction_Force
    - Fracture_Perienergy_Bond
    - Fracture_Perienergy_Total
    - Total_Energy
    - Total_Fracture_Perienergy_Bond
    - Total_Fracture_Perienergy_Total
"""

FORCE_BC = """Force_BC:
  Sets: 1
  Set_1:
    Location:
      Rectangle: [-0.001, -0.001, 0.0105, 0.032]
    Direction: [1, 2]
    Time_Function:
      Type: linear
      Parameters: [2.5e10]
    Spatial_Function:
      Type: constant
      Parameters: [4.0]
"""


def unpack_bits(row_ptr, by):
    bits = np.zeros(int(row_ptr[-1]), np.uint8)
    off = 0
    for i in range(len(row_ptr) - 1):
        m = int(row_ptr[i + 1] - row_ptr[i])
        for jl in range(m):
            bits[row_ptr[i] + jl] = (by[off + jl // 8] >> (jl % 8)) & 1
        off += (m + 7) // 8
    return bits


def test_postprocessing_tags_match_oracle(host_bin, oracle, tmp_path):
    """Work done, kinetic, strain and fracture energies, reaction force and
    their totals as model::Output writes them (model/util.cpp:74-164), from
    computePostProcFields (fDModel.cpp:652-778) and computeForce's reaction
    force (fDModel.cpp:548-605), against the oracle on the GPU's own state
    at both outputs (fracture energies keep their last value where Z has
    dropped below the threshold).  Total_Reaction_Force accumulates in int
    as the reference's std::accumulate(..., 0) does."""
    d = tmp_path / "run"
    d.mkdir()
    (tmp_path / "m.yaml").write_text(PLATE_MESH)
    run_tool(host_bin / "nlm_mesh", "-i", "m.yaml", "-d", "2", cwd=tmp_path)
    deck = PLATE_DECK.format(time="central_difference",
                             mesh=tmp_path / "mesh.vtu", extra=FORCE_BC)
    deck = deck.replace("    - Neighbors\n", ENERGY_TAGS)
    (d / "input.yaml").write_text(deck)
    run_deck(host_bin, d)
    p = plate_problem(tmp_path)
    p = pr.with_(p, fbc=[pr.BC("rectangle", [1, 2], "linear", [2.5e10],
                               "constant", [4.0], x1=-0.001, x2=0.0105,
                               y1=-0.001, y2=0.032)])
    o = oracle.OracleFD(p)
    rp = o.row_ptr
    x = p.x
    box = (min(0.0, x[:, 0].min()), x[:, 0].max(), x[:, 1].max())
    fE = np.zeros(p.n)
    for k, steps in ((1, 20), (2, 40)):
        out = read_vtu(str(d / f"output_{k}.vtu"))
        u, v = out["Displacement"], out["Velocity"]
        o.set_state(u=u, v=v, f=np.zeros_like(u))
        o.set_bonds(unpack_bits(rp, out["Bond_State"]))
        Z, _, e = o.postproc()
        t = steps * p.dt
        W = np.einsum("ij,ij->i", u, o.external_force(t))
        KE = 0.5 * 1200.0 * np.einsum("ij,ij->i", v, v) * p.vol
        # definitelyGreaterThan(Z, 1 - 1e-10) on the stored (float) Z
        hit = np.array([oracle.definitely_greater(float(z), 1.0 - 1.0e-10)
                        for z in Z])
        fE = np.where(hit, e, fE)
        R, L = o.reaction_force(box, p.horizon)
        acc = 0
        for val in L:
            acc = int(acc + val)
        area = (abs(box[1] - box[0]) - p.horizon) * \
            (abs(box[2] - 0.0) - p.horizon)
        assert np.abs(W).max() > 0
        np.testing.assert_allclose(out["Work_Done"], W, rtol=1e-12,
                                   atol=1e-12 * np.abs(W).max())
        np.testing.assert_allclose(out["Strain_Energy"], e, rtol=1e-12,
                                   atol=1e-300)
        np.testing.assert_allclose(out["Fracture_Perienergy_Total"], fE,
                                   rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(out["Fracture_Perienergy_Bond"], fE,
                                   rtol=1e-12, atol=1e-300)
        assert out["Total_Energy"][0] == pytest.approx(
            e.sum() - W.sum() + KE.sum(), rel=1e-10)
        assert out["Total_Fracture_Perienergy_Total"][0] == pytest.approx(
            fE.sum(), rel=1e-10, abs=1e-300)
        sr = max(np.abs(R).max(), 1e-300)
        assert np.abs(out["Reaction_Force"] - R).max() <= 1e-12 * sr
        assert out["Total_Reaction_Force"][0] == pytest.approx(acc * area,
                                                               rel=1e-12)
    assert (fE > 0).sum() > 0
    o.close()


def _copy_decks(tmp_path, pairs):
    src = GOLDEN / "decks"
    for a, b in pairs:
        shutil.copy(src / a, tmp_path / b)


def test_reference_restart_ctest(host_bin, oracle, tmp_path):
    """explicit.restart.{mesh,run,restart,compare} (test/CMakeLists.txt:
    103-133) with the reference's own decks (examples/fdModel/restart):
    uniform triangle mesh, 5 velocity-Verlet steps under a linear force
    ramp, restart from output_4.vtu at step 4, dc Displacement L2 < 1e-7
    between output_5.vtu and restart_output_5.vtu.  The full run is also
    checked against the oracle on the same mesh (linear influence function,
    Check_Sc_Factor 10, force BC), and the restarted output equals the
    uninterrupted one bit for bit (outputs carry the broken-bond bits)."""
    _copy_decks(tmp_path, (("restart_input_mesh.yaml", "input_mesh.yaml"),
                           ("restart_input.yaml", "input.yaml"),
                           ("restart_restart.yaml", "restart.yaml"),
                           ("restart_input_compare.yaml",
                            "input_compare.yaml")))
    progress("[gpu] reference restart decks")
    run_tool(host_bin / "nlm_mesh", "-i", "input_mesh.yaml", "-d", "2",
             cwd=tmp_path)
    run_tool(host_bin / "nlm_fd", "-i", "input.yaml", cwd=tmp_path)
    run_tool(host_bin / "nlm_fd", "-i", "restart.yaml", cwd=tmp_path)
    rc, _ = run_tool(host_bin / "nlm_dc", "-i", "input_compare.yaml", "-k",
                     "fd_simple", cwd=tmp_path, check=False)
    lines = (tmp_path / "compare.txt").read_text().splitlines()
    progress(f"[gpu]   dc: {lines[-1]}")
    l2 = float(lines[-1].split(",")[0])
    assert rc == 0 and l2 < 1e-7
    full = read_vtu(str(tmp_path / "output_5.vtu"))
    rest = read_vtu(str(tmp_path / "restart_output_5.vtu"))
    for k in ("Displacement", "Velocity", "Force", "Bond_State"):
        assert np.array_equal(full[k], rest[k]), k
    # the full run against the oracle: the same deck with Node_Volume added
    # to the output tags (the nodal volumes the GPU computed from the
    # triangles), run beside it
    vdir = tmp_path / "vol"
    vdir.mkdir()
    deck = (tmp_path / "input.yaml").read_text()
    deck = deck.replace("File: mesh.vtu", f"File: {tmp_path / 'mesh.vtu'}")
    deck = deck.replace("    - Damage_Z\n", "    - Damage_Z\n    - Node_Volume\n")
    assert "Node_Volume" in deck
    (vdir / "input.yaml").write_text(deck)
    run_tool(host_bin / "nlm_fd", "-i", "input.yaml", cwd=vdir)
    tagged = read_vtu(str(vdir / "output_5.vtu"))
    assert np.array_equal(tagged["Displacement"], full["Displacement"])
    m = read_vtu(str(tmp_path / "mesh.vtu"))
    x = m["Points"].astype(np.float64)
    vol = tagged["Node_Volume"]
    fbc = [pr.BC("rectangle", [2], "linear", [-5.0e9], "constant", [1.0],
                 x1=0.0, y1=0.0, x2=0.1, y2=0.002)]
    mat = pr.rnp_material(dim=2, E=3.24e9, Gc=500.0, density=1200.0,
                          check_sc=10.0, inf_type=1, inf_params=())
    p = pr.Problem(name="restart", dim=2, x=x, vol=np.asarray(vol, float),
                   horizon=0.002, h=0.002 / 4, dt=1e-7 / 5, nsteps=5,
                   material=mat, fbc=fbc)
    o = oracle.OracleFD(p)
    o.init()
    o.step(5, scheme="verlet")
    u_ref, v_ref, f_ref = o.get_state()
    bytes_ref = o.get_bond_bytes()
    o.close()
    u = full["Displacement"]
    err = np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)
    progress(f"[gpu]   full run vs oracle: rel L2 {err:.2e}")
    assert err <= 1e-9
    assert np.array_equal(full["Bond_State"], bytes_ref)


def test_reference_crack_propagation_deck(host_bin, oracle, tmp_path):
    """examples/fdModel/crack_propagation run by nlm_fd, shortened to its
    first 500 of 35,000 steps (same dt = 4e-9 s), outputs at steps 250 and
    500 (Node_Volume added to the tags, which the output writes with post-
    processing on, as model/util.cpp does), checked against the oracle at both
    outputs: relative L2 displacement error <= 1e-9 and identical broken-
    bond bytes.  The deck exercises the uniform triangle mesh, the linear
    influence function without parameters, Check_Sc_Factor 10, a vertical
    (Orientation -1) pre-crack, a fixed top layer and bottom layers pulled
    apart at -/+1 m/s, and the 'constant' spatial function without
    parameters (multiplier 1; the reference reads params[0] there,
    uLoading.cpp:206-208)."""
    _copy_decks(tmp_path, (("crack_propagation_mesh.yaml", "input_mesh.yaml"),
                           ("crack_propagation.yaml", "input.yaml")))
    deck = (tmp_path / "input.yaml").read_text()
    deck = (deck.replace("Final_Time: 0.000140", "Final_Time: 0.000002")
                .replace("Time_Steps: 35000", "Time_Steps: 500")
                .replace("Output_Interval: 500", "Output_Interval: 250")
                .replace("Enable_PostProcessing: false",
                         "Enable_PostProcessing: true")
                .replace("    - Damage_Z\n", "    - Damage_Z\n    - Node_Volume\n"))
    assert "Time_Steps: 500 " in deck or "Time_Steps: 500\n" in deck
    (tmp_path / "input.yaml").write_text(deck)
    progress("[gpu] reference crack_propagation deck (500 steps)")
    run_tool(host_bin / "nlm_mesh", "-i", "input_mesh.yaml", "-d", "2",
             cwd=tmp_path)
    run_tool(host_bin / "nlm_fd", "-i", "input.yaml", cwd=tmp_path)
    outs = [read_vtu(str(tmp_path / f"output_{k}.vtu")) for k in (1, 2)]
    m = read_vtu(str(tmp_path / "mesh.vtu"))
    x = m["Points"].astype(np.float64)
    vol = np.asarray(outs[0]["Node_Volume"], np.float64)
    ubc = [pr.BC("rectangle", [1, 2], "constant", [0.0], "constant", [1.0],
                 x1=0.0, y1=0.098, x2=0.1, y2=0.1),
           pr.BC("rectangle", [1], "linear", [-1.0], "constant", [1.0],
                 x1=0.0, y1=0.0, x2=0.05, y2=0.002),
           pr.BC("rectangle", [1], "linear", [1.0], "constant", [1.0],
                 x1=0.050125, y1=0.0, x2=0.1, y2=0.002)]
    cracks = [pr.CrackSpec(-1, (0.050005, 0.0), (0.050005, 0.02))]
    mat = pr.rnp_material(dim=2, E=3.75e10, Gc=500.0, density=1200.0,
                          check_sc=10.0, inf_type=1, inf_params=())
    p = pr.Problem(name="crack_propagation", dim=2, x=x, vol=vol,
                   horizon=0.002, h=0.002 / 4, dt=0.000002 / 500,
                   nsteps=500, material=mat, ubc=ubc, cracks=cracks)
    o = oracle.OracleFD(p)
    o.init()
    assert o.get_bonds().sum() > 0          # the pre-crack
    for k, out in enumerate(outs):
        o.step(250)
        u_ref, _, _ = o.get_state()
        err = np.linalg.norm(out["Displacement"] - u_ref) / \
            np.linalg.norm(u_ref)
        progress(f"[gpu]   step {250 * (k + 1)}: rel L2 {err:.2e}, "
                 f"{int(o.get_bonds().sum())} broken bonds")
        assert err <= 1e-9
        assert np.array_equal(out["Bond_State"], o.get_bond_bytes())
    o.close()


@pytest.mark.parametrize("case", ["qs1d", "qs2d"])
def test_reference_quasi_static_ctest(host_bin, tmp_path, case):
    """quasistatic.{1D,2D}.elastic.{mesh,run,compare} (test/CMakeLists.txt:
    11-54) with the reference's own decks, unchanged: nlm_mesh -d 1|2,
    nlm_fd -i input.yaml, then nlm_dc -i input_compare.yaml -k fd_simple,
    which compares Displacement, Force, Strain_Energy, Strain_Tensor and
    Stress_Tensor against the reference-held output_res.vtu with L2 < 1e-7
    (its exit status says whether every tag passed)."""
    src = GOLDEN / case
    for f in ("input.yaml", "input_mesh.yaml", "input_compare.yaml",
              "output_res.vtu"):
        shutil.copy(src / f, tmp_path / f)
    progress(f"[gpu] reference quasi-static CTest {case}")
    run_tool(host_bin / "nlm_mesh", "-i", "input_mesh.yaml", "-d",
             "1" if case == "qs1d" else "2", cwd=tmp_path)
    out = run_deck(host_bin, tmp_path)
    progress("[gpu]   " + " | ".join(l for l in out.splitlines()
                                     if l.startswith("Load step")))
    rc, dc = run_tool(host_bin / "nlm_dc", "-i", "input_compare.yaml", "-k",
                      "fd_simple", cwd=tmp_path, check=False)
    lines = (tmp_path / "compare.txt").read_text().splitlines()
    progress(f"[gpu]   dc: {lines[-1]}")
    vals = [float(v) for v in lines[-1].split(",")]
    l2 = vals[0::2]
    assert len(l2) == 5 and max(l2) < 1e-7, l2
    assert rc == 0


def read_legacy_vtk(path):
    """Points and point data of an ASCII legacy .vtk unstructured grid:
    {"Points": (n, 3), name: (n,) / (n, k) for SCALARS, (n, 3, 3) for
    TENSORS}."""
    lines = Path(path).read_text().splitlines()
    out, n, i = {}, 0, 0
    while i < len(lines):
        t = lines[i].split()
        if t and t[0] == "POINTS":
            n = int(t[1])
            out["Points"] = np.loadtxt(lines[i + 1:i + 1 + n], ndmin=2)
            i += 1 + n
        elif t and t[0] == "SCALARS":
            a = np.loadtxt(lines[i + 2:i + 2 + n], ndmin=2)
            out[t[1]] = a[:, 0] if int(t[3]) == 1 else a
            i += 2 + n
        elif t and t[0] == "TENSORS":
            out[t[1]] = np.loadtxt(lines[i + 1:i + 1 + 3 * n],
                                   ndmin=2).reshape(n, 3, 3)
            i += 1 + 3 * n
        else:
            i += 1
    return out


def test_reference_legacy_vtk_ctest(host_bin, tmp_path):
    """quasistatic.1D.elastic_legacy_vtk (test/CMakeLists.txt:88-91): the
    1D quasi-static deck with File_Format: legacy_vtk, unchanged, writes
    output_1.vtk in the ASCII legacy format (points x + u, SCALARS, TENSORS
    for the stress and strain tensors, numbers at the stream's default 6
    significant digits), whose fields equal the reference-held
    output_res.vtu to that precision: Displacement, Force, Strain_Energy,
    the tensors (the .vtu's (xx, yy, zz, yz, xz, xy) as full matrices),
    Fixity and Neighbors."""
    src = GOLDEN / "qs1d"
    for f in ("input_mesh.yaml", "input_legacy_vtk.yaml", "output_res.vtu"):
        shutil.copy(src / f, tmp_path / f)
    progress("[gpu] reference legacy vtk CTest (qs1d)")
    run_tool(host_bin / "nlm_mesh", "-i", "input_mesh.yaml", "-d", "1",
             cwd=tmp_path)
    run_tool(host_bin / "nlm_fd", "-i", "input_legacy_vtk.yaml", cwd=tmp_path)
    assert not list(tmp_path.glob("output_[0-9]*.vtu"))
    got = read_legacy_vtk(tmp_path / "output_1.vtk")
    ref = read_vtu(str(tmp_path / "output_res.vtu"))

    def close(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return np.allclose(a, b, rtol=1e-5, atol=1e-5 * np.abs(b).max())

    # (both files hold the deformed positions x + u)
    assert close(got["Points"], ref["Points"])
    for tag in ("Displacement", "Force"):
        assert close(got[tag], ref[tag]), tag
    assert close(got["Strain_Energy"], ref["Strain_Energy"])
    for tag in ("Strain_Tensor", "Stress_Tensor"):
        v = np.asarray(ref[tag], np.float64).reshape(-1, 6)
        m = v[:, [0, 5, 4, 5, 1, 3, 4, 3, 2]].reshape(-1, 3, 3)
        assert close(got[tag], m), tag
    assert np.array_equal(got["Fixity"], np.asarray(ref["Fixity"]))
    assert np.array_equal(got["Neighbors"], np.asarray(ref["Neighbors"]))


# ------------------------------------------------------------ gmsh meshes
GMSH = GOLDEN / "gmsh"


@pytest.mark.parametrize("case", ["quad", "triangle"])
def test_reference_gmsh_read_ctest(host_bin, tmp_path, case):
    """mesh.2d.read.{quad,triangle}.gmsh (test/CMakeLists.txt:72-82): nlm_fd
    -i input_2d_gmsh_<case>.yaml on the reference's deck and gmsh 4.1 mesh,
    unchanged (a state-based quasi-static model without loads, one load
    step, Node_Volume output): the run succeeds and the written volumes sum
    to the 5 x 5 domain's area."""
    deck = f"input_2d_gmsh_{case}.yaml"
    shutil.copy(GMSH / deck, tmp_path / deck)
    shutil.copy(GMSH / f"mesh_2d_{case}.msh", tmp_path / f"mesh_2d_{case}.msh")
    progress(f"[gpu] reference gmsh {case} deck")
    run_tool(host_bin / "nlm_fd", "-i", deck, cwd=tmp_path)
    outs = sorted(tmp_path.glob("output_*.vtu"))
    assert outs
    v = np.asarray(read_vtu(str(outs[-1]))["Node_Volume"], np.float64)
    assert v.size == (36 if case == "quad" else 45)
    assert abs(v.sum() - 25.0) <= 1e-12 * 25.0


def read_msh22(path):
    """Nodes and 3-node triangles of a gmsh 2.2 ASCII file (node tags 1..N,
    the triangles in file order)."""
    lines = Path(path).read_text().splitlines()
    i = lines.index("$Nodes")
    n = int(lines[i + 1])
    x = np.zeros((n, 3))
    for ln in lines[i + 2:i + 2 + n]:
        t = ln.split()
        x[int(t[0]) - 1] = [float(t[1]), float(t[2]), float(t[3])]
    j = lines.index("$Elements")
    tri = []
    for ln in lines[j + 2:j + 2 + int(lines[j + 1])]:
        t = [int(a) for a in ln.split()]
        if t[1] == 2:
            tri.append([a - 1 for a in t[3 + t[2]:]])
    return x, np.asarray(tri, np.int64)


def test_reference_logo_deck(host_bin, oracle, tmp_path):
    """examples/fdModel/logo (its gmsh 2.2 mesh of 5107 nodes and 9486
    triangles, PDBond = RNPBond, linear influence function, Check_Sc_Factor
    10, a body force 1.5e9 t sin(1.090909 x) in x over the whole logo)
    run by nlm_fd, shortened to its first 300 of 50,000 steps (same dt =
    2e-7 s), outputs at steps 150 and 300 written to ./ with Node_Volume
    added: nodal volumes equal the triangles' areas split in thirds, and the
    displacements match the oracle's on the same mesh (relative L2 <= 1e-9)
    with identical broken-bond bytes."""
    shutil.copy(GMSH / "logo_input.yaml", tmp_path / "input.yaml")
    shutil.copy(GMSH / "logo_mesh.msh", tmp_path / "mesh.msh")
    deck = (tmp_path / "input.yaml").read_text()
    deck = (deck.replace("Final_Time: 0.010000", "Final_Time: 0.000060")
                .replace("Time_Steps: 50000", "Time_Steps: 300")
                .replace("Output_Interval: 500", "Output_Interval: 150")
                .replace("Path: ../out/", "Path: ./")
                .replace("    - Force\n", "    - Force\n    - Node_Volume\n"))
    assert "Time_Steps: 300\n" in deck and "Node_Volume" in deck
    (tmp_path / "input.yaml").write_text(deck)
    progress("[gpu] reference logo deck (300 steps)")
    run_tool(host_bin / "nlm_fd", "-i", "input.yaml", cwd=tmp_path)
    outs = [read_vtu(str(tmp_path / f"output_{k}.vtu")) for k in (1, 2)]
    x, tri = read_msh22(tmp_path / "mesh.msh")
    a, b, c = x[tri[:, 0], :2], x[tri[:, 1], :2], x[tri[:, 2], :2]
    area = 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) -
                        (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1]))
    vol = np.zeros(x.shape[0])
    for k in range(3):
        np.add.at(vol, tri[:, k], area / 3.0)
    assert np.allclose(np.asarray(outs[0]["Node_Volume"], np.float64), vol,
                       rtol=1e-12, atol=0)
    fbc = [pr.BC("rectangle", [1], "linear", [1.5e9], "sin_x", [1.090909],
                 x1=0.0, y1=0.0, x2=2.75, y2=1.5)]
    mat = pr.rnp_material(dim=2, E=88000.0, Gc=999.802380,
                          density=1011.203677, check_sc=10.0, inf_type=1,
                          inf_params=())
    p = pr.Problem(name="logo", dim=2, x=x, vol=vol, horizon=0.1,
                   h=0.1 / 4, dt=0.000060 / 300, nsteps=300, material=mat,
                   fbc=fbc)
    o = oracle.OracleFD(p)
    o.init()
    for k, out in enumerate(outs):
        o.step(150)
        u_ref, _, _ = o.get_state()
        err = np.linalg.norm(out["Displacement"] - u_ref) / \
            np.linalg.norm(u_ref)
        progress(f"[gpu]   step {150 * (k + 1)}: rel L2 {err:.2e}, "
                 f"{int(o.get_bonds().sum())} broken bonds")
        assert err <= 1e-9
        assert np.array_equal(out["Bond_State"], o.get_bond_bytes())
    o.close()

