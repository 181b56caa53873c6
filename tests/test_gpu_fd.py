"""GPU parity of the explicit FD path against the CPU oracle.

Bars (all through the C-ABI, libnlmech_amd.so):
  - neighbor lists: identical CSR (sorted rows) -- integer, bit-exact;
  - bond bits (pre-crack and every break decision): bit-exact;
  - per-node force from an identical state, FP64 modes:
        |f_gpu - f_ref|_inf <= 1e-12 * sum_j |F_ij^ref|   (per node)
    FP32-force mode: <= 1e-5 * sum_j |F_ij^ref| (FORCE_TOL);
  - trajectories, FP64 modes: relative L2 displacement error <= 1e-9 and
    identical broken-bond sets; FP32-force mode: <= 1e-5 and < 0.1 %
    broken-bond mismatch.
Device exp differs from glibc exp in the last ulp, so FP64 trajectories are
compared with a tolerance, not bitwise.
"""
import numpy as np
import pytest

from conftest import progress
from nlmech_amd import problems as pr
from nlmech_amd.capi import (NLM_PREC_FP32_FORCE, NLM_PREC_FP64,
                             NLM_PREC_FP64_EXACT, NLM_TIME_VELOCITY_VERLET)

pytestmark = pytest.mark.gpu

MODES = [NLM_PREC_FP64_EXACT, NLM_PREC_FP64, NLM_PREC_FP32_FORCE]
# FP32-force mode: a bond's position difference comes from tile-relative
# FP32 coordinates (within ~25 h of their origin), each rounded by at most
# 2^-24 of its size: up to 2^-23 * 25 h / r relative, 3e-6 for the shortest
# bonds (r >= h).  With ~15 FP32 roundings and the rsq / exp2 approximations
# per bond force, the per-node bar is the mode's accuracy target, 1e-5 of
# sum_j |F_ij| (five random states of the force-parity test read 0.5-4.6e-6
# at their worst node).
FORCE_TOL = {NLM_PREC_FP64_EXACT: 1e-12, NLM_PREC_FP64: 1e-12,
             NLM_PREC_FP32_FORCE: 1e-5}


@pytest.fixture(scope="module")
def engine():
    from nlmech_amd import engine
    return engine


def rel_l2(a, b):
    d = np.linalg.norm(a - b)
    s = np.linalg.norm(b)
    return d / s if s > 0 else d


def force_err(f_gpu, f_ref, absF):
    d = np.abs(f_gpu - f_ref).max(axis=1)
    scale = np.where(absF > 0, absF, 1.0)
    return (d / scale).max()


# ------------------------------------------------------------ neighbors
@pytest.mark.parametrize("case", ["2d_3h", "2d_4h", "3d_3h", "cloud"])
def test_neighbor_lists_match_oracle(engine, oracle, case):
    progress(f"[gpu] neighbor lists {case}")
    if case.startswith("2d"):
        ratio = 3 if case == "2d_3h" else 4
        p = pr.plate2d(61, 47, 1e-3, ratio, 1e-7, 1, notch_len=0.0, pull=0.0)
        x, horizon, dim = p.x, p.horizon, 2
    elif case == "3d_3h":
        g = np.arange(13) * 1e-3
        z, y, xx = np.meshgrid(g, g, g, indexing="ij")
        x = np.stack([xx.ravel(), y.ravel(), z.ravel()], 1)
        x += pr.uniform01(np.arange(x.size, dtype=np.uint64), 11).reshape(x.shape) * 1e-9
        horizon, dim = 3e-3, 3
    else:
        rng = np.random.default_rng(5)
        x = np.zeros((3000, 3))
        x[:, :2] = rng.random((3000, 2))
        horizon, dim = 0.04, 2
    rp_ref, col_ref = oracle.neighbors(x, horizon, "brute")
    lib = engine.capi.load()
    import ctypes as C
    h = C.c_void_p()
    assert lib.nlm_ctx_create(C.byref(h), 0) == 0
    try:
        xx = np.ascontiguousarray(x)
        nnz = C.c_int64()
        rc = lib.nlm_neighbor_build(h, dim, x.shape[0],
                                    engine.capi.ptr(xx, C.c_double), horizon,
                                    C.byref(nnz))
        assert rc == 0, lib.nlm_last_error(h)
        rp = np.zeros(x.shape[0] + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        assert lib.nlm_neighbor_get(h, engine.capi.ptr(rp, C.c_int64),
                                    engine.capi.ptr(col, C.c_int32)) == 0
    finally:
        lib.nlm_ctx_destroy(h)
    assert np.array_equal(rp, rp_ref)
    assert np.array_equal(col, col_ref)


# ------------------------------------------------------------ setup
def test_precrack_bits_and_packing_match(engine, oracle):
    progress("[gpu] pre-crack bits")
    p = pr.plate2d(40, 40, 1e-3, 3, 1e-7, 1)
    o = oracle.OracleFD(p)
    with engine.GpuFD(p) as g:
        rp, col = g.neighbors()
        assert np.array_equal(rp, o.row_ptr) and np.array_equal(col, o.col)
        assert np.array_equal(g.get_bonds(), o.get_bonds())
        assert np.array_equal(g.get_bond_bytes(), o.get_bond_bytes())
        assert np.array_equal(g.get_fixity(), o.get_fixity())
        assert o.get_bonds().sum() > 0


def test_angled_and_vertical_cracks(engine, oracle):
    progress("[gpu] angled/vertical cracks")
    p = pr.plate2d(30, 30, 1e-3, 3, 1e-7, 1, notch_len=0.0, pull=0.0)
    p.cracks = [pr.CrackSpec(-1, (0.0145, -0.001), (0.0145, 0.012)),
                pr.CrackSpec(0, (0.003, 0.0205), (0.02, 0.0265), theta=0.34)]
    o = oracle.OracleFD(p)
    with engine.GpuFD(p) as g:
        assert np.array_equal(g.get_bonds(), o.get_bonds())
        assert o.get_bonds().sum() > 0


# ------------------------------------------------------------ forces
def _deformed_state(p, seed, sigma):
    """Random displacements of standard deviation sigma [m]."""
    rng = np.random.default_rng(seed)
    u = np.zeros((p.n, 3))
    u[:, :2] = rng.normal(scale=sigma, size=(p.n, 2))
    v = np.zeros((p.n, 3))
    v[:, :2] = rng.normal(scale=1e-2, size=(p.n, 2))
    return u, v


@pytest.mark.parametrize("mode", MODES)
def test_force_parity_identical_state(engine, oracle, mode):
    """One computeForces from the same (u, bits): per-node forces within the
    condition-aware bar and identical break decisions, with bonds near and
    across the break threshold."""
    progress(f"[gpu] force parity mode={mode}")
    p = pr.plate2d(50, 50, 1e-3, 3, 1e-7, 1)
    o = oracle.OracleFD(p)
    # bond stretches ~ sqrt(2) sigma / r ~ the critical stretch rbar/sqrt(r)
    u, v = _deformed_state(p, 1, sigma=4e-6)
    with engine.GpuFD(p, precision=mode) as g:
        f0 = np.zeros((p.n, 3))
        g.set_state(u=u, v=v, f=f0)
        o.set_state(u=u, v=v, f=f0)
        absF = o.force_abs_sum()
        o.compute_forces()
        g.compute_forces()
        _, _, fr = o.get_state()
        _, _, fg = g.get_state()
        br, bg = o.get_bonds(), g.get_bonds()
        assert br.sum() > 2000  # the state breaks bonds beyond the notch
        assert np.array_equal(bg, br)
        assert force_err(fg, fr, absF) <= FORCE_TOL[mode]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("inf", [1, 2], ids=["linear", "gaussian"])
def test_force_parity_influence_functions(engine, oracle, mode, inf):
    """The same from an identical deformed state with the linear and the
    Gaussian influence functions (default parameters, influenceFn.h), whose
    value the fast paths evaluate from the bond length (the staged
    FP32-force kernel from its scaled length)."""
    progress(f"[gpu] force parity influence={inf} mode={mode}")
    p = pr.plate2d(40, 40, 1e-3, 3, 1e-7, 1, inf_type=inf, inf_params=())
    o = oracle.OracleFD(p)
    u, v = _deformed_state(p, 3, sigma=4e-6)
    with engine.GpuFD(p, precision=mode) as g:
        f0 = np.zeros((p.n, 3))
        g.set_state(u=u, v=v, f=f0)
        o.set_state(u=u, v=v, f=f0)
        absF = o.force_abs_sum()
        o.compute_forces()
        g.compute_forces()
        fr = o.get_state()[2]
        fg = g.get_state()[2]
        br, bg = o.get_bonds(), g.get_bonds()
        assert br.sum() > 1000
        assert np.array_equal(bg, br)
        assert force_err(fg, fr, absF) <= FORCE_TOL[mode]


# ------------------------------------------------------------ trajectories
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("pull", [0.5, 30.0])
def test_trajectory_parity_small(engine, oracle, mode, pull):
    progress(f"[gpu] trajectory 40x40 pull={pull} mode={mode}")
    steps = 80
    p = pr.plate2d(40, 40, 1e-3, 3, 1e-7, steps, pull=pull)
    o = oracle.OracleFD(p)
    o.init()
    o.step(steps)
    ur, vr, fr = o.get_state()
    br = o.get_bonds()
    with engine.GpuFD(p, precision=mode) as g:
        g.init()
        g.step(steps // 2)
        g.step(steps - steps // 2)  # two calls: the state carries over
        ug, vg, fg = g.get_state()
        bg = g.get_bonds()
        t, n = g.time()
    to, no = o.time()
    assert n == no and t == to
    if mode == NLM_PREC_FP32_FORCE:
        assert rel_l2(ug, ur) <= 1e-5
        mism = np.count_nonzero(bg != br)
        assert mism <= 1e-3 * max(br.sum(), 1)
    else:
        assert rel_l2(ug, ur) <= 1e-9
        assert rel_l2(vg, vr) <= 1e-8
        assert np.array_equal(bg, br)
    if pull > 1:
        assert br.sum() > 1000


def test_trajectory_plate2d_1e4_exact(engine, oracle):
    """100x100 plate, horizon 3h, notch, +/-0.5 m/s pulls, 1000 steps of
    1e-7 s in FP64 exact mode."""
    progress("[gpu] plate2d_1e4 1000 steps exact")
    p = pr.workload("plate2d_1e4")
    o = oracle.OracleFD(p)
    o.init()
    o.step(p.nsteps)
    ur, vr, fr = o.get_state()
    with engine.GpuFD(p, precision=NLM_PREC_FP64_EXACT) as g:
        assert g.nnz == 234036
        g.init()
        g.step(p.nsteps)
        ug, vg, fg = g.get_state()
        assert np.array_equal(g.get_bonds(), o.get_bonds())
    assert rel_l2(ug, ur) <= 1e-9
    assert rel_l2(fg, fr) <= 1e-7


def test_stepwise_parity_breaking(engine, oracle):
    """Per-step parity judged from identical input state along a breaking
    trajectory: at each checkpoint the GPU state is loaded into the oracle,
    one step is taken on both, bits must match exactly and forces within
    1e-12 of sum |F_ij|."""
    progress("[gpu] stepwise parity breaking")
    p = pr.plate2d(40, 40, 1e-3, 3, 1e-7, 1, pull=30.0)
    o = oracle.OracleFD(p)
    with engine.GpuFD(p, precision=NLM_PREC_FP64) as g:
        g.init()
        for chk in range(6):
            g.step(15)
            u, v, f = g.get_state()
            b = g.get_bonds()
            t, n = g.time()
            o.set_state(u=u, v=v, f=f)
            o.set_bonds(b)
            o.set_time(t, n)
            g.step(1)
            o.step(1)
            ur, vr, fr = o.get_state()
            absF = o.force_abs_sum()
            ug, vg, fg = g.get_state()
            assert np.array_equal(g.get_bonds(), o.get_bonds())
            assert force_err(fg, fr, absF) <= 1e-12
            assert rel_l2(ug, ur) <= 1e-14
        assert o.get_bonds().sum() > 1000


def test_postproc_parity(engine, oracle):
    progress("[gpu] postproc")
    p = pr.plate2d(40, 40, 1e-3, 3, 1e-7, 60, pull=30.0)
    o = oracle.OracleFD(p)
    o.init()
    o.step(60)
    with engine.GpuFD(p) as g:
        g.init()
        g.step(60)
        u, v, f = g.get_state()
        o.set_state(u=u, v=v, f=f)
        o.set_bonds(g.get_bonds())
        Zr, phir, er = o.postproc()
        Zg, phig, eg = g.postproc()
        assert np.array_equal(g.get_bonds(), o.get_bonds())
    assert np.array_equal(phig, phir)
    assert np.max(np.abs(Zg - Zr) / np.maximum(Zr, 1e-30)) <= 1e-6
    assert np.max(np.abs(eg - er) / np.maximum(np.abs(er), 1e-300)) <= 1e-12


@pytest.mark.parametrize("scale", ["metre", "millimetre"])
def test_reaction_and_external_force_parity(engine, oracle, scale):
    """Reaction force of computeForce / is_reaction_force and the external
    force of FLoading::apply (constant and line_load sets) against the
    oracle on the same state (GPU state
    copied into the oracle): metre-sized plate, where the reference's int
    orientation test selects the bonds crossing the line, and millimetre
    plate, where its truncation leaves only the onLine endpoint cases."""
    progress(f"[gpu] reaction force ({scale})")
    if scale == "metre":
        p = pr.plate2d(30, 30, 0.5, 3, 1e-5, 20, notch_len=0.0, pull=0.5)
        box = (0.0, 14.5, 14.5)
    else:
        p = pr.plate2d(30, 30, 1e-3, 3, 1e-7, 20, pull=10.0)
        box = (0.0, 0.029, 0.029)
    h = p.h
    top = (p.shape[1] - 1) * h
    fb = [pr.BC("rectangle", [1, 2], "linear", [2.5], "constant", [4.0],
                x1=-h, x2=0.3 * top, y1=-h, y2=top + h),
          pr.BC("rectangle", [2], "linear", [1.5], "line_load", [-1.0],
                x1=0.4 * top, x2=top + 0.5 * h, y1=-h, y2=0.5 * top)]
    p = pr.with_(p, fbc=fb)
    o = oracle.OracleFD(p)
    with engine.GpuFD(p) as g:
        g.init()
        g.step(20)
        u, v, f = g.get_state()
        o.set_state(u=u, v=v, f=f)
        o.set_bonds(g.get_bonds())
        Rg, Lg = g.reaction_force(box, p.horizon)
        fg = g.external_force(7.0e-6)
    Rr, Lr = o.reaction_force(box, p.horizon)
    fr = o.external_force(7.0e-6)
    scale_r = max(np.abs(Rr).max(), 1e-300)
    assert np.array_equal(Rg != 0, Rr != 0)
    assert np.abs(Rg - Rr).max() <= 1e-12 * scale_r
    assert np.abs(Lg - Lr).max() <= 1e-12 * scale_r
    if scale == "metre":
        assert np.count_nonzero(Lr) >= 30
    assert np.array_equal(fg, fr) and np.abs(fr).max() > 0


@pytest.mark.parametrize("mode", MODES)
def test_verlet_parity(engine, oracle, mode):
    """velocity_verlet (fDModel.cpp:446-525: half-step velocity, position,
    force-only kernel, second half-step) against the oracle's Verlet run;
    FP64 modes to 1e-9 with identical bonds, FP32-force to 1e-5."""
    progress(f"[gpu] velocity verlet mode={mode}")
    p = pr.plate2d(30, 30, 1e-3, 3, 1e-7, 40, pull=10.0)
    o = oracle.OracleFD(p)
    o.init()
    o.step(40, scheme="verlet")
    ur, vr, _ = o.get_state()
    br = o.get_bonds()
    with engine.GpuFD(p, precision=mode,
                      scheme=NLM_TIME_VELOCITY_VERLET) as g:
        g.init()
        g.step(40)
        ug, vg, _ = g.get_state()
        bg = g.get_bonds()
    if mode == NLM_PREC_FP32_FORCE:
        assert rel_l2(ug, ur) <= 1e-5
        assert np.count_nonzero(bg != br) <= 1e-3 * max(br.sum(), 1)
    else:
        assert np.array_equal(bg, br)
        assert rel_l2(ug, ur) <= 1e-9 and rel_l2(vg, vr) <= 1e-8


@pytest.mark.parametrize("mode", [NLM_PREC_FP64_EXACT, NLM_PREC_FP64,
                                  NLM_PREC_FP32_FORCE])
def test_force_bc_nofail_contact_timed_crack(engine, oracle, mode):
    """Force BCs (hat / linear / sin time), a no-fail region, the contact
    force of broken bonds and a crack activated during stepping, in every
    precision mode (the fast modes take the no-fail and contact branches
    through their exact evaluation): exact and FP64 modes to 1e-9 with
    identical bonds, FP32-force within its 1e-5 bound and at most 0.1 % of
    the broken bonds differing."""
    progress(f"[gpu] fbc + no-fail + contact + timed crack, mode {mode}")
    p = pr.plate2d(36, 36, 1e-3, 3, 1e-7, 50, pull=20.0, no_penetration=True)
    p.fbc = [pr.BC("rectangle", [1], "linear", [4e9], "hat_y", [1.0],
                   x1=-1e-3, x2=2.5e-3, y1=-1e-3, y2=0.036),
             pr.BC("rectangle", [2], "sin", [2e9, 1e4], "constant", [1.0],
                   x1=0.03, x2=0.036, y1=0.01, y2=0.02)]
    p.cracks = list(p.cracks) + [pr.CrackSpec(-1, (0.03, 0.005), (0.03, 0.02),
                                              activation_time=2.05e-6)]
    lo, hi = np.zeros(3), np.zeros(3)
    interior = np.ones(p.n, np.uint8)
    interior[(p.x[:, 0] > 0.025) & (p.x[:, 1] > 0.025)] = 0
    p.interior = interior
    o = oracle.OracleFD(p)
    o.init()
    o.step(50)
    ur, vr, fr = o.get_state()
    with engine.GpuFD(p, precision=mode) as g:
        g.init()
        g.step(50)
        ug, vg, fg = g.get_state()
        bg = g.get_bonds()
    br = o.get_bonds()
    assert br.sum() > 0
    if mode == NLM_PREC_FP32_FORCE:
        assert rel_l2(ug, ur) <= 1e-5
        assert np.count_nonzero(bg != br) <= 1e-3 * br.sum()
    else:
        assert np.array_equal(bg, br)
        assert rel_l2(ug, ur) <= 1e-9


@pytest.mark.parametrize("mode", [NLM_PREC_FP64, NLM_PREC_FP32_FORCE])
def test_autotune_and_unroll_do_not_change_results(engine, mode):
    """The chunk width changes the schedule only: every width gives the same
    bits and bitwise the same displacements, and autotuning leaves the state
    untouched."""
    progress(f"[gpu] autotune invariance mode={mode}")
    p = pr.plate2d(48, 40, 1e-3, 3, 1e-7, 1, pull=30.0)
    outs = []
    for unroll in (None, 2, 4, 8):
        with engine.GpuFD(p, precision=mode) as g:
            g.init()
            if unroll is None:
                chosen, us = g.autotune()
                assert chosen in (2, 4, 8) and us > 0
            else:
                g.set_unroll(unroll)
            g.step(30)
            u, v, f = g.get_state()
            outs.append((u, v, f, g.get_bonds()))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [NLM_PREC_FP64_EXACT, NLM_PREC_FP64,
                                  NLM_PREC_FP32_FORCE])
def test_tiles_and_joint_autotune(engine, mode):
    """The LDS-staged kernel's tile (128, 256 or 512 nodes per workgroup and
    halo block) changes the data movement only: in the FP64 modes every tile
    gives the same bits; in the FP32-force mode the block-relative FP32
    positions take their origin from the tile, so displacements agree to
    1e-5 (relative L2) and the broken-bond sets to 1e-3 of the broken bonds.
    The joint tile x chunk autotuner leaves the state untouched and times
    every pair the mode runs."""
    progress(f"[gpu] tiles mode={mode}")
    p = pr.plate2d(160, 120, 1e-3, 4, 1e-7, 1, pull=30.0)
    outs = {}
    for tile in (256, 128, 512, "auto"):
        with engine.GpuFD(p, precision=mode) as g:
            g.init()
            if tile == "auto":
                t, u, tab = g.autotune_tiles()
                progress(f"[gpu]   autotune: tile {t} chunk {u}\n{tab}")
                assert t in (128, 256, 512) and u in (2, 4, 8)
                cols = [1] if mode == NLM_PREC_FP64_EXACT else [0, 1, 2]
                assert np.all(tab[:, cols] > 0)
                tile = t
            else:
                g.set_tile(tile)
            g.step(40)
            u, v, f = g.get_state()
            outs.setdefault(tile, []).append((u, v, f, g.get_bonds()))
    ref = outs[256][0]
    broken = int((~ref[3]).sum())
    assert broken > 0
    for tile, runs in outs.items():
        for o in runs:
            if mode != NLM_PREC_FP32_FORCE:
                for a, b in zip(ref, o):
                    assert np.array_equal(a, b), tile
            else:
                assert rel_l2(o[0], ref[0]) <= 1e-5, tile
                assert int((o[3] != ref[3]).sum()) <= 1e-3 * broken + 2, tile


def test_tiles_on_tiny_models(engine):
    """Models smaller than one tile (a 6 x 6 plate: one slice of 36 nodes;
    a 10 x 10 plate: two slices): every tile gives the FP64 result of the
    default tile bit for bit, and the joint autotuner runs."""
    progress("[gpu] tiles on tiny models")
    for nx in (6, 10):
        p = pr.plate2d(nx, nx, 1e-3, 3, 1e-7, 1, notch_len=0.0, pull=30.0,
                       layers=1)
        outs = []
        for tile in (256, 128, 512, "auto"):
            with engine.GpuFD(p, precision=NLM_PREC_FP64) as g:
                g.init()
                if tile == "auto":
                    t, u, tab = g.autotune_tiles()
                    assert t in (128, 256, 512) and u in (2, 4, 8)
                else:
                    g.set_tile(tile)
                g.step(20)
                outs.append(g.get_state() + (g.get_bonds(),))
        for o in outs[1:]:
            for a, b in zip(outs[0], o):
                assert np.array_equal(a, b), nx


def ragged_cloud(steps):
    """A 40 x 40 plate (horizon 3h) whose nodes are jittered by up to 0.3 h
    (rows of every length around the mean, bond lengths spread across the
    volume-correction edge), with a cluster of 200 extra nodes in a 5h x 5h
    square (rows up to several times the mean length) and 130 isolated
    nodes far from the plate (empty rows; whole slices of them have no
    slot) that fly at 1 m/s from the start."""
    h = 1e-3
    p = pr.plate2d(40, 40, h, 3, 1e-7, steps, pull=30.0, jitter=0.3, seed=3)
    rng = np.random.default_rng(11)
    xc = np.zeros((200, 3))
    xc[:, :2] = 0.012 + 0.005 * rng.random((200, 2))
    xi = np.zeros((130, 3))
    xi[:, 0] = 1.0 + np.arange(130) * 5 * p.horizon
    xi[:, 1] = 0.5
    x = np.concatenate([p.x, xc, xi])
    vol = np.concatenate([p.vol, np.full(200, 0.25 * h * h),
                          np.full(130, h * h)])
    v0 = np.zeros_like(x)
    v0[-130:, 0] = 1.0
    return pr.with_(p, x=x, vol=vol, v0=v0, name="ragged_cloud")


@pytest.mark.parametrize("mode", MODES)
def test_ragged_and_empty_rows(engine, oracle, mode):
    """Irregular model (ragged_cloud) through the LDS-staged step kernel
    against the oracle's trajectory: FP64 modes to 1e-9 with identical bond
    sets, FP32-force to 1e-5; isolated nodes move freely (u = v0 t); in the
    FP64 mode every tile gives the same bits."""
    progress(f"[gpu] ragged and empty rows mode={mode}")
    steps = 60
    p = ragged_cloud(steps)
    o = oracle.OracleFD(p)
    o.init()
    o.step(steps)
    ur, vr, fr = o.get_state()
    br = o.get_bonds()
    rowlen = np.diff(o.row_ptr)
    assert (rowlen[-130:] == 0).all() and rowlen.max() > 3 * rowlen[:1600].mean()
    outs = []
    for tile in ((256, 128, 512) if mode == NLM_PREC_FP64 else (256,)):
        with engine.GpuFD(p, precision=mode) as g:
            g.init()
            g.set_tile(tile)  # raises unless the staged kernel is in use
            g.step(steps)
            outs.append(g.get_state() + (g.get_bonds(),))
    ug, vg, fg, bg = outs[0]
    t = steps * p.dt
    assert np.allclose(ug[-130:, 0], t, rtol=1e-12, atol=0)
    assert np.all(ug[-130:, 1:] == 0)
    if mode == NLM_PREC_FP32_FORCE:
        assert rel_l2(ug, ur) <= 1e-5
        assert np.count_nonzero(bg != br) <= 1e-3 * max(br.sum(), 1)
    else:
        assert rel_l2(ug, ur) <= 1e-9
        assert rel_l2(vg, vr) <= 1e-8
        assert np.array_equal(bg, br)
    assert br.sum() > 1000
    for o2 in outs[1:]:
        for a, b in zip(outs[0], o2):
            assert np.array_equal(a, b)


def fp32_underflow_floor(o, p, csr):
    """Per-node force that the FP32-force mode may lose to underflow: its
    bond arithmetic (du, dx.du, S) runs in FP32, whose smallest normal
    number is 1.18e-38, so a bond whose displacement difference is of that
    order relative to its length (the undisturbed region ahead of a wave,
    where FP64 displacements are ~1e-40 m) carries a force of at most
    c4 V_j FLT_MIN / r^2 (c4 = 4 C beta J / invFactor, the linear force per
    unit stretch, rnpBond.cpp:277-279) that FP32 cannot resolve; a factor 8
    covers the product chain."""
    m = o.mat
    c4 = 4.0 * m.C * m.beta * m.inf_a / m.inv_factor
    rp, col = csr
    rows = np.repeat(np.arange(p.n), np.diff(rp))
    r2 = ((p.x[col] - p.x[rows]) ** 2).sum(1)
    per = np.bincount(rows, p.vol[col] / r2, minlength=p.n)
    return 8.0 * np.finfo(np.float32).tiny * c4 * per


# ------------------------------------------------------------ full size
@pytest.mark.parametrize("mode", [NLM_PREC_FP64, NLM_PREC_FP32_FORCE],
                         ids=["fp64", "fp32"])
def test_plate2d_1e6_properties(engine, mode):
    """1000x1000 plate at horizon 4h: B = 43,856,108 directed bonds,
    symmetric sorted lists, and sum_i V_i f_i = 0 to rounding after 20
    steps, in the FP64 and the FP32-force modes."""
    progress(f"[gpu] plate2d_1e6 properties mode {mode}")
    p = pr.workload("plate2d_1e6")
    with engine.GpuFD(p, precision=mode) as g:
        assert g.nnz == 43856108
        rp, col = g.neighbors()
        rows = np.repeat(np.arange(p.n, dtype=np.int64), np.diff(rp))
        assert np.all(np.diff(col.astype(np.int64))[np.diff(rows) == 0] > 0)
        fwd = rows * p.n + col
        bwd = col.astype(np.int64) * p.n + rows
        assert np.array_equal(np.sort(fwd), np.sort(bwd))
        del rows, fwd, bwd
        g.init()
        g.step(20)
        u, v, f = g.get_state()
        nb = g.get_bonds().sum()
    assert np.isfinite(u).all()
    assert nb > 0  # the notch
    # bond (i, j) and (j, i) carry opposite forces up to rounding: FP64 to
    # ~1e-16 relative each, FP32-force to ~1e-7 (each bond force rounded
    # once from its own FP32 differences), so the bound is per mode
    tot = (p.vol[:, None] * f).sum(0)
    tol = 1e-10 if mode == NLM_PREC_FP64 else 1e-6
    assert np.all(np.abs(tot) <= tol * (p.vol[:, None] * np.abs(f)).sum())


@pytest.mark.parametrize("mode", [NLM_PREC_FP64, NLM_PREC_FP32_FORCE],
                         ids=["fp64", "fp32"])
def test_plate2d_1e6_stepwise_breaking(engine, oracle, mode):
    """Full-size per-step parity on the 1000x1000 plate (horizon 4h) pulled
    at +/-30 m/s, where the loaded layers tear off (about 1.3e5 bonds break
    in the first 40 steps): the GPU's neighbor lists equal the oracle's,
    and at three checkpoints the GPU state is loaded into the oracle and one
    step is taken on both -- bond bits identical (including the bonds that
    break in that step) and per-node force within 1e-12 (FP64) / 1e-5
    (FP32-force, FORCE_TOL) of sum_j |F_ij|, plus in FP32 the underflow floor of nodes
    the waves have not reached (fp32_underflow_floor)."""
    progress(f"[gpu] plate2d_1e6 stepwise breaking mode {mode}")
    p = pr.plate2d(1000, 1000, 1e-4, 4, 1e-8, 100, pull=30.0)
    csr = oracle.neighbors(p.x, p.horizon, "cells")
    o = oracle.OracleFD(p, csr)
    floor = fp32_underflow_floor(o, p, csr) if mode == NLM_PREC_FP32_FORCE \
        else 0.0
    new_total = 0
    with engine.GpuFD(p, precision=mode) as g:
        rp, col = g.neighbors()
        assert np.array_equal(rp, csr[0]) and np.array_equal(col, csr[1])
        g.init()
        done = 0
        for chk in (3, 12, 25):
            g.step(chk - done)
            done = chk + 1
            u, v, f = g.get_state()
            b = g.get_bonds()
            t, n = g.time()
            o.set_state(u=u, v=v, f=f)
            o.set_bonds(b)
            o.set_time(t, n)
            g.step(1)
            o.step(1)
            ur, vr, fr = o.get_state()
            br = o.get_bonds()
            absF = o.force_abs_sum()
            ug, vg, fg = g.get_state()
            bg = g.get_bonds()
            progress(f"  step {done}: {int(br.sum() - b.sum())} bonds broke")
            new_total += int(br.sum() - b.sum())
            assert np.array_equal(bg, br)
            d = np.abs(fg - fr).max(axis=1)
            excess = d - (FORCE_TOL[mode] * absF + floor)
            k = int(excess.argmax())
            assert excess[k] <= 0.0, (f"node {k}: |df| {d[k]:.3e} "
                                      f"sum|F| {absF[k]:.3e}")
    o.close()
    assert new_total > 1000
