"""GPU graph-network preconditioner against the numpy restatement
(oracle/gnn_oracle.py, BF16 rounding emulated at the same points), through
the C-ABI.

Bars:
  - zero weights: M_ii = I / s_i exactly, no off-diagonal blocks;
  - forward: preconditioner blocks within 2e-2 (norm-wise) of the oracle's
    (FP32 accumulation order differs, so a BF16-rounded intermediate can
    land one BF16 step apart);
  - loss within 1e-3 and every parameter block's gradient within 5e-2
    norm-wise (cosine > 0.998) of the oracle's;
  - training on the solver's own residual / error pairs lowers the BiCGSTAB
    iteration count of a held-out specimen below the untrained network's and
    below point Jacobi's;
  - on the 300 x 300 workload (two load steps, Newton and BiCGSTAB to 1e-8):
    the training loss falls to below 0.4 of its initial value, and the
    learned preconditioner takes at most 0.8x (one level) and 0.77x (two
    level) of point Jacobi's BiCGSTAB iterations, with no failed solve.
"""
import numpy as np
import pytest

from conftest import progress
from nlmech_amd import problems as pr
from nlmech_amd.capi import (NLM_PREC_FP64_EXACT, NLM_QS_PRECOND_BLOCKS,
                             NLM_QS_PRECOND_JACOBI)
from oracle import gnn_oracle as go

pytestmark = pytest.mark.gpu


def setup(nx=20, seed=3, strain=1e-3):
    from nlmech_amd.engine import GpuFD
    p = pr.qs_plate(nx, jitter=0.05)
    g = GpuFD(p, precision=NLM_PREC_FP64_EXACT)
    rng = np.random.default_rng(seed)
    u = np.zeros((p.n, 3))
    u[:, :2] = 1e-7 * rng.standard_normal((p.n, 2))
    u[:, 1] += strain * p.x[:, 1]
    z = np.zeros_like(u)
    g.set_state(u=u, v=z, f=z)
    g.qs_assemble(0)
    return p, g


def oracle_graph(p, g):
    blk, dg = g.qs_tangent()
    from oracle import pyoracle as po
    rp, col = po.neighbors(p.x, p.horizon, "brute")
    fix = g.get_fixity()
    return go.Graph(p.x, p.horizon, rp, col, blk.reshape(-1, 4),
                    dg.reshape(-1, 4), fix)


def pairs(p, g, K=3, seed=7):
    rng = np.random.default_rng(seed)
    fix = g.get_fixity()
    e = np.zeros((K, p.n, 3))
    e[:, :, :2] = rng.standard_normal((K, p.n, 2))
    e[:, fix & 1 != 0, 0] = 0.0
    e[:, fix & 2 != 0, 1] = 0.0
    r = np.stack([g.qs_apply(ek) for ek in e])
    return r, e


def test_gnn_zero_weights():
    progress("[gpu] gnn zero weights")
    p, g = setup(12)
    g.gnn_create(1)
    g.gnn_set_weights(np.zeros(go.NPARAMS, np.float32))
    g.gnn_predict()
    mb, md = g.qs_precond()
    _, dg = g.qs_tangent()
    s = 0.5 * (dg[:, 0, 0] + dg[:, 1, 1])
    fix = g.get_fixity()
    assert np.all(mb == 0.0)
    ref = np.zeros_like(md)
    ref[:, 0, 0] = np.where(fix & 1, 0.0, 1.0)
    ref[:, 1, 1] = np.where(fix & 2, 0.0, 1.0)
    isq = (1.0 / np.sqrt(s)).astype(np.float32).astype(np.float64)
    ref *= (isq * isq)[:, None, None]
    # the preconditioner blocks are stored in FP32
    assert np.array_equal(md, ref.astype(np.float32).astype(np.float64))
    g.close()


def test_gnn_forward_matches_oracle():
    progress("[gpu] gnn forward vs oracle")
    p, g = setup(20)
    g.gnn_create(11)
    w = g.gnn_weights()
    g.gnn_predict()
    mb, md = g.qs_precond()
    gr = oracle_graph(p, g)
    Me, Mn, _ = go.Net(w, bf16_mode=True).forward(gr)
    eb = np.linalg.norm(mb.reshape(-1, 4) - Me) / np.linalg.norm(Me)
    ed = np.linalg.norm(md.reshape(-1, 4) - Mn) / np.linalg.norm(Mn)
    progress(f"[gpu]   forward rel err edges {eb:.2e} nodes {ed:.2e}")
    assert eb <= 2e-2 and ed <= 2e-2
    g.close()


def test_gnn_gradient_matches_oracle():
    progress("[gpu] gnn gradient vs oracle")
    p, g = setup(20)
    g.gnn_create(13)
    w = g.gnn_weights()
    r, e = pairs(p, g)
    loss, grad = g.gnn_loss_grad(r, e, want_grad=True)
    gr = oracle_graph(p, g)
    lref, gref = go.Net(w, bf16_mode=True).loss_grad(gr, r[:, :, :2],
                                                     e[:, :, :2])
    progress(f"[gpu]   loss {loss:.6e} oracle {lref:.6e}")
    assert abs(loss - lref) <= 1e-3 * lref
    blocks = {"enc": (0, go.OFF_LAYER)}
    for l in range(go.L):
        o = go.OFF_LAYER + l * go.LAYER
        blocks[f"layer{l}"] = (o, o + go.LAYER)
    blocks["dec"] = (go.OFF_WOE, go.NPARAMS)
    for name, (a, b) in blocks.items():
        x, y = grad[a:b].astype(np.float64), gref[a:b]
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        cos = x @ y / (np.linalg.norm(x) * np.linalg.norm(y))
        progress(f"[gpu]   grad {name}: rel {rel:.2e} cos {cos:.6f}")
        assert rel <= 5e-2 and cos >= 0.998, name
    g.close()


def test_gnn_training_lowers_iterations():
    """Adam on residual / error pairs of a plain-solver run; the held-out
    specimen's BiCGSTAB iterations drop below Jacobi's."""
    progress("[gpu] gnn training gain")
    from nlmech_amd import gnn_train
    net = gnn_train.train(nx=32, specimens=6, load_steps=3, epochs=12,
                          seed=5, log=progress)
    res = gnn_train.compare(net, nx=40, seed=99, log=progress)
    progress(f"[gpu]   {res}")
    assert res["learned_iters"] < res["untrained_iters"]
    assert res["learned_iters"] < res["jacobi_iters"]
    assert res["learned_coarse_iters"] < res["jacobi_coarse_iters"]


def test_gnn_gain_on_plate300():
    """The network against point Jacobi on qs_plate300, one and two level
    (QuasiStaticModel.cpp:457-524 is the reference's linear solve of each
    Newton iteration; here BiCGSTAB preconditioned by point Jacobi or the
    network's sparse approximate inverse, each optionally with the
    two-level coarse correction)."""
    progress("[gpu] gnn gain on qs_plate300")
    from nlmech_amd import gnn_train
    from nlmech_amd.engine import GpuFD
    nets = {}
    for coarse in (False, True):
        net = gnn_train.train(nx=48, specimens=40, load_steps=5, epochs=4,
                              seed=1 + int(coarse), coarse=coarse)
        progress(f"[gpu]   coarse={coarse}: loss {net['loss_initial']:.4f} "
                 f"-> {net['loss'][-1]:.4f}")
        assert net["loss"][-1] <= 0.4 * net["loss_initial"]
        nets[coarse] = net
    p = pr.workload("qs_plate300", nsteps=2)
    iters = {}
    for name, pre in (("jacobi", NLM_QS_PRECOND_JACOBI),
                      ("learned", NLM_QS_PRECOND_BLOCKS),
                      ("jacobi2", NLM_QS_PRECOND_JACOBI | 16),
                      ("learned2", NLM_QS_PRECOND_BLOCKS | 16)):
        g = GpuFD(p, precision=NLM_PREC_FP64_EXACT)
        if pre & 15 == NLM_QS_PRECOND_BLOCKS:
            g.gnn_create(1)
            g.gnn_set_weights(nets[bool(pre & 16)]["weights"])
        st = g.qs_solve(2, tol_rel=1e-8, precond=pre, lin_tol=1e-8,
                        lin_max_iter=50000, gnn_refresh=2)
        g.close()
        progress(f"[gpu]   {name}: {st['linear_iters']} iterations, "
                 f"{st['ms_linear']:.1f} ms linear")
        assert st["linear_failures"] == 0
        assert st["residual"] <= 1e-8 * st["residual0"]
        iters[name] = st["linear_iters"]
    assert iters["learned"] <= 0.8 * iters["jacobi"]
    assert iters["learned2"] <= 0.77 * iters["jacobi2"]
