"""GPU quasi-static path (bond-based, 2D) against the oracle, through the
C-ABI.

Bars:
  - tangent blocks (exact and same-component) from an identical state:
    max |A_gpu - A_ref| <= 1e-13 max |A_ref| (same formula, FP64);
  - y = A_ff x against the dense oracle matrix: <= 1e-13 relative;
  - BiCGSTAB (none / Jacobi / explicit blocks): ||A x - b|| <= 1e-8 ||b||
    checked independently with the dense matrix;
  - Newton: the converged displacement of every load step equals the
    oracle's (dense direct solves) to 1e-9 relative, predictor and reference
    order, exact and same-component tangent.
"""
import numpy as np
import pytest

from conftest import progress
from nlmech_amd import problems as pr
from nlmech_amd.capi import (NLM_PREC_FP64_EXACT, NLM_QS_PRECOND_BLOCKS,
                             NLM_QS_PRECOND_JACOBI, NLM_QS_PRECOND_NONE,
                             NLM_QS_TANGENT_COMPONENT, NLM_QS_TANGENT_FULL)
from oracle import qs_oracle as qo

pytestmark = pytest.mark.gpu


def gpu(p, csr):
    from nlmech_amd.engine import GpuFD
    return GpuFD(p, precision=NLM_PREC_FP64_EXACT, csr=csr)


def strained_state(p, seed=1):
    rng = np.random.default_rng(seed)
    u = np.zeros((p.n, 3))
    u[:, :2] = 1e-6 * rng.standard_normal((p.n, 2))
    u[:, 1] += 1.2e-3 * p.x[:, 1]
    return u


@pytest.mark.parametrize("tangent", [NLM_QS_TANGENT_FULL,
                                     NLM_QS_TANGENT_COMPONENT])
def test_qs_tangent_blocks(tangent):
    progress(f"[gpu] qs tangent {tangent}")
    p = pr.qs_plate(14, jitter=0.05)
    q = qo.OracleQS(p)
    u = strained_state(p)
    J = q.jacobian_analytic(u, component_only=tangent == 1)
    bits = q.fd.get_bonds()
    csr = (q.fd.row_ptr, q.fd.col)
    g = gpu(p, csr)
    g.set_bonds(bits)
    z = np.zeros_like(u)
    g.set_state(u=u, v=z, f=z)
    g.qs_assemble(tangent)
    blk, dg = g.qs_tangent()
    A = -J
    rows, cols = q.rows, q.cols
    ref = np.zeros((len(rows), 2, 2))
    for a in range(2):
        for b in range(2):
            ref[:, a, b] = A[2 * rows + a, 2 * cols + b]
    refd = np.zeros((p.n, 2, 2))
    for a in range(2):
        for b in range(2):
            refd[:, a, b] = A[2 * np.arange(p.n) + a, 2 * np.arange(p.n) + b]
    scale = np.abs(ref).max()
    assert np.abs(blk - ref).max() <= 1e-13 * scale
    assert np.abs(dg - refd).max() <= 1e-13 * np.abs(refd).max()
    # y = A_ff x, from the bond scalars (g dx dx^T, dx from the jittered
    # positions) and from the stored blocks
    rng = np.random.default_rng(3)
    x = np.zeros((p.n, 3))
    x[:, :2] = rng.standard_normal((p.n, 2))
    free = q.free
    xf = x[:, :2].reshape(-1) * free
    yref = (A @ xf) * free
    for bond_scalar in (True, False):
        g.qs_set_launch(graph=True, bond_scalar=bond_scalar)
        y = g.qs_apply(x)
        assert np.abs(y[:, :2].reshape(-1) - yref).max() <= \
            1e-13 * np.abs(yref).max(), bond_scalar
    g.close()
    q.close()


@pytest.mark.parametrize("precond", [NLM_QS_PRECOND_NONE,
                                     NLM_QS_PRECOND_JACOBI,
                                     NLM_QS_PRECOND_BLOCKS,
                                     NLM_QS_PRECOND_JACOBI | 16,
                                     NLM_QS_PRECOND_BLOCKS | 16])
def test_qs_bicgstab(precond):
    progress(f"[gpu] qs bicgstab precond {precond}")
    p = pr.qs_plate(40)
    q = qo.OracleQS(p)
    u = strained_state(p, 2)
    A = -q.jacobian_analytic(u)
    bits = q.fd.get_bonds()
    g = gpu(p, (q.fd.row_ptr, q.fd.col))
    g.set_bonds(bits)
    z = np.zeros_like(u)
    g.set_state(u=u, v=z, f=z)
    g.qs_assemble(NLM_QS_TANGENT_FULL)
    fr = q.free
    if precond & 15 == NLM_QS_PRECOND_BLOCKS:
        # block Jacobi written as explicit blocks: inverse diagonal blocks,
        # zero off-diagonal blocks
        blk, dg = g.qs_tangent()
        inv = np.zeros_like(dg)
        for i in range(p.n):
            m = dg[i].copy()
            fx = ~fr[2 * i:2 * i + 2]
            m[fx, :] = 0.0
            m[:, fx] = 0.0
            m[fx, fx] = 1.0
            inv[i] = np.linalg.inv(m)
            inv[i][fx, :] = 0.0
            inv[i][:, fx] = 0.0
        g.qs_set_precond(np.zeros_like(blk), inv)
    rng = np.random.default_rng(4)
    b = np.zeros((p.n, 3))
    b[:, :2] = rng.standard_normal((p.n, 2))
    x, it, rr = g.qs_linear_solve(b, precond=precond, tol=1e-8,
                                  max_iter=4000)
    progress(f"[gpu]   iterations {it}")
    assert 0 < it < 4000 and rr <= 1e-8
    Af = A[np.ix_(fr, fr)]
    bf = b[:, :2].reshape(-1)[fr]
    xf = x[:, :2].reshape(-1)
    assert np.all(xf[~fr] == 0.0)
    res = np.linalg.norm(Af @ xf[fr] - bf) / np.linalg.norm(bf)
    assert res <= 2e-8, res
    g.close()
    q.close()


@pytest.mark.parametrize("case", ["full", "component", "reference_order"])
def test_qs_newton_matches_oracle(case):
    progress(f"[gpu] qs newton {case}")
    p = pr.qs_plate(16, pull=2e-6)
    steps = 3
    jac = "component" if case == "component" else "analytic"
    pred = case != "reference_order"
    q = qo.OracleQS(p)
    ref = q.solve(steps, tol_rel=1e-11, max_newton=300, jacobian=jac,
                  predictor=pred)
    csr = (q.fd.row_ptr, q.fd.col)
    q.close()
    g = gpu(p, csr)
    us = []
    for _ in range(steps):
        st = g.qs_solve(1, tol_rel=1e-11, max_newton=300,
                        tangent=1 if case == "component" else 0,
                        precond=NLM_QS_PRECOND_JACOBI, lin_tol=1e-12,
                        lin_max_iter=20000, predictor=pred)
        assert st["residual"] <= 1e-11 * st["residual0"] * 1.0001
        assert st["linear_failures"] == 0
        us.append(g.get_state()[0])
    t, n = g.time()
    g.close()
    assert n == steps and t == pytest.approx(steps * 1.0)
    for (u_ref, _, _), u in zip(ref, us):
        err = np.abs(u - u_ref).max() / np.abs(u_ref).max()
        assert err <= 1e-9, err


def test_qs_newton_two_level_matches_jacobi():
    progress("[gpu] qs newton two-level")
    from nlmech_amd.engine import GpuFD
    p = pr.qs_plate(60, pull=2e-6)
    us = []
    for pre in (NLM_QS_PRECOND_JACOBI, NLM_QS_PRECOND_JACOBI | 16):
        g = GpuFD(p, precision=NLM_PREC_FP64_EXACT)
        st = g.qs_solve(3, tol_rel=1e-10, precond=pre, lin_tol=1e-10,
                        lin_max_iter=20000)
        assert st["linear_failures"] == 0
        progress(f"[gpu]   precond {pre}: {st['linear_iters']} iterations")
        us.append(g.get_state()[0])
        g.close()
    assert np.abs(us[0] - us[1]).max() <= 1e-8 * np.abs(us[0]).max()


def test_qs_plate300_jacobi_runs():
    """The 300 x 300 workload: two load steps converge to 1e-8."""
    progress("[gpu] qs plate300 jacobi")
    from nlmech_amd.engine import GpuFD
    p = pr.workload("qs_plate300")
    g = GpuFD(p, precision=NLM_PREC_FP64_EXACT)
    st = g.qs_solve(2, tol_rel=1e-8, precond=NLM_QS_PRECOND_JACOBI,
                    lin_tol=1e-8, lin_max_iter=20000)
    assert st["linear_failures"] == 0
    assert st["residual"] <= 1e-8 * st["residual0"]
    progress(f"[gpu]   {st}")
    g.close()


@pytest.mark.parametrize("precond", [NLM_QS_PRECOND_JACOBI,
                                     NLM_QS_PRECOND_JACOBI | 16,
                                     NLM_QS_PRECOND_NONE,
                                     NLM_QS_PRECOND_BLOCKS,
                                     NLM_QS_PRECOND_BLOCKS | 16])
def test_qs_bicgstab_graph_equals_eager(precond):
    """Batches replayed from the captured HIP graph give the eager launches'
    result bit for bit, on every replay, with eager work in between, with
    the tangent's products from its bond scalars and from its blocks; an
    iteration cap inside a batch stops both at the same iteration."""
    progress(f"[gpu] qs bicgstab graph vs eager, precond {precond}")
    p = pr.qs_plate(40)
    q = qo.OracleQS(p)
    u = strained_state(p, 3)
    g = gpu(p, (q.fd.row_ptr, q.fd.col))
    z = np.zeros_like(u)
    g.set_state(u=u, v=z, f=z)
    g.qs_assemble(NLM_QS_TANGENT_FULL)
    if precond & 15 == NLM_QS_PRECOND_BLOCKS:
        # a sparse approximate inverse with off-diagonal blocks: scaled
        # tangent blocks around the inverse diagonal
        blk, dg = g.qs_tangent()
        inv = np.zeros_like(dg)
        for i in range(p.n):
            if abs(np.linalg.det(dg[i])) > 0:
                inv[i] = np.linalg.inv(dg[i])
        sc = 0.5 / np.abs(dg[:, 0, 0]).max()
        g.qs_set_precond(-sc * 0.1 * blk, inv)
    rng = np.random.default_rng(5)
    b = np.zeros((p.n, 3))
    b[:, :2] = rng.standard_normal((p.n, 2))
    for bond_scalar in (True, False):
        for max_iter, tol in ((4000, 1e-8), (13, 1e-30)):
            g.qs_set_launch(graph=False, bond_scalar=bond_scalar)
            x0, it0, rr0 = g.qs_linear_solve(b, precond=precond, tol=tol,
                                             max_iter=max_iter)
            g.qs_set_launch(graph=True, bond_scalar=bond_scalar)
            for rep in range(3):
                x, it, rr = g.qs_linear_solve(b, precond=precond, tol=tol,
                                              max_iter=max_iter)
                assert it == it0 and np.array_equal(x, x0) and rr == rr0, \
                    (bond_scalar, rep)
                g.qs_apply(x)          # eager work between replays
            if max_iter == 13:
                assert it0 == 13
    g.close()
    q.close()
