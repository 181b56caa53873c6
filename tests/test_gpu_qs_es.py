"""Quasi-static solve with the state-based material on the GPU: the
finite-difference tangent (the reference's Jacobian,
QuasiStaticModel.cpp:363-454: central differences of the force, eps =
Perturbation * h, same-component entries on the neighbor sparsity).

Bars:
  - every tangent entry equals the oracle's central difference of the
    state-based force at the same state with the same eps (perturbing one
    dof at a time, as the reference does), to 1e-7 of the largest entry;
    cross-component entries are zero;
  - the Newton solve of examples/qsModel/2D's loading reproduces the
    reference-held output_res.vtu displacement to < 1e-7 in L2 (the
    reference test's dc tolerance; the deck itself runs in
    test_gpu_host.py).
"""
import numpy as np
import pytest

from conftest import progress
from nlmech_amd import problems as pr
from nlmech_amd.capi import (NLM_PREC_FP64_EXACT, NLM_QS_PRECOND_JACOBI,
                             NLM_QS_TANGENT_FD)
import es_cases

pytestmark = pytest.mark.gpu


def qs2d_problem():
    p, g = es_cases.qs2d()
    p.ubc = [pr.BC("line", [1], "constant", [0.0], "constant", [1.0],
                   x1=1.55, x2=1.65),
             pr.BC("line", [2], "constant", [0.0], "constant", [1.0],
                   x1=1.55, x2=1.65)]
    p.fbc = [pr.BC("line", [1], "linear", [1.0], "constant", [-250.0],
                   x1=-0.1, x2=0.05)]
    p.dt = 1.0
    return p, g


def test_es_fd_tangent_matches_oracle(oracle):
    progress("[gpu] es fd tangent vs oracle")
    from nlmech_amd.engine import GpuFD
    p, gold = qs2d_problem()
    o = oracle.OracleES(p)
    u = gold["Displacement"].copy()
    z = np.zeros_like(u)
    eps = 1e-7 * p.h
    with GpuFD(p, precision=NLM_PREC_FP64_EXACT,
               csr=(o.row_ptr, o.col)) as g:
        g.set_state(u=u, v=z, f=z)
        g.qs_assemble(NLM_QS_TANGENT_FD)
        blk, dg = g.qs_tangent()
    rp, col = o.row_ptr, o.col
    assert np.all(blk[:, 0, 1] == 0.0) and np.all(blk[:, 1, 0] == 0.0)
    scale = max(np.abs(blk).max(), np.abs(dg).max())
    rng = np.random.default_rng(3)
    for j in rng.choice(p.n, 12, replace=False):
        for r in range(2):
            fs = []
            for sgn in (1.0, -1.0):
                up = u.copy()
                up[j, r] += sgn * eps
                o.set_state(u=up, v=z, f=z)
                o.compute_forces()
                fs.append(o.get_state()[2])
            col_j = -(fs[0] - fs[1])[:, r] / (2.0 * eps)
            # entry (i, j) sits in row i's list at j's position
            for i in list(col[rp[j]:rp[j + 1]]) + [j]:
                if i == j:
                    got = dg[j, r, r]
                else:
                    k = rp[i] + int(np.flatnonzero(col[rp[i]:rp[i + 1]]
                                                   == j)[0])
                    got = blk[k, r, r]
                assert abs(got - col_j[i]) <= 1e-7 * scale, (i, j, r)
    o.close()


def test_es_newton_reproduces_reference_result():
    progress("[gpu] es quasi-static newton vs output_res.vtu")
    from nlmech_amd.engine import GpuFD
    p, gold = qs2d_problem()
    with GpuFD(p, precision=NLM_PREC_FP64_EXACT) as g:
        st = g.qs_solve(1, dt=1.0, tol_abs=1e-9, tol_rel=0.0, max_newton=200,
                        tangent=NLM_QS_TANGENT_FD,
                        precond=NLM_QS_PRECOND_JACOBI, lin_tol=1e-10,
                        lin_max_iter=5000, predictor=False, perturbation=1e-7)
        u, _, _ = g.get_state()
    progress(f"[gpu]   {st}")
    assert st["residual"] <= 1e-9
    d = np.linalg.norm(u - gold["Displacement"])
    progress(f"[gpu]   displacement L2 vs output_res.vtu {d:.3e}")
    assert d < 1e-7
