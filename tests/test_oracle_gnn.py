"""Pins the numpy restatement of the graph-network preconditioner.

- In float64 mode (no BF16 rounding) its gradient equals central finite
  differences of its loss, parameter by parameter, for a sample of
  parameters from every block (encoders, the four layers, decoders).
- With all weights zero the network's preconditioner is point scaling by the
  node diagonal scale, M_ii = I / s_i, with no off-diagonal blocks.
- BF16 rounding is round-to-nearest-even.
"""
from __future__ import annotations

import numpy as np
import pytest

from nlmech_amd import problems as pr
from oracle import gnn_oracle as go
from oracle import qs_oracle as qo


def tiny_graph(nx=7, seed=0):
    p = pr.qs_plate(nx, jitter=0.05)
    q = qo.OracleQS(p)
    u = np.zeros((p.n, 3))
    u[:, 1] = 1e-3 * p.x[:, 1]
    J = q.jacobian_analytic(u)
    A = -J
    rows, cols = q.rows, q.cols
    blk = np.stack([A[2 * rows + a, 2 * cols + b] for a in range(2)
                    for b in range(2)], 1)
    ar = np.arange(p.n)
    dg = np.stack([A[2 * ar + a, 2 * ar + b] for a in range(2)
                   for b in range(2)], 1)
    g = go.Graph(p.x, p.horizon, q.fd.row_ptr, q.fd.col, blk, dg, q.fix)
    q.close()
    rng = np.random.default_rng(seed)
    K = 3
    e = rng.standard_normal((K, p.n, 2))
    e[:, q.fix & 1 != 0, 0] = 0.0
    e[:, q.fix & 2 != 0, 1] = 0.0
    Af = A.copy()
    r = np.einsum("ij,kj->ki", Af, e.reshape(K, -1)).reshape(K, p.n, 2)
    r[:, q.fix & 1 != 0, 0] = 0.0
    r[:, q.fix & 2 != 0, 1] = 0.0
    return g, r, e


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.001953125, -2.5e-3,
                     3.0e38, 0.0], np.float32)
    out = go.bf16(vals)
    assert out[0] == 1.0
    assert out[1] == 1.0078125 or out[1] == 1.0  # tie -> even
    assert out[3] == 1.0                           # tie below -> even (1.0)
    ref = np.array([np.float32(v) for v in vals])
    assert np.all(np.abs(out - ref) <= np.abs(ref) * 2 ** -8)


def test_zero_weights_give_diagonal_scaling():
    g, r, e = tiny_graph()
    net = go.Net(np.zeros(go.NPARAMS), bf16_mode=False)
    Me, Mn, _ = net.forward(g)
    assert np.all(Me == 0.0)
    s = 0.5 * (g.diag[:, 0] + g.diag[:, 3])
    mk = g.mask(np.arange(g.n), np.arange(g.n))
    ref = np.array([1.0, 0, 0, 1.0])[None] * mk / s[:, None]
    assert np.allclose(Mn, ref, rtol=1e-6, atol=0)


def test_gradient_equals_finite_differences():
    g, r, e = tiny_graph(seed=2)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(go.NPARAMS) * 0.15
    net = go.Net(w, bf16_mode=False)
    loss, grad = net.loss_grad(g, r, e)
    assert np.isfinite(loss) and loss > 0
    blocks = [(go.OFF_WN, 512), (go.OFF_BN, 64), (go.OFF_WE, 512),
              (go.OFF_BE, 64), (go.OFF_WOE, 256), (go.OFF_BOE, 4),
              (go.OFF_WON, 256), (go.OFF_BON, 4)]
    for l in range(go.L):
        o = go.OFF_LAYER + l * go.LAYER
        blocks += [(o + go.L_WA, 4096), (o + go.L_WB, 4096),
                   (o + go.L_WC, 4096), (o + go.L_BM, 64),
                   (o + go.L_WD, 4096), (o + go.L_WF, 4096),
                   (o + go.L_BU, 64)]
    eps = 1e-6
    bad = []
    for off, size in blocks:
        for t in rng.choice(size, 2, replace=False):
            k = off + int(t)
            wp = w.copy()
            wp[k] += eps
            wm = w.copy()
            wm[k] -= eps
            lp, _ = go.Net(wp, False).loss_grad(g, r, e)
            lm, _ = go.Net(wm, False).loss_grad(g, r, e)
            fd = (lp - lm) / (2 * eps)
            scale = max(abs(fd), abs(grad[k]), 1e-3 * np.abs(grad).max())
            if abs(fd - grad[k]) > 1e-4 * scale:
                bad.append((off, k, fd, grad[k]))
    assert not bad, bad[:5]
