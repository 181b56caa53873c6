"""Training data and training of the graph-network preconditioner.

Data: the plain (point-Jacobi) quasi-static solver is run on specimens with
randomized pre-notches and load amplitudes.  At the equilibrium of every load
step the tangent A is assembled; its residual / error pairs (r, e) with
A e = r are white-noise probes e on the free dofs (r = A e), so that
sum_k ||M r_k - e_k||^2 / ||e_k||^2 is a sampled ||I - M A||_F^2 (the
sparse-approximate-inverse objective).  The load step's own increment
(e = du, r = A du) can be added as a pair (increment=True), but it is a smooth
field: (I - M A) du stays near du for every local M (about 0.9 of ||du||^2 for
point Jacobi and for the optimal sparse approximate inverse alike on a 48 x 48
plate), so that pair dominates the loss and its gradient without telling the
network anything a local preconditioner can act on.  Training runs Adam on
the GPU through the C-ABI (nlm_gnn_loss_grad / nlm_gnn_adam_step).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import problems as pr
from .capi import (NLM_PREC_FP64_EXACT, NLM_QS_PRECOND_BLOCKS,
                   NLM_QS_PRECOND_COARSE, NLM_QS_PRECOND_JACOBI)

PROBES = 4


def specimen(nx: int, seed: int, pull: float = 1e-5) -> pr.Problem:
    """Plate of the quasi-static workload with a random notch from the left
    edge (length U[0.1, 0.3] L, angle U[-30, 30] degrees) and a random pull
    amplitude (U[0.5, 1.0] of pull)."""
    rng = np.random.default_rng(seed)
    L = (nx - 1) * 1e-3
    return pr.qs_plate(nx, pull=pull * rng.uniform(0.5, 1.0),
                       notch_len=rng.uniform(0.1, 0.3) * L,
                       notch_angle=math.radians(rng.uniform(-30.0, 30.0)),
                       seed=seed + 1)


def collect(nx: int, specimens: int, load_steps: int, seed: int = 0,
            probes: int = PROBES, log=None, coarse: bool = False,
            increment: bool = False):
    """List of samples (bonds, u, r[K, n, 3], e[K, n, 3]).  coarse: the
    targets are e_k - C r_k with C the two-level coarse correction, so the
    network learns the part of A^-1 the coarse level leaves (P = M + C)."""
    from .engine import GpuFD
    out = []
    rng = np.random.default_rng(seed + 4242)
    for sidx in range(specimens):
        p = specimen(nx, seed * 1000 + sidx)
        g = GpuFD(p, precision=NLM_PREC_FP64_EXACT)
        u_prev = g.get_state()[0]
        for _ in range(load_steps):
            st = g.qs_solve(1, tol_rel=1e-8, precond=NLM_QS_PRECOND_JACOBI,
                            lin_tol=1e-10, lin_max_iter=20000)
            u = g.get_state()[0]
            if (st["linear_failures"] or not np.isfinite(u).all() or
                    not st["residual"] <= 1e-6 * max(st["residual0"], 1e-300)):
                break  # the specimen failed (fracture): keep what it gave
            du = u - u_prev
            u_prev = u
            g.qs_assemble(0)
            fix = g.get_fixity()
            du[fix & 1 != 0, 0] = 0.0
            du[fix & 2 != 0, 1] = 0.0
            es = [du] if increment else []
            for _ in range(probes):
                v = np.zeros_like(du)
                v[:, :2] = rng.standard_normal((p.n, 2))
                v[fix & 1 != 0, 0] = 0.0
                v[fix & 2 != 0, 1] = 0.0
                es.append(v)
            rs = [g.qs_apply(e) for e in es]
            if coarse:
                es = [e - g.qs_coarse_apply(r) for e, r in zip(es, rs)]
            r_, e_ = np.stack(rs), np.stack(es)
            if np.isfinite(r_).all() and np.isfinite(e_).all():
                out.append((g.get_bonds(), u, r_, e_))
        g.close()
        if log:
            log(f"[gnn] specimen {sidx + 1}/{specimens}: {len(out)} samples")
    return out


def train(nx: int = 48, specimens: int = 20, load_steps: int = 4,
          epochs: int = 10, lr: float = 1e-3, seed: int = 0, log=None,
          samples=None, coarse: bool = False):
    """Trains a network on freshly collected samples; returns its FP32
    weights and the loss history."""
    from .engine import GpuFD
    t0 = time.perf_counter()
    if samples is None:
        samples = collect(nx, specimens, load_steps, seed, log=log,
                          coarse=coarse)
    t_data = time.perf_counter() - t0
    base = pr.qs_plate(nx, notch_len=0.0)
    g = GpuFD(base, precision=NLM_PREC_FP64_EXACT)
    g.gnn_create(seed + 17)
    rng = np.random.default_rng(seed)

    def load(sample):
        bonds, u, r, e = sample
        g.set_bonds(bonds)
        z = np.zeros_like(u)
        g.set_state(u=u, v=z, f=z)
        g.qs_assemble(0)
        return r, e

    t1 = time.perf_counter()
    # the loss of the initial weights (close to point Jacobi) over all samples
    loss0 = float(np.mean([g.gnn_loss_grad(*load(sm))[0] for sm in samples]))
    g.gnn_zero_grad()
    if log:
        log(f"[gnn] initial weights: mean loss {loss0:.4f}")
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for idx in order:
            r, e = load(samples[idx])
            loss, _ = g.gnn_loss_grad(r, e)
            if not np.isfinite(loss):
                raise RuntimeError("non-finite training loss")
            g.gnn_adam_step(lr)
            tot += loss
        hist.append(tot / len(samples))
        if log:
            log(f"[gnn] epoch {ep + 1}/{epochs}: mean loss {hist[-1]:.4f}")
    w = g.gnn_weights()
    g.close()
    return {"weights": w, "loss": hist, "loss_initial": loss0,
            "pairs": len(samples) * len(samples[0][2]),
            "s_data": t_data, "s_train": time.perf_counter() - t1}


def compare(net, nx: int = 64, seed: int = 99, load_steps: int = 2, log=None):
    """BiCGSTAB iterations on a held-out specimen's tangent system b = A du:
    point Jacobi, untrained network, trained network."""
    from .engine import GpuFD
    p = specimen(nx, seed)
    g = GpuFD(p, precision=NLM_PREC_FP64_EXACT)
    u_prev = g.get_state()[0]
    for _ in range(load_steps):
        g.qs_solve(1, tol_rel=1e-8, precond=NLM_QS_PRECOND_JACOBI,
                   lin_tol=1e-10, lin_max_iter=20000)
    u = g.get_state()[0]
    du = u - u_prev
    g.qs_assemble(0)
    fix = g.get_fixity()
    du[fix & 1 != 0, 0] = 0.0
    du[fix & 2 != 0, 1] = 0.0
    b = g.qs_apply(du)
    res = {}
    J, B, C = NLM_QS_PRECOND_JACOBI, NLM_QS_PRECOND_BLOCKS, NLM_QS_PRECOND_COARSE
    _, res["jacobi_iters"], _ = g.qs_linear_solve(b, J, 1e-8, 20000)
    _, res["jacobi_coarse_iters"], _ = g.qs_linear_solve(b, J | C, 1e-8, 20000)
    g.gnn_create(12345)
    g.gnn_predict()
    _, res["untrained_iters"], _ = g.qs_linear_solve(b, B, 1e-8, 20000)
    g.gnn_set_weights(net["weights"])
    g.gnn_predict()
    _, res["learned_iters"], _ = g.qs_linear_solve(b, B, 1e-8, 20000)
    _, res["learned_coarse_iters"], _ = g.qs_linear_solve(b, B | C, 1e-8,
                                                          20000)
    g.close()
    return res
