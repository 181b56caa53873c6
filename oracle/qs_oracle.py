"""Quasi-static Newton solver of NLMech restated on the CPU (small cases).

TEST INFRASTRUCTURE ONLY: the checker of the GPU quasi-static path.

Follows model::QuasiStaticModel<T> (src/model/quasistatic/QuasiStaticModel.cpp)
with the bond-based material of the oracle (OracleFD, RNPBond::getBondEF):
  - solver() :527-592: load step n at t = n dt applies FLoading (body force
    b) and ULoading (prescribed displacements), then Newton iterations
    u += delta until the residual norm meets the tolerance;
  - computeResidual() :597-646: R = f_int + b on free dofs, 0 on prescribed;
  - assembly_jacobian_matrix_part() :363-454: central differences of the
    force with eps = perturbation * h, keeping only the entries with equal
    components (r == s) and no row for nodes of a displacement BC set;
  - newton_step() :457-525: rows and columns of prescribed dofs removed,
    J delta = -R solved (the reference uses BlazeIterative; here a dense
    direct solve, since only converged solutions are compared).
Deliberate differences, shared with the GPU engine:
  - predictor (default): the first Newton iteration of a load step is
    linearized at the previous equilibrium with the prescribed displacement
    increment as Dirichlet data, K_ff delta_f = -R_f - K_fp du_p, so the
    prescribed jump is spread over the specimen before the first force
    evaluation (the reference evaluates the force right after moving the
    prescribed nodes, which breaks every bond across the loaded layer's edge
    once the jump exceeds the critical stretch); predictor=False gives the
    reference's order;
  - the first residual of a load step is evaluated with the forces of the
    current displacement (the reference reuses the previous step's forces,
    :555-560);
  - convergence may be relative: ||R|| <= tol_rel ||R_0||, R_0 the step's
    first (linearized, with the predictor) residual.
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as po


class OracleQS:
    def __init__(self, p, csr=None):
        self.p = p
        self.fd = po.OracleFD(p, csr)
        self.n = p.n
        self.dim = p.dim
        self.fix = self.fd.get_fixity()
        d = self.dim
        self.free = np.ones(self.n * d, bool)
        for a in range(d):
            self.free[a::d] = (self.fix >> a) & 1 == 0
        self.bc_node = self.fix != 0
        self.rows = np.repeat(np.arange(self.n), np.diff(self.fd.row_ptr))
        self.cols = self.fd.col.astype(np.int64)

    def close(self):
        self.fd.close()

    # -- state ---------------------------------------------------------------
    def loads_at(self, u, t):
        """(u with ULoading applied at t, FLoading body force b at t)
        (QuasiStaticModel.cpp:541-548); bond bits untouched."""
        bits = self.fd.get_bonds()
        z = np.zeros_like(u)
        self.fd.set_state(u=u, v=z, f=z)
        self.fd.set_time(t, 0)
        self.fd.lib.ora_fd_init(self.fd.h)  # BCs at t, then f = b + f_int
        u2, _, f_all = self.fd.get_state()
        self.fd.set_bonds(bits)
        f_int = self.fd.internal_force(u2)
        self.fd.set_bonds(bits)
        return u2, f_all - f_int

    def residual(self, u):
        f = self.fd.internal_force(u)
        R = (f + self.b)[:, : self.dim].reshape(-1).copy()
        R[~self.free] = 0.0
        return R

    # -- Jacobians -----------------------------------------------------------
    def jacobian_fd(self, u, perturbation=1e-7):
        """assembly_jacobian_matrix_part: central differences, r == s
        entries, no rows for BC nodes (dense, small cases only)."""
        d, n = self.dim, self.n
        eps = perturbation * self.p.h
        J = np.zeros((n * d, n * d))
        for i in range(n):
            if self.bc_node[i]:
                continue
            trav = [i] + list(self.cols[self.fd.row_ptr[i]:self.fd.row_ptr[i + 1]])
            for j in trav:
                for r in range(d):
                    up = u.copy()
                    up[j, r] += eps
                    fp = self.fd.internal_force(up)[i]
                    um = u.copy()
                    um[j, r] -= eps
                    fm = self.fd.internal_force(um)[i]
                    J[i * d + r, j * d + r] = (fp[r] - fm[r]) / (2.0 * eps)
        return J

    def jacobian_analytic(self, u, component_only=False):
        """Dense J = df_int/du from the per-bond tangent scalars."""
        d, n = self.dim, self.n
        self.fd.internal_force(u)  # bond bits at u
        z = np.zeros_like(u)
        self.fd.set_state(u=u, v=z, f=z)
        k = self.fd.tangent()
        dx = self.p.x[self.cols, :d] - self.p.x[self.rows, :d]
        e = dx / np.linalg.norm(dx, axis=1)[:, None]
        B = k[:, None, None] * e[:, :, None] * e[:, None, :]
        if component_only:
            B = B * np.eye(d)[None]
        J = np.zeros((n * d, n * d))
        for a in range(d):
            for c in range(d):
                np.add.at(J, (self.rows * d + a, self.cols * d + c), B[:, a, c])
                np.add.at(J, (self.rows * d + a, self.rows * d + c), -B[:, a, c])
        return J

    # -- solver --------------------------------------------------------------
    def jacobian(self, u, kind):
        if kind == "fd":
            return self.jacobian_fd(u)
        return self.jacobian_analytic(u, component_only=kind == "component")

    def solve(self, nsteps, dt=1.0, tol_abs=0.0, tol_rel=1e-8, max_newton=50,
              jacobian="analytic", predictor=True):
        """Returns per load step (u, newton iterations, residual history);
        the predictor solve counts as the first iteration."""
        out = []
        d = self.dim
        fr = self.free
        u = self.fd.get_state()[0]
        for step in range(1, nsteps + 1):
            t = step * dt
            u_new, self.b = self.loads_at(u, t)
            it = 0
            if predictor:
                du_p = (u_new - u)[:, :d].reshape(-1)
                du_p[fr] = 0.0
                R = self.residual(u)
                J = self.jacobian(u, jacobian)
                rhs = -R - J @ du_p
                r0 = np.linalg.norm(rhs[fr])
                delta = du_p.copy()
                delta[fr] = np.linalg.solve(J[np.ix_(fr, fr)], rhs[fr])
                u = u.copy()
                u[:, :d] += delta.reshape(self.n, d)
                it = 1
                R = self.residual(u)
                hist = [r0, np.linalg.norm(R)]
            else:
                u = u_new
                R = self.residual(u)
                r0 = np.linalg.norm(R)
                hist = [r0]
            while hist[-1] > max(tol_abs, tol_rel * r0) and it < max_newton:
                J = self.jacobian(u, jacobian)
                delta = np.zeros(self.n * d)
                delta[fr] = np.linalg.solve(J[np.ix_(fr, fr)], -R[fr])
                u = u.copy()
                u[:, :d] += delta.reshape(self.n, d)
                R = self.residual(u)
                hist.append(np.linalg.norm(R))
                it += 1
            z = np.zeros_like(u)
            self.fd.set_state(u=u, v=z, f=z)
            out.append((u.copy(), it, hist))
        return out
