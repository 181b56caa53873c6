"""Python driver of the C-ABI: one explicit FD model on one GPU.

``GpuFD`` mirrors the part of model::FDModel<RNPBond> that the library
replaces (src/model/fd/fDModel.cpp): setup of mesh, neighbor lists, fracture,
loading and material (initHObjects, :134-189), the integrate() prologue
(:318-333), the step loop (:344-373) and post-processing (:652-778).  All
compute happens in libnlmech_amd.so; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import (NLM_MATERIAL_ELASTIC_STATE, NLM_MATERIAL_RNP_BOND,
                   NLM_PREC_FP64_EXACT, NLM_TIME_CENTRAL_DIFFERENCE, Bc, Crack,
                   EsDeck, EsParams, FdProblem, MaterialDeck, NlmError,
                   RnpParams, ptr)
from .problems import BC, CrackSpec, Problem


def _check(lib, ctx, rc, what):
    if rc != 0:
        msg = lib.nlm_last_error(ctx)
        raise NlmError(f"{what} failed ({rc}): "
                       f"{msg.decode() if msg else ''}")


def material_deck(p: Problem) -> MaterialDeck:
    m = p.material
    d = MaterialDeck()
    d.dim = p.dim
    d.horizon = p.horizon
    d.h = p.h
    d.density = m["density"]
    d.E = m.get("E", -1.0)
    d.K = m.get("K", -1.0)
    d.Gc = m.get("Gc", -1.0)
    d.KIc = m.get("KIc", -1.0)
    d.plane_strain = int(m.get("plane_strain", 0))
    d.compute_from_classical = int(m.get("compute_from_classical", 1))
    bp = list(m.get("bond_params", []))
    d.n_bond_params = len(bp)
    for k, v in enumerate(bp[:2]):
        d.bond_params[k] = v
    d.inf_type = int(m.get("inf_type", 0))
    ip = list(m.get("inf_params", []))
    d.n_inf_params = len(ip)
    for k, v in enumerate(ip[:2]):
        d.inf_params[k] = v
    d.check_sc_factor = m.get("check_sc_factor", 1.0)
    d.irreversible = int(m.get("irreversible", 1))
    d.no_penetration = int(m.get("no_penetration", 0))
    return d


def rnp_params(p: Problem) -> RnpParams:
    lib = capi.load()
    d = material_deck(p)
    out = RnpParams()
    rc = lib.nlm_rnp_params_from_deck(C.byref(d), C.byref(out))
    if rc != 0:
        raise NlmError(f"nlm_rnp_params_from_deck failed ({rc})")
    return out


def is_state_based(p: Problem) -> bool:
    return p.material.get("model") == "elastic_state"


def es_params(p: Problem) -> EsParams:
    """ElasticState constants (nlm_es_params_from_deck)."""
    lib = capi.load()
    m = p.material
    d = EsDeck()
    d.dim = p.dim
    d.horizon = p.horizon
    d.h = p.h
    d.density = m["density"]
    d.E = m.get("E", -1.0)
    d.K = m.get("K", -1.0)
    d.G = m.get("G", -1.0)
    d.plane_strain = int(m.get("plane_strain", 0))
    d.critical_stretch = m.get("sc", -1.0)
    out = EsParams()
    rc = lib.nlm_es_params_from_deck(C.byref(d), C.byref(out))
    if rc != 0:
        raise NlmError(f"nlm_es_params_from_deck failed ({rc})")
    return out


def bc_struct(b: BC) -> Bc:
    s = Bc()
    s.region = capi.REGION[b.region]
    for k in ("x1", "x2", "y1", "y2", "z1", "z2", "r1", "r2", "theta"):
        setattr(s, k, getattr(b, k))
    s.ndir = len(b.dirs)
    for k, d in enumerate(b.dirs):
        s.dir[k] = d
    s.time_fn = capi.TIME_FN[b.time_fn]
    s.n_time_params = len(b.time_params)
    for k, v in enumerate(b.time_params):
        s.time_params[k] = v
    s.space_fn = capi.SPACE_FN[b.space_fn]
    s.n_space_params = len(b.space_params)
    for k, v in enumerate(b.space_params):
        s.space_params[k] = v
    return s


def crack_struct(c: CrackSpec) -> Crack:
    s = Crack()
    s.o = c.o
    s.pb[0], s.pb[1] = c.pb
    s.pt[0], s.pt[1] = c.pt
    s.theta = c.theta
    s.activation_time = c.activation_time
    return s


class GpuFD:
    """One FDModel<T> instance on the GPU, T = RNPBond or (explicit)
    ElasticState as the problem's material says."""

    def __init__(self, p: Problem, precision: int = NLM_PREC_FP64_EXACT,
                 scheme: int = NLM_TIME_CENTRAL_DIFFERENCE, device: int = 0,
                 csr: tuple | None = None):
        self.lib = lib = capi.load()
        self.p = p
        h = C.c_void_p()
        rc = lib.nlm_ctx_create(C.byref(h), device)
        if rc != 0:
            raise NlmError(f"nlm_ctx_create failed ({rc}): "
                           f"{lib.nlm_last_error(None).decode()}")
        self.ctx = h
        self._x = np.ascontiguousarray(p.x, dtype=np.float64)
        self._vol = np.ascontiguousarray(p.vol, dtype=np.float64)
        if csr is None:
            nnz = C.c_int64()
            self._call("nlm_neighbor_build", p.dim, p.n, ptr(self._x, C.c_double),
                       p.horizon, C.byref(nnz))
            self.nnz = nnz.value
            row_ptr = col = None
        else:
            row_ptr = np.ascontiguousarray(csr[0], dtype=np.int64)
            col = np.ascontiguousarray(csr[1], dtype=np.int32)
            self.nnz = int(row_ptr[-1])
        prob = FdProblem()
        prob.dim = p.dim
        prob.n = p.n
        prob.x = ptr(self._x, C.c_double)
        prob.vol = ptr(self._vol, C.c_double)
        prob.row_ptr = ptr(row_ptr, C.c_int64)
        prob.col = ptr(col, C.c_int32)
        self._interior = (None if p.interior is None else
                          np.ascontiguousarray(p.interior, dtype=np.uint8))
        prob.interior = ptr(self._interior, C.c_uint8)
        prob.dt = p.dt
        prob.precision = precision
        prob.time_scheme = scheme
        if is_state_based(p):
            prob.material = NLM_MATERIAL_ELASTIC_STATE
            prob.es = es_params(p)
            prob.mat.density = p.material["density"]
            self.mat = prob.es
        else:
            prob.material = NLM_MATERIAL_RNP_BOND
            prob.mat = rnp_params(p)
            self.mat = prob.mat
        self._call("nlm_fd_setup", C.byref(prob))
        if p.ubc:
            arr = (Bc * len(p.ubc))(*[bc_struct(b) for b in p.ubc])
            self._call("nlm_fd_set_ubc", len(p.ubc), arr)
        if p.fbc:
            arr = (Bc * len(p.fbc))(*[bc_struct(b) for b in p.fbc])
            self._call("nlm_fd_set_fbc", len(p.fbc), arr)
        if p.cracks:
            arr = (Crack * len(p.cracks))(*[crack_struct(c) for c in p.cracks])
            self._call("nlm_fd_add_cracks", len(p.cracks), arr)
        if getattr(p, "node_bc", None) is not None:
            self.set_node_ubc(*p.node_bc)
        if getattr(p, "v0", None) is not None:
            self.set_state(v=p.v0)

    def _node_vec(self, a, name):
        """A per-node vector field as the C-ABI reads it: double[n][3]."""
        if a is None:
            return None
        a = np.ascontiguousarray(a, np.float64)
        if a.shape != (self.p.n, 3):
            raise ValueError(f"{name}: expected shape ({self.p.n}, 3), got "
                             f"{a.shape}")
        return a

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(self.ctx, *args)
        _check(self.lib, self.ctx, rc, name)

    def close(self):
        if self.ctx:
            self.lib.nlm_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- model ------------------------------------------------------------
    def init(self):
        self._call("nlm_fd_init")

    def step(self, nsteps: int):
        self._call("nlm_fd_step", int(nsteps))

    def compute_forces(self):
        self._call("nlm_fd_compute_forces")

    def autotune(self):
        u = C.c_int()
        us = C.c_double()
        self._call("nlm_fd_autotune", C.byref(u), C.byref(us))
        return u.value, us.value

    def set_unroll(self, unroll: int):
        self._call("nlm_fd_set_unroll", int(unroll))

    def autotune_tiles(self):
        """(tile, chunk, table): the chosen staged-kernel tile and chunk
        width and the 3 x 3 table of microseconds per step (rows: tiles 128,
        256, 512; columns: chunks 2, 4, 8; 0 = not run)."""
        t, u = C.c_int(), C.c_int()
        us = np.zeros(9)
        self._call("nlm_fd_autotune_tiles", C.byref(t), C.byref(u),
                   ptr(us, C.c_double))
        return t.value, u.value, us.reshape(3, 3)

    def set_tile(self, tile: int):
        self._call("nlm_fd_set_tile", int(tile))

    def timed_steps(self, nsteps: int):
        tot = C.c_double()
        per = C.c_double()
        self._call("nlm_fd_timed_steps", int(nsteps), C.byref(tot),
                   C.byref(per))
        return tot.value, per.value

    def set_node_ubc(self, nodes, dofs, rates):
        """Node-list displacement BCs: u = rate t on the masked dofs."""
        nodes = np.ascontiguousarray(nodes, np.int64).ravel()
        dofs = np.ascontiguousarray(dofs, np.uint8).ravel()
        rates = np.ascontiguousarray(rates, np.float64)
        k = nodes.size
        if dofs.size != k or rates.size != 3 * k:
            raise ValueError(f"set_node_ubc: {k} nodes need {k} dof masks and "
                             f"{k} x 3 rates, got {dofs.size} and "
                             f"{rates.shape}")
        if k and (nodes.min() < 0 or nodes.max() >= self.p.n):
            raise ValueError("set_node_ubc: node index out of range")
        rates = rates.reshape(-1, 3)
        self._call("nlm_fd_set_node_ubc", int(nodes.size), ptr(nodes, C.c_int64),
                   ptr(dofs, C.c_uint8), ptr(rates, C.c_double))

    def step_graph(self, nsteps: int, chunk: int = 50):
        self._call("nlm_fd_step_graph", int(nsteps), int(chunk))

    def timed_steps_graph(self, nsteps: int, chunk: int = 50):
        tot = C.c_double()
        per = C.c_double()
        self._call("nlm_fd_timed_steps_graph", int(nsteps), int(chunk),
                   C.byref(tot), C.byref(per))
        return tot.value, per.value

    def neighbors(self):
        rp = np.zeros(self.p.n + 1, np.int64)
        col = np.zeros(max(self.nnz, 1), np.int32)
        self._call("nlm_neighbor_get", ptr(rp, C.c_int64), ptr(col, C.c_int32))
        return rp, col[: self.nnz]

    def get_state(self):
        n = self.p.n
        u = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 3))
        self._call("nlm_fd_get_state", ptr(u, C.c_double), ptr(v, C.c_double),
                   ptr(f, C.c_double))
        return u, v, f

    def set_state(self, u=None, v=None, f=None):
        u, v, f = (self._node_vec(a, nm) for a, nm in ((u, "u"), (v, "v"),
                                                        (f, "f")))
        self._call("nlm_fd_set_state", ptr(u, C.c_double), ptr(v, C.c_double),
                   ptr(f, C.c_double))

    def get_bonds(self) -> np.ndarray:
        b = np.zeros(max(self.nnz, 1), np.uint8)
        self._call("nlm_fd_get_bonds", ptr(b, C.c_uint8))
        return b[: self.nnz]

    def set_bonds(self, b: np.ndarray):
        b = np.ascontiguousarray(b, np.uint8)
        if b.size != self.nnz:
            raise ValueError(f"set_bonds: {b.size} bond states given, the "
                             f"model has {self.nnz} directed bonds")
        self._call("nlm_fd_set_bonds", ptr(b, C.c_uint8))

    def get_bond_bytes(self) -> np.ndarray:
        sz = C.c_int64()
        self._call("nlm_fd_bond_bytes_size", C.byref(sz))
        b = np.zeros(max(sz.value, 1), np.uint8)
        self._call("nlm_fd_get_bond_bytes", ptr(b, C.c_uint8))
        return b[: sz.value]

    def get_fixity(self) -> np.ndarray:
        f = np.zeros(self.p.n, np.uint8)
        self._call("nlm_fd_get_fixity", ptr(f, C.c_uint8))
        return f

    def time(self):
        t = C.c_double()
        s = C.c_int64()
        self._call("nlm_fd_get_time", C.byref(t), C.byref(s))
        return t.value, s.value

    # -- quasi-static model (2D bond-based) ---------------------------------
    def qs_solve(self, nsteps, dt=1.0, tol_abs=0.0, tol_rel=1e-8,
                 max_newton=50, tangent=0, precond=1, lin_tol=1e-8,
                 lin_max_iter=5000, predictor=True, gnn_refresh=1,
                 perturbation=0.0):
        """nsteps load steps of the quasi-static Newton solver; returns the
        run statistics as a dict."""
        o = capi.QsOptions(nsteps=int(nsteps), dt=float(dt),
                           tol_abs=float(tol_abs), tol_rel=float(tol_rel),
                           max_newton=int(max_newton), tangent=int(tangent),
                           precond=int(precond), lin_tol=float(lin_tol),
                           lin_max_iter=int(lin_max_iter),
                           predictor=int(bool(predictor)),
                           gnn_refresh=int(gnn_refresh),
                           perturbation=float(perturbation))
        st = capi.QsStats()
        self._call("nlm_qs_solve", C.byref(o), C.byref(st))
        return {f: getattr(st, f) for f, _ in capi.QsStats._fields_}

    def qs_set_graph(self, on: bool):
        """BiCGSTAB batches from a captured HIP graph (on) or launched
        kernel by kernel (off)."""
        self._call("nlm_qs_set_graph", int(bool(on)))

    def qs_set_launch(self, graph: bool, bond_scalar: bool = True):
        """BiCGSTAB: batches from graph replays or kernel by kernel; products
        with the analytic bond tangent from one scalar per bond (True) or
        from the stored 2x2 blocks (False)."""
        self._call("nlm_qs_set_launch", int(bool(graph)),
                   int(bool(bond_scalar)))

    def qs_assemble(self, tangent=0):
        self._call("nlm_qs_assemble", int(tangent))

    def qs_tangent(self):
        """(blocks[nnz, 2, 2] in reference CSR order, diag[n, 2, 2])."""
        b = np.zeros((max(self.nnz, 1), 4))
        d = np.zeros((self.p.n, 4))
        self._call("nlm_qs_get_tangent", ptr(b, C.c_double), ptr(d, C.c_double))
        return b[: self.nnz].reshape(-1, 2, 2), d.reshape(-1, 2, 2)

    def qs_apply(self, x):
        x = self._node_vec(x, "x")
        y = np.zeros((self.p.n, 3))
        self._call("nlm_qs_apply", ptr(x, C.c_double), ptr(y, C.c_double))
        return y

    def qs_set_precond(self, blocks, diag):
        b = np.ascontiguousarray(np.asarray(blocks).reshape(-1, 4), np.float64)
        d = np.ascontiguousarray(np.asarray(diag).reshape(-1, 4), np.float64)
        if b.shape[0] != self.nnz or d.shape[0] != self.p.n:
            raise ValueError(f"qs_set_precond: need {self.nnz} bond blocks and "
                             f"{self.p.n} node blocks, got {b.shape[0]} and "
                             f"{d.shape[0]}")
        self._call("nlm_qs_set_precond", ptr(b, C.c_double), ptr(d, C.c_double))

    def qs_coarse_apply(self, r):
        r = self._node_vec(r, "r")
        z = np.zeros((self.p.n, 3))
        self._call("nlm_qs_coarse_apply", ptr(r, C.c_double), ptr(z, C.c_double))
        return z

    def qs_precond(self):
        """(blocks[nnz, 2, 2], diag[n, 2, 2]) of the preconditioner."""
        b = np.zeros((max(self.nnz, 1), 4))
        d = np.zeros((self.p.n, 4))
        self._call("nlm_qs_get_precond", ptr(b, C.c_double), ptr(d, C.c_double))
        return b[: self.nnz].reshape(-1, 2, 2), d.reshape(-1, 2, 2)

    # -- learned preconditioner -------------------------------------------
    def gnn_create(self, seed=1):
        self._call("nlm_gnn_create", C.c_uint64(seed))

    def gnn_weights(self):
        w = np.zeros(self.lib.nlm_gnn_num_params(), np.float32)
        self._call("nlm_gnn_get_weights", ptr(w, C.c_float))
        return w

    def gnn_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != self.lib.nlm_gnn_num_params():
            raise ValueError(f"gnn_set_weights: {w.size} weights given, the "
                             f"network has {self.lib.nlm_gnn_num_params()}")
        self._call("nlm_gnn_set_weights", ptr(w, C.c_float))

    def gnn_predict(self):
        self._call("nlm_gnn_predict")

    def gnn_loss_grad(self, r, e, want_grad=False):
        """r, e: [K, n, 3] residual / error pairs; returns (loss, grad)."""
        r = np.ascontiguousarray(r, np.float64)
        e = np.ascontiguousarray(e, np.float64)
        K = r.shape[0]
        if r.shape != (K, self.p.n, 3) or e.shape != r.shape:
            raise ValueError(f"gnn_loss_grad: r and e must be ({K}, "
                             f"{self.p.n}, 3), got {r.shape} and {e.shape}")
        loss = C.c_double()
        g = np.zeros(self.lib.nlm_gnn_num_params(), np.float32) if want_grad else None
        self._call("nlm_gnn_loss_grad", int(K), ptr(r, C.c_double),
                   ptr(e, C.c_double), C.byref(loss), ptr(g, C.c_float))
        return loss.value, g

    def gnn_adam_step(self, lr):
        self._call("nlm_gnn_adam_step", float(lr))

    def gnn_zero_grad(self):
        self._call("nlm_gnn_zero_grad")

    def qs_linear_solve(self, b, precond=1, tol=1e-8, max_iter=5000):
        b = self._node_vec(b, "b")
        x = np.zeros((self.p.n, 3))
        it = C.c_int()
        rr = C.c_double()
        self._call("nlm_qs_linear_solve", ptr(b, C.c_double), ptr(x, C.c_double),
                   int(precond), float(tol), int(max_iter), C.byref(it),
                   C.byref(rr))
        return x, it.value, rr.value

    def node_fields(self):
        """State-based material: (dilatation, weighted volume) per node."""
        th = np.zeros(self.p.n)
        mw = np.zeros(self.p.n)
        self._call("nlm_fd_get_node_fields", ptr(th, C.c_double),
                   ptr(mw, C.c_double))
        return th, mw

    def es_fields(self, energy=None):
        """State-based material: (strain energy, strain[n, 6], stress[n, 6])
        of QuasiStaticModel::computeForces(true) at the current state; the
        strain energy is added in float to `energy` (zeros if None)."""
        n = self.p.n
        e = (np.zeros(n, np.float32) if energy is None
             else np.ascontiguousarray(energy, np.float32).copy())
        if e.shape != (n,):
            raise ValueError(f"es_fields: energy must have {n} entries")
        st = np.zeros((n, 6))
        ss = np.zeros((n, 6))
        self._call("nlm_fd_es_fields", ptr(e, C.c_float), ptr(st, C.c_double),
                   ptr(ss, C.c_double))
        return e, st, ss

    def postproc(self):
        n = self.p.n
        Z = np.zeros(n, np.float32)
        phi = np.zeros(n, np.float32)
        e = np.zeros(n)
        self._call("nlm_fd_postproc", ptr(Z, C.c_float), ptr(phi, C.c_float),
                   ptr(e, C.c_double))
        return Z, phi, e

    def external_force(self, t: float):
        """External force density (n, 3) at time t (FLoading::apply)."""
        f = np.zeros((self.p.n, 3))
        self._call("nlm_fd_external_force", float(t), ptr(f, C.c_double))
        return f

    def reaction_force(self, box, horizon: float):
        """Reaction force (n, 3) and its per-node length; box = (min_x,
        max_x, max_y) of the mesh bounding box."""
        b = np.ascontiguousarray(box, np.float64)
        R = np.zeros((self.p.n, 3))
        L = np.zeros(self.p.n)
        self._call("nlm_fd_reaction_force", ptr(b, C.c_double), float(horizon),
                   ptr(R, C.c_double), ptr(L, C.c_double))
        return R, L
