"""ctypes binding of the CPU oracle (oracle/nlm_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libnlm_oracle.so"

REGION = {"rectangle": 0, "angled_rectangle": 1, "circle": 2, "torus": 3,
          "line": 4, "cuboid": 5}
TIME_FN = {"constant": 0, "linear": 1, "quadratic": 2, "sin": 3,
           "linear_step": 4, "linear_slow_fast": 5}
SPACE_FN = {"constant": 0, "sin_x": 1, "sin_y": 2, "sin_z": 3, "linear_x": 4,
            "linear_y": 5, "linear_z": 6, "hat_x": 7, "hat_y": 8, "hat_z": 9,
            "line_load": 10}


class OraMatDeck(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("E", C.c_double), ("K", C.c_double),
        ("Gc", C.c_double), ("KIc", C.c_double), ("plane_strain", C.c_int),
        ("compute_from_classical", C.c_int), ("n_bond_params", C.c_int),
        ("bond_params", C.c_double * 2), ("inf_type", C.c_int),
        ("n_inf_params", C.c_int), ("inf_params", C.c_double * 2),
        ("check_sc_factor", C.c_double), ("irreversible", C.c_int),
        ("no_penetration", C.c_int),
    ]


class OraRnp(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("C", C.c_double), ("beta", C.c_double),
        ("rbar", C.c_double), ("inv_factor", C.c_double),
        ("factor_sc", C.c_double), ("irreversible", C.c_int),
        ("inf_type", C.c_int), ("inf_a", C.c_double), ("inf_b", C.c_double),
        ("inf_moment", C.c_double), ("contact", C.c_int),
        ("contact_kn", C.c_double), ("contact_rc", C.c_double),
        ("E", C.c_double), ("K", C.c_double), ("nu", C.c_double),
        ("lambda_", C.c_double), ("G", C.c_double), ("Gc", C.c_double),
        ("KIc", C.c_double),
    ]


class OraBc(C.Structure):
    _fields_ = [
        ("region", C.c_int), ("x1", C.c_double), ("x2", C.c_double),
        ("y1", C.c_double), ("y2", C.c_double), ("z1", C.c_double),
        ("z2", C.c_double), ("r1", C.c_double), ("r2", C.c_double),
        ("theta", C.c_double), ("ndir", C.c_int), ("dir", C.c_int * 3),
        ("time_fn", C.c_int), ("n_time_params", C.c_int),
        ("time_params", C.c_double * 4), ("space_fn", C.c_int),
        ("n_space_params", C.c_int), ("space_params", C.c_double * 4),
    ]


class OraEsDeck(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("E", C.c_double), ("K", C.c_double),
        ("G", C.c_double), ("plane_strain", C.c_int), ("sc", C.c_double),
    ]


class OraEs(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("horizon", C.c_double), ("h", C.c_double),
        ("density", C.c_double), ("E", C.c_double), ("K", C.c_double),
        ("G", C.c_double), ("nu", C.c_double), ("mu", C.c_double),
        ("factor2D", C.c_double), ("plane_strain", C.c_int),
        ("sc", C.c_double),
    ]


class OraCrack(C.Structure):
    _fields_ = [
        ("o", C.c_int), ("pb", C.c_double * 2), ("pt", C.c_double * 2),
        ("theta", C.c_double), ("activation_time", C.c_double),
    ]


_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_U8 = C.POINTER(C.c_uint8)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_P = C.c_void_p
SIGNATURES = {
    "ora_definitely_greater": (C.c_int, [C.c_double, C.c_double]),
    "ora_definitely_less": (C.c_int, [C.c_double, C.c_double]),
    "ora_rnp_params": (C.c_int, [C.POINTER(OraMatDeck), C.POINTER(OraRnp)]),
    "ora_bond_ef": (None, [C.POINTER(OraRnp), _D, _D, _D, _D, C.c_double,
                           C.c_double, C.c_int, C.c_int,
                           C.POINTER(C.c_int), _D, _D]),
    "ora_neighbor_bruteforce": (_P, [C.c_int64, _D, C.c_double]),
    "ora_neighbor_cells": (_P, [C.c_int64, _D, C.c_double]),
    "ora_nbr_nnz": (C.c_int64, [_P]),
    "ora_nbr_copy": (None, [_P, _I64, _I32]),
    "ora_nbr_free": (None, [_P]),
    "ora_fd_create": (_P, [C.c_int, C.c_int64, _D, _D, _I64, _I32,
                           C.POINTER(OraRnp), C.c_double, C.c_int,
                           C.POINTER(OraBc), C.c_int, C.POINTER(OraBc),
                           C.c_int, C.POINTER(OraCrack), _U8]),
    "ora_fd_free": (None, [_P]),
    "ora_fd_init": (None, [_P]),
    "ora_fd_step_cd": (None, [_P, C.c_int]),
    "ora_fd_step_verlet": (None, [_P, C.c_int]),
    "ora_fd_compute_forces": (None, [_P]),
    "ora_fd_get_state": (None, [_P, _D, _D, _D]),
    "ora_fd_set_state": (None, [_P, _D, _D, _D]),
    "ora_fd_get_bonds": (None, [_P, _U8]),
    "ora_fd_set_bonds": (None, [_P, _U8]),
    "ora_fd_get_bond_bytes": (None, [_P, _U8]),
    "ora_fd_bond_bytes_size": (C.c_int64, [_P]),
    "ora_fd_get_fixity": (None, [_P, _U8]),
    "ora_fd_time": (C.c_double, [_P]),
    "ora_fd_step_count": (C.c_int64, [_P]),
    "ora_fd_postproc": (None, [_P, _F, _F, _D]),
    "ora_fd_external_force": (None, [_P, C.c_double, _D]),
    "ora_fd_reaction_force": (None, [_P, _D, C.c_double, _D, _D]),
    "ora_fd_force_abs_sum": (None, [_P, _D]),
    "ora_fd_set_time": (None, [_P, C.c_double, C.c_int64]),
    "ora_interior_flags": (None, [C.c_int64, _D, _D, _D, C.c_double, C.c_int,
                                  _D, _U8]),
    "ora_bbox": (None, [C.c_int64, _D, _D, _D]),
    "ora_num_threads": (C.c_int, []),
    "ora_es_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ora_es_create": (_P, [C.c_int, C.c_int64, _D, _D, _I64, _I32,
                           C.c_void_p, C.c_double, C.c_int,
                           C.POINTER(OraBc), C.c_int, C.POINTER(OraBc)]),
    "ora_es_free": (None, [_P]),
    "ora_es_init": (None, [_P]),
    "ora_es_step_cd": (None, [_P, C.c_int]),
    "ora_es_compute_forces": (None, [_P]),
    "ora_es_get_state": (None, [_P, _D, _D, _D]),
    "ora_es_set_state": (None, [_P, _D, _D, _D]),
    "ora_es_get_bonds": (None, [_P, _U8]),
    "ora_es_set_bonds": (None, [_P, _U8]),
    "ora_es_get_node_fields": (None, [_P, _D, _D]),
    "ora_es_time": (C.c_double, [_P]),
    "ora_es_set_time": (None, [_P, C.c_double]),
    "ora_es_fields": (None, [_P, C.POINTER(C.c_float), _D, _D]),
    "ora_es_extension_rounding_bound": (None, [_P, _D]),
    "ora_fd_set_faithful": (None, [_P, C.c_int]),
    "ora_es_force_abs_sum": (None, [_P, _D]),
    "ora_fd_tangent": (None, [_P, _D]),
}

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def mat_deck(p) -> OraMatDeck:
    m = p.material
    d = OraMatDeck()
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


def rnp_params(p) -> OraRnp:
    lib = load()
    out = OraRnp()
    rc = lib.ora_rnp_params(C.byref(mat_deck(p)), C.byref(out))
    if rc != 0:
        raise ValueError(f"ora_rnp_params failed ({rc})")
    return out


def bc_struct(b) -> OraBc:
    s = OraBc()
    s.region = REGION[b.region]
    for k in ("x1", "x2", "y1", "y2", "z1", "z2", "r1", "r2", "theta"):
        setattr(s, k, getattr(b, k))
    s.ndir = len(b.dirs)
    for k, d in enumerate(b.dirs):
        s.dir[k] = d
    s.time_fn = TIME_FN[b.time_fn]
    s.n_time_params = len(b.time_params)
    for k, v in enumerate(b.time_params):
        s.time_params[k] = v
    s.space_fn = SPACE_FN[b.space_fn]
    s.n_space_params = len(b.space_params)
    for k, v in enumerate(b.space_params):
        s.space_params[k] = v
    return s


def neighbors(x: np.ndarray, horizon: float, method: str = "brute"):
    lib = load()
    x = np.ascontiguousarray(x, np.float64)
    fn = lib.ora_neighbor_bruteforce if method == "brute" else lib.ora_neighbor_cells
    h = fn(x.shape[0], _p(x, C.c_double), horizon)
    nnz = lib.ora_nbr_nnz(h)
    rp = np.zeros(x.shape[0] + 1, np.int64)
    col = np.zeros(max(nnz, 1), np.int32)
    lib.ora_nbr_copy(h, _p(rp, C.c_int64), _p(col, C.c_int32))
    lib.ora_nbr_free(h)
    return rp, col[:nnz]


class OracleFD:
    """FDModel<RNPBond> restated on the CPU (see nlm_oracle.cpp)."""

    def __init__(self, p, csr=None):
        self.lib = lib = load()
        self.p = p
        if csr is None:
            csr = neighbors(p.x, p.horizon, "cells" if p.n > 4000 else "brute")
        self.row_ptr = np.ascontiguousarray(csr[0], np.int64)
        self.col = np.ascontiguousarray(csr[1], np.int32)
        self.nnz = int(self.row_ptr[-1])
        self._x = np.ascontiguousarray(p.x, np.float64)
        self._vol = np.ascontiguousarray(p.vol, np.float64)
        self.mat = rnp_params(p)
        ub = (OraBc * max(len(p.ubc), 1))(*[bc_struct(b) for b in p.ubc])
        fb = (OraBc * max(len(p.fbc), 1))(*[bc_struct(b) for b in p.fbc])
        cr = (OraCrack * max(len(p.cracks), 1))()
        for k, c in enumerate(p.cracks):
            cr[k].o = c.o
            cr[k].pb[0], cr[k].pb[1] = c.pb
            cr[k].pt[0], cr[k].pt[1] = c.pt
            cr[k].theta = c.theta
            cr[k].activation_time = c.activation_time
        self._interior = (None if p.interior is None else
                          np.ascontiguousarray(p.interior, np.uint8))
        self.h = lib.ora_fd_create(
            p.dim, p.n, _p(self._x, C.c_double), _p(self._vol, C.c_double),
            _p(self.row_ptr, C.c_int64), _p(self.col, C.c_int32),
            C.byref(self.mat), p.dt, len(p.ubc), ub, len(p.fbc), fb,
            len(p.cracks), cr, _p(self._interior, C.c_uint8))
        # initial velocity, as the GPU engine applies it at setup
        if getattr(p, "v0", None) is not None:
            self.set_state(v=p.v0)

    def close(self):
        if self.h:
            self.lib.ora_fd_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        self.lib.ora_fd_init(self.h)

    def step(self, n, scheme="cd"):
        if scheme == "cd":
            self.lib.ora_fd_step_cd(self.h, int(n))
        else:
            self.lib.ora_fd_step_verlet(self.h, int(n))

    def compute_forces(self):
        self.lib.ora_fd_compute_forces(self.h)

    def set_faithful(self, on=True):
        """Reference-faithful timing variant: the per-node output-tag
        lookups and per-bond is_reaction_force of FDModel::computeForce and
        virtual interior-flag reads (same forces, more work per bond)."""
        self.lib.ora_fd_set_faithful(self.h, int(bool(on)))

    def get_state(self):
        n = self.p.n
        u, v, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self.lib.ora_fd_get_state(self.h, _p(u, C.c_double), _p(v, C.c_double),
                                  _p(f, C.c_double))
        return u, v, f

    def set_state(self, u=None, v=None, f=None):
        def c(a):
            return None if a is None else np.ascontiguousarray(a, np.float64)
        u, v, f = c(u), c(v), c(f)
        self.lib.ora_fd_set_state(self.h, _p(u, C.c_double), _p(v, C.c_double),
                                  _p(f, C.c_double))

    def get_bonds(self):
        b = np.zeros(max(self.nnz, 1), np.uint8)
        self.lib.ora_fd_get_bonds(self.h, _p(b, C.c_uint8))
        return b[: self.nnz]

    def set_bonds(self, b):
        b = np.ascontiguousarray(b, np.uint8)
        self.lib.ora_fd_set_bonds(self.h, _p(b, C.c_uint8))

    def get_bond_bytes(self):
        sz = self.lib.ora_fd_bond_bytes_size(self.h)
        b = np.zeros(max(sz, 1), np.uint8)
        self.lib.ora_fd_get_bond_bytes(self.h, _p(b, C.c_uint8))
        return b[:sz]

    def get_fixity(self):
        f = np.zeros(self.p.n, np.uint8)
        self.lib.ora_fd_get_fixity(self.h, _p(f, C.c_uint8))
        return f

    def time(self):
        return self.lib.ora_fd_time(self.h), self.lib.ora_fd_step_count(self.h)

    def set_time(self, t, step):
        self.lib.ora_fd_set_time(self.h, float(t), int(step))

    def force_abs_sum(self):
        out = np.zeros(self.p.n)
        self.lib.ora_fd_force_abs_sum(self.h, _p(out, C.c_double))
        return out

    def tangent(self):
        """Per-bond k of dF_ij/du_j = k e e^T at the current state."""
        k = np.zeros(max(self.nnz, 1))
        self.lib.ora_fd_tangent(self.h, _p(k, C.c_double))
        return k[: self.nnz]

    def internal_force(self, u):
        """f_int(u) with the current bond bits (which the evaluation may
        update, as computeForces does)."""
        z = np.zeros((self.p.n, 3))
        self.set_state(u=u, v=z, f=z)
        self.compute_forces()
        return self.get_state()[2]

    def postproc(self):
        n = self.p.n
        Z = np.zeros(n, np.float32)
        phi = np.zeros(n, np.float32)
        e = np.zeros(n)
        self.lib.ora_fd_postproc(self.h, _p(Z, C.c_float), _p(phi, C.c_float),
                                 _p(e, C.c_double))
        return Z, phi, e

    def external_force(self, t):
        f = np.zeros((self.p.n, 3))
        self.lib.ora_fd_external_force(self.h, float(t), _p(f, C.c_double))
        return f

    def reaction_force(self, box, delta):
        b = np.ascontiguousarray(box, np.float64)
        R = np.zeros((self.p.n, 3))
        L = np.zeros(self.p.n)
        self.lib.ora_fd_reaction_force(self.h, _p(b, C.c_double), float(delta),
                                       _p(R, C.c_double), _p(L, C.c_double))
        return R, L


def bond_ef(mat, xi, ui, xj, uj, voli, volj, int_i=1, int_j=1, fs=0):
    lib = load()
    a = [np.ascontiguousarray(v, np.float64) for v in (xi, ui, xj, uj)]
    F = np.zeros(3)
    E = np.zeros(1)
    b = C.c_int(fs)
    lib.ora_bond_ef(C.byref(mat), *[_p(v, C.c_double) for v in a], voli, volj,
                    int_i, int_j, C.byref(b), _p(F, C.c_double),
                    _p(E, C.c_double))
    return F, float(E[0]), b.value


def definitely_greater(a, b):
    return bool(load().ora_definitely_greater(a, b))


def definitely_less(a, b):
    return bool(load().ora_definitely_less(a, b))


def es_params(p) -> OraEs:
    """ElasticState constants of a problem whose material is elastic_state."""
    m = p.material
    d = OraEsDeck()
    d.dim = p.dim
    d.horizon = p.horizon
    d.h = p.h
    d.density = m["density"]
    d.E = m.get("E", -1.0)
    d.K = m.get("K", -1.0)
    d.G = m.get("G", -1.0)
    d.plane_strain = int(m.get("plane_strain", 0))
    d.sc = m.get("sc", -1.0)
    out = OraEs()
    rc = load().ora_es_params(C.byref(d), C.byref(out))
    if rc != 0:
        raise ValueError(f"ora_es_params failed ({rc})")
    return out


class OracleES:
    """Explicit linear state-based model (ElasticState force, QS scatter
    form, integrateCD update) restated on the CPU (see nlm_oracle.cpp)."""

    def __init__(self, p, csr=None):
        self.lib = lib = load()
        self.p = p
        if csr is None:
            csr = neighbors(p.x, p.horizon, "cells" if p.n > 4000 else "brute")
        self.row_ptr = np.ascontiguousarray(csr[0], np.int64)
        self.col = np.ascontiguousarray(csr[1], np.int32)
        self.nnz = int(self.row_ptr[-1])
        self._x = np.ascontiguousarray(p.x, np.float64)
        self._vol = np.ascontiguousarray(p.vol, np.float64)
        self.mat = es_params(p)
        ub = (OraBc * max(len(p.ubc), 1))(*[bc_struct(b) for b in p.ubc])
        fb = (OraBc * max(len(p.fbc), 1))(*[bc_struct(b) for b in p.fbc])
        self.h = lib.ora_es_create(
            p.dim, p.n, _p(self._x, C.c_double), _p(self._vol, C.c_double),
            _p(self.row_ptr, C.c_int64), _p(self.col, C.c_int32),
            C.byref(self.mat), p.dt, len(p.ubc), ub, len(p.fbc), fb)
        if getattr(p, "v0", None) is not None:
            self.set_state(v=p.v0)

    def close(self):
        if self.h:
            self.lib.ora_es_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        self.lib.ora_es_init(self.h)

    def step(self, n):
        self.lib.ora_es_step_cd(self.h, int(n))

    def compute_forces(self):
        self.lib.ora_es_compute_forces(self.h)

    def get_state(self):
        n = self.p.n
        u, v, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self.lib.ora_es_get_state(self.h, _p(u, C.c_double), _p(v, C.c_double),
                                  _p(f, C.c_double))
        return u, v, f

    def set_state(self, u=None, v=None, f=None):
        def c(a):
            return None if a is None else np.ascontiguousarray(a, np.float64)
        u, v, f = c(u), c(v), c(f)
        self.lib.ora_es_set_state(self.h, _p(u, C.c_double), _p(v, C.c_double),
                                  _p(f, C.c_double))

    def get_bonds(self):
        b = np.zeros(max(self.nnz, 1), np.uint8)
        self.lib.ora_es_get_bonds(self.h, _p(b, C.c_uint8))
        return b[: self.nnz]

    def set_bonds(self, b):
        b = np.ascontiguousarray(b, np.uint8)
        self.lib.ora_es_set_bonds(self.h, _p(b, C.c_uint8))

    def node_fields(self):
        """(dilatation theta_i, weighted volume m_i) of the last force pass."""
        th = np.zeros(self.p.n)
        mw = np.zeros(self.p.n)
        self.lib.ora_es_get_node_fields(self.h, _p(th, C.c_double),
                                        _p(mw, C.c_double))
        return th, mw

    def time(self):
        return self.lib.ora_es_time(self.h)

    def set_time(self, t):
        self.lib.ora_es_set_time(self.h, float(t))

    def fields(self, energy=None):
        """(strain energy, strain[n, 6], stress[n, 6]) at the current state:
        the output fields of QuasiStaticModel::computeForces(true); the
        strain energy is added in float to `energy` (zeros if None)."""
        n = self.p.n
        e = (np.zeros(n, np.float32) if energy is None
             else np.ascontiguousarray(energy, np.float32).copy())
        st = np.zeros((n, 6))
        ss = np.zeros((n, 6))
        self.lib.ora_es_fields(self.h, _p(e, C.c_float), _p(st, C.c_double),
                               _p(ss, C.c_double))
        return e, st, ss

    def extension_rounding_bound(self):
        """Per-node bound (3D) on the force difference that the reference's
        rounding of e = |Y| - |X| alone can cause at the current state."""
        out = np.zeros(self.p.n)
        self.lib.ora_es_extension_rounding_bound(self.h, _p(out, C.c_double))
        return out

    def force_abs_sum(self):
        out = np.zeros(self.p.n)
        self.lib.ora_es_force_abs_sum(self.h, _p(out, C.c_double))
        return out
