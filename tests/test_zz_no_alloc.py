"""The allocation-free post-processing entry point, nlm_fd_postproc_noalloc,
against nlm_fd_postproc through the C-ABI: Z, phi and strain energy bit for
bit, on two calls in a row (the second on a smaller model, so the shared
static staging still holds the first call's larger contents), and the
clear error beyond the staging limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import progress
from nlmech_amd import capi
from nlmech_amd import problems as pr
from nlmech_amd.capi import NLM_PREC_FP64, NLM_PREC_FP32_FORCE, ptr

pytestmark = pytest.mark.gpu

MAX_NODES = 1048576  # NLM_NOALLOC_MAX_NODES
NLM_ERR_INVALID = -1


def postproc_both(g, noalloc_first):
    lib = capi.load()
    n = g.p.n
    out = {}
    order = ("noalloc", "plain") if noalloc_first else ("plain", "noalloc")
    for which in order:
        Z = np.full(n, np.nan, np.float32)
        phi = np.full(n, np.nan, np.float32)
        e = np.full(n, np.nan)
        fn = (lib.nlm_fd_postproc_noalloc if which == "noalloc"
              else lib.nlm_fd_postproc)
        rc = fn(g.ctx, ptr(Z, C.c_float), ptr(phi, C.c_float),
                ptr(e, C.c_double))
        assert rc == 0, lib.nlm_last_error(g.ctx)
        out[which] = (Z, phi, e)
    return out


def check_equal(out):
    for a, b in zip(out["noalloc"], out["plain"]):
        assert not np.isnan(a).any()
        assert np.array_equal(a, b)


def test_noalloc_postproc_matches_twice_then_smaller():
    from nlmech_amd.engine import GpuFD
    progress("[gpu] noalloc postproc: 300x300 plate, then 70x70")
    big = pr.plate2d(300, 300, 1e-3, 3, 1e-7, 1, notch_len=0.06, pull=30.0)
    with GpuFD(big, precision=NLM_PREC_FP64) as g:
        g.init()
        g.step(60)
        out = postproc_both(g, noalloc_first=False)
        check_equal(out)
        assert (out["plain"][1] > 0).sum() > 0      # damaged nodes
        assert out["plain"][0].max() > 0
        # a second call on the same model, noalloc first
        check_equal(postproc_both(g, noalloc_first=True))
    small = pr.plate2d(70, 70, 1e-3, 3, 1e-7, 1, notch_len=0.02, pull=20.0)
    with GpuFD(small, precision=NLM_PREC_FP32_FORCE) as g:
        g.init()
        g.step(40)
        out = postproc_both(g, noalloc_first=True)
        check_equal(out)
        assert out["noalloc"][0].size == small.n


def test_noalloc_postproc_null_outputs_and_limit():
    from nlmech_amd.engine import GpuFD
    lib = capi.load()
    progress("[gpu] noalloc postproc: null outputs, size limit")
    p = pr.plate2d(50, 50, 1e-3, 3, 1e-7, 1, notch_len=0.01, pull=10.0)
    with GpuFD(p, precision=NLM_PREC_FP64) as g:
        g.init()
        g.step(10)
        phi = np.zeros(p.n, np.float32)
        assert lib.nlm_fd_postproc_noalloc(g.ctx, None, ptr(phi, C.c_float),
                                           None) == 0
        ref = postproc_both(g, noalloc_first=False)["plain"][1]
        assert np.array_equal(phi, ref)
    n_side = 1025                       # 1,050,625 nodes > the limit
    assert n_side * n_side > MAX_NODES
    p = pr.plate2d(n_side, n_side, 1e-4, 3, 1e-8, 1, notch_len=0.0, pull=0.0)
    with GpuFD(p, precision=NLM_PREC_FP32_FORCE) as g:
        Z = np.zeros(p.n, np.float32)
        rc = lib.nlm_fd_postproc_noalloc(g.ctx, ptr(Z, C.c_float), None, None)
        assert rc == NLM_ERR_INVALID
        assert b"staging limit" in lib.nlm_last_error(g.ctx)
