"""Batched ensembles of uneven plates and graph-captured stepping, through
the C-ABI.

Bars:
  - node-list BCs reproduce the equivalent rectangle BC sets bit for bit;
  - the packed ensemble matches the oracle (FP64 exact: relative L2
    displacement error <= 1e-9, identical broken bonds), the oracle given
    the same plates with rectangle BC sets;
  - a plate stepped inside the ensemble equals the plate stepped alone, bit
    for bit (rows keep the reference order, so per-node sums are the same);
  - graph-captured stepping equals direct stepping bit for bit, including a
    remainder of steps outside whole graph chunks, also on the full
    256-plate workload with the jointly tuned tile and chunk width, where
    the pulled rows follow their rates exactly, no bond re-forms and each
    plate's internal forces sum to zero to rounding.
"""
import numpy as np
import pytest

from conftest import progress
from nlmech_amd import problems as pr
from nlmech_amd.capi import NLM_PREC_FP32_FORCE, NLM_PREC_FP64_EXACT

pytestmark = pytest.mark.gpu


def run(p, mode, steps, graph_chunk=0, csr=None):
    from nlmech_amd.engine import GpuFD
    g = GpuFD(p, precision=mode, csr=csr)
    g.init()
    if graph_chunk:
        g.step_graph(steps, graph_chunk)
    else:
        g.step(steps)
    u, v, f = g.get_state()
    b = g.get_bonds()
    t = g.time()
    g.close()
    return u, v, f, b, t


def test_node_bc_equals_bc_sets():
    progress("[gpu] ensemble node-list BCs vs sets")
    pn = pr.ensemble(count=4, lo=20, hi=32, dt=1e-7, pull=5.0)
    ps = pr.ensemble(count=4, lo=20, hi=32, dt=1e-7, pull=5.0, as_sets=True)
    a = run(pn, NLM_PREC_FP64_EXACT, 30)
    b = run(ps, NLM_PREC_FP64_EXACT, 30)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)


def test_ensemble_matches_oracle(oracle):
    progress("[gpu] ensemble vs oracle")
    ps = pr.ensemble(count=5, lo=18, hi=30, dt=1e-7, pull=20.0,
                     as_sets=True)
    o = oracle.OracleFD(ps)
    o.init()
    o.step(40)
    u_ref, _, _ = o.get_state()
    b_ref = o.get_bonds()
    csr = (o.row_ptr, o.col)
    o.close()
    pn = pr.ensemble(count=5, lo=18, hi=30, dt=1e-7, pull=20.0)
    u, _, _, b, _ = run(pn, NLM_PREC_FP64_EXACT, 40, csr=csr)
    assert b_ref.sum() > 0
    assert np.linalg.norm(u - u_ref) <= 1e-9 * np.linalg.norm(u_ref)
    assert np.array_equal(b, b_ref)


def test_member_alone_equals_member_in_ensemble():
    progress("[gpu] ensemble member alone vs packed")
    p = pr.ensemble(count=6, lo=20, hi=40, dt=1e-7, pull=5.0)
    u, _, _, b, _ = run(p, NLM_PREC_FP64_EXACT, 25)
    k = 3
    m = pr.ensemble_member(p, k)
    um, _, _, bm, _ = run(m, NLM_PREC_FP64_EXACT, 25)
    a = p.specimens[k]["offset"]
    assert np.array_equal(u[a:a + m.n], um)


@pytest.mark.parametrize("mode", [NLM_PREC_FP64_EXACT, NLM_PREC_FP32_FORCE])
def test_graph_steps_equal_direct_steps(mode):
    progress(f"[gpu] graph steps mode {mode}")
    p = pr.ensemble(count=4, lo=30, hi=60, dt=1e-7, pull=20.0)
    for steps, chunk in ((21, 4), (40, 6)):
        a = run(p, mode, steps)
        b = run(p, mode, steps, graph_chunk=chunk)
        for x, y in zip(a[:4], b[:4]):
            assert np.array_equal(x, y)
        assert a[4] == b[4]


def test_ensemble256_graph_equals_direct_tuned():
    """ensemble256 (256 plates, 15.8M nodes, 376M bonds), FP32-force, tile
    and chunk width from the joint autotuner: 9 steps replayed from graphs of
    4 (plus a remainder of 1) equal 9 direct steps bit for bit."""
    from nlmech_amd.engine import GpuFD
    progress("[gpu] ensemble256 graph vs direct, tuned")
    p = pr.workload("ensemble256")
    steps = 9
    with GpuFD(p, precision=NLM_PREC_FP32_FORCE) as g:
        g.init()
        b0 = g.get_bonds()
        tile, unroll, tab = g.autotune_tiles()
        progress(f"[gpu]   tuned: tile {tile} chunk {unroll}\n{tab}")
        g.step(steps)
        a = g.get_state() + (g.get_bonds(), g.time())
    with GpuFD(p, precision=NLM_PREC_FP32_FORCE) as g:
        g.set_tile(tile)
        g.set_unroll(unroll)
        g.init()
        g.step_graph(steps, 4)
        b = g.get_state() + (g.get_bonds(), g.time())
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    assert a[4] == b[4]
    u, v, f, bits, (t, n) = a
    assert n == steps and np.isfinite(u).all() and np.isfinite(v).all()
    assert not np.any(bits & ~b0)  # no bond re-forms
    # pulled rows: u_y = +/- pull t exactly (node-list BCs, linear in time)
    ids, rates = p.node_bc[0], np.asarray(p.node_bc[2]).reshape(-1, 3)
    assert np.allclose(u[ids, 1], rates[:, 1] * t, rtol=1e-12, atol=0.0)
    # per plate: sum_i V_i f_i = 0 to the FP32-force rounding
    for sp in p.specimens[::37]:
        sl = slice(sp["offset"], sp["offset"] + sp["side"] ** 2)
        vf = p.vol[sl, None] * f[sl]
        assert np.all(np.abs(vf.sum(0)) <= 1e-6 * np.abs(vf).sum()), sp
