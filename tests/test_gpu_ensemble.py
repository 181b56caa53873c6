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
    remainder of steps outside whole graph chunks.
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
