"""CPU checks of the C-ABI library: it loads, exports every entry point
include/nlmech_amd.h declares, and its host-side logic (RNPBond constants)
equals the oracle's.  No GPU work is issued here."""
import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT
from nlmech_amd import capi
from nlmech_amd import problems as pr


def declared_functions():
    text = (ROOT / "include" / "nlmech_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nlm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers the header
    assert set(names) <= set(capi.SIGNATURES)


@pytest.mark.parametrize("variant", ["default", "plane_strain", "linear",
                                     "gauss", "params", "contact", "K_only"])
def test_rnp_params_equal_oracle(oracle, variant):
    from nlmech_amd.engine import material_deck
    p = pr.plate2d(4, 4, 1e-3, 4, 1e-8, 1)
    mat = pr.rnp_material()
    if variant == "plane_strain":
        mat = pr.rnp_material(plane_strain=True)
    elif variant == "linear":
        mat = pr.rnp_material(inf_type=1, inf_params=())
    elif variant == "gauss":
        mat = pr.rnp_material(inf_type=2, inf_params=())
    elif variant == "params":
        mat.update(compute_from_classical=0, bond_params=[900.0, 3e7])
    elif variant == "contact":
        mat = pr.rnp_material(no_penetration=True)
    elif variant == "K_only":
        mat.update(E=-1.0, K=2.0e9, Gc=-1.0, KIc=1.2e6)
    p = pr.with_(p, material=mat)
    lib = capi.load()
    out = capi.RnpParams()
    assert lib.nlm_rnp_params_from_deck(C.byref(material_deck(p)),
                                        C.byref(out)) == 0
    ref = oracle.rnp_params(p)
    for f, _ in capi.RnpParams._fields_:
        a, b = getattr(out, f), getattr(ref, f)
        assert a == b or (np.isnan(a) and np.isnan(b)), f


def test_rnp_params_rejects_bad_deck():
    from nlmech_amd.engine import material_deck
    lib = capi.load()
    p = pr.plate2d(4, 4, 1e-3, 3, 1e-7, 1)
    bad = pr.with_(p, material=dict(pr.rnp_material(), E=-1.0, K=-1.0))
    out = capi.RnpParams()
    assert lib.nlm_rnp_params_from_deck(C.byref(material_deck(bad)),
                                        C.byref(out)) != 0
    bad = pr.with_(p, material=dict(pr.rnp_material(), inf_type=7))
    assert lib.nlm_rnp_params_from_deck(C.byref(material_deck(bad)),
                                        C.byref(out)) != 0
    # explicit bond parameters [C, beta] need beta > 0
    for beta in (0.0, -1.0):
        bad = pr.with_(p, material=dict(pr.rnp_material(),
                                        compute_from_classical=0,
                                        bond_params=[100.0, beta]))
        assert lib.nlm_rnp_params_from_deck(C.byref(material_deck(bad)),
                                            C.byref(out)) != 0
    good = pr.with_(p, material=dict(pr.rnp_material(),
                                     compute_from_classical=0,
                                     bond_params=[100.0, 2.0e7]))
    assert lib.nlm_rnp_params_from_deck(C.byref(material_deck(good)),
                                        C.byref(out)) == 0


@pytest.mark.parametrize("case", ["qs2d", "qs1d", "cube", "plane_strain"])
def test_es_params_equal_oracle(oracle, case):
    from es_cases import qs1d, qs2d, small_cube
    from nlmech_amd.engine import es_params
    if case == "qs2d":
        p, _ = qs2d()
    elif case == "qs1d":
        p, _ = qs1d()
    elif case == "cube":
        p = small_cube()
    else:
        p, _ = qs2d()
        p = pr.with_(p, material=dict(p.material, plane_strain=1))
    out = es_params(p)
    ref = oracle.es_params(p)
    for f, _ in capi.EsParams._fields_:
        g = "sc" if f == "critical_stretch" else f
        assert getattr(out, f) == getattr(ref, g), f


def test_es_params_reject_bad_deck():
    from es_cases import qs2d
    from nlmech_amd.engine import es_params
    p, _ = qs2d()
    for bad in (dict(K=-1.0), dict(G=-1.0), dict(E=1e3)):
        with pytest.raises(capi.NlmError):
            es_params(pr.with_(p, material=dict(p.material, **bad)))


def test_engine_rejects_misshapen_arrays():
    """The driver checks array shapes before it hands raw pointers to the
    C-ABI (which reads n x 3 doubles, nnz bond bytes, ...): a wrong shape is
    a ValueError, never an out-of-bounds host read."""
    from nlmech_amd.engine import GpuFD
    p = pr.plate2d(6, 6, 1e-3, 3, 1e-7, 1)
    g = GpuFD.__new__(GpuFD)       # no context: nothing reaches the library
    g.p, g.nnz, g.ctx = p, 100, None
    g.lib = None
    with pytest.raises(ValueError):
        g.set_state(u=np.zeros((p.n, 2)))
    with pytest.raises(ValueError):
        g.set_state(v=np.zeros((p.n - 1, 3)))
    with pytest.raises(ValueError):
        g.set_bonds(np.zeros(99, np.uint8))
    with pytest.raises(ValueError):
        g.set_node_ubc([0, 1], [1], np.zeros((2, 3)))
    with pytest.raises(ValueError):
        g.set_node_ubc([0, p.n], [1, 1], np.zeros((2, 3)))
    with pytest.raises(ValueError):
        g.qs_apply(np.zeros(3 * p.n))
