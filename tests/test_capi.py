"""CPU tests of the C ABI library: it loads, exports every entry point include/femorax_amd.h
declares, and its host compile step (fx_compile_model) gives the oracle's arrays bit for bit
(connectivity, colouring and DOF numbering exact). No compute call
touches a GPU here."""
import os
import re

import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
from oracle import fxo
from tests import meshes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ARRAYS = ["node_order", "node_rank", "elem_order", "node_eq", "coords", "node", "eq", "param", "param_group",
          "param_region", "color_start", "color_elements"]


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(fx_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = core.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python mirror binds exactly the declared set
    assert names == {s[0] for s in core.SIGNATURES}


MODELS = {
    "c1_box16": lambda: S.config_c1(),
    "c2_layered16": lambda: S.config_c2(n=16),
    "distorted6": lambda: meshes.distorted_box(6),
    "column24": lambda: meshes.column(24),
    "column_2x2": lambda: S.soil_column(10, 1.0, params=[dict(E=1e8, nu=0.3)], nx=2, ny=2),
}


@pytest.mark.parametrize("name", sorted(MODELS))
@pytest.mark.parametrize("renumber", [True, False])
def test_compile_step_matches_oracle(name, renumber):
    m = MODELS[name]()
    pm = core.Mesh(m, renumber=renumber)
    om = fxo.OracleMesh(m, renumber=renumber)
    assert (pm.neq, pm.padded, pm.ncolor, pm.nparam) == (om.neq, om.padded, om.ncolor, om.nparam)
    for a in ARRAYS:
        np.testing.assert_array_equal(getattr(pm, a), getattr(om, a), err_msg=a)


def test_compile_step_rejects_bad_models():
    m = S.config_c1()
    bad = dict(m)
    bad["conn"] = m["conn"].copy()
    bad["conn"][5, 3] = m["nnode"] + 10
    with pytest.raises(core.FxError) as e:
        core.Mesh(bad)
    assert e.value.code == -1
    bad = dict(m)
    bad["fix"] = np.array([[0, 3]], np.int32)
    with pytest.raises(core.FxError):
        core.Mesh(bad)


def test_padding_is_inert():
    m = S.config_c1()
    pm = core.Mesh(m, simd=64)
    assert pm.padded % 64 == 0
    pm1 = core.Mesh(m, simd=1)
    assert pm1.padded == m["nelem"]
    P, P1 = pm.padded, pm1.padded
    np.testing.assert_array_equal(pm.eq.reshape(24, P)[:, :P1], pm1.eq.reshape(24, P1))
    assert np.all(pm.eq.reshape(24, P)[:, P1:] == -1)
    assert np.all(pm.node.reshape(8, P)[:, P1:] == -1)


def test_column_pool_refuses_substep_counts_beyond_int32():
    """The pool counts a substep's position in units of num_substeps^level; descriptors whose
    deepest level would exceed 2^31 - 1 units are refused before any device work."""
    import ctypes as C

    d = core.ColumnPoolDesc(max_elements=64, max_columns=1, max_materials=1, max_record=64, gamma=0.5, beta=0.25,
                            dt=0.005, test=1, tol=1e-10, max_iter=10, cg_rtol=1e-12, cg_max_it=100, flags=0,
                            inner_rtol=0.0, max_retries=1, num_sublevels=6, num_substeps=1500)
    out = C.c_void_p()
    L = core.lib()
    assert L.fx_column_pool_create(None, C.byref(d), C.byref(out)) == core.FX_EINVAL
    assert b"num_substeps^num_sublevels" in L.fx_last_error(None)
    assert not out.value
    d.num_substeps, d.num_sublevels = 2, 6  # 64 units: passes the descriptor checks, then needs a context
    assert L.fx_column_pool_create(None, C.byref(d), C.byref(out)) == core.FX_EINVAL
    assert b"null context" in L.fx_last_error(None)
