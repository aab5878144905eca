"""The oracle's PressureIndependMultiYield (the J2 soil return map) against the reference's
OpenSees goldens: 34 strain paths (stage switches, discarded trials, commits), stress and
tangent at every step within 1e-14 of the file's largest value (test_pimy_reference.cpp:97-140),
and against the reference's own header (oracle/_ref) on seeded random paths. Also the
material properties the reference's test_pimy.cpp pins without data: elastic stage linear,
the update is a pure function of the committed state."""
import ctypes as C

import numpy as np
import pytest

from oracle import fxo
from tests.golden_data import files

REL = 1e-14


PimyParams = fxo.PimyParams  # one ctypes class for the library's signatures


def lib():
    L = fxo.lib()
    L.fxo_pimy_update.argtypes = [C.POINTER(PimyParams)] + [C.c_void_p] * 5
    return L


def params_from(ref):
    L = lib()
    inp = np.array([ref.param(k) for k in ("rho", "refShearModul", "refBulkModul", "cohesi", "peakShearStra",
                                           "frictionAng", "refPress", "pressDependCoe")])
    pairs = np.array(ref.values("pairs"))
    p = PimyParams()
    rc = L.fxo_pimy_setup(inp.ctypes.data, int(ref.param("noYieldSurf")), pairs.ctypes.data if len(pairs) else None,
                          C.byref(p))
    assert rc == 0
    return p, inp, pairs


def replay(ref):
    L = lib()
    p, _, _ = params_from(ref)
    ns = L.fxo_pimy_nstate()
    Cs, Ts = np.zeros(ns), np.zeros(ns)
    L.fxo_pimy_initial_state(C.byref(p), Cs.ctypes.data)
    L.fxo_pimy_initial_state(C.byref(p), Ts.ctypes.data)
    cst, cpre, cpre0, ceps = ref.col("stage"), ref.col("pre"), ref.col("pre0"), ref.col("eps0")
    out = []
    sig, D = np.zeros(6), np.zeros(36)
    for row in ref.data:
        p.load_stage = int(row[cst])
        if row[cpre] != 0.0:
            e = np.ascontiguousarray(row[cpre0:cpre0 + 6])
            L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, Ts.ctypes.data, e.ctypes.data, sig.ctypes.data,
                              D.ctypes.data)
        e = np.ascontiguousarray(row[ceps:ceps + 6])
        L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, Ts.ctypes.data, e.ctypes.data, sig.ctypes.data, D.ctypes.data)
        out.append(np.concatenate([sig, D]))
        Cs, Ts = Ts, Cs  # commit = swap
    return np.array(out)


def pimy_files():
    return files("pimy.npz")


def test_pimy_goldens_present():
    assert len(pimy_files()) == 34


@pytest.mark.parametrize("ref", pimy_files(), ids=lambda r: r.name.replace("PressureIndependMultiYield_", ""))
def test_pimy_matches_opensees(ref):
    got = replay(ref)
    c = ref.col("sig0")
    want = ref.data[:, c:c + 42]
    s_scale = np.max(np.abs(want[:, :6]))
    d_scale = np.max(np.abs(want[:, 6:]))
    err_s = np.max(np.abs(got[:, :6] - want[:, :6])) / s_scale
    err_d = np.max(np.abs(got[:, 6:] - want[:, 6:])) / d_scale
    assert err_s <= REL and err_d <= REL, (err_s, err_d)
    # the restatement keeps the operation order: nearly every value is bit-identical
    assert np.mean(got == want) >= 0.99


@pytest.mark.skipif(fxo.ref() is None, reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("case", ["clay", "frictional"])
def test_pimy_matches_reference_header_random_paths(seed, case):
    """Seeded random strain walks with stage switches, against the reference's own
    PressureIndependMultiYield<20> compiled from its header: every state value bit-identical."""
    L, R = lib(), fxo.ref()
    if case == "clay":
        inp = np.array([2.0, 6.0e4, 1.6e5, 30.0, 0.1, 0.0, 100.0, 0.0])
    else:
        inp = np.array([2.0, 6.0e4, 1.6e5, 5.0, 0.1, 30.0, 80.0, 0.5])
    p = PimyParams()
    assert L.fxo_pimy_setup(inp.ctypes.data, 20, None, C.byref(p)) == 0
    nbytes = 8 * 300
    rp = np.zeros(300)
    assert R.fxref_pimy_setup(inp.ctypes.data, 20, None, rp.ctypes.data) == 0
    ns = L.fxo_pimy_nstate()
    Co, To, Cr, Tr = np.zeros(ns), np.zeros(ns), np.zeros(ns), np.zeros(ns)
    L.fxo_pimy_initial_state(C.byref(p), Co.ctypes.data)
    R.fxref_pimy_initial_state(rp.ctypes.data, Cr.ctypes.data)
    np.testing.assert_array_equal(Co, Cr)
    u = fxo.lcg_sequence(1000 + seed, 6 * 400)
    eps = np.zeros(6)
    so, Do, sr, Dr = np.zeros(6), np.zeros(36), np.zeros(6), np.zeros(36)
    for step in range(400):
        stage = 0 if step < 20 else 1
        p.load_stage = stage
        if step < 20:
            eps = eps + np.array([-2e-5, -2e-5, -4e-5, 0, 0, 0])
        else:
            eps = eps + 2e-3 * u[6 * step:6 * step + 6] * np.array([0.1, 0.1, 0.1, 1, 1, 1])
        e = np.ascontiguousarray(eps)
        L.fxo_pimy_update(C.byref(p), Co.ctypes.data, To.ctypes.data, e.ctypes.data, so.ctypes.data, Do.ctypes.data)
        R.fxref_pimy_update(rp.ctypes.data, stage, Cr.ctypes.data, Tr.ctypes.data, e.ctypes.data, sr.ctypes.data,
                            Dr.ctypes.data)
        np.testing.assert_array_equal(so, sr)
        np.testing.assert_array_equal(Do, Dr)
        np.testing.assert_array_equal(To, Tr)
        Co, To = To, Co
        Cr, Tr = Tr, Cr
    assert nbytes > 0


def test_pimy_elastic_stage_is_linear_and_pure():
    """Stage 0 is linear elastic in the strain; an update is a pure function of the committed
    buffer (repeating it, or running a discarded trial first, gives the same bits)."""
    L = lib()
    inp = np.array([2.0, 6.0e4, 1.6e5, 30.0, 0.1, 0.0, 100.0, 0.0])
    p = PimyParams()
    L.fxo_pimy_setup(inp.ctypes.data, 20, None, C.byref(p))
    ns = L.fxo_pimy_nstate()
    Cs, Ts = np.zeros(ns), np.zeros(ns)
    L.fxo_pimy_initial_state(C.byref(p), Cs.ctypes.data)
    s1, s2, D = np.zeros(6), np.zeros(6), np.zeros(36)
    e = np.array([1e-4, -2e-4, 3e-4, 1e-4, 0, -5e-5])
    L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, Ts.ctypes.data, e.ctypes.data, s1.ctypes.data, D.ctypes.data)
    e2 = 2 * e
    L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, Ts.ctypes.data, e2.ctypes.data, s2.ctypes.data, D.ctypes.data)
    np.testing.assert_allclose(s2, 2 * s1, rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(D.reshape(6, 6) @ e, s1, rtol=1e-12)
    p.load_stage = 1
    a, b = np.zeros(6), np.zeros(6)
    T1, T2 = np.zeros(ns), np.zeros(ns)
    e3 = np.array([0, 0, 0, 3e-3, 1e-3, 0])
    L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, T1.ctypes.data, e3.ctypes.data, a.ctypes.data, D.ctypes.data)
    L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, T2.ctypes.data, e2.ctypes.data, b.ctypes.data, D.ctypes.data)
    L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, T2.ctypes.data, e3.ctypes.data, b.ctypes.data, D.ctypes.data)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(T1, T2)
