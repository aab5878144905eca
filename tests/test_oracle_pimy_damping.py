"""The oracle's Rayleigh damping of stdBrick<PressureIndependMultiYield> (std_brick.hpp:409-464
point_update_shape / point_node_contribution; element_operator.hpp:45-75 factors and effective
tangent) and the material's committed tangent (pressure_independ_multi_yield.hpp:162-168).

Pinned by:
- the committed tangent at rest is the elastic tangent of the reference moduli (closed form);
  after an update and its commit it is the update's own tangent (the reference's definition:
  getTangent after commitState), bit for bit, on random walks and on the 34 OpenSees strain
  paths, where it also matches the recorded tangent within 1e-14;
- the damping force is the damping operator: f(u, v) - f(u, 0) = (betaK K_t + betaKinit K_0 +
  betaKcomm K_c + alphaM M) v, the right side from the operator with the coefficients
  (stiffness 0, damping 1, mass 0), to rounding;
- where oracle/_ref is built: the element force with damping, the operator and its diagonal,
  bit for bit against the reference's own headers on distorted elements and random walks.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import fxo
from tests.golden_data import files

NS = fxo.PIMY_NSTATE
P = C.POINTER


def lib():
    L = fxo.lib()
    L.fxo_pimy_update.argtypes = [P(fxo.PimyParams)] + [C.c_void_p] * 5
    L.fxo_brick_force_pimy.argtypes = [P(fxo.PimyElem), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
    L.fxo_brick_operator_apply_gp.argtypes = [P(fxo.PimyElem), P(fxo.Coeffs), C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p]
    L.fxo_brick_operator_diagonal_gp.argtypes = [P(fxo.PimyElem), P(fxo.Coeffs), C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p]
    return L


CLAY = dict(rho=2000.0, G=60e6, K=160e6, cohesion=40e3, gamma_peak=0.1, phi=0.0, pref=100e3, pdc=0.0)


def elem(m=CLAY, stage=1, ray=(0.0, 0.0, 0.0, 0.0), mass_type=0, b=(0.0, 0.0, -9.81 * 2000.0)):
    return fxo.pimy_elems([dict(m, rayleigh=ray, mass_type=mass_type, b=b)], stage=stage)


def geometry(seed, h=1.0):
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    return np.ascontiguousarray(h * (base + 0.12 * fxo.lcg_sequence(seed, 24).reshape(8, 3)))


def ptrs(H):
    return (C.c_void_p * 8)(*[H[g].ctypes.data for g in range(8)])


def force(pe, xe, u, v, Cs, Ts):
    """element force at u (and v); Cs, Ts: [8, NS] committed / trial histories; returns fe, D[8, 36]"""
    fe, D = np.zeros(24), np.zeros((8, 36))
    lib().fxo_brick_force_pimy(C.byref(pe[0]), xe.ctypes.data, np.ascontiguousarray(u).ctypes.data,
                               None if v is None else np.ascontiguousarray(v).ctypes.data, ptrs(Cs), ptrs(Ts),
                               fe.ctypes.data, D.ctypes.data)
    return fe, D


def operator(pe, k, xe, D, Cs, v, eq=None, diagonal=False):
    out = np.zeros(24)
    if diagonal:
        lib().fxo_brick_operator_diagonal_gp(C.byref(pe[0]), C.byref(k), xe.ctypes.data, D.ctypes.data, ptrs(Cs),
                                             np.ascontiguousarray(eq, np.int32).ctypes.data, out.ctypes.data)
    else:
        lib().fxo_brick_operator_apply_gp(C.byref(pe[0]), C.byref(k), xe.ctypes.data, D.ctypes.data, ptrs(Cs),
                                          np.ascontiguousarray(v).ctypes.data, out.ctypes.data)
    return out


def committed_tangent(pe, Cp):
    D = np.zeros(36)
    fxo.lib().fxo_pimy_committed_tangent(C.byref(pe[0].mat), np.ascontiguousarray(Cp), D)
    return D


def initial_histories(pe):
    Cs = np.zeros((8, NS))
    for g in range(8):
        fxo.lib().fxo_pimy_initial_state(C.byref(pe[0].mat), Cs[g].ctypes.data)
    return Cs, Cs.copy()


def walk(seed, nstep, scale=2e-3):
    """element displacement and velocity sequences driving the points well into the plastic range"""
    r = fxo.lcg_sequence(seed, 48 * nstep).reshape(nstep, 2, 24)
    u = np.cumsum(scale * r[:, 0], axis=0)
    return u, 0.3 * r[:, 1]


def test_committed_tangent_at_rest_is_the_elastic_tangent():
    pe = elem()
    Cs, _ = initial_histories(pe)
    G, K = CLAY["G"], CLAY["K"]
    D = np.zeros((6, 6))
    D[:3, :3] = K - 2.0 * G / 3.0
    D[np.arange(3), np.arange(3)] += 2.0 * G
    D[np.arange(3, 6), np.arange(3, 6)] = G
    np.testing.assert_allclose(committed_tangent(pe, Cs[0]).reshape(6, 6), D, rtol=1e-15, atol=0)


@pytest.mark.parametrize("seed", [3, 4])
def test_committed_tangent_is_the_last_updates_tangent(seed):
    pe = elem()
    xe = geometry(seed)
    Cs, Ts = initial_histories(pe)
    U, _ = walk(seed, 60)
    plastic = 0
    for k in range(len(U)):
        _, D = force(pe, xe, U[k], None, Cs, Ts)
        Cs, Ts = Ts, Cs  # commit
        for g in range(8):
            np.testing.assert_array_equal(committed_tangent(pe, Cs[g]), D[g])
        plastic += int(np.sum(Cs[:, 26] > 0))
    assert plastic > 0, "the walk reached the plastic range"


def test_committed_tangent_on_the_opensees_paths():
    """After each committed step of the 34 OpenSees strain paths the committed tangent is the
    recorded tangent of that step (1e-14 of the file's largest entry)."""
    L = lib()
    for ref in files("pimy.npz"):
        inp = np.array([ref.param(k) for k in ("rho", "refShearModul", "refBulkModul", "cohesi", "peakShearStra",
                                               "frictionAng", "refPress", "pressDependCoe")])
        pairs = np.array(ref.values("pairs"))
        p = fxo.PimyParams()
        assert L.fxo_pimy_setup(inp.ctypes.data, int(ref.param("noYieldSurf")),
                                pairs.ctypes.data if len(pairs) else None, C.byref(p)) == 0
        Cs, Ts = np.zeros(NS), np.zeros(NS)
        L.fxo_pimy_initial_state(C.byref(p), Cs.ctypes.data)
        L.fxo_pimy_initial_state(C.byref(p), Ts.ctypes.data)
        cst, cpre, cpre0, ceps, csig = ref.col("stage"), ref.col("pre"), ref.col("pre0"), ref.col("eps0"), ref.col("sig0")
        sig, D = np.zeros(6), np.zeros(36)
        scale = np.max(np.abs(ref.data[:, csig + 6:csig + 42]))
        for row in ref.data:
            p.load_stage = int(row[cst])
            if row[cpre] != 0.0:
                e = np.ascontiguousarray(row[cpre0:cpre0 + 6])
                L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, Ts.ctypes.data, e.ctypes.data, sig.ctypes.data,
                                  D.ctypes.data)
            e = np.ascontiguousarray(row[ceps:ceps + 6])
            L.fxo_pimy_update(C.byref(p), Cs.ctypes.data, Ts.ctypes.data, e.ctypes.data, sig.ctypes.data, D.ctypes.data)
            Cs, Ts = Ts, Cs
            Dc = np.zeros(36)
            fxo.lib().fxo_pimy_committed_tangent(C.byref(p), Cs, Dc)
            np.testing.assert_array_equal(Dc, D)
            assert np.max(np.abs(Dc - row[csig + 6:csig + 42])) <= 1e-14 * scale


@pytest.mark.parametrize("mass_type", [0, 1])
def test_damping_force_is_the_damping_operator(mass_type):
    ray = (0.7, 3e-3, 2e-3, 1.5e-3)
    pe = elem(ray=ray, mass_type=mass_type)
    xe = geometry(11)
    Cs, Ts = initial_histories(pe)
    U, V = walk(11, 30)
    k = fxo.coeffs(0.0, 1.0, 0.0)
    for step in range(len(U)):
        f0, D0 = force(pe, xe, U[step], None, Cs, Ts)
        T0 = Ts.copy()
        f1, D1 = force(pe, xe, U[step], V[step], Cs, Ts)
        np.testing.assert_array_equal(D0, D1)
        np.testing.assert_array_equal(T0, Ts)  # the velocity does not reach the history
        fd = operator(pe, k, xe, D1, Cs, V[step])
        assert np.max(np.abs((f1 - f0) - fd)) <= 1e-12 * max(np.max(np.abs(fd)), np.max(np.abs(f0)))
        Cs, Ts = Ts, Cs


def _ref_params(pe_m, R):
    inp = np.array([pe_m["rho"], pe_m["G"], pe_m["K"], pe_m["cohesion"], pe_m["gamma_peak"], pe_m["phi"], pe_m["pref"],
                    pe_m["pdc"]])
    rp = np.zeros(300)
    assert R.fxref_pimy_setup(inp.ctypes.data, 20, None, rp.ctypes.data) == 0
    return rp


REF = pytest.mark.skipif(fxo.ref() is None, reason="oracle/_ref not built (needs /root/reference)")
RAYS = [(0.0, 2e-3, 0.0, 0.0), (0.0, 0.0, 3e-3, 0.0), (0.0, 0.0, 0.0, 2.5e-3), (0.9, 0.0, 0.0, 0.0),
        (0.6, 3e-3, 2e-3, 1.5e-3)]


@REF
@pytest.mark.parametrize("ray", RAYS)
@pytest.mark.parametrize("mass_type", [0, 1])
def test_damped_force_matches_the_reference_header(ray, mass_type):
    R = fxo.ref()
    R.fxref_pimy_brick_force.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 1 + [C.c_int] + [C.c_void_p] * 7
    rp = _ref_params(CLAY, R)
    b = np.array([0.0, 0.0, -9.81 * 2000.0])
    pe = elem(ray=ray, mass_type=mass_type, b=tuple(b))
    xe = geometry(21 + mass_type)
    Cs, Ts = initial_histories(pe)
    Cr, Tr = Cs.copy(), Ts.copy()
    rr = np.array(ray, float)
    U, V = walk(21 + mass_type, 40)
    for step in range(len(U)):
        fo, _ = force(pe, xe, U[step], V[step], Cs, Ts)
        fr = np.zeros(24)
        u, v = np.ascontiguousarray(U[step]), np.ascontiguousarray(V[step])
        R.fxref_pimy_brick_force(rp.ctypes.data, 1, b.ctypes.data, mass_type, rr.ctypes.data, xe.ctypes.data,
                                 u.ctypes.data, v.ctypes.data, Cr.ctypes.data, Tr.ctypes.data, fr.ctypes.data)
        np.testing.assert_array_equal(fo, fr)
        np.testing.assert_array_equal(Ts, Tr)
        Cs, Ts = Ts, Cs
        Cr, Tr = Tr, Cr
    assert np.max(Cs[:, 26]) > 0


@REF
@pytest.mark.parametrize("ray", RAYS)
@pytest.mark.parametrize("coef", [(1.0, 0.0, 0.0, 0), (1.0, 400.0, 8e4, 0), (0.0, 1.0, 0.0, 0), (1.0, 400.0, 8e4, 1)])
def test_damped_operator_matches_the_reference_header(ray, coef):
    R = fxo.ref()
    R.fxref_pimy_operator.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int] + \
        [C.c_void_p] * 5 + [C.c_int, C.c_void_p]
    rp = _ref_params(CLAY, R)
    pe = elem(ray=ray, mass_type=0)
    xe = geometry(31)
    Cs, Ts = initial_histories(pe)
    U, V = walk(31, 25)
    k = fxo.coeffs(coef[0], coef[1], coef[2], initial=bool(coef[3]))
    kk = np.array(coef[:3], float)
    rr = np.array(ray, float)
    eq = np.arange(24, dtype=np.int32)
    eq[3:6] = eq[0:3]  # two nodes tied: the diagonal's merged-DOF terms
    for step in range(len(U)):
        _, D = force(pe, xe, U[step], None, Cs, Ts)
        Cs, Ts = Ts, Cs
        _, D = force(pe, xe, U[step] * 1.01, None, Cs, Ts)  # the stored tangents of a trial beyond the commit
        v = np.ascontiguousarray(V[step])
        for diag in (False, True):
            yo = operator(pe, k, xe, D, Cs, v, eq=eq, diagonal=diag)
            yr = np.zeros(24)
            R.fxref_pimy_operator(rp.ctypes.data, 1, 0, rr.ctypes.data, kk.ctypes.data, coef[3], xe.ctypes.data,
                                  D.ctypes.data, np.ascontiguousarray(Cs).ctypes.data, v.ctypes.data, eq.ctypes.data,
                                  int(diag), yr.ctypes.data)
            np.testing.assert_array_equal(yo, yr)
