"""The oracle's SSPbrick + ElasticIsotropic against the reference's OpenSees goldens.

Mirrors the reference's test_ssp_brick_reference.cpp:60-131: on 4 geometries x 2 materials x 4
(u, v) fields: internal force with body force, tangent K, initial tangent K0 (OpenSees returns
the tangent), Rayleigh damping force, and the mass matrix, each within 1e-12 of the largest
entry. Every geometry is compared with OpenSees' polyJac mass formula; the exact nodal volumes
(the element's production mass) on the parallelepipeds, where the two agree. Then the element
operator and its diagonal against K v and diag(K) built from the same goldens, and the
properties the mathematics fixes: rigid translations and rotations give no force, a linear
displacement field gives the constant strain exactly (patch test), and Kstab is symmetric
positive semi-definite.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import fxo
from tests.golden_data import files

REL = 1e-12


def cases():
    return files("ssp_brick.npz")


def params_of(ref, body=True, damp=True):
    b = ref.values("body_force") if body else (0.0, 0.0, 0.0)
    r = ref.values("rayleigh") if damp else (0.0, 0.0, 0.0, 0.0)
    return fxo.Elastic(ref.param("E"), ref.param("nu"), ref.param("rho"), (C.c_double * 3)(*b), 1,
                       fxo.Rayleigh(*r), 1)


def close(got, want, what):
    scale = np.max(np.abs(want))
    err = np.max(np.abs(got - want)) / scale if scale > 0 else np.max(np.abs(got))
    assert err <= REL, f"{what}: {err:.3e} of the largest entry"


def parallelepiped(ref):
    return ref.header["geometry"][0] in ("cube", "soil_column")


def lib():
    L = fxo.lib()
    vp = C.c_void_p
    L.fxo_ssp_force.argtypes = [vp, fxo.f64p, C.c_int, fxo.f64p, fxo.f64p, fxo.f64p, vp]
    L.fxo_ssp_mass.argtypes = [vp, fxo.f64p, C.c_int, fxo.f64p]
    L.fxo_ssp_element.argtypes = [vp, fxo.f64p, C.c_int, fxo.f64p, fxo.f64p]
    L.fxo_ssp_operator_apply.argtypes = [vp, vp, fxo.f64p, fxo.f64p, fxo.f64p]
    L.fxo_ssp_operator_diagonal.argtypes = [vp, vp, fxo.f64p, fxo.i32p, fxo.f64p]
    L.fxo_ssp_max_omega_sq.argtypes = [vp, fxo.f64p]
    L.fxo_ssp_max_omega_sq.restype = C.c_double
    return L


def test_ssp_goldens_present():
    names = [r.name for r in cases()]
    assert len([n for n in names if n.endswith("_state")]) == 8
    assert len([n for n in names if n.endswith("_mass")]) == 8


@pytest.mark.parametrize("ref", [r for r in cases() if r.name.endswith("_state")], ids=lambda r: r.name)
def test_ssp_force_tangent_damping(ref):
    L = lib()
    xe = np.array(ref.values("coords")).reshape(8, 3)
    cu, cv, cf = ref.col("u0"), ref.col("v0"), ref.col("f0")
    cK, cK0, cd = ref.col("K0"), ref.col("K0_0"), ref.col("d0")
    assert ref.data.shape[0] == 4
    formulas = (1, 0) if parallelepiped(ref) else (1,)
    for row in ref.data:
        u = np.ascontiguousarray(row[cu:cu + 24])
        v = np.ascontiguousarray(row[cv:cv + 24])
        for om in formulas:
            fe = np.zeros(24)
            Ke = np.zeros(576)
            p = params_of(ref, body=True, damp=False)
            L.fxo_ssp_force(C.byref(p), xe, om, u, v, fe, Ke.ctypes.data)
            close(fe, row[cf:cf + 24], f"f (mass formula {om})")
            close(Ke, row[cK:cK + 576], "K")
            close(Ke, row[cK0:cK0 + 576], "K0")
            p = params_of(ref, body=False, damp=True)
            L.fxo_ssp_force(C.byref(p), xe, om, np.zeros(24), v, fe, None)
            close(fe, row[cd:cd + 24], f"damping (mass formula {om})")


@pytest.mark.parametrize("ref", [r for r in cases() if r.name.endswith("_mass")], ids=lambda r: r.name)
def test_ssp_mass(ref):
    L = lib()
    xe = np.array(ref.values("coords")).reshape(8, 3)
    row = ref.data[0]
    want = row[ref.col("M0"):ref.col("M0") + 576]
    Me = np.zeros(576)
    p = params_of(ref)
    L.fxo_ssp_mass(C.byref(p), xe, 1, Me)
    close(Me, want, "M (OpenSees formula)")
    L.fxo_ssp_mass(C.byref(p), xe, 0, Me)
    if parallelepiped(ref):
        close(Me, want, "M (exact)")
    # the exact weights: total mass rho V on every shape (V from the stdBrick Gauss volume)
    m8 = np.zeros(8)
    L.fxo_brick_lumped_mass(ref.param("rho"), xe, m8)
    assert np.allclose(np.diag(Me.reshape(24, 24))[::3], m8, rtol=1e-15, atol=0)


def operator(L, p, xe, k):
    """The dense 24 x 24 K_eff of the element operator (unit vectors through the apply)."""
    A = np.zeros((24, 24))
    for b in range(24):
        e = np.zeros(24)
        e[b] = 1.0
        y = np.zeros(24)
        L.fxo_ssp_operator_apply(C.byref(p), C.byref(k), xe, e, y)
        A[:, b] = y
    return A


@pytest.mark.parametrize("ref", [r for r in cases() if r.name.endswith("_state")], ids=lambda r: r.name)
def test_ssp_operator_against_goldens(ref):
    """element_operator_apply with an element term: K v from the golden K; with k = damping, the
    damping force C v (on parallelepipeds, where the production mass equals OpenSees'); the
    diagonal equals the operator's."""
    L = lib()
    xe = np.array(ref.values("coords")).reshape(8, 3)
    row = ref.data[0]
    K = row[ref.col("K0"):ref.col("K0") + 576].reshape(24, 24)
    p = params_of(ref, body=True, damp=True)
    for kk in [(1.0, 0.0, 0.0), (2.5, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (1.3, 0.7, 4.0)]:
        k = fxo.coeffs(*kk)
        A = operator(L, p, xe, k)
        r = ref.values("rayleigh")
        beta = r[1] + r[2] + r[3]
        Mw = np.zeros(576)
        L.fxo_ssp_mass(C.byref(p), xe, 0, Mw)
        want = (kk[0] + kk[1] * beta) * K + (kk[2] + kk[1] * r[0]) * Mw.reshape(24, 24)
        close(A.ravel(), want.ravel(), f"K_eff {kk}")
        d = np.zeros(24)
        L.fxo_ssp_operator_diagonal(C.byref(p), C.byref(k), xe, np.arange(24, dtype=np.int32), d)
        close(d, np.diag(A), f"diagonal {kk}")
        if kk == (0.0, 1.0, 0.0) and parallelepiped(ref):
            v = np.ascontiguousarray(row[ref.col("v0"):ref.col("v0") + 24])
            close(A @ v, row[ref.col("d0"):ref.col("d0") + 24], "C v")


def test_ssp_diagonal_with_merged_equations():
    """Local DOFs that share one equation (an element tied to itself): the assembled diagonal,
    P^T K P's diagonal, on a distorted element."""
    L = lib()
    rng = np.random.default_rng(3)
    xe = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    xe = xe * [1.3, 0.8, 1.1] + rng.uniform(-0.1, 0.1, (8, 3))
    p = fxo.Elastic(2e5, 0.3, 1.8, (C.c_double * 3)(0, 0, 0), 1, fxo.Rayleigh(0.2, 1e-3, 2e-4, 3e-4), 1)
    k = fxo.coeffs(1.1, 0.6, 3.0)
    A = operator(L, p, xe, k)
    eq = np.arange(24, dtype=np.int32) + 10
    eq[[0, 3, 9, 21]] = 0   # x DOFs of nodes 0, 1, 3, 7 on one equation
    eq[[10, 13]] = 1        # y of node 3 and y of node 4 on one equation
    eq[[5]] = -1
    d = np.zeros(24)
    L.fxo_ssp_operator_diagonal(C.byref(p), C.byref(k), xe, eq, d)
    # summing de into the equations gives the assembled diagonal sum_{a, b on q} A[a, b]
    qs = sorted(set(eq.tolist()) - {-1})
    got = np.array([d[eq == q].sum() for q in qs])
    want = np.array([A[np.ix_(eq == q, eq == q)].sum() for q in qs])
    close(got, want, "merged diagonal")


def test_ssp_rigid_modes_and_patch_test():
    """No force under rigid translations and small rotations; under u = G x the mean strain is
    sym(G) exactly, Kstab u = 0 and the force is mVol Bnot^T D sym(G)."""
    L = lib()
    rng = np.random.default_rng(11)
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    for trial in range(6):
        xe = np.ascontiguousarray(base * rng.uniform(0.5, 2.0, 3) + rng.uniform(-0.15, 0.15, (8, 3)))
        p = fxo.Elastic(3e5, 0.25 + 0.04 * trial, 2.0, (C.c_double * 3)(0, 0, 0), 1, fxo.Rayleigh(0, 0, 0, 0), 1)
        K = np.zeros(576)
        fe = np.zeros(24)
        L.fxo_ssp_force(C.byref(p), xe, 0, np.zeros(24), np.zeros(24), fe, K.ctypes.data)
        K = K.reshape(24, 24)
        scale = np.abs(K).max()
        for d in range(3):
            t = np.zeros(24)
            t[d::3] = 1.0
            assert np.abs(K @ t).max() <= 1e-12 * scale * 3
        W = rng.standard_normal((3, 3))
        W = W - W.T
        r = (xe @ W.T).ravel()
        assert np.abs(K @ r).max() <= 1e-11 * scale * np.abs(r).max()
        # patch test: Kstab alone (K minus the mean-strain part) vanishes on linear fields
        geom = np.zeros(33)
        cst = np.zeros(110)
        L.fxo_ssp_element(C.byref(p), xe, 0, geom, cst)
        G = rng.standard_normal((3, 3)) * 1e-3
        u = (xe @ G.T).ravel()
        B = np.zeros((6, 24))
        for i in range(8):
            bx, by, bz = geom[3 * i:3 * i + 3]
            B[:, 3 * i] = [bx, 0, 0, by, 0, bz]
            B[:, 3 * i + 1] = [0, by, 0, bx, bz, 0]
            B[:, 3 * i + 2] = [0, 0, bz, 0, by, bx]
        eps = B @ u
        S = 0.5 * (G + G.T)
        want = np.array([S[0, 0], S[1, 1], S[2, 2], 2 * S[0, 1], 2 * S[1, 2], 2 * S[0, 2]])
        assert np.abs(eps - want).max() <= 1e-13 * np.abs(want).max()
        gam = cst[:32].reshape(4, 8)
        assert np.abs(gam @ xe).max() <= 1e-13 * np.abs(xe).max()  # hourglass vectors orthogonal to x, y, z
        assert np.abs(gam.sum(axis=1)).max() <= 1e-14                # and to constants
        D = np.zeros(36)
        L.fxo_elastic_tangent(p.E, p.nu, D)
        f_mean = geom[24] * B.T @ (D.reshape(6, 6) @ eps)
        L.fxo_ssp_force(C.byref(p), xe, 0, u, np.zeros(24), fe, None)
        assert np.abs(fe - f_mean).max() <= 1e-11 * np.abs(f_mean).max()
        # Kstab = K - mVol B^T D B is symmetric positive semi-definite with the 4 x 3 hourglass rank
        Kst = K - geom[24] * B.T @ D.reshape(6, 6) @ B
        w = np.linalg.eigvalsh(0.5 * (Kst + Kst.T))
        assert w.min() >= -1e-10 * w.max()
        assert np.sum(w > 1e-8 * w.max()) == 12


def test_ssp_volume_and_critical_step():
    """mVol equals the element volume (the Gauss volume of stdBrick) on a parallelepiped, and the
    largest omega^2 of the lumped-mass eigenproblem equals the dense eigenvalue of M^-1/2 K M^-1/2."""
    L = lib()
    xe = np.array([[0, 0, 0], [2, 0, 0], [2.3, 1, 0], [0.3, 1, 0], [0, 0.2, 1.5], [2, 0.2, 1.5], [2.3, 1.2, 1.5],
                   [0.3, 1.2, 1.5]], float)
    p = fxo.Elastic(1e5, 0.3, 1.7, (C.c_double * 3)(0, 0, 0), 1, fxo.Rayleigh(0, 0, 0, 0), 1)
    geom = np.zeros(33)
    cst = np.zeros(110)
    L.fxo_ssp_element(C.byref(p), xe, 0, geom, cst)
    vol = abs(np.linalg.det(np.array([xe[1] - xe[0], xe[3] - xe[0], xe[4] - xe[0]])))
    assert abs(geom[24] - vol) <= 1e-14 * vol
    assert abs(geom[25:33].sum() - vol) <= 1e-14 * vol
    K = np.zeros(576)
    fe = np.zeros(24)
    L.fxo_ssp_force(C.byref(p), xe, 0, np.zeros(24), np.zeros(24), fe, K.ctypes.data)
    m = np.repeat(p.rho * geom[25:33], 3)
    A = K.reshape(24, 24) / np.sqrt(np.outer(m, m))
    w = np.linalg.eigvalsh(0.5 * (A + A.T)).max()
    got = L.fxo_ssp_max_omega_sq(C.byref(p), xe)
    assert abs(got - w) <= 1e-10 * w
