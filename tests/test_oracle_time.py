"""Pins of the oracle's internal force, stable time step and CentralDifference loop (CPU).

- Force (internal_force.cpp:354-377): for the elastic stdBrick, f(u, v) = K u + C v - F_body, so
  the force assembly is checked against the operator (pinned by a densely assembled K) and the
  body load, on a distorted multi-material mesh with Rayleigh damping and on a tied column.
- Stable step (std_brick.hpp:658-694, sym_eigen.hpp, explicit_stability.hpp): the cyclic-Jacobi
  largest eigenvalue against numpy's symmetric eigensolver on M_l^-1/2 K M_l^-1/2 built from
  golden-pinned element matrices, and against the reference's own headers (oracle/_ref).
- CentralDifference loop (explicit_analysis.cpp:87-117, 712-722): against the textbook
  recurrence u(n+1) = 2 u(n) - u(n-1) + dt^2 M^-1 (P(n) - K u(n)) with a dense K (undamped), and
  (M/dt^2 + C/(2dt)) u(n+1) = ... with a diagonal C.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import fxo
from tests import meshes


def compiled(model):
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    return om, op


def damped_distorted():
    model = meshes.distorted_box(4)
    model["params"] = [dict(p, rayleigh=(0.6, 5.8e-4, 1e-4, 2e-5)) for p in model["params"]]
    return model


@pytest.mark.parametrize("name", ["distorted4_damped", "column8"])
def test_force_equals_operator_and_body_load(name):
    model = damped_distorted() if name == "distorted4_damped" else meshes.column(8)
    om, op = compiled(model)
    u = meshes.random_vector(om.neq, 1) * 1e-3
    v = meshes.random_vector(om.neq, 2) * 1e-2
    f0 = om.force(op, np.zeros(om.neq))
    np.testing.assert_allclose(f0, -om.body_load(op), rtol=0, atol=1e-12 * np.max(np.abs(f0)))
    fu = om.force(op, u) - f0
    Ku = om.apply(op, u)
    assert np.max(np.abs(fu - Ku)) <= 1e-11 * np.max(np.abs(Ku))
    fv = om.force(op, np.zeros(om.neq), v) - f0
    Cv = om.apply(op, v, fxo.coeffs(0.0, 1.0, 0.0))
    assert np.max(np.abs(fv - Cv)) <= 1e-11 * np.max(np.abs(Cv))


def element_matrix(p, xe):
    Ke = np.zeros(576)
    fxo.lib().fxo_brick_force_elastic(C.byref(p), xe, np.zeros(24), np.zeros(24), np.zeros(24), Ke.ctypes.data)
    return Ke.reshape(24, 24)


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_max_omega_sq_against_numpy_and_reference(seed):
    rng = np.random.default_rng(seed)
    E, nu, rho = rng.uniform(1e7, 3e8), rng.uniform(0.05, 0.45), rng.uniform(1500, 2500)
    p = fxo.elastic_params([dict(E=E, nu=nu, rho=rho)])[0]
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    xe = np.ascontiguousarray(base * rng.uniform(0.5, 3.0, 3) + rng.uniform(-0.15, 0.15, (8, 3)))
    K = element_matrix(p, xe)
    m = np.zeros(8)
    fxo.lib().fxo_brick_lumped_mass(rho, xe, m)
    s = np.repeat(1.0 / np.sqrt(m), 3)
    want = np.linalg.eigvalsh(K * s[:, None] * s[None, :])[-1]
    got = fxo.lib().fxo_brick_max_omega_sq(C.byref(p), xe)
    assert abs(got - want) <= 1e-12 * want
    dt = fxo.lib().fxo_stable_dt(0, C.byref(p.damp), got)
    assert dt == 2.0 / np.sqrt(got)
    R = fxo.ref()
    if R is not None:
        R.fxref_brick_critical_dt.restype = C.c_double
        R.fxref_brick_critical_dt.argtypes = [C.c_void_p] * 2 + [C.c_int] + [C.c_void_p] * 2
        mat = np.array([E, nu, rho])
        ray = np.zeros(4)
        om2 = C.c_double()
        rdt = R.fxref_brick_critical_dt(mat.ctypes.data, ray.ctypes.data, 0, xe.ctypes.data, C.byref(om2))
        # K is summed per Gauss point in the reference (point_block_tangent) and via the force
        # path here: the same to rounding
        assert abs(om2.value - got) <= 1e-14 * got
        assert abs(rdt - dt) <= 1e-14 * dt


def test_stable_dt_explicit_difference_closed_form():
    """ExplicitDifference: the positive root of w^2 dt^2 + 2 c dt - 4 = 0."""
    d = fxo.Rayleigh(0.3, 1e-3, 0.0, 0.0)
    w2 = 4.0e4
    dt = fxo.lib().fxo_stable_dt(1, C.byref(d), w2)
    c = 0.3 + 1e-3 * w2
    assert abs(w2 * dt * dt + 2 * c * dt - 4.0) <= 1e-12
    assert np.isinf(fxo.lib().fxo_stable_dt(0, C.byref(d), 0.0))


def test_critical_dt_is_the_smallest_element_step():
    model = meshes.distorted_box(3)
    om, op = compiled(model)
    dts = []
    L = fxo.lib()
    xyz = model["xyz"]
    for e in range(model["nelem"]):
        p = fxo.elastic_params([model["params"][model["group"][e]]])[0]
        xe = np.ascontiguousarray(xyz[model["conn"][e]], np.float64)
        dts.append(2.0 / np.sqrt(L.fxo_brick_max_omega_sq(C.byref(p), xe)))
    assert om.critical_dt(op) == min(dts)


@pytest.mark.parametrize("damping", ["none", "rayleigh_alphaM", "dashpots"])
def test_central_difference_against_the_recurrence(damping):
    """Rayleigh alphaM with lumped element mass: the element force's alphaM M v at the garbage
    velocity and the lumped alphaM mass in the tangent make the textbook damped recurrence with
    C = alphaM M_l; a diagonal damper outside the elements does the same."""
    model = meshes.column(6)
    for p in model["params"]:
        p["b"] = (0.0, 0.0, 0.0)
        if damping == "rayleigh_alphaM":
            p["rayleigh"] = (3.0, 0.0, 0.0, 0.0)
    om, op = compiled(model)
    n = om.neq
    K = np.column_stack([om.apply(op, e) for e in np.eye(n)])
    m, am = om.lumped_mass(op)
    dash = np.abs(meshes.random_vector(n, 8)) * m * 5.0 if damping == "dashpots" else None
    cm = am if damping == "rayleigh_alphaM" else dash
    dt = 0.5 * om.critical_dt(op)
    nsteps = 60
    load_eq = np.array([n - 1, n - 3], np.int32)
    t = np.arange(nsteps) * dt
    hist = np.stack([1e5 * np.sin(40 * t), -2e4 * np.cos(25 * t)], axis=1)
    u = meshes.random_vector(n, 5) * 1e-4
    um, v, a = u.copy(), np.zeros(n), np.zeros(n)
    U, Um, V, A = u.copy(), um.copy(), v.copy(), a.copy()
    om.cd_run(op, m, cm, dt, nsteps, u, um, v, a, load_eq, hist.ravel(), dash=dash)
    c = np.zeros(n) if cm is None else cm
    for k in range(nsteps):
        P = np.zeros(n)
        P[load_eq] = hist[k]
        rhs = P - K @ U + m / dt**2 * (2 * U - Um) + c / (2 * dt) * Um
        Un = rhs / (m / dt**2 + c / (2 * dt))
        # the recorded velocity is OpenSees CentralDifference::update's one-sided second-order
        # difference at t(n+1): (3 u(n+1) - 4 u(n) + u(n-1)) / (2 dt)
        Vn = (3 * Un - 4 * U + Um) / (2 * dt)
        A = (Vn - V) / dt
        Um, U, V = U, Un, Vn
    scale = np.max(np.abs(U))
    assert np.max(np.abs(u - U)) <= 1e-10 * scale
    assert np.max(np.abs(um - Um)) <= 1e-10 * scale
    assert np.max(np.abs(v - V)) <= 1e-8 * np.max(np.abs(V))
    assert np.max(np.abs(a - A)) <= 1e-6 * np.max(np.abs(A))


def test_explicit_difference_matches_the_closed_form_leapfrog():
    """ExplicitDifference (explicit_analysis.cpp:60-85) started in one undamped mode of a dense
    system: the predictor and update form the leapfrog recurrence u(n+1) - 2 u(n) + u(n-1) =
    dt^2 a(n), whose solution in a mode of circular frequency w is u(n) = cos(n t) + c sin(n t)
    with cos t = 1 - (w dt)^2 / 2 and c = (cos t - 1) / sin t (u(0) = 1, vh(0) = 0,
    a(0) = -w^2); the recorded velocity is vh + dt/8 (3 a(n+1) + a(n))."""
    model = meshes.column(6)
    om, op = compiled(model)
    n = om.neq
    m, _ = om.lumped_mass(op)
    K = np.column_stack([om.apply(op, e) for e in np.eye(n)])
    # the mass-normalised modes of M^-1/2 K M^-1/2
    Mh = 1.0 / np.sqrt(m)
    w2, Q = np.linalg.eigh(Mh[:, None] * K * Mh[None, :])
    for mode in (0, 3):
        w = np.sqrt(w2[mode])
        phi = Mh * Q[:, mode]
        dt = 0.4 * 2.0 / np.sqrt(w2[-1])  # 0.4 of the stable step
        nsteps = 40
        u = phi.copy()
        vh = np.zeros(n)
        a = -w * w * phi
        v = np.zeros(n)
        # no body force: the column's own gravity would add a static part
        op0 = fxo.elastic_params([dict(model["params"][g], b=(0.0, 0.0, 0.0)) for g in om.param_group])
        om.explicit_run(op0, m, None, dt, nsteps, u, vh, v, a, scheme=1)
        th = np.arccos(1.0 - 0.5 * (w * dt) ** 2)
        cc = (np.cos(th) - 1.0) / np.sin(th)
        want = np.cos(nsteps * th) + cc * np.sin(nsteps * th)
        scale = np.max(np.abs(phi))
        assert np.max(np.abs(u - want * phi)) <= 1e-9 * scale, (mode, np.max(np.abs(u - want * phi)))
        # the element velocity: vh(n) = (u(n) - u(n-1)) / dt
        prev = np.cos((nsteps - 1) * th) + cc * np.sin((nsteps - 1) * th)
        assert np.max(np.abs(vh - (want - prev) / dt * phi)) <= 1e-8 * w * scale
        # a(n) = -w^2 u(n), and the recorded velocity from it
        assert np.max(np.abs(a + w * w * want * phi)) <= 1e-8 * w * w * scale
        a_prev = -w * w * prev * phi
        assert np.max(np.abs(v - (vh + dt / 8 * (3 * a + a_prev)))) <= 1e-8 * w * scale


@pytest.mark.parametrize("scheme", [0, 1])
def test_explicit_ground_motion_is_a_rigid_base_translation(scheme):
    """A uniform excitation's load -m a_g (ground_load, explicit_analysis.cpp:48-58) on a column
    whose base is fixed: for a constant a_g applied from rest, the relative motion is that of the
    force -M r a_g, i.e. the same run as a point-load history -m a_g at every equation of the
    direction (a closed-form equivalence of the two load paths); checked for both schemes."""
    model = meshes.column(6)
    om, op = compiled(model)
    op0 = fxo.elastic_params([dict(model["params"][g], b=(0.0, 0.0, 0.0)) for g in om.param_group])
    n = om.neq
    m, _ = om.lumped_mass(op0)
    dt, nsteps = 2e-4, 30
    ag = 0.3 * 9.81 * np.sin(np.arange(nsteps) * 0.3)
    x_eqs = np.array(sorted({q for i, q in enumerate(om.node_eq) if q >= 0 and i % 3 == 0}), np.int32)
    s1 = [np.zeros(n) for _ in range(4)]
    om.explicit_run(op0, m, None, dt, nsteps, *s1, scheme=scheme, ground_dir=0, ground_accel=ag)
    # the same loads as point loads: at most 32 equations per run in the device form, so compare
    # with the oracle's own point-load path over all of them
    hist = np.array([[-m[q] * ag[k] for q in x_eqs] for k in range(nsteps)])
    s2 = [np.zeros(n) for _ in range(4)]
    om.explicit_run(op0, m, None, dt, nsteps, *s2, scheme=scheme, load_eq=x_eqs, load_hist=hist)
    for a1, a2 in zip(s1, s2):
        assert np.max(np.abs(a1 - a2)) <= 1e-12 * max(np.max(np.abs(a2)), 1e-300)
    assert np.max(np.abs(s1[0])) > 0
