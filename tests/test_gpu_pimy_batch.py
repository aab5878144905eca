"""GPU parity of stdBrick<PressureIndependMultiYield> batches against the oracle: the internal
force with the return map at every Gauss point (trial history), the operator and its diagonal
on the stored tangents, commits, stage switches, CG on the tangent operator, the explicit
loop with the history committed every step, and the compacted return map (predict / compact /
correct / replay) against the fused one, bit for bit.

Tolerances: forces, operator, diagonal 1e-10 of the largest entry; histories 1e-10 of each
field's largest value; CG 1e-8 (the reference's time-history tolerance).
"""
import ctypes as C

import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
from oracle import fxo
from tests import meshes

pytestmark = pytest.mark.gpu
NS = core.PIMY_NSTATE


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def rel(got, want):
    return np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300)


# fields of the history grouped as the material uses them: stress tensor, strain tensor, the four
# scalars, surface sizes, moduli, centres; differences are measured against each group's magnitude
HISTORY_GROUPS = [(0, 13), (13, 26), (26, 27), (27, 28), (28, 29), (29, 30), (30, 50), (50, 70), (70, 190)]


def history_error(got, want):
    worst = 0.0
    for a, b in HISTORY_GROUPS:
        scale = max(np.max(np.abs(want[:, a:b])), 1e-300)
        worst = max(worst, np.max(np.abs(got[:, a:b] - want[:, a:b])) / scale)
    return worst


def soil(seed, nlayer):
    rng = S.Lcg(seed)
    out = []
    for _ in range(nlayer):
        G = rng.uniform(10e6, 300e6)
        out.append(dict(rho=2000.0, G=G, K=2.17 * G, cohesion=rng.uniform(20e3, 150e3), gamma_peak=0.1, phi=0.0,
                        pref=100e3, pdc=0.0, nsurf=20, b=(0.0, 0.0, -2000.0 * 9.81), mass_type=1))
    return out


def pimy_model(name):
    if name == "column12_tied":
        m = meshes.column(12)
    else:
        m = meshes.distorted_box(3, layers=3)
    mats = soil(20261206, 3)
    return m, mats


def setup(ctx, name):
    model, mats = pimy_model(name)
    pm = core.Mesh(model)
    per_set = [mats[g] for g in pm.param_group]
    keep = []
    dev_mats = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"],
                                   m["pdc"], nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep)
                for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev_mats)
    om = fxo.OracleMesh(model)
    return model, pm, batch, om, per_set


@pytest.mark.parametrize("name", ["column12_tied", "distorted3"])
def test_force_tangents_operator_and_commits(ctx, name):
    model, pm, batch, om, per_set = setup(ctx, name)
    n = pm.neq
    pe = fxo.pimy_elems(per_set, stage=0)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    np.testing.assert_array_equal(batch.pimy_state(True), Co)
    dir_ = meshes.random_vector(n, 3)
    du, df = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    x = meshes.random_vector(n, 4)
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, n)
    for step in range(8):
        stage = 0 if step < 2 else 1
        if step == 2:
            batch.set_stage(1)
            pe = fxo.pimy_elems(per_set, stage=1)
        u = dir_ * (2e-3 * (step + 1) * (1 if step % 3 else -0.5))
        du.upload(u)
        # a discarded trial first, then the one that is committed
        batch.force(du, df)
        f_o, D_o = om.pimy_force(pe, Co, To, u, want_D=True)
        assert rel(df.get(), f_o) <= 1e-10, (step, rel(df.get(), f_o))
        batch.store_tangents(du, df)
        assert rel(df.get(), f_o) <= 1e-10
        T_dev = batch.pimy_state(False)
        assert history_error(T_dev, To) <= 1e-10, step
        for kk in ((1.0, 0.0, 0.0, False), (1.0, 0.0, 400.0, False), (1.0, 0.0, 0.0, True)):
            batch.apply(dx, dy, core.coeffs(*kk))
            want = om.apply_gp(pe, D_o, x, fxo.coeffs(*kk))
            assert rel(dy.get(), want) <= 1e-10, (step, kk)
            batch.diagonal(dy, core.coeffs(*kk))
            want = om.diagonal_gp(pe, D_o, fxo.coeffs(*kk))
            assert rel(dy.get(), want) <= 1e-10, (step, kk)
        batch.commit()
        Co, To = To, Co
    active = T_dev[:, 26]
    assert np.any(active > 0), "the walk reached the plastic range"


def test_cg_on_the_tangent_operator(ctx):
    model, pm, batch, om, per_set = setup(ctx, "column12_tied")
    n = pm.neq
    batch.set_stage(1)
    pe = fxo.pimy_elems(per_set, stage=1)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    u = meshes.random_vector(n, 9) * 5e-3
    du = core.DeviceArray.from_host(ctx, u)
    batch.store_tangents(du)
    _, D_o = om.pimy_force(pe, Co, To, u, want_D=True)
    k = (1.0, 0.0, 1.6e3, False)  # Newmark-like c3 M term keeps it definite
    b = meshes.random_vector(n, 10)
    db, dx = core.DeviceArray.from_host(ctx, b), core.DeviceArray(ctx, n)
    dx.zero()
    r = batch.cg(db, dx, rtol=1e-12, k=core.coeffs(*k))
    assert r["converged"]
    A = np.column_stack([om.apply_gp(pe, D_o, e, fxo.coeffs(*k)) for e in np.eye(n)])
    xd = np.linalg.solve(A, b)
    assert rel(dx.get(), xd) <= 1e-8


def test_explicit_loop_commits_the_history(ctx):
    model, pm, batch, om, per_set = setup(ctx, "column12_tied")
    n = pm.neq
    batch.set_stage(1)
    pe = fxo.pimy_elems(per_set, stage=1)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    m, _ = om.lumped_mass(pe_as_elastic(per_set))
    dt = 2e-4
    nsteps = 40
    u0 = np.zeros(n)
    v0 = meshes.random_vector(n, 12) * 0.2
    um0 = u0 - v0 * dt
    load_eq = np.array([n - 1], np.int32)
    hist = (2e5 * np.sin(np.arange(nsteps) * dt * 60.0))[:, None]
    # oracle loop with the PIMY force and a commit per step
    u, um, v, a = u0.copy(), um0.copy(), v0.copy(), np.zeros(n)
    c2, c3 = 0.5 / dt, 1.0 / dt**2
    for k in range(nsteps):
        f = om.pimy_force(pe, Co, To, u)
        f[load_eq] -= hist[k]
        ag = u * (-2.0 * c3) + um * c3
        R = -f - m * ag
        U = R / (m * c3)
        Ud = (U * 3.0 + u * -4.0 + um) * c2
        a = (Ud - v) / dt
        um, u, v = u, U, Ud
        Co, To = To, Co
    du, dum = core.DeviceArray.from_host(ctx, u0), core.DeviceArray.from_host(ctx, um0)
    dv, da = core.DeviceArray.from_host(ctx, v0), core.DeviceArray(ctx, n)
    da.zero()
    batch.explicit_cd(dt, nsteps, du, dum, dv, da, load_eq=load_eq, load_hist=hist.ravel())
    assert rel(du.get(), u) <= 1e-9
    assert rel(dv.get(), v) <= 1e-8
    assert history_error(batch.pimy_state(True), Co) <= 1e-9


def test_explicit_difference_commits_the_history(ctx):
    """ExplicitDifference (explicit_analysis.cpp:60-85) on a PressureIndependMultiYield column under
    a ground acceleration: the oracle's loop of predictor, return-map force at u, a = R / m with
    R = -f - m a_g at the x equations, recorded velocity, commit."""
    model, pm, batch, om, per_set = setup(ctx, "column12_tied")
    n = pm.neq
    batch.set_stage(1)
    pe = fxo.pimy_elems(per_set, stage=1)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    m, _ = om.lumped_mass(pe_as_elastic(per_set))
    dt, nsteps = 2e-4, 40
    ag = 8.0 * np.sin(np.arange(1, nsteps + 1) * dt * 90.0)
    xdir = np.zeros(n, bool)
    for i, q in enumerate(pm.node_eq):
        if q >= 0 and i % 3 == 0:
            xdir[q] = True
    u, vh, v, a = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    for k in range(nsteps):
        vh = vh + a * dt
        u = u + vh * dt
        f = om.pimy_force(pe, Co, To, u)
        R = -f + np.where(xdir, -m * ag[k], 0.0)
        an = R / m
        v = vh + (3.0 * an + a) * (dt * 0.125)
        a = an
        Co, To = To, Co
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for x in d:
        x.zero()
    batch.explicit_cd(dt, nsteps, d[0], d[1], d[2], d[3], scheme=1, ground_dir=0, ground_accel=ag)
    assert rel(d[0].get(), u) <= 1e-9
    assert rel(d[1].get(), vh) <= 1e-9
    assert rel(d[2].get(), v) <= 1e-8
    assert history_error(batch.pimy_state(True), Co) <= 1e-9
    assert np.max(Co[:, 26]) > 0, "the shaking reached the plastic range"


def pe_as_elastic(per_set):
    """The lumped mass only needs rho (elastic parameter records with the PIMY sets' rho)."""
    return fxo.elastic_params([dict(E=1.0, nu=0.3, rho=m["rho"], mass_type=m["mass_type"]) for m in per_set])


def test_compacted_return_map_gives_the_same_bits(ctx):
    """FX_BATCH_COMPACT (compacted_kernels.hpp:85-202): forces, stored tangents (through the
    operator) and trial histories equal to the fused kernel's bit for bit, and to the oracle within
    1e-10, along a walk with elastic and plastic points, commits between steps."""
    model, mats = pimy_model("distorted3")
    pm = core.Mesh(model)
    per_set = [mats[g] for g in pm.param_group]
    keep = []

    def dev():
        return [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"],
                                   m["pdc"], nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep)
                for m in per_set]

    fused, comp = core.Batch.pimy(ctx, pm, dev()), core.Batch.pimy(ctx, pm, dev(), compact=True)
    n = pm.neq
    dir_ = meshes.random_vector(n, 31)
    du = core.DeviceArray(ctx, n)
    f1, f2 = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    xh = meshes.random_vector(n, 32)
    x = core.DeviceArray.from_host(ctx, xh)
    y1, y2 = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    # the oracle along the same walk: the compacted results are also checked against it
    om = fxo.OracleMesh(model)
    pe = fxo.pimy_elems(per_set, stage=0)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    plastic = 0
    for step in range(7):
        if step == 1:
            fused.set_stage(1)
            comp.set_stage(1)
            pe = fxo.pimy_elems(per_set, stage=1)
        u = dir_ * (3e-3 * (step + 1) * (1 if step % 3 else -0.6))
        du.upload(u)
        fused.store_tangents(du, f1)
        comp.store_tangents(du, f2)
        np.testing.assert_array_equal(f2.get(), f1.get())
        np.testing.assert_array_equal(comp.pimy_state(False), fused.pimy_state(False))
        f_o, D_o = om.pimy_force(pe, Co, To, u, want_D=True)
        assert rel(f2.get(), f_o) <= 1e-10, step
        assert history_error(comp.pimy_state(False), To) <= 1e-10, step
        fused.apply(x, y1)
        comp.apply(x, y2)
        np.testing.assert_array_equal(y2.get(), y1.get())
        assert rel(y2.get(), om.apply_gp(pe, D_o, xh)) <= 1e-10, step
        fused.force(du, f1)
        comp.force(du, f2)
        np.testing.assert_array_equal(f2.get(), f1.get())
        plastic = max(plastic, int(np.sum(fused.pimy_state(False)[:, 26] > 0)))
        fused.commit()
        comp.commit()
        Co, To = To, Co
    assert plastic > 0


DAMP = (0.6, 2e-3, 1.5e-3, 1e-3)  # alphaM, betaK, betaKinit, betaKcomm: every term of the damping


def setup_damped(ctx, name, mass_type=None):
    model, mats = pimy_model(name)
    pm = core.Mesh(model)
    per_set = [dict(mats[g]) for g in pm.param_group]
    for m in per_set:
        m["rayleigh"] = DAMP
        if mass_type is not None:
            m["mass_type"] = mass_type
    keep = []
    dev_mats = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"],
                                   m["pdc"], nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep)
                for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev_mats, rayleigh=[DAMP] * len(per_set))
    return model, pm, batch, fxo.OracleMesh(model), per_set


@pytest.mark.parametrize("name,mass_type", [("column12_tied", None), ("distorted3", 0)])
def test_damped_force_operator_and_diagonal(ctx, name, mass_type):
    """Rayleigh damping on PIMY batches (std_brick.hpp:409-464, element_operator.hpp:45-75): the
    force at (u, v), the operator and diagonal with the Newmark-like and the damping-only
    coefficients (the committed tangent of betaKcomm from the committed history), against the
    oracle along a walk into the plastic range."""
    model, pm, batch, om, per_set = setup_damped(ctx, name, mass_type)
    n = pm.neq
    batch.set_stage(1)
    pe = fxo.pimy_elems(per_set, stage=1)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    dir_ = meshes.random_vector(n, 41)
    du, dv, df = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    x = meshes.random_vector(n, 42)
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, n)
    for step in range(6):
        u = dir_ * (2.5e-3 * (step + 1) * (1 if step % 3 else -0.5))
        v = meshes.random_vector(n, 50 + step) * 0.4
        du.upload(u)
        dv.upload(v)
        batch.force(du, df, dv)
        f_o = om.pimy_force(pe, Co, To, u, v=v)
        assert rel(df.get(), f_o) <= 1e-10, (step, rel(df.get(), f_o))
        batch.store_tangents(du)
        _, D_o = om.pimy_force(pe, Co, To, u, want_D=True)
        for kk in ((1.0, 400.0, 8e4, False), (0.0, 1.0, 0.0, False), (1.0, 400.0, 8e4, True)):
            batch.apply(dx, dy, core.coeffs(*kk))
            want = om.apply_gp(pe, D_o, x, fxo.coeffs(*kk), committed=Co)
            assert rel(dy.get(), want) <= 1e-10, (step, kk)
            batch.diagonal(dy, core.coeffs(*kk))
            want = om.diagonal_gp(pe, D_o, fxo.coeffs(*kk), committed=Co)
            assert rel(dy.get(), want) <= 1e-10, (step, kk)
        batch.commit()
        Co, To = To, Co
    assert np.any(Co[:, 26] > 0), "the walk reached the plastic range"


def test_damped_newmark_stage(ctx):
    """Newmark + Newton on a damped PIMY column: the unbalance carries the damping force at the
    trial velocities, the tangent operator the damping terms; against the oracle's stage."""
    model, pm, batch, om, per_set = setup_damped(ctx, "column12_tied")
    for m in per_set:
        m["b"] = (0.0, 0.0, 0.0)
    keep = []
    dev_mats = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"],
                                   m["pdc"], nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep)
                for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev_mats, rayleigh=[DAMP] * len(per_set))
    batch.set_stage(1)
    pe = fxo.pimy_elems(per_set, stage=1)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    n = pm.neq
    r = np.zeros(n)
    q = om.node_eq.reshape(-1, 3)[:, 0]
    r[q[q >= 0]] = 1.0
    h, nsteps = 0.005, 16
    ag = 6.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 2)
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, h, nsteps, pimy=pe, pimy_state=[Co, To], ground_dir=r, ground_accel=ag, test=0,
                        tol=1e-3)
    assert rc == 0 and st["steps"] == nsteps
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for x in d:
        x.zero()
    res = batch.newmark(*d, h, nsteps, ground_dir=0, ground_accel=ag, test=0, tol=1e-3)
    assert res["failed"] == 0 and res["steps"] == nsteps
    assert abs(res["newton_iterations"] - st["newton_iterations"]) <= 2
    for g, w, tol in zip((x.get() for x in d[:3]), (u, v, a), (1e-8, 1e-7, 1e-7)):
        assert rel(g, w) <= tol
    assert history_error(batch.pimy_state(True), Co) <= 1e-7
    assert np.max(Co[:, 26]) > 0


def test_damped_explicit_loop(ctx):
    """CentralDifference on a damped PIMY column (explicit_analysis.cpp:87-117, 712-722): the
    element force at the garbage velocity -c2 u(n-1) with its damping terms, the lumped alphaM
    mass on the diagonal."""
    model, pm, batch, om, per_set = setup_damped(ctx, "column12_tied")
    n = pm.neq
    batch.set_stage(1)
    pe = fxo.pimy_elems(per_set, stage=1)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    m, am = om.lumped_mass(fxo.elastic_params([dict(E=1.0, nu=0.3, rho=s["rho"], mass_type=s["mass_type"],
                                                    rayleigh=DAMP) for s in per_set]))
    dt, nsteps = 2e-4, 40
    u0 = np.zeros(n)
    v0 = meshes.random_vector(n, 13) * 0.2
    um0 = u0 - v0 * dt
    u, um, v = u0.copy(), um0.copy(), v0.copy()
    c2, c3 = 0.5 / dt, 1.0 / dt**2
    for k in range(nsteps):
        f = om.pimy_force(pe, Co, To, u, v=um * -c2)
        ag = u * (-2.0 * c3) + um * c3
        R = -f - m * ag
        U = R / (m * c3 + am * c2)
        Ud = (U * 3.0 + u * -4.0 + um) * c2
        um, u, v = u, U, Ud
        Co, To = To, Co
    du, dum = core.DeviceArray.from_host(ctx, u0), core.DeviceArray.from_host(ctx, um0)
    dv, da = core.DeviceArray.from_host(ctx, v0), core.DeviceArray(ctx, n)
    da.zero()
    batch.explicit_cd(dt, nsteps, du, dum, dv, da)
    assert rel(du.get(), u) <= 1e-9
    assert rel(dv.get(), v) <= 1e-8
    assert history_error(batch.pimy_state(True), Co) <= 1e-9


@pytest.mark.skipif(fxo.ref() is None, reason="oracle/_ref not built")
def test_stable_step_from_the_committed_stability_tangent(ctx):
    """fx_critical_dt of a PIMY batch (min_critical_dt, element_kernels.hpp:453-474): every
    point's stability tangent from its committed state (stability_moduli,
    pressure_independ_multi_yield.hpp:174-203; a frictional soil scales its moduli by the
    pressure while the switch to the plastic stage is pending). Against the reference's own
    StdBrick<PressureIndependMultiYield<20>>::critical_dt (oracle/_ref) element by element, on
    the committed histories read back from the device; 1e-12 (two eigenvalue solvers)."""
    from tests.test_oracle_pimy_damping import _ref_params
    R = fxo.ref()
    R.fxref_pimy_critical_dt.restype = C.c_double
    R.fxref_pimy_critical_dt.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    model = meshes.distorted_box(3, layers=1)
    pm = core.Mesh(model)
    m = dict(rho=1900.0, G=60e6, K=160e6, cohesion=5e3, gamma_peak=0.1, phi=30.0, pref=80e3, pdc=0.5, nsurf=20,
             b=(0.0, 0.0, -1900.0 * 9.81), mass_type=0)
    keep = []
    dev_mats = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"],
                                   m["pdc"], nsurf=20, b=m["b"], mass_type=0, keep=keep) for _ in range(pm.nparam)]
    batch = core.Batch.pimy(ctx, pm, dev_mats)
    rp = _ref_params(m, R)
    ray = np.zeros(4)
    coords = np.asarray(pm.coords).reshape(3, -1)
    node = np.asarray(pm.node).reshape(8, -1)

    def reference_dt(stage, scheme):
        Cs = np.ascontiguousarray(batch.pimy_state(True))
        best = np.inf
        for e in range(pm.nelem):
            xe = np.ascontiguousarray(coords[:, node[:, e]].T)
            Ce = np.ascontiguousarray(Cs[8 * e:8 * e + 8])
            best = min(best, R.fxref_pimy_critical_dt(rp.ctypes.data, stage, ray.ctypes.data, scheme, xe.ctypes.data,
                                                      Ce.ctypes.data))
        return best

    # compress the box (stage 0), commit, switch to the plastic stage: the pending switch scales
    n = pm.neq
    du, df = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    du.upload(meshes.random_vector(n, 71) * 2e-4 - 1e-3)
    batch.force(du, df)
    batch.commit()
    at_rest = batch.critical_dt(0)
    batch.set_stage(1)
    for scheme in (0, 1):
        got, want = batch.critical_dt(scheme), reference_dt(1, scheme)
        assert abs(got - want) <= 1e-12 * want, (scheme, got, want)
    assert batch.critical_dt(0) != at_rest, "the pending switch changes the moduli"
    # after a plastic-stage update and its commit the points carry their own moduli
    batch.force(du, df)
    batch.commit()
    got, want = batch.critical_dt(0), reference_dt(1, 0)
    assert abs(got - want) <= 1e-12 * want, (got, want)


def test_large_batch_operator_one_element_per_thread(ctx):
    """A batch of 33^3 elements (above the size where the operator and its diagonal run one
    thread per element): the operator with the Newmark-like coefficients and damping terms, and
    the diagonal, on tangents stored at a yielding state, against the oracle."""
    model = S.elastic_box(33, 33.0, layers=3, body=True, seed=17)
    pm = core.Mesh(model)
    assert pm.nelem >= 32768
    mats = soil(20261215, 3)
    per_set = [dict(mats[g % 3], rayleigh=DAMP) for g in range(pm.nparam)]
    keep = []
    dev_mats = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"],
                                   m["pdc"], nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep)
                for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev_mats, rayleigh=[DAMP] * len(per_set))
    batch.set_stage(1)
    om = fxo.OracleMesh(model)
    pe = fxo.pimy_elems([per_set[g] for g in range(pm.nparam)], stage=1)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    n = pm.neq
    u = meshes.random_vector(n, 61) * 2e-3
    du = core.DeviceArray.from_host(ctx, u)
    batch.store_tangents(du)
    _, D_o = om.pimy_force(pe, Co, To, u, want_D=True)
    x = meshes.random_vector(n, 62)
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, n)
    for kk in ((1.0, 400.0, 8e4, False), (1.0, 0.0, 0.0, True)):
        batch.apply(dx, dy, core.coeffs(*kk))
        assert rel(dy.get(), om.apply_gp(pe, D_o, x, fxo.coeffs(*kk), committed=Co)) <= 1e-10, kk
        batch.diagonal(dy, core.coeffs(*kk))
        assert rel(dy.get(), om.diagonal_gp(pe, D_o, fxo.coeffs(*kk), committed=Co)) <= 1e-10, kk


def test_graph_replayed_newmark_keeps_the_committed_history(ctx):
    """Graph-replayed CG inside a Newmark stage on a damped PIMY column (betaKcomm: the operator
    reads the committed history, whose buffer swaps with the trial one at every commit): the
    same bits as the launched iterations over several committed steps."""
    results = []
    for graphs in (False, True):
        model, pm, batch, om, per_set = setup_damped(ctx, "column12_tied")
        batch.set_stage(1)
        n = pm.neq
        h, nsteps = 0.005, 6
        ag = 6.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 2)
        d = [core.DeviceArray(ctx, n) for _ in range(4)]
        for x in d:
            x.zero()
        ctx.set_graphs(graphs)
        try:
            res = batch.newmark(*d, h, nsteps, ground_dir=0, ground_accel=ag, test=0, tol=1e-3)
        finally:
            ctx.set_graphs(False)
        assert res["failed"] == 0 and res["steps"] == nsteps, res
        results.append(([x.get() for x in d[:3]], batch.pimy_state(True), res))
        batch.close()
    (u0, s0, r0), (u1, s1, r1) = results
    assert r0["krylov_iterations"] == r1["krylov_iterations"], (r0, r1)
    for a, b in zip(u0, u1):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(s0, s1)
    assert np.max(s0[:, 26]) > 0, "the stage reached the plastic range"
