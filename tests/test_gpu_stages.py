"""GPU parity of the implicit stage forms beyond Newmark + Newton (implicit_stage.cpp:118-402)
against the oracle: static LoadControl stages (fixed and adaptive increments), the Linear and
ModifiedNewton (-initial, -factoronce) algorithms, and the solver picked from the operator's
symmetry (matrix_free_operator.cpp:5-16, krylov.cpp:10-22): GMRES from the start for a
PressureIndependMultiYield batch in a static stage, CG otherwise.

Tolerances: displacements 1e-9 of the largest (elastic; the solves stop at rtol 1e-12), 1e-7 for
the plastic column (the reference's GPU plastic tolerances are looser still, 5e-4 implicit);
counters (steps, commits, Newton iterations) equal, time equal to rounding.
"""
import numpy as np
import pytest

from femorax_amd import core
from oracle import fxo
from tests import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def rel(got, want):
    return np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300)


def ground_dir(om, d):
    r = np.zeros(om.neq)
    q = om.node_eq.reshape(-1, 3)[:, d]
    r[q[q >= 0]] = 1.0
    return r


def device_run(ctx, batch, n, h, nsteps, v0=None, a0=None, **kw):
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for x in d:
        x.zero()
    if v0 is not None:
        d[1].upload(v0)
        d[2].upload(a0)
    r = batch.newmark(*d, h, nsteps, **kw)
    return r, tuple(x.get() for x in d[:3])


STATIC_FORMS = [dict(algorithm=0), dict(algorithm=1), dict(algorithm=2), dict(algorithm=2, modified_initial=True),
                dict(algorithm=2, factor_once=True)]


@pytest.mark.parametrize("form", STATIC_FORMS)
def test_static_elastic_stage(ctx, form):
    model = meshes.column(12)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, dynamic=True)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    n = pm.neq
    v0, a0 = meshes.random_vector(n, 3), meshes.random_vector(n, 4)
    kw = dict(static=True, load_increment=0.25, test=0, tol=1e-6, time=1.0, **form)
    u, v, a, du = np.zeros(n), v0.copy(), a0.copy(), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, 1.0, 4, elastic=op, **kw)
    assert rc == 0
    r, (gu, gv, ga) = device_run(ctx, batch, n, 1.0, 4, v0, a0, **kw)
    print(f"static elastic {form}: {r}", flush=True)
    assert r["failed"] == 0 and r["steps"] == 4 and r["commits"] == st["commits"]
    assert r["newton_iterations"] == st["newton_iterations"]
    assert abs(r["time"] - st["time"]) <= 1e-12 and abs(r["time"] - 2.0) <= 1e-12
    assert rel(gu, u) <= 1e-9
    np.testing.assert_array_equal(gv, v0)
    np.testing.assert_array_equal(ga, a0)


def test_load_control_adaptive_increment(ctx):
    model = meshes.column(8)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, dynamic=True)
    n = pm.neq
    kw = dict(static=True, load_increment=0.1, load_min_increment=0.05, load_max_increment=0.5, load_num_iter=2,
              test=0, tol=1e-6)
    r, _ = device_run(ctx, batch, n, 1.0, 5, **kw)
    assert r["steps"] == 5 and abs(r["time"] - 1.7) <= 1e-12, r  # 0.1 + 0.2 + 0.4 + 0.5 + 0.5


@pytest.mark.parametrize("form", STATIC_FORMS[1:])
def test_transient_algorithms_on_a_damped_column(ctx, form):
    model = meshes.column(16)
    model["params"] = [dict(p, rayleigh=(0.4, 2e-4, 0.0, 0.0)) for p in model["params"]]
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, dynamic=True)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    n = pm.neq
    h, nsteps = 0.004, 30
    ag = 2.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 3)
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, h, nsteps, elastic=op, ground_dir=ground_dir(om, 0), ground_accel=ag, test=1,
                        tol=1e-10, **form)
    assert rc == 0
    r, got = device_run(ctx, batch, n, h, nsteps, ground_dir=0, ground_accel=ag, test=1, tol=1e-10, **form)
    print(f"transient elastic {form}: {r}", flush=True)
    assert r["failed"] == 0 and r["steps"] == nsteps
    assert abs(r["newton_iterations"] - st["newton_iterations"]) <= nsteps // 10 + 1, (r, st)
    for g, w, tol in zip(got, (u, v, a), (1e-9, 1e-8, 1e-8)):
        assert rel(g, w) <= tol


def pimy_column(ctx, gmres=True):
    model = meshes.column(10)
    pm = core.Mesh(model)
    mats = [dict(rho=2000.0, G=G, K=2.17 * G, cohesion=c, gamma_peak=0.1, phi=0.0, pref=100e3, pdc=0.0,
                 b=(0.0, 0.0, -2000.0 * 9.81), mass_type=1) for G, c in ((40e6, 30e3), (80e6, 60e3), (150e6, 90e3))]
    per_set = [mats[g] for g in pm.param_group]
    keep = []
    dev = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"],
                              nsurf=20, b=m["b"], mass_type=m["mass_type"], keep=keep) for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev, gmres=gmres)
    return model, pm, per_set, batch


@pytest.mark.parametrize("form", [dict(algorithm=0), dict(algorithm=2), dict(algorithm=2, modified_initial=True)])
def test_pimy_gravity_stage_runs_gmres(ctx, form):
    """The gravity stage of a site-response run: a static stage on a PressureIndependMultiYield
    column under its body load, elastic stage then plastic, solved by GMRES (the operator is only
    semi-definite without mass)."""
    model, pm, per_set, batch = pimy_column(ctx)
    assert batch.info()["tangent_symmetry"] == core.FX_TANGENT_SPSD
    om = fxo.OracleMesh(model)
    n = pm.neq
    u = np.zeros(n)
    gu = core.DeviceArray(ctx, n)
    dv, da, ddu = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    for x in (gu, dv, da, ddu):
        x.zero()
    du = np.zeros(n)
    Cs = Ts = None
    for stage in (0, 1):
        pe = fxo.pimy_elems(per_set, stage=stage)
        if Cs is None:
            Cs = om.pimy_initial_state(pe)
            Ts = Cs.copy()
        batch.set_stage(stage)
        kw = dict(static=True, load_increment=1.0, test=0, tol=1e-3, max_iter=60, **form)
        rc, st = om.newmark(u, np.zeros(n), np.zeros(n), du, 1.0, 2, pimy=pe, pimy_state=[Cs, Ts], **kw)
        assert rc == 0
        r = batch.newmark(gu, dv, da, ddu, 1.0, 2, **kw)
        print(f"gravity stage {stage} {form}: {r} oracle {st}", flush=True)
        assert r["failed"] == 0 and r["steps"] == 2
        assert r["newton_iterations"] == st["newton_iterations"], (r, st)
        assert rel(gu.get(), u) <= 1e-7, rel(gu.get(), u)
    assert np.max(Cs[:, 26]) > 0, "the body load reached the plastic range"


def test_solver_choice_follows_the_operator_symmetry(ctx):
    """fx_krylov: GMRES from the start on the PIMY batch's stiffness alone (symmetric
    semi-definite), CG with a mass term (definite); a static PIMY stage without the GMRES
    workspace is refused, as the Linear algorithm on a history material is."""
    model, pm, per_set, batch = pimy_column(ctx)
    n = pm.neq
    om = fxo.OracleMesh(model)
    batch.set_stage(0)
    u = core.DeviceArray(ctx, n)
    u.zero()
    batch.store_tangents(u)
    b = core.DeviceArray.from_host(ctx, meshes.random_vector(n, 8))
    x = core.DeviceArray(ctx, n)
    x.zero()
    r_k = batch.krylov(b, x, rtol=1e-10, k=core.coeffs(1.0, 0.0, 0.0))
    xk = x.get()
    x.zero()
    r_g = batch.gmres(b, x, rtol=1e-10, k=core.coeffs(1.0, 0.0, 0.0))
    assert r_k["iterations"] == r_g["iterations"] and np.array_equal(xk, x.get())  # GMRES, not CG first
    assert r_k["indefinite"] == 0
    x.zero()
    r_k = batch.krylov(b, x, rtol=1e-10, k=core.coeffs(1.0, 0.0, 1e4))
    xk = x.get()
    x.zero()
    r_c = batch.cg(b, x, rtol=1e-10, k=core.coeffs(1.0, 0.0, 1e4))
    assert r_k["iterations"] == r_c["iterations"] and np.array_equal(xk, x.get())  # CG
    del om
    _, _, _, nog = pimy_column(ctx, gmres=False)
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for a in d:
        a.zero()
    with pytest.raises(core.FxError, match="FX_BATCH_GMRES"):
        nog.newmark(*d, 1.0, 1, static=True, load_increment=1.0)
    with pytest.raises(core.FxError, match="Linear"):
        nog.newmark(*d, 0.01, 1, algorithm=1)
