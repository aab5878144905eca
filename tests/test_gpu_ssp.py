"""GPU parity of the SSPbrick<ElasticIsotropic> batch (FX_BATCH_SSPBRICK) against the oracle, whose
SSPbrick is pinned by the reference's OpenSees goldens (tests/test_oracle_ssp.py).

On the structured-lattice path (modal stabilisation per parameter set) and the general path
(per-element constants), on boxes, layered soil, a distorted mesh and a tied column: the operator
(element_operator_apply with SSPbrick's element term) for stiffness, damping, mass and mixed
coefficients, its diagonal, fx_run_operator's slot and plain forms and the unit-direction mass
product, the internal force with damping and body load, the lumped mass, the stable step, CG
against the oracle's CG and a dense solve, the FP32 operator, and a CentralDifference run.

Tolerances: the reference's (docs/deviations.md): one apply, diagonal, force within 1e-12 of the
largest entry; FP32 within 2e-6; stable step 1e-12 relative; CG at rtol 1e-13 within 1e-10;
time history 1e-10 (displacements) and 1e-8 (velocities, accelerations).
"""
import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
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


def damped(model):
    """One Rayleigh region per element group, each with its own factors."""
    model["region"] = model["group"].copy()
    model["rayleigh"] = [(0.6 + 0.1 * g, 5.8e-4, 1e-4 * (g % 2), 2e-5) for g in range(len(model["params"]))]
    return model


CASES = {
    "box16": (lambda: S.config_c1(), (False, True)),
    "layered16_damped": (lambda: damped(S.config_c2(n=16)), (False, True)),
    "c3_small": (lambda: S.config_c3(12, 10, 8, 5.0, layers=4), (False,)),
    "distorted6_damped": (lambda: damped(meshes.distorted_box(6)), (True,)),
    "column24_tied": (lambda: meshes.column(24), (True,)),
}
PATHS = [(n, g) for n in sorted(CASES) for g in CASES[n][1]]
COEFFS = [(1.0, 0.0, 0.0, False), (1.0, 40.0, 1.6e3, False), (0.0, 0.0, 1.0, False), (0.0, 1.0, 0.0, False),
          (1.0, 20.0, 4e2, True)]


def path_id(p):
    return f"{p[0]}-{'general' if p[1] else 'lattice'}"


def setup(ctx, name, general, **kw):
    model = CASES[name][0]()
    pm = core.Mesh(model)
    params = [dict(p, element="SSPbrick") for p in pm.params_for(model)]
    batch = core.Batch(ctx, pm, params, general=general, element="SSPbrick", **kw)
    om = fxo.OracleMesh(model)
    return model, pm, batch, om, fxo.elastic_params(params)


@pytest.mark.parametrize("case", PATHS, ids=path_id)
def test_ssp_apply_and_diagonal(ctx, case):
    name, general = case
    model, pm, batch, om, op = setup(ctx, name, general)
    info = batch.info()
    tied = name == "column24_tied"
    assert info["lattice"] == (0 if general or tied else 1)
    x = meshes.random_vector(pm.neq, 31)
    dx, dy, dd = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, pm.neq), core.DeviceArray(ctx, pm.neq)
    for kk in COEFFS:
        k = core.coeffs(*kk)
        batch.apply(dx, dy, k)
        assert rel(dy.get(), om.apply(op, x, fxo.coeffs(*kk))) <= 1e-12, kk
        batch.diagonal(dd, k)
        assert rel(dd.get(), om.diagonal(op, fxo.coeffs(*kk))) <= 1e-12, kk
    # deterministic
    batch.apply(dx, dy)
    a = dy.get()
    for _ in range(2):
        batch.apply(dx, dy)
        np.testing.assert_array_equal(dy.get(), a)


@pytest.mark.parametrize("case", [("layered16_damped", False), ("distorted6_damped", True)], ids=path_id)
def test_ssp_run_operator_forms(ctx, case):
    """fx_run_operator: the per-element slots gathered on the host, the plain form added into y,
    the diagonal's slots, and the unit-direction mass product (mass_times_ground)."""
    name, general = case
    model, pm, batch, om, op = setup(ctx, name, general, slots=True)
    P, n = pm.padded, pm.neq
    x = meshes.random_vector(n, 5)
    k = (1.0, 3.0, 200.0, False)
    dx = core.DeviceArray.from_host(ctx, x)
    ds = core.DeviceArray(ctx, 24 * P)
    eq = np.asarray(pm.eq).reshape(24, P)

    def gather(s):
        y = np.zeros(n)
        s = s.reshape(24, P)
        for a in range(24):
            m = eq[a, :pm.nelem] >= 0
            np.add.at(y, eq[a, :pm.nelem][m], s[a, :pm.nelem][m])
        return y

    batch.run_operator(core.FX_OUT_SLOTS, x=dx, slots=ds, k=core.coeffs(*k))
    assert rel(gather(ds.get()), om.apply(op, x, fxo.coeffs(*k))) <= 1e-12
    base = meshes.random_vector(n, 6)
    dy = core.DeviceArray.from_host(ctx, base)
    batch.run_operator(core.FX_OUT_PLAIN, x=dx, y=dy, k=core.coeffs(*k))
    assert rel(dy.get() - base, om.apply(op, x, fxo.coeffs(*k))) <= 1e-12
    batch.run_operator(core.FX_OUT_SLOTS, slots=ds, diagonal=True, k=core.coeffs(*k))
    assert rel(gather(ds.get()), om.diagonal(op, fxo.coeffs(*k))) <= 1e-12
    for d in range(3):
        dm = core.DeviceArray(ctx, n)
        dm.zero()
        batch.mass_times_ground(d, dm)
        assert rel(dm.get(), om.mass_times_ground(d, elastic=op)) <= 1e-12, d


@pytest.mark.parametrize("case", PATHS, ids=path_id)
def test_ssp_force_mass_and_stable_step(ctx, case):
    name, general = case
    model, pm, batch, om, op = setup(ctx, name, general)
    n = pm.neq
    m, am = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    batch.lumped_mass(m, am)
    om_m, om_am = om.lumped_mass(op)
    assert rel(m.get(), om_m) <= 1e-12
    if np.any(om_am):
        assert rel(am.get(), om_am) <= 1e-12
    fb = core.DeviceArray(ctx, n)
    batch.body_load(fb)
    want_b = om.body_load(op)
    if np.any(want_b):
        assert rel(fb.get(), want_b) <= 1e-12
    u = meshes.random_vector(n, 41) * 1e-3
    v = meshes.random_vector(n, 42) * 1e-2
    du, dv, df = core.DeviceArray.from_host(ctx, u), core.DeviceArray.from_host(ctx, v), core.DeviceArray(ctx, n)
    batch.force(du, df, dv)
    assert rel(df.get(), om.force(op, u, v)) <= 1e-12
    for scheme in (0, 1):
        got, want = batch.critical_dt(scheme), om.critical_dt(op, scheme)
        assert abs(got - want) <= 1e-12 * want, (scheme, got, want)


@pytest.mark.parametrize("case", [("box16", False), ("distorted6_damped", True), ("column24_tied", True)],
                         ids=path_id)
def test_ssp_cg_matches_oracle(ctx, case):
    name, general = case
    model, pm, batch, om, op = setup(ctx, name, general)
    b = om.body_load(op)
    if np.max(np.abs(b)) == 0:
        b = meshes.random_vector(pm.neq, 3)
    db = core.DeviceArray.from_host(ctx, b)
    dx = core.DeviceArray(ctx, pm.neq)
    dx.zero()
    r = batch.cg(db, dx, rtol=1e-8)
    xo, ro = om.cg(op, b, rtol=1e-8, lanes=64)
    assert r["converged"] and ro["converged"]
    assert abs(r["iterations"] - ro["iterations"]) <= 2, (r["iterations"], ro["iterations"])
    dx.zero()
    batch.cg(db, dx, rtol=1e-13)
    xo2, _ = om.cg(op, b, rtol=1e-13, lanes=64)
    x = dx.get()
    assert rel(x, xo2) <= 1e-10
    if pm.neq <= 4000:
        A = np.column_stack([om.apply(op, e) for e in np.eye(pm.neq)])
        assert rel(x, np.linalg.solve(A, b)) <= 1e-9


@pytest.mark.parametrize("case", [("layered16_damped", False), ("distorted6_damped", True)], ids=path_id)
def test_ssp_fp32_apply(ctx, case):
    name, general = case
    model, pm, batch, om, op = setup(ctx, name, general, fp32=True)
    x = meshes.random_vector(pm.neq, 9)
    dx = core.DeviceArray.from_host(ctx, x.astype(np.float32), dtype=np.float32)
    dy = core.DeviceArray(ctx, pm.neq, dtype=np.float32)
    for kk in COEFFS[:3]:
        batch.apply_f32(dx, dy, core.coeffs(*kk))
        want = om.apply(op, x.astype(np.float32).astype(np.float64), fxo.coeffs(*kk))
        assert rel(dy.get().astype(np.float64), want) <= 2e-6, kk


@pytest.mark.parametrize("case", [("c3_small", False), ("distorted6_damped", True)], ids=path_id)
def test_ssp_central_difference(ctx, case):
    name, general = case
    model, pm, batch, om, op = setup(ctx, name, general)
    n, nsteps = pm.neq, 60
    dt = 0.5 * om.critical_dt(op)
    m, am = om.lumped_mass(op)
    load_eq = np.array([n - 1, n // 2], np.int32)
    t = np.arange(nsteps) * dt
    hist = np.stack([1e4 * np.sin(60 * t), -3e3 * np.cos(35 * t)], axis=1)
    u0 = meshes.random_vector(n, 7) * 1e-5
    v0 = meshes.random_vector(n, 8) * 1e-4
    um0 = u0 - v0 * dt
    u, um, v, a = u0.copy(), um0.copy(), v0.copy(), np.zeros(n)
    om.cd_run(op, m, am if np.any(am) else None, dt, nsteps, u, um, v, a, load_eq, hist.ravel())
    du, dum = core.DeviceArray.from_host(ctx, u0), core.DeviceArray.from_host(ctx, um0)
    dv, da = core.DeviceArray.from_host(ctx, v0), core.DeviceArray(ctx, n)
    da.zero()
    batch.explicit_cd(dt, nsteps, du, dum, dv, da, load_eq=load_eq, load_hist=hist.ravel())
    assert np.max(np.abs(u - u0)) > 1e-3 * np.max(np.abs(u0))
    assert rel(du.get(), u) <= 1e-10
    assert rel(dv.get(), v) <= 1e-8
    assert rel(da.get(), a) <= 1e-8
