"""GPU parity of the internal force, lumped mass, body load, stable step and the explicit
CentralDifference loop against the oracle (explicit_analysis.cpp:87-117, 712-722;
internal_force.cpp:354-377; std_brick.hpp:609-694).

Tolerances: force, lumped vectors 1e-12 of the largest entry; stable step 1e-12 relative;
CentralDifference displacements 1e-10 of the largest after 60-120 steps (the reference's
time-history tolerance, docs/deviations.md), velocities and accelerations 1e-8.
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


def damped_distorted():
    m = meshes.distorted_box(4)
    m["params"] = [dict(p, rayleigh=(0.6, 5.8e-4, 1e-4, 2e-5)) for p in m["params"]]
    m["region"] = m["group"].copy()
    m["rayleigh"] = [(0.6, 5.8e-4, 1e-4, 2e-5), (0.9, 3e-4, 0.0, 2e-5)][:len(m["params"])]
    return m


def c3_small():
    return S.config_c3(12, 10, 8, 5.0, layers=4)


MODELS = {
    "c3_small_lattice": (c3_small, False),
    "c3_small_general": (c3_small, True),
    "distorted4_damped": (damped_distorted, True),
    "column10_tied": (lambda: meshes.column(10), True),
}


def setup(ctx, name):
    make, general = MODELS[name]
    model = make()
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, general=general)
    om = fxo.OracleMesh(model)
    return model, pm, batch, om, fxo.elastic_params(params)


@pytest.mark.parametrize("name", sorted(MODELS))
def test_lumped_mass_body_load_and_force(ctx, name):
    model, pm, batch, om, op = setup(ctx, name)
    n = pm.neq
    m, am = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    batch.lumped_mass(m, am)
    om_m, om_am = om.lumped_mass(op)
    assert rel(m.get(), om_m) <= 1e-12
    if np.any(om_am):
        assert rel(am.get(), om_am) <= 1e-12
    else:
        assert not np.any(am.get())
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
    batch.force(du, df)
    assert rel(df.get(), om.force(op, u)) <= 1e-12


@pytest.mark.parametrize("name", sorted(MODELS))
def test_critical_dt(ctx, name):
    model, pm, batch, om, op = setup(ctx, name)
    for scheme in (0, 1):
        got, want = batch.critical_dt(scheme), om.critical_dt(op, scheme)
        assert abs(got - want) <= 1e-12 * want, (scheme, got, want)


@pytest.mark.parametrize("name,nsteps", [("c3_small_lattice", 120), ("c3_small_general", 61),
                                         ("distorted4_damped", 41), ("column10_tied", 60)])
def test_central_difference_matches_oracle(ctx, name, nsteps):
    model, pm, batch, om, op = setup(ctx, name)
    n = pm.neq
    dt = 0.5 * om.critical_dt(op)
    m, am = om.lumped_mass(op)
    dash = pm.to_equations(model["dashpots"]) if "dashpots" in model else None
    if "source_node" in model:
        src = model["source_node"]
        load_eq = np.array([pm.node_eq[3 * pm.node_rank[src] + 2]], np.int32)
        hist = S.ricker(np.arange(nsteps) * dt, f=5.0, amplitude=1e6)[:, None]
    else:
        load_eq = np.array([n - 1, n // 2], np.int32)
        t = np.arange(nsteps) * dt
        hist = np.stack([1e4 * np.sin(60 * t), -3e3 * np.cos(35 * t)], axis=1)
    u0 = meshes.random_vector(n, 7) * 1e-5
    v0 = meshes.random_vector(n, 8) * 1e-4
    um0 = u0 - v0 * dt
    # oracle
    u, um, v, a = u0.copy(), um0.copy(), v0.copy(), np.zeros(n)
    cm = None
    if np.any(am) or dash is not None:
        cm = (am if np.any(am) else 0.0) + (dash if dash is not None else 0.0)
    om.cd_run(op, m, cm, dt, nsteps, u, um, v, a, load_eq, hist.ravel(), dash=dash)
    # device
    du, dum = core.DeviceArray.from_host(ctx, u0), core.DeviceArray.from_host(ctx, um0)
    dv, da = core.DeviceArray.from_host(ctx, v0), core.DeviceArray(ctx, n)
    da.zero()
    ddash = core.DeviceArray.from_host(ctx, dash) if dash is not None else None
    batch.explicit_cd(dt, nsteps, du, dum, dv, da, dash=ddash, load_eq=load_eq, load_hist=hist.ravel())
    assert np.max(np.abs(u - u0)) > 1e-3 * np.max(np.abs(u0))  # the run moved
    assert rel(du.get(), u) <= 1e-10
    assert rel(dum.get(), um) <= 1e-10
    assert rel(dv.get(), v) <= 1e-8
    assert rel(da.get(), a) <= 1e-8


@pytest.mark.parametrize("name,nsteps", [("c3_small_lattice", 80), ("distorted4_damped", 41), ("column10_tied", 60)])
@pytest.mark.parametrize("scheme", [0, 1])
def test_explicit_schemes_under_ground_motion(ctx, name, nsteps, scheme):
    """CentralDifference and ExplicitDifference (explicit_analysis.cpp:60-117, 687-722) with a
    uniform excitation's load -m a_g (:48-58) from a time series evaluated at each step's load
    time (t(n) for CentralDifference, t(n+1) for ExplicitDifference), point loads and dampers
    where the model has them: against the oracle's loop (per-step accelerations)."""
    model, pm, batch, om, op = setup(ctx, name)
    n = pm.neq
    dt = 0.5 * om.critical_dt(op, scheme=scheme)
    m, am = om.lumped_mass(op)
    dash = pm.to_equations(model["dashpots"]) if "dashpots" in model else None
    series = dict(values=np.concatenate([[0.0], 2.5 * np.sin(np.arange(1, nsteps + 4) * 0.21)]), dt=dt * 0.9,
                  factor=1.3)
    t0 = 0.37 * dt
    times = t0 + dt * (np.arange(nsteps) + (1 if scheme == 1 else 0))
    ag = np.array([fxo.series_value(series["values"], series["dt"], t, factor=series["factor"], fact=0.8)
                   for t in times])
    load_eq = np.array([n - 1, n // 2], np.int32)
    t = np.arange(nsteps) * dt
    hist = np.stack([1e3 * np.sin(60 * t), -3e2 * np.cos(35 * t)], axis=1)
    u0 = meshes.random_vector(n, 7) * 1e-5
    v0 = meshes.random_vector(n, 8) * 1e-4
    um0 = u0 - v0 * dt if scheme == 0 else v0.copy()  # ED: the element velocity starts at v
    u, um, v, a = u0.copy(), um0.copy(), v0.copy(), np.zeros(n)
    cm = None
    if scheme == 0 and (np.any(am) or dash is not None):
        cm = (am if np.any(am) else 0.0) + (dash if dash is not None else 0.0)
    om.explicit_run(op, m, cm, dt, nsteps, u, um, v, a, scheme=scheme, ground_dir=0, ground_accel=ag,
                    load_eq=load_eq, load_hist=hist.ravel(), dash=dash)
    du, dum = core.DeviceArray.from_host(ctx, u0), core.DeviceArray.from_host(ctx, um0)
    dv, da = core.DeviceArray.from_host(ctx, v0), core.DeviceArray(ctx, n)
    da.zero()
    ddash = core.DeviceArray.from_host(ctx, dash) if dash is not None else None
    print(f"explicit {name} scheme {scheme}: {nsteps} steps", flush=True)
    batch.explicit_cd(dt, nsteps, du, dum, dv, da, dash=ddash, load_eq=load_eq, load_hist=hist.ravel(), scheme=scheme,
                      ground_dir=0, series=series, ground_factor=0.8, time=t0)
    assert np.max(np.abs(u - u0)) > 1e-3 * np.max(np.abs(u0))
    assert rel(du.get(), u) <= 1e-10
    assert rel(dum.get(), um) <= 1e-10
    assert rel(dv.get(), v) <= 1e-8
    assert rel(da.get(), a) <= 1e-8


def test_explicit_difference_needs_masses_everywhere(ctx):
    model = meshes.column(6)
    model["params"] = [dict(p, rho=0.0) for p in model["params"]]
    pm = core.Mesh(model)
    batch = core.Batch(ctx, pm, pm.params_for(model))
    d = [core.DeviceArray(ctx, pm.neq) for _ in range(4)]
    for x in d:
        x.zero()
    with pytest.raises(core.FxError, match="positive lumped mass"):
        batch.explicit_cd(1e-4, 2, *d, scheme=1)
