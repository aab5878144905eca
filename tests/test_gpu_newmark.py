"""GPU parity of the implicit Newmark stage with Newton (implicit_stage.cpp) against the oracle's
restatement: elastic tied column with Rayleigh damping and a ground acceleration (general
path), elastic box with body force (lattice path), and a PressureIndependMultiYield column
shaken into the plastic range (stored tangents, commits).

Tolerances: displacements 1e-9 of the largest, velocities and accelerations 1e-8 (the solves
stop at rtol 1e-12, so the two runs' iterates differ by rounding only); histories 1e-8.
"""
import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
from oracle import fxo
from tests import meshes
from tests.test_gpu_pimy_batch import history_error, soil

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


def run_both(ctx, om, batch, n, h, nsteps, oracle_kw, ag, test=1, tol=1e-10):
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, h, nsteps, ground_dir=ground_dir(om, 0), ground_accel=ag, test=test, tol=tol,
                        **oracle_kw)
    assert rc == 0 and st["steps"] == nsteps
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for x in d:
        x.zero()
    r = batch.newmark(*d, h, nsteps, ground_dir=0, ground_accel=ag, test=test, tol=tol)
    assert r["failed"] == 0 and r["steps"] == nsteps, r
    assert abs(r["newton_iterations"] - st["newton_iterations"]) <= nsteps // 10 + 1, (r, st)
    return (u, v, a), tuple(x.get() for x in d[:3])


def test_elastic_column_with_damping_and_ground_motion(ctx):
    model = meshes.column(16)
    model["params"] = [dict(p, rayleigh=(0.4, 2e-4, 0.0, 0.0)) for p in model["params"]]
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, dynamic=True)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    h, nsteps = 0.004, 40
    ag = 2.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 3)
    want, got = run_both(ctx, om, batch, pm.neq, h, nsteps, dict(elastic=op), ag)
    for g, w, tol in zip(got, want, (1e-9, 1e-8, 1e-8)):
        assert rel(g, w) <= tol


def test_elastic_box_lattice_path(ctx):
    model = S.elastic_box(8, 8.0, layers=2, body=True, seed=5)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, dynamic=True)
    assert batch.info()["lattice"] == 1
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    h, nsteps = 0.002, 25
    ag = 1.5 * np.cos(np.arange(1, nsteps + 1) * h * 2 * np.pi * 5)
    want, got = run_both(ctx, om, batch, pm.neq, h, nsteps, dict(elastic=op), ag)
    for g, w, tol in zip(got, want, (1e-9, 1e-8, 1e-8)):
        assert rel(g, w) <= tol


def test_pimy_column_shaken_into_plasticity(ctx):
    model = meshes.column(12)
    pm = core.Mesh(model)
    mats = soil(20261206, 3)
    per_set = [mats[g] for g in pm.param_group]
    for m in per_set:
        m["b"] = (0.0, 0.0, 0.0)
    keep = []
    dev = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"],
                              nsurf=20, b=m["b"], mass_type=m["mass_type"], keep=keep) for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev)
    batch.set_stage(1)
    om = fxo.OracleMesh(model)
    pe = fxo.pimy_elems(per_set, stage=1)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    h, nsteps = 0.005, 20
    ag = 6.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 2)
    want, got = run_both(ctx, om, batch, pm.neq, h, nsteps, dict(pimy=pe, pimy_state=[Cs, Ts]), ag, test=0, tol=1e-3)
    for g, w, tol in zip(got, want, (1e-8, 1e-7, 1e-7)):
        assert rel(g, w) <= tol
    assert history_error(batch.pimy_state(True), Cs) <= 1e-7
    assert np.max(Cs[:, 26]) > 0, "the shaking reached the plastic range"


def test_energy_incr_test(ctx):
    """EnergyIncr (0.5 |dU . B|, implicit_stage.cpp:239-246) on the elastic damped column: the
    same steps, iterations and state as the oracle."""
    model = meshes.column(16)
    model["params"] = [dict(p, rayleigh=(0.4, 2e-4, 0.0, 0.0)) for p in model["params"]]
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, dynamic=True)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    h, nsteps = 0.004, 20
    ag = 2.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 3)
    want, got = run_both(ctx, om, batch, pm.neq, h, nsteps, dict(elastic=op), ag, test=2, tol=1e-12)
    for g, w, tol in zip(got, want, (1e-9, 1e-8, 1e-8)):
        assert rel(g, w) <= tol


def test_pimy_steps_retried_and_substepped(ctx):
    """Recovery (implicit_stage.cpp:385-402) on a PIMY column with Newton capped at 2 iterations:
    steps that need 3 at full size fail, are retried once and split into halves (quarters where
    a half fails), the excitation taken from its time series at every attempt's end time. The
    same commits, retries, substeps, failed attempts and time as the oracle; the state to the
    plastic-column tolerance."""
    from tests.test_oracle_newmark import _pimy_column_case
    model, om, per_set, pe, h, nsteps, series = _pimy_column_case()
    pm = core.Mesh(model)
    keep = []
    # per_set follows the oracle mesh's parameter groups, which the product mesh shares
    assert list(pm.param_group) == list(om.param_group)
    dev = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"],
                              nsurf=20, b=m["b"], mass_type=m["mass_type"], keep=keep) for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev)
    batch.set_stage(1)
    n = pm.neq
    kw = dict(test=0, tol=1e-5, max_iter=2, max_retries=1, num_sublevels=2, num_substeps=2)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, h, nsteps, pimy=pe, pimy_state=[Cs, Ts], ground_dir=ground_dir(om, 0),
                        series=series, **kw)
    assert rc == 0 and st["substepped"] > 0
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for x in d:
        x.zero()
    r = batch.newmark(*d, h, nsteps, ground_dir=0, series=series, **kw)
    assert r["failed"] == 0 and r["steps"] == nsteps, r
    for key in ("commits", "retries", "substepped", "failed_attempts", "newton_iterations", "max_newton_iterations"):
        assert r[key] == st[key], (key, r, st)
    assert r["time"] == st["time"]
    for g, w, tol in zip((x.get() for x in d[:3]), (u, v, a), (1e-8, 1e-7, 1e-7)):
        assert rel(g, w) <= tol
    assert history_error(batch.pimy_state(True), Cs) <= 1e-7


def test_pimy_steps_under_a_time_array_series(ctx):
    """A PathTimeSeries excitation (ground_motion.cpp:43-74; times off the step grid) on the
    recovering PIMY column: a failed attempt restores the time, so the series' interval walk
    goes back as well as forward. The same commits, retries, substeps and time as the oracle,
    the state to the plastic-column tolerance, and the same final interval in loc."""
    from tests.test_oracle_newmark import _pimy_column_case
    model, om, per_set, pe, h, nsteps, _ = _pimy_column_case()
    tm = np.concatenate([[0.0], np.cumsum(h * (0.6 + 0.5 * np.abs(np.sin(np.arange(1, 3 * nsteps)))))])
    vals = 12.0 * np.sin(tm * 2 * np.pi * 2)
    pm = core.Mesh(model)
    keep = []
    dev = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"],
                              nsurf=20, b=m["b"], mass_type=m["mass_type"], keep=keep) for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev)
    batch.set_stage(1)
    n = pm.neq
    kw = dict(test=0, tol=1e-5, max_iter=2, max_retries=1, num_sublevels=2, num_substeps=2, ground_factor=0.5)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    ser_o = fxo.TimeArraySeries(vals, tm)
    rc, st = om.newmark(u, v, a, du, h, nsteps, pimy=pe, pimy_state=[Cs, Ts], ground_dir=ground_dir(om, 0),
                        series=ser_o, **kw)
    assert rc == 0 and st["substepped"] > 0 and st["failed_attempts"] > 0
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for x in d:
        x.zero()
    loc = np.zeros(1, np.int32)
    r = batch.newmark(*d, h, nsteps, ground_dir=0, series=dict(values=vals, time=tm, loc=loc), **kw)
    assert r["failed"] == 0 and r["steps"] == nsteps, r
    for key in ("commits", "retries", "substepped", "failed_attempts", "newton_iterations", "max_newton_iterations"):
        assert r[key] == st[key], (key, r, st)
    assert r["time"] == st["time"]
    assert loc[0] == ser_o.loc.value > 0
    for g, w, tol in zip((x.get() for x in d[:3]), (u, v, a), (1e-8, 1e-7, 1e-7)):
        assert rel(g, w) <= tol
    assert history_error(batch.pimy_state(True), Cs) <= 1e-7


def test_time_array_series_is_validated(ctx):
    model = meshes.column(4)
    pm = core.Mesh(model)
    batch = core.Batch(ctx, pm, pm.params_for(model), dynamic=True)
    d = [core.DeviceArray(ctx, pm.neq) for _ in range(4)]
    for x in d:
        x.zero()
    for bad in (dict(values=[0.0, 1.0], time=[0.0, np.nan]), dict(values=[0.0, 1.0, 2.0], time=[0.0, 0.2, 0.1]),
                dict(values=[0.0, 1.0], time=[0.0, 0.1], loc=np.array([2], np.int32))):
        with pytest.raises(core.FxError):
            batch.newmark(*d, 0.01, 2, ground_dir=0, series=bad)
    r = batch.newmark(*d, 0.01, 2, ground_dir=0, series=dict(values=[0.0, 1.0], time=[0.0, 0.1]))
    assert r["steps"] == 2


def test_substeps_under_excitation_need_a_series(ctx):
    model = meshes.column(4)
    pm = core.Mesh(model)
    batch = core.Batch(ctx, pm, pm.params_for(model), dynamic=True)
    d = [core.DeviceArray(ctx, pm.neq) for _ in range(4)]
    for x in d:
        x.zero()
    with pytest.raises(core.FxError):
        batch.newmark(*d, 0.01, 2, ground_dir=0, ground_accel=np.zeros(2), num_sublevels=1)
