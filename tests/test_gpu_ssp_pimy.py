"""GPU parity of SSPbrick<PressureIndependMultiYield> batches (fx_pimy_batch_create with
FX_BATCH_SSPBRICK) against the oracle, whose SSPbrick + PIMY matches OpenSees' recorders of the
reference's explicit soil-column system cases to 1e-13 (tests/test_oracle_system.py):

- the internal force with the return map at each element's point, with and without the Rayleigh
  damping velocities (every damping term), the trial history, the stored tangent's operator and
  diagonal for stiffness, damping (betaKcomm on the committed tangent), mass and mixed
  coefficients, and the ground-load vector, along a walk into the plastic range with commits and
  the stage switch, on a distorted box and a tied column;
- the three explicit SSPbrick system cases end to end (fx_newmark gravity and re-equilibration,
  fx_explicit_cd ExplicitDifference shaking): against OpenSees within the reference's GPU
  tolerances (5e-6, 5e-6, 1e-6) and against the oracle within its host tolerances.

Tolerances: forces, operator, diagonal 1e-10 of the largest entry; histories 1e-10.
"""
import numpy as np
import pytest

from femorax_amd import core
from oracle import fxo
from tests import meshes
from tests import system_cases as SC
from tests.test_gpu_pimy_batch import history_error
from tests.test_oracle_system import run_oracle_explicit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def rel(got, want):
    return np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300)


def dev_materials(sets, keep):
    return [core.pimy_material(s["rho"], s["G"], s["K"], s["cohesion"], s["gamma_peak"], s["phi"], s["pref"], s["pdc"],
                               nsurf=s["nsurf"], b=s["b"], mass_type=1, keep=keep) for s in sets]


def ssp_soil(group_count, damped):
    names = ("frictional", "clay", "frictional")
    out = []
    for g in range(group_count):
        m = dict(SC.SOILS[names[g % 3]])
        m.update(nsurf=20, b=(0.0, 0.0, -9.81 * m["rho"]), mass_type=1, element="SSPbrick",
                 rayleigh=(0.3 + 0.1 * g, 2e-3, 1e-3 * (g % 2), 5e-4) if damped else (0.0, 0.0, 0.0, 0.0))
        out.append(m)
    return out


@pytest.mark.parametrize("name", ["distorted4", "column10_tied"])
def test_ssp_pimy_force_operator_and_commits(ctx, name):
    model = meshes.distorted_box(4, layers=2) if name == "distorted4" else meshes.column(10)
    pm = core.Mesh(model)
    ngroup = int(np.max(model["group"])) + 1
    mats = ssp_soil(ngroup, damped=True)
    per_set = [mats[g] for g in pm.param_group]
    keep = []
    batch = core.Batch.pimy(ctx, pm, dev_materials(per_set, keep), rayleigh=[m["rayleigh"] for m in per_set],
                            element="SSPbrick")
    om = fxo.OracleMesh(model)
    n = pm.neq
    pe = fxo.pimy_elems(per_set, stage=0)
    Co = om.pimy_initial_state(pe)
    To = Co.copy()
    dir_ = meshes.random_vector(n, 13)
    vel = meshes.random_vector(n, 14) * 0.05
    du, dv, df = core.DeviceArray(ctx, n), core.DeviceArray.from_host(ctx, vel), core.DeviceArray(ctx, n)
    x = meshes.random_vector(n, 15)
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, n)
    for step in range(6):
        if step == 2:
            batch.set_stage(1)
            for p in pe:
                p.mat.load_stage = 1
        u = dir_ * (3e-3 * (step + 1) * (1 if step % 3 else -0.5))
        du.upload(u)
        batch.force(du, df, dv)  # a discarded trial with the damping force
        f_v = om.pimy_force(pe, Co, To, u, v=vel)
        assert rel(df.get(), f_v) <= 1e-10, (step, rel(df.get(), f_v))
        batch.store_tangents(du, df)
        f_o, D_o = om.pimy_force(pe, Co, To, u, want_D=True)
        assert rel(df.get(), f_o) <= 1e-10, step
        assert history_error(batch.pimy_state(False), To) <= 1e-10, step
        for kk in ((1.0, 0.0, 0.0, False), (1.0, 40.0, 1.6e3, False), (0.0, 0.0, 1.0, False), (1.0, 20.0, 4e2, True)):
            batch.apply(dx, dy, core.coeffs(*kk))
            want = om.apply_gp(pe, D_o, x, fxo.coeffs(*kk), committed=Co)
            assert rel(dy.get(), want) <= 1e-10, (step, kk)
            dd = core.DeviceArray(ctx, n)
            batch.diagonal(dd, core.coeffs(*kk))
            assert rel(dd.get(), om.diagonal_gp(pe, D_o, fxo.coeffs(*kk), committed=Co)) <= 1e-10, (step, kk)
        for d in range(3):
            dm = core.DeviceArray(ctx, n)
            dm.zero()
            batch.mass_times_ground(d, dm)
            assert rel(dm.get(), om.mass_times_ground(d, pimy=pe)) <= 1e-12, d
        batch.commit()
        Co, To = To.copy(), Co.copy()
        assert history_error(batch.pimy_state(True), Co) <= 1e-10
    assert np.max(Co[:, 26]) > 0, "the walk reached the plastic range"
    with pytest.raises(core.FxError):
        batch.critical_dt(1)


def run_device_explicit(ctx, case):
    m = SC.model(case)
    pm = core.Mesh(m)
    keep = []
    sets = SC.soil_of_set(case, pm.param_group)
    batch = core.Batch.pimy(ctx, pm, dev_materials(sets, keep), rayleigh=[SC.rayleigh()] * len(sets),
                            element="SSPbrick")
    n = pm.neq
    u, v, a, du = (core.DeviceArray(ctx, n) for _ in range(4))
    for x in (u, v, a, du):
        x.zero()
    batch.set_stage(0)
    r = batch.newmark(u, v, a, du, 1.0, 30, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30, time=0.0)
    assert r["failed"] == 0 and r["steps"] == 30, r
    batch.set_stage(1)
    newton = SC.EXPLICIT_CASES[case]["equilibration"] == "newton"
    r = batch.newmark(u, v, a, du, 0.01, 100, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30,
                      algorithm=0 if newton else 2, factor_once=not newton, time=r["time"])
    assert r["failed"] == 0 and r["steps"] == 100, r
    t, acc = SC.motion()
    series = dict(values=acc, time=t, factor=SC.GRAVITY * SC.AMP)
    nodes = SC.recorded_nodes(SC.EXPLICIT_CASES[case]["nz"])
    vh = core.DeviceArray.from_host(ctx, v.get())  # the element velocity starts at the committed one
    dt = SC.EXPLICIT_DT
    rows = [[], [], []]
    time = 0.0
    for _ in range(SC.shaking_steps(dt)):
        batch.explicit_cd(dt, 1, u, vh, v, a, scheme=1, ground_dir=0, series=series, time=time)
        time += dt
        for rr, x in zip(rows, (u, v, a)):
            rr.append(np.concatenate([[time], SC.record(pm.node_eq, pm.node_rank, nodes, x.get())]))
    return tuple(np.array(rr) for rr in rows)


@pytest.mark.parametrize("case", sorted(SC.EXPLICIT_CASES))
def test_ssp_pimy_explicit_system_case(ctx, case):
    got = run_device_explicit(ctx, case)
    want = SC.opensees(case)
    oracle, _ = run_oracle_explicit(case)
    for g, w, o in zip(got, want, oracle):
        assert g.shape == w.shape
        np.testing.assert_allclose(g[:, 0], w[:, 0], rtol=0, atol=1e-9)
        assert SC.relative_difference(g, w) <= SC.EXPLICIT_CASES[case]["tol_gpu"], SC.relative_difference(g, w)
        assert SC.relative_difference(g, o) <= max(SC.EXPLICIT_CASES[case]["tol_host"], 1e-9), \
            SC.relative_difference(g, o)
    print(case, "vs OpenSees", [f"{SC.relative_difference(g, w):.2e}" for g, w in zip(got, want)],
          "vs oracle", [f"{SC.relative_difference(g, o):.2e}" for g, o in zip(got, oracle)])
