"""The oracle's implicit stage path against OpenSees on the reference's system cases.

The oracle runs each case's whole analysis (tests/system_cases.py: gravity with the soils
elastic, the plastic re-equilibration by ModifiedNewton -factoronce, 100 steps of Newmark +
Newton shaking under the uniform excitation) and its recorded displacements, velocities and
accelerations must match OpenSees' node recorders (tests/golden/system_pimy.npz) within the
reference's host tolerance for the case. This pins, end to end, the oracle's Newmark step and
Newton loop, both convergence tests, ModifiedNewton's kept tangent, the material-stage switch,
PressureIndependMultiYield inside stdBrick, Rayleigh damping, the laminar ties and the
PathTimeSeries excitation.
"""
import numpy as np
import pytest

from oracle import fxo
from tests import system_cases as SC


def oracle_stages(case):
    """Gravity (soils elastic) and the plastic re-equilibration on the oracle: the mesh, the
    elements, the history pair and u, v, a."""
    m = SC.model(case)
    om = fxo.OracleMesh(m)
    pe = fxo.pimy_elems(SC.soil_of_set(case, om.param_group), stage=0)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    n = om.neq
    u, v, a, du = (np.zeros(n) for _ in range(4))
    kw = dict(pimy=pe, pimy_state=[Cs, Ts], nthreads=1)  # a few elements: one thread
    rc, st = om.newmark(u, v, a, du, 1.0, 30, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30, time=0.0, **kw)
    assert rc == 0 and st["steps"] == 30, st
    for p in pe:
        p.mat.load_stage = 1
    newton = SC.case_of(case).get("equilibration", "modified") == "newton"
    rc, st = om.newmark(u, v, a, du, 0.01, 100, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30,
                        algorithm=0 if newton else 2, factor_once=not newton, time=st["time"], **kw)
    assert rc == 0 and st["steps"] == 100, st
    return om, pe, Cs, Ts, u, v, a, du, kw


def run_oracle(case):
    """The case's analysis on the oracle: the recorded rows (time, disp), (time, vel), (time, accel)."""
    om, pe, Cs, Ts, u, v, a, du, kw = oracle_stages(case)
    t, acc = SC.motion()
    series = fxo.TimeArraySeries(acc, t, factor=SC.GRAVITY * SC.AMP)
    nodes = SC.recorded_nodes(SC.CASES[case]["nz"])
    rows = [[], [], []]
    time = 0.0
    for _ in range(SC.shaking_steps()):
        rc, st = om.newmark(u, v, a, du, SC.DT, 1, test=1, tol=1e-10, max_iter=25, ground_dir=0, series=series,
                            time=time, **kw)
        assert rc == 0 and st["steps"] == 1, st
        time = st["time"]
        for r, x in zip(rows, (u, v, a)):
            r.append(np.concatenate([[time], SC.record(om.node_eq, om.node_rank, nodes, x)]))
    return tuple(np.array(r) for r in rows), Cs


@pytest.mark.parametrize("case", sorted(SC.CASES))
def test_oracle_matches_opensees_system_case(case):
    (disp, vel, acc), Cs = run_oracle(case)
    want = SC.opensees(case)
    for got, w in zip((disp, vel, acc), want):
        assert got.shape == w.shape
        np.testing.assert_allclose(got[:, 0], w[:, 0], rtol=0, atol=1e-12)
    errs = [SC.relative_difference(g, w) for g, w in zip((disp, vel, acc), want)]
    assert max(errs) <= SC.CASES[case]["tol_host"], errs
    if case == "column4_pimy_std_implicit":  # the frictional layer yields (surfaces up to 12 active)
        assert np.max(Cs[:, 26]) > 0


def test_shaking_loop_steps():
    assert SC.shaking_steps() == 100


def run_oracle_explicit(case):
    """An SSPbrick case: the implicit stages, then ExplicitDifference shaking (explicit_analysis.cpp:
    60-85, 646-700): from the committed u, v, a, the predictor vh += dt a, u += dt vh, the force at u
    with the elements' damping at vh, a = R / m with R = -f - m a_g at the x equations, the recorded
    velocity vh + dt/8 (3 a(n+1) + a(n)), the history committed every step."""
    om, pe, Cs, Ts, u, v, a, du, kw = oracle_stages(case)
    groups = om.param_group
    sets = SC.soil_of_set(case, groups)
    mass_params = fxo.elastic_params([dict(E=1.0, nu=0.3, rho=s["rho"], element="SSPbrick") for s in sets])
    m, _ = om.lumped_mass(mass_params)
    xdir = np.zeros(om.neq, bool)
    for i, q in enumerate(om.node_eq):
        if q >= 0 and i % 3 == 0:
            xdir[q] = True
    t_rec, acc = SC.motion()
    series = fxo.TimeArraySeries(acc, t_rec, factor=SC.GRAVITY * SC.AMP)
    nodes = SC.recorded_nodes(SC.case_of(case)["nz"])
    dt = SC.EXPLICIT_DT
    vh = v.copy()
    rows = [[], [], []]
    time = 0.0
    for _ in range(SC.shaking_steps(dt)):
        vh += a * dt
        u += vh * dt
        time += dt
        ag = series(time)
        f = om.pimy_force(pe, Cs, Ts, u, v=vh, nthreads=1)
        R = -f + np.where(xdir, -m * ag, 0.0)
        an = R / m
        v = vh + (3.0 * an + a) * (dt * 0.125)
        a = an
        Cs[...], Ts[...] = Ts.copy(), Cs.copy()
        for r, x in zip(rows, (u, v, a)):
            r.append(np.concatenate([[time], SC.record(om.node_eq, om.node_rank, nodes, x)]))
    return tuple(np.array(r) for r in rows), Cs


@pytest.mark.parametrize("case", sorted(SC.EXPLICIT_CASES))
def test_oracle_matches_opensees_explicit_ssp_case(case):
    (disp, vel, acc), Cs = run_oracle_explicit(case)
    want = SC.opensees(case)
    errs = []
    for got, w in zip((disp, vel, acc), want):
        assert got.shape == w.shape, (got.shape, w.shape)
        np.testing.assert_allclose(got[:, 0], w[:, 0], rtol=0, atol=1e-12)
        errs.append(SC.relative_difference(got, w))
    assert max(errs) <= SC.EXPLICIT_CASES[case]["tol_host"], errs
