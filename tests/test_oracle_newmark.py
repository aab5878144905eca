"""Pins of the oracle's implicit Newmark stage with Newton (implicit_stage.cpp) on the CPU.

- Elastic: the textbook Newmark recurrence (average acceleration, gamma 1/2, beta 1/4) with the
  dense K, M and C = alphaM M + betaK K assembled from the oracle's operator (which a densely
  assembled K pins), body force and a ground acceleration: the same displacements,
  velocities and accelerations.
- PressureIndependMultiYield: every converged step is in equilibrium, -f(u) - M a - a_g M r = 0
  to the Newton tolerance, checked with the oracle's force (pinned by the PIMY goldens and the
  element goldens), and the histories commit once per step.
"""
import numpy as np

from oracle import fxo
from tests import meshes


def ground_dir(om, d):
    r = np.zeros(om.neq)
    q = om.node_eq.reshape(-1, 3)[:, d]
    r[q[q >= 0]] = 1.0
    return r


def test_elastic_newmark_matches_the_textbook_recurrence():
    model = meshes.column(8)
    model["params"] = [dict(p, rayleigh=(0.4, 2e-4, 0.0, 0.0)) for p in model["params"]]
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    n = om.neq
    E = np.eye(n)
    K = np.column_stack([om.apply(op, e) for e in E])
    M = np.column_stack([om.apply(op, e, fxo.coeffs(0.0, 0.0, 1.0)) for e in E])
    Cd = np.column_stack([om.apply(op, e, fxo.coeffs(0.0, 1.0, 0.0)) for e in E])
    Fb = om.body_load(op)
    r = ground_dir(om, 0)
    h, nsteps, g, b = 0.004, 40, 0.5, 0.25
    ag = 2.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 3)
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, h, nsteps, elastic=op, ground_dir=r, ground_accel=ag, tol=1e-12, test=1)
    assert rc == 0 and st["steps"] == nsteps
    # textbook: M a1 + C v1 + K u1 = Fb - ag M r, with the Newmark kinematics
    U, V, A = np.zeros(n), np.zeros(n), np.zeros(n)
    Keff = K + g / (b * h) * Cd + M / (b * h * h)
    for k in range(nsteps):
        P = Fb - ag[k] * (M @ r)
        rhs = P + M @ (U / (b * h * h) + V / (b * h) + (0.5 / b - 1) * A) + \
            Cd @ (g / (b * h) * U + (g / b - 1) * V + h * (0.5 * g / b - 1) * A)
        U1 = np.linalg.solve(Keff, rhs)
        A1 = (U1 - U) / (b * h * h) - V / (b * h) - (0.5 / b - 1) * A
        V1 = V + h * ((1 - g) * A + g * A1)
        U, V, A = U1, V1, A1
    assert np.max(np.abs(u - U)) <= 1e-9 * np.max(np.abs(U))
    assert np.max(np.abs(v - V)) <= 1e-8 * np.max(np.abs(V))
    assert np.max(np.abs(a - A)) <= 1e-7 * np.max(np.abs(A))


def test_pimy_newmark_steps_end_in_equilibrium():
    model = meshes.column(10)
    om = fxo.OracleMesh(model)
    mats = [dict(rho=2000.0, G=G, K=2.17 * G, cohesion=c, gamma_peak=0.1, phi=0.0, pref=100e3, pdc=0.0,
                 b=(0.0, 0.0, 0.0), mass_type=1) for G, c in ((40e6, 30e3), (80e6, 60e3), (150e6, 90e3))]
    sets = [mats[g] for g in om.param_group]
    pe = fxo.pimy_elems(sets, stage=1)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    n = om.neq
    r = ground_dir(om, 0)
    h, nsteps = 0.005, 30
    ag = 4.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 2)
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    op_mass = fxo.elastic_params([dict(E=1.0, nu=0.3, rho=m["rho"], mass_type=1) for m in sets])
    for k in range(nsteps):
        rc, st = om.newmark(u, v, a, du, h, 1, pimy=pe, pimy_state=[Cs, Ts], ground_dir=r, ground_accel=ag[k:k + 1],
                            test=0, tol=1e-4, max_iter=25)
        assert rc == 0 and st["steps"] == 1
        # equilibrium at the committed state: the trial buffer is free, so evaluate on a copy
        T2 = Cs.copy()
        C_prev = Ts  # after the commit, Ts holds the previous committed history
        f = om.pimy_force(pe, C_prev, T2, u)
        Ma = om.apply(op_mass, a, fxo.coeffs(0.0, 0.0, 1.0))
        Mr = om.apply(op_mass, r, fxo.coeffs(0.0, 0.0, 1.0))
        res = -f - Ma - ag[k] * Mr
        assert np.linalg.norm(res) <= 1e-4 * 1.0001, (k, np.linalg.norm(res))
        np.testing.assert_array_equal(T2, Cs)  # the committed history is the state at u
    assert np.max(Cs[:, 26]) > 0, "the excitation reached the plastic range"


def test_path_series_values():
    """PathSeries::getFactor with a uniform step (ground_motion.cpp:17-33), hand-derived: linear
    interpolation between samples, 0 before the start and at or past the last sample unless
    use_last, 0 before time 0, the pattern factor on top."""
    v = [0.0, 1.0, 3.0]
    assert fxo.series_value(v, 0.5, 0.25, factor=2.0) == 1.0
    assert fxo.series_value(v, 0.5, 0.75, factor=2.0) == 4.0
    assert fxo.series_value(v, 0.5, 1.0, factor=2.0) == 0.0
    assert fxo.series_value(v, 0.5, 1.0, factor=2.0, use_last=True) == 6.0
    assert fxo.series_value(v, 0.5, 7.0, use_last=True) == 3.0
    assert fxo.series_value(v, 0.5, -0.1, use_last=True) == 0.0
    assert fxo.series_value(v, 0.5, 0.2, start=0.25) == 0.0
    assert fxo.series_value(v, 0.5, 0.5, start=0.25) == 0.5
    assert fxo.series_value(v, 0.5, 0.25, fact=3.0) == 1.5
    # samples at the step ends reproduce the per-step accelerations
    ag = [0.3, -0.7, 1.1]
    assert [fxo.series_value([0.0] + ag, 0.25, 0.25 * (k + 1)) for k in range(2)] == ag[:2]


# the series of the reference's own unit test (tests/unit/test_ground_motion.cpp:40-55)
REF_TIME_SERIES = dict(values=[0.0, 0.3, -0.2, 0.5, 0.1], time=[0.0, 0.01, 0.035, 0.05, 0.08], factor=9.81)


def _interp(t, s, fact):
    """Linear interpolation between the samples (numpy's), 0 outside the data."""
    tm, v = np.array(s["time"]), np.array(s["values"])
    return 0.0 if t < tm[0] or t > tm[-1] else fact * s["factor"] * float(np.interp(t, tm, v))


def test_path_time_series_walk_matches_linear_interpolation():
    """PathTimeSeries (ground_motion.cpp:43-74) on the reference test's series and pattern factor
    (test_ground_motion.cpp:95-110): increasing times including the data points, then decreasing
    times (the walk backwards), against numpy's interpolation, 0 before the first and after the
    last time and before time 0; a time equal to the walk's lower end gives that sample exactly."""
    tol = 1e-15 * 9.81 * 1.5
    s = fxo.TimeArraySeries(REF_TIME_SERIES["values"], REF_TIME_SERIES["time"], REF_TIME_SERIES["factor"])
    ts = sorted(list(np.arange(-0.01, 0.1, 0.0025)) + REF_TIME_SERIES["time"])
    for t in ts:
        assert abs(s(t, 1.5) - _interp(t, REF_TIME_SERIES, 1.5)) <= tol, t
    assert s.loc.value == 3  # the walk stopped in the last interval
    for t in reversed(ts):
        if t >= 0.0:
            assert abs(s(t, 1.5) - _interp(t, REF_TIME_SERIES, 1.5)) <= tol, t
    assert s.loc.value == 0
    # exact samples where the walk's lower end is hit
    s.loc.value = 2
    assert s(0.035, 1.0) == 9.81 * -0.2
    assert s.loc.value == 2
    s.loc.value = 0
    assert s(0.0, 1.0) == 0.0 and s(0.01, 2.0) == 2.0 * 9.81 * 0.3
    # outside the data, and a series that starts after time 0
    assert s(0.09, 1.0) == 0.0 and s(-0.001, 1.0) == 0.0
    late = fxo.TimeArraySeries([1.0, 2.0], [0.02, 0.04])
    assert late(0.01) == 0.0 and late(0.02) == 1.0 and abs(late(0.03) - 1.5) <= 1e-15
    one = fxo.TimeArraySeries([4.0], [0.5])
    assert one(0.4) == 0.0 and one(0.5) == 4.0 and one(0.6) == 0.0


def test_newmark_with_a_time_array_series_uses_it_at_each_step_end():
    """The elastic Newmark run with a PathTimeSeries excitation (times not on the step grid)
    equals the run with the accelerations that series gives at the accumulated step-end times
    (the same walk, evaluated ahead), bit for bit; the run leaves its walk in loc."""
    model = meshes.column(6)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    n = om.neq
    h, nsteps = 0.004, 30
    tm = np.concatenate([[0.0], np.cumsum(0.003 + 0.004 * np.abs(np.sin(np.arange(1, 60))))])
    vals = 2.0 * np.sin(tm * 2 * np.pi * 5)
    ahead = fxo.TimeArraySeries(vals, tm, 1.0)
    t, ag = 0.0, []
    for _ in range(nsteps):
        t += h
        ag.append(ahead(t, 0.7))
    states = []
    for kw in (dict(series=fxo.TimeArraySeries(vals, tm, 1.0), ground_factor=0.7), dict(ground_accel=np.array(ag))):
        u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        rc, st = om.newmark(u, v, a, du, h, nsteps, elastic=op, ground_dir=0, tol=1e-10, test=1, **kw)
        assert rc == 0 and st["steps"] == nsteps
        states.append((u, v, a))
        if "series" in kw:
            assert kw["series"].loc.value == ahead.loc.value > 0
    for x, y in zip(*states):
        np.testing.assert_array_equal(x, y)
    assert np.max(np.abs(states[0][0])) > 0


def test_energy_incr_is_half_the_last_increments_work():
    """EnergyIncr (implicit_stage.cpp:239-246): 0.5 |dU . B| with B the unbalance after the
    update; B recomputed from the oracle's force, mass and ground load at the final state."""
    model = meshes.column(8)
    model["params"] = [dict(p, rayleigh=(0.4, 2e-4, 0.0, 0.0)) for p in model["params"]]
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    n = om.neq
    r = ground_dir(om, 0)
    h, nsteps = 0.004, 6
    ag = 2.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 3)
    u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, h, nsteps, elastic=op, ground_dir=r, ground_accel=ag, test=2, tol=1e-12,
                        max_iter=10)
    assert rc == 0 and st["steps"] == nsteps
    B = -om.force(op, u, v) - om.apply(op, a, fxo.coeffs(0.0, 0.0, 1.0)) - ag[-1] * om.apply(op, r, fxo.coeffs(0.0, 0.0, 1.0))
    assert st["last_test"] == 0.5 * abs(fxo.dot(du, B, 64))
    assert st["last_test"] <= 1e-12


def _pimy_column_case():
    from tests.test_gpu_pimy_batch import soil
    model = meshes.column(12)
    om = fxo.OracleMesh(model)
    mats = soil(20261206, 3)
    per_set = [dict(mats[g], b=(0.0, 0.0, 0.0)) for g in om.param_group]
    pe = fxo.pimy_elems(per_set, stage=1)
    h, nsteps = 0.01, 12
    series = dict(values=np.concatenate([[0.0], 6.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 2)]), dt=h)
    return model, om, per_set, pe, h, nsteps, series


def test_retries_and_substeps_compose_single_attempts():
    """recover (implicit_stage.cpp:385-402) inside the oracle's stage loop equals the same
    recursion driven from here with single attempts (a one-step call without recovery is one
    attempt that restores u, v, a, the initial guess and the time when it fails): the same bits,
    commits, retries, substeps and times. The case (a PIMY column, Newton capped at 2
    iterations) has steps that need 3 iterations at full size."""
    model, om, per_set, pe, h, nsteps, series = _pimy_column_case()
    n = om.neq
    r = ground_dir(om, 0)
    kw = dict(pimy=pe, ground_dir=r, series=series, test=0, tol=1e-5, max_iter=2)
    R, L, S = 1, 2, 2
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    A = [np.zeros(n) for _ in range(4)]
    rc, st = om.newmark(*A, h, nsteps, pimy_state=[Cs, Ts], max_retries=R, num_sublevels=L, num_substeps=S, **kw)
    assert rc == 0 and st["steps"] == nsteps
    assert st["substepped"] > 0 and st["retries"] > 0 and st["commits"] > nsteps
    # the same recursion, one attempt per call
    Cb = om.pimy_initial_state(pe)
    Tb = Cb.copy()
    B = [np.zeros(n) for _ in range(4)]
    count = dict(commits=0, retries=0, substepped=0, failed_attempts=0, newton_iterations=0)
    clock = [0.0]

    def attempt(dt):
        rc1, s1 = om.newmark(*B, dt, 1, pimy_state=[Cb, Tb], time=clock[0], **kw)
        count["newton_iterations"] += s1["newton_iterations"]
        if rc1 == 0:
            count["commits"] += 1
            clock[0] = s1["time"]
            return True
        assert s1["time"] == clock[0]
        count["failed_attempts"] += 1
        return False

    def recover(dt, level):
        if attempt(dt):
            return True
        for _ in range(R):
            count["retries"] += 1
            if attempt(dt):
                return True
        if level >= L:
            return False
        count["substepped"] += 1
        return all(recover(dt / S, level + 1) for _ in range(S))

    for _ in range(nsteps):
        assert recover(h, 0)
    for x, y in zip(A, B):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(Cs, Cb)
    for key, val in count.items():
        assert st[key] == val, key
    assert st["time"] == clock[0]


def test_ground_load_is_the_mass_times_a_rigid_translation():
    """mass_times_ground (internal_force.cpp:601-626): M r with r the unit translation at every
    element DOF, fixed ones included. A consistent element mass times a rigid translation is its
    row sums rho sum_gp N_j dvol (the shape functions sum to one), so M r is the lumped mass at the
    direction's free equations; the free-DOF product misses the coupling to the fixed base."""
    model = meshes.distorted_box(3, layers=1)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    assert all(model["params"][g].get("mass_type", 0) == 0 for g in om.param_group)
    m, _ = om.lumped_mass(op)
    for d in range(3):
        r = ground_dir(om, d)
        mg = om.mass_times_ground(d, elastic=op)
        assert np.max(np.abs(mg - m * r)) <= 1e-14 * np.max(m)
        free_only = om.apply(op, r, fxo.coeffs(0.0, 0.0, 1.0))
        assert np.max(np.abs(free_only - mg)) >= 1e-2 * np.max(m)


def _dense(om, op):
    E = np.eye(om.neq)
    K = np.column_stack([om.apply(op, e) for e in E])
    M = np.column_stack([om.apply(op, e, fxo.coeffs(0.0, 0.0, 1.0)) for e in E])
    return K, M


def test_static_stage_solves_the_body_load_and_keeps_velocities():
    """A static LoadControl stage (implicit_stage.cpp:118-402, analysis Static) on an elastic
    column: K u = F_body by a dense solve (the body force does not scale with the load factor),
    Newton converging in one iteration per step, Linear and ModifiedNewton (plain, -initial,
    -factoronce) giving the same u, v and a untouched, the time advancing by the increment."""
    model = meshes.column(8)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    n = om.neq
    K, _ = _dense(om, op)
    U = np.linalg.solve(K, om.body_load(op))
    v0 = meshes.random_vector(n, 3)
    a0 = meshes.random_vector(n, 4)
    for algorithm, modified_initial, factor_once in ((0, False, False), (1, False, False), (2, False, False),
                                                     (2, True, False), (2, False, True)):
        u, v, a, du = np.zeros(n), v0.copy(), a0.copy(), np.zeros(n)
        rc, st = om.newmark(u, v, a, du, 1.0, 3, elastic=op, static=True, algorithm=algorithm,
                            modified_initial=modified_initial, factor_once=factor_once, load_increment=0.1, test=0,
                            tol=1e-6, time=2.0)
        assert rc == 0 and st["steps"] == 3 and st["commits"] == 3, (algorithm, st)
        assert np.max(np.abs(u - U)) <= 1e-9 * np.max(np.abs(U)), algorithm
        np.testing.assert_array_equal(v, v0)
        np.testing.assert_array_equal(a, a0)
        assert abs(st["time"] - 2.3) <= 1e-12
        if algorithm == 1:
            assert st["newton_iterations"] == 3  # one solve per step
        else:
            # step 1: one iteration; steps 2-3 start in equilibrium: one iteration each
            assert st["newton_iterations"] == 3 and st["max_newton_iterations"] == 1, st


def test_load_control_increment_adapts_to_the_iterations():
    """LoadControl with dlambdaMin, dlambdaMax and numIter (implicit_stage.cpp:262-269, 277-283,
    351): dlambda *= numIter / (the last attempt's iterations), clamped; hand-derived sequence for
    an elastic column whose steps take one Newton iteration each: 0.1, 0.2, 0.4, 0.5, 0.5."""
    model = meshes.column(6)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    n = om.neq
    v, a = np.zeros(n), np.zeros(n)
    # one call carries the LoadControl state: run the five steps together
    u, du = np.zeros(n), np.zeros(n)
    rc, st = om.newmark(u, v, a, du, 1.0, 5, elastic=op, static=True, load_increment=0.1, load_min_increment=0.05,
                        load_max_increment=0.5, load_num_iter=2, test=0, tol=1e-6)
    assert rc == 0 and abs(st["time"] - (0.1 + 0.2 + 0.4 + 0.5 + 0.5)) <= 1e-12, st["time"]


def test_linear_and_modified_newton_transient_match_the_textbook_recurrence():
    """Transient Linear and ModifiedNewton (plain, -initial, -factoronce) on an elastic column:
    the same textbook Newmark recurrence as Newton (the system is linear: one solve per step)."""
    model = meshes.column(8)
    model["params"] = [dict(p, rayleigh=(0.4, 2e-4, 0.0, 0.0)) for p in model["params"]]
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params([model["params"][g] for g in om.param_group])
    n = om.neq
    h, nsteps = 0.004, 12
    ag = 2.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 3)
    r = ground_dir(om, 0)
    ref = None
    for algorithm, mi, fo in ((0, False, False), (1, False, False), (2, False, False), (2, True, False),
                              (2, False, True)):
        u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        rc, st = om.newmark(u, v, a, du, h, nsteps, elastic=op, ground_dir=r, ground_accel=ag, tol=1e-12, test=1,
                            algorithm=algorithm, modified_initial=mi, factor_once=fo)
        assert rc == 0 and st["steps"] == nsteps
        if ref is None:
            ref = (u.copy(), v.copy(), a.copy())
        else:
            assert np.max(np.abs(u - ref[0])) <= 1e-9 * np.max(np.abs(ref[0])), algorithm
            assert np.max(np.abs(a - ref[2])) <= 1e-7 * np.max(np.abs(ref[2])), algorithm


def test_pimy_static_stage_by_gmres_ends_in_equilibrium():
    """A PressureIndependMultiYield column in a static stage: the operator is symmetric
    semi-definite without a mass term, so the solver is GMRES (matrix_free_operator.cpp:5-16,
    krylov.cpp:10-22); Newton, ModifiedNewton and ModifiedNewton -initial each end every step in
    equilibrium, -f(u) = 0 to the test tolerance, under the body load."""
    model = meshes.column(10)
    om = fxo.OracleMesh(model)
    mats = [dict(rho=2000.0, G=G, K=2.17 * G, cohesion=c, gamma_peak=0.1, phi=0.0, pref=100e3, pdc=0.0,
                 b=(0.0, 0.0, -2000.0 * 9.81), mass_type=1) for G, c in ((40e6, 30e3), (80e6, 60e3), (150e6, 90e3))]
    sets = [mats[g] for g in om.param_group]
    n = om.neq
    for algorithm, mi in ((0, False), (2, False), (2, True)):
        for stage in (0, 1):
            pe = fxo.pimy_elems(sets, stage=stage)
            Cs = om.pimy_initial_state(pe)
            Ts = Cs.copy()
            u, v, a, du = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
            rc, st = om.newmark(u, v, a, du, 1.0, 2, pimy=pe, pimy_state=[Cs, Ts], static=True, algorithm=algorithm,
                                modified_initial=mi, load_increment=1.0, test=0, tol=1e-3, max_iter=60)
            assert rc == 0 and st["steps"] == 2, (algorithm, mi, stage, st)
            T2 = Cs.copy()
            f = om.pimy_force(pe, Ts, T2, u)
            assert np.linalg.norm(f) <= 1e-3 * 1.0001, (algorithm, mi, stage, np.linalg.norm(f))
            assert np.max(np.abs(u)) > 0
            if stage == 1:
                assert np.max(Cs[:, 26]) > 0, "the body load reached the plastic range"
