"""GPU parity of the batched soil-column pool (fx_column_*): each column's Newmark + Newton +
PCG analysis against (a) the oracle's implicit stage on the same column model (eight Gauss
points per element, the reference's numbering and ties), (b) the per-model GPU path
(fx_newmark on a stdBrick<PressureIndependMultiYield> batch), for columns of every solve
workgroup size; the mixed-precision solve (FP32 inner PCG in FP64 refinement) against the
FP64 one; the compacted return map (predict / compact / correct) against the fused one, bit for
bit; the independence of a column's result from the other columns in the pool (bit for
bit, admitted at a different time and place); admission limits and the failure path.

Tolerances: the pool evaluates one material point per element (the strain of a tied box
element is uniform) and sums in its own order, so it agrees with the eight-point models to
rounding: displacements, velocities, accelerations and histories within 1e-10 of the largest
value of the history (measured: about 1e-15 for u, 1e-13 for a; the differences are printed).
"""
import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
from oracle import fxo
from tests.test_gpu_pimy_batch import history_error

pytestmark = pytest.mark.gpu

DT = 0.005


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def dev_mats(mats, keep):
    return [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"],
                               nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep) for m in mats]


def column_case(n, seed, nlayer=3, mass_type=1, body=(0.0, 0.0, 0.0)):
    rng = S.Lcg(seed)
    mats = S.soil_layers(rng, nlayer)
    for m in mats:
        m["mass_type"] = mass_type
        m["b"] = body
    return mats, S.column_layer_of(n, nlayer)


def column_model(n, nlayer, mat_of):
    dummy = [dict(E=1e8, nu=0.3, rho=2000.0) for _ in range(nlayer)]
    return S.soil_column(n, 1.0, params=dummy, layer_of=lambda k: mat_of[k])


def level_values(pm, x, n):
    """Equation vector of a column model -> [n, 3] values of levels 1..n (the side node tied
    to the others at each level is node 4 k)."""
    return pm.to_nodes(x)[4 * np.arange(1, n + 1)]


def run_pool(ctx, cases, nsteps, ag, max_iter=25, tol=1e-10, iters_per_call=4, pool=None, mixed=False,
             compact=False, test=1, tune=False, waves=0):
    own = pool is None
    if own:
        pool = core.ColumnPool(ctx, max_elements=sum(len(mo) for _, mo in cases) + 64, max_columns=len(cases) + 2,
                               dt=DT, max_record=10 * nsteps * (len(cases) + 1) + 16, max_iter=max_iter, tol=tol,
                               mixed=mixed, compact=compact, test=test, tune=tune, waves=waves)
    keep = []
    ids = [pool.add(dev_mats(m, keep), mo, nsteps, ag) for m, mo in cases]
    assert all(i is not None for i in ids)
    for _ in range(100000):
        if pool.run(iters_per_call) == 0:
            break
    out = [dict(pool.fetch(i), status=pool.status(i), state=pool.state(i)) for i in ids]
    if own:
        pool.close()
    return out


def rel(got, want):
    return np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300)


def compare(res, hist_u, hist_v, hist_a, u, v, a, label):
    h = res["history"]
    errs = (rel(h[:, 0:3], hist_u), rel(h[:, 3:6], hist_v), rel(h[:, 6:9], hist_a), rel(res["u"], u),
            rel(res["v"], v), rel(res["a"], a))
    print(label, "errors u v a (history), u v a (final):", ["%.2e" % e for e in errs])
    assert max(errs) <= 1e-10, errs


def oracle_column(mats, mat_of, nsteps, ag, test=1, tol=1e-10):
    """The oracle's eight-point implicit stage on the column model: the top level's u, v, a per
    step, the levels' final u, v, a and the committed histories at Gauss point 0 of each element
    (bottom to top)."""
    n = len(mat_of)
    model = column_model(n, len(mats), mat_of)
    pm = core.Mesh(model)
    om = fxo.OracleMesh(model)
    pe = fxo.pimy_elems([mats[g] for g in pm.param_group], stage=1)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    N = om.neq
    u, v, a, du = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N)
    gd = np.zeros(N)
    q = om.node_eq.reshape(-1, 3)[:, 0]
    gd[q[q >= 0]] = 1.0
    hist = []
    for k in range(nsteps):
        rc, st = om.newmark(u, v, a, du, DT, 1, pimy=pe, pimy_state=[Cs, Ts], ground_dir=gd, ground_accel=ag[k:k + 1],
                            test=test, tol=tol)
        assert rc == 0 and st["steps"] == 1
        hist.append(np.concatenate([level_values(pm, w, n)[-1] for w in (u, v, a)]))
    final = [level_values(pm, w, n) for w in (u, v, a)]
    return np.array(hist), final, Cs[8 * np.argsort(pm.elem_order)]


@pytest.mark.parametrize("n,mass_type,test", [(6, 1, 1), (9, 0, 1), (8, 1, 0)])
def test_columns_match_the_oracle(ctx, n, mass_type, test):
    """test 1: NormDispIncr 1e-10; test 0: NormUnbalance 1e-3 (N)."""
    nsteps = 16
    tol = 1e-10 if test == 1 else 1e-3
    mats, mat_of = column_case(n, 101 + n, mass_type=mass_type, body=(0.0, 0.0, -1e3) if mass_type == 0 else (0, 0, 0))
    ag = S.ground_motion(nsteps, DT, 7, pga=4.0)
    res = run_pool(ctx, [(mats, mat_of)], nsteps, ag, test=test, tol=tol)[0]
    assert res["status"]["status"] == 1 and res["status"]["steps"] == nsteps, res["status"]
    model = column_model(n, len(mats), mat_of)
    pm = core.Mesh(model)
    om = fxo.OracleMesh(model)
    per_set = [mats[g] for g in pm.param_group]
    pe = fxo.pimy_elems(per_set, stage=1)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    N = om.neq
    u, v, a, du = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N)
    gd = np.zeros(N)
    q = om.node_eq.reshape(-1, 3)[:, 0]
    gd[q[q >= 0]] = 1.0
    hu, hv, ha = [], [], []
    for k in range(nsteps):
        rc, st = om.newmark(u, v, a, du, DT, 1, pimy=pe, pimy_state=[Cs, Ts], ground_dir=gd, ground_accel=ag[k:k + 1],
                            test=test, tol=tol)
        assert rc == 0 and st["steps"] == 1
        hu.append(level_values(pm, u, n)[-1])
        hv.append(level_values(pm, v, n)[-1])
        ha.append(level_values(pm, a, n)[-1])
    compare(res, np.array(hu), np.array(hv), np.array(ha), level_values(pm, u, n), level_values(pm, v, n),
            level_values(pm, a, n), f"oracle n={n}")
    # the committed histories of the column's points against Gauss point 0 of each element
    elem_of_level = np.argsort(pm.elem_order)  # model element k -> batch element
    want = Cs[8 * elem_of_level]
    err = history_error(res["state"], want)
    print("history error", err)
    assert err <= 1e-10
    assert np.max(res["state"][:, 26]) > 0, "the record drives the column into the plastic range"


def test_a_tall_column_matches_the_oracle_over_many_plastic_steps(ctx):
    """A column of 400 elements (the top of C4's height range, the 8-wave solve workgroup) over
    80 steps of a strong record, against the oracle's eight-point stage step by step: top-level
    history, final state and committed histories within 1e-10, with points past their first
    surface in at least 50 of the steps."""
    n, nsteps = 400, 80
    mats, mat_of = column_case(n, 4242, nlayer=10)
    ag = S.ground_motion(nsteps, DT, 13, pga=6.0)
    res = run_pool(ctx, [(mats, mat_of)], nsteps, ag)[0]
    assert res["status"]["status"] == 1 and res["status"]["steps"] == nsteps, res["status"]
    model = column_model(n, len(mats), mat_of)
    pm = core.Mesh(model)
    om = fxo.OracleMesh(model)
    per_set = [mats[g] for g in pm.param_group]
    pe = fxo.pimy_elems(per_set, stage=1)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    N = om.neq
    u, v, a, du = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N)
    gd = np.zeros(N)
    q = om.node_eq.reshape(-1, 3)[:, 0]
    gd[q[q >= 0]] = 1.0
    hu, hv, ha = [], [], []
    plastic_steps = 0
    for k in range(nsteps):
        rc, st = om.newmark(u, v, a, du, DT, 1, pimy=pe, pimy_state=[Cs, Ts], ground_dir=gd, ground_accel=ag[k:k + 1],
                            test=1, tol=1e-10)
        assert rc == 0 and st["steps"] == 1
        plastic_steps += int(np.any(Cs[:, 26] > 0))
        hu.append(level_values(pm, u, n)[-1])
        hv.append(level_values(pm, v, n)[-1])
        ha.append(level_values(pm, a, n)[-1])
    compare(res, np.array(hu), np.array(hv), np.array(ha), level_values(pm, u, n), level_values(pm, v, n),
            level_values(pm, a, n), f"oracle n={n}, {nsteps} steps")
    elem_of_level = np.argsort(pm.elem_order)
    err = history_error(res["state"], Cs[8 * elem_of_level])
    print("history error", err, "plastic steps", plastic_steps)
    assert err <= 1e-10
    assert plastic_steps >= 50, plastic_steps


@pytest.mark.parametrize("n", [40, 100, 200, 330])
def test_columns_match_the_per_model_gpu_path(ctx, n):
    nsteps = 12
    mats, mat_of = column_case(n, 500 + n, nlayer=4)
    ag = S.ground_motion(nsteps, DT, 11, pga=3.0)
    res = run_pool(ctx, [(mats, mat_of)], nsteps, ag)[0]
    assert res["status"]["status"] == 1, res["status"]
    model = column_model(n, len(mats), mat_of)
    pm = core.Mesh(model)
    per_set = [mats[g] for g in pm.param_group]
    keep = []
    batch = core.Batch.pimy(ctx, pm, dev_mats(per_set, keep))
    batch.set_stage(1)
    d = [core.DeviceArray(ctx, pm.neq) for _ in range(4)]
    for x in d:
        x.zero()
    hu, hv, ha = [], [], []
    newton = 0
    for k in range(nsteps):
        r = batch.newmark(*d, DT, 1, ground_dir=0, ground_accel=ag[k:k + 1])
        assert r["failed"] == 0 and r["steps"] == 1
        newton += r["newton_iterations"]
        u, v, a = (level_values(pm, x.get(), n) for x in d[:3])
        hu.append(u[-1])
        hv.append(v[-1])
        ha.append(a[-1])
    compare(res, np.array(hu), np.array(hv), np.array(ha), u, v, a, f"gpu n={n}")
    assert abs(res["status"]["newton_iterations"] - newton) <= nsteps // 4 + 1, (res["status"], newton)


def test_a_column_does_not_depend_on_the_others(ctx):
    nsteps = 20
    ag = S.ground_motion(nsteps, DT, 13, pga=3.5)
    target = column_case(75, 901)
    alone = run_pool(ctx, [target], nsteps, ag)[0]
    others = [column_case(n, 1000 + n) for n in (30, 150, 260, 8, 64)]
    pool = core.ColumnPool(ctx, max_elements=2000, max_columns=12, dt=DT, max_record=20000)
    keep = []
    ids = [pool.add(dev_mats(m, keep), mo, nsteps, ag) for m, mo in others[:3]]
    pool.run(7)  # the target joins while the others are mid-analysis, in another slot
    pool.remove(ids[1])
    ids.append(pool.add(dev_mats(others[3][0], keep), others[3][1], nsteps, ag))
    tid = pool.add(dev_mats(target[0], keep), target[1], nsteps, ag)
    ids.append(pool.add(dev_mats(others[4][0], keep), others[4][1], nsteps, ag))
    while pool.run(3):
        pass
    mixed = dict(pool.fetch(tid), status=pool.status(tid), state=pool.state(tid))
    pool.close()
    for key in ("u", "v", "a", "history", "state"):
        np.testing.assert_array_equal(mixed[key], alone[key], err_msg=key)
    assert mixed["status"] == alone["status"]


def test_admission_limits_and_reuse(ctx):
    pool = core.ColumnPool(ctx, max_elements=100, max_columns=3, dt=DT, max_record=1000)
    keep = []
    ag = np.zeros(5)
    m, mo = column_case(80, 3)
    a = pool.add(dev_mats(m, keep), mo, 5, ag)
    assert a is not None
    m2, mo2 = column_case(30, 4)
    assert pool.add(dev_mats(m2, keep), mo2, 5, ag) is None  # 80 + 30 > 100 element slots
    pool.remove(a)
    b = pool.add(dev_mats(m2, keep), mo2, 5, ag)
    assert b is not None
    with pytest.raises(core.FxError):
        pool.add(dev_mats(m2, keep), np.full(30, 7), 5, ag)  # parameter set out of range
    with pytest.raises(core.FxError):
        big, mob = column_case(core.FX_COLUMN_MAX_ELEMS + 1, 5)
        pool.add(dev_mats(big, keep), mob, 5, ag)
    while pool.run(4):
        pass
    assert pool.finished() == [b]
    st = pool.status(b)
    assert st["status"] == 1 and st["steps"] == 5
    r = pool.fetch(b)
    assert np.all(r["u"] == 0.0)  # no load, no motion
    pool.close()


def test_a_step_that_does_not_converge_fails_and_keeps_its_start(ctx):
    nsteps = 4
    ag = S.ground_motion(nsteps, DT, 17, pga=3.0)
    case = column_case(20, 77)
    res = run_pool(ctx, [case], nsteps, ag, max_iter=1, tol=1e-300)[0]
    st = res["status"]
    assert st["status"] == 2 and st["steps"] == 0 and st["newton_iterations"] == 1, st
    assert np.all(res["u"] == 0.0) and np.all(res["v"] == 0.0) and np.all(res["a"] == 0.0)


def test_failed_steps_are_retried_and_substepped_as_the_oracle_stage(ctx):
    """Recovery inside the pool (implicit_stage.cpp:385-402) with Newton capped at 3 iterations:
    a step that needs more fails, is retried once, then split into halves, each half's ground
    acceleration the record read as a path series with a leading zero. Against the oracle's stage
    with the same recovery and that series: the same commits, failed attempts, retries and
    substeps, the state to 1e-8 (the two evaluate the series' fraction differently)."""
    n, nsteps = 10, 10
    mats, mat_of = column_case(n, 302)
    ag = S.ground_motion(nsteps, DT, 7, pga=6.0)
    kw = dict(max_retries=1, num_sublevels=3, num_substeps=2)
    pool = core.ColumnPool(ctx, max_elements=n + 64, max_columns=2, dt=DT, max_record=10 * nsteps + 16, max_iter=3,
                           tol=1e-10, test=1, **kw)
    res = run_pool(ctx, [(mats, mat_of)], nsteps, ag, pool=pool)[0]
    pool.close()
    st = res["status"]
    model = column_model(n, len(mats), mat_of)
    pm = core.Mesh(model)
    om = fxo.OracleMesh(model)
    pe = fxo.pimy_elems([mats[g] for g in pm.param_group], stage=1)
    Cs = om.pimy_initial_state(pe)
    Ts = Cs.copy()
    N = om.neq
    u, v, a, du = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N)
    gd = np.zeros(N)
    q = om.node_eq.reshape(-1, 3)[:, 0]
    gd[q[q >= 0]] = 1.0
    rc, ost = om.newmark(u, v, a, du, DT, nsteps, pimy=pe, pimy_state=[Cs, Ts], ground_dir=gd,
                         series=dict(values=np.concatenate([[0.0], ag]), dt=DT), test=1, tol=1e-10, max_iter=3, **kw)
    assert rc == 0 and ost["substepped"] > 0 and ost["retries"] > 0, ost
    assert st["status"] == 1 and st["steps"] == nsteps, st
    for key in ("commits", "failed_attempts", "retries", "substepped"):
        assert st[key] == ost[key], (key, st, ost)
    for got, want in zip((res["u"], res["v"], res["a"]), (u, v, a)):
        assert rel(got, level_values(pm, want, n)) <= 1e-8
    assert history_error(res["state"], Cs[8 * np.argsort(pm.elem_order)]) <= 1e-8


def test_mixed_precision_columns_match_fp64(ctx):
    """FP32 inner PCG on FP32 blocks inside FP64 refinement to the same relative residual: the
    error against the all-FP64 pool is stated per result (printed) and bounded by 1e-10."""
    nsteps = 16
    ag = S.ground_motion(nsteps, DT, 19, pga=3.5)
    cases = [column_case(n, 300 + n) for n in (12, 90, 180, 400)]
    f64 = run_pool(ctx, cases, nsteps, ag)
    f32 = run_pool(ctx, cases, nsteps, ag, mixed=True)
    for (m, mo), x, y in zip(cases, f64, f32):
        assert y["status"]["status"] == 1 and y["status"]["steps"] == nsteps
        errs = [rel(y[k], x[k]) for k in ("u", "v", "a", "history")] + [history_error(y["state"], x["state"])]
        print(f"n={len(mo)} mixed vs FP64 (u v a history state):", ["%.2e" % e for e in errs],
              "newton", y["status"]["newton_iterations"], x["status"]["newton_iterations"],
              "cg", y["status"]["krylov_iterations"], x["status"]["krylov_iterations"])
        assert max(errs) <= 1e-10, errs


def test_compacted_return_map_gives_the_same_bits(ctx):
    """Predict / compact / correct (compacted_kernels.hpp:85-202): the corrector redoes the whole
    update, so every result equals the fused kernel's bit for bit, with elastic and plastic
    points in the same iterations; the compacted columns also match the oracle (1e-10)."""
    nsteps = 14
    ag = S.ground_motion(nsteps, DT, 23, pga=4.0)
    cases = [column_case(n, 700 + n) for n in (7, 64, 65, 250, 257, 512)]
    fused = run_pool(ctx, cases, nsteps, ag)
    comp = run_pool(ctx, cases, nsteps, ag, compact=True)
    plastic = 0
    for x, y in zip(fused, comp):
        for key in ("u", "v", "a", "history", "state"):
            np.testing.assert_array_equal(y[key], x[key], err_msg=key)
        assert y["status"] == x["status"]
        plastic += int(np.sum(x["state"][:, 26] > 0))
    assert plastic > 0
    # and the compacted columns against the oracle's eight-point stage (two of the sizes)
    for (mats, mat_of), y in list(zip(cases, comp))[:2]:
        hist, final, state = oracle_column(mats, mat_of, nsteps, ag)
        compare(y, hist[:, 0:3], hist[:, 3:6], hist[:, 6:9], *final, f"compacted vs oracle n={len(mat_of)}")
        assert history_error(y["state"], state) <= 1e-10


@pytest.mark.parametrize("compact", [False, True])
def test_autotuned_register_budget_gives_the_same_bits(ctx, compact, tmp_path, monkeypatch):
    """FX_COLUMN_TUNE: the return maps are timed at each register budget on the first running
    columns (repeats of a pure function of the columns' state) and the fastest kept; every
    result equals the fixed-budget pool's bit for bit, with the cache empty and again with the
    choice read back from it."""
    monkeypatch.setenv("FX_TUNE_CACHE", str(tmp_path / "autotune.txt"))
    nsteps = 10
    ag = S.ground_motion(nsteps, DT, 29, pga=4.0)
    cases = [column_case(n, 900 + n) for n in (9, 70, 130)]
    fixed = run_pool(ctx, cases, nsteps, ag, compact=compact, waves=3)
    for attempt in range(2):
        tuned = run_pool(ctx, cases, nsteps, ag, compact=compact, tune=True)
        for x, y in zip(fixed, tuned):
            for key in ("u", "v", "a", "history", "state"):
                np.testing.assert_array_equal(y[key], x[key], err_msg=key)
            assert y["status"] == x["status"]
        text = (tmp_path / "autotune.txt").read_text()
        assert "column_pool" in text, text

