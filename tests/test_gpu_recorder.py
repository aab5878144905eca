"""Node recorders on the device (fx_recorder_*, fed by fx_newmark and fx_explicit_cd after every
commit): the reference's system cases with OpenSees' recorders, written by the library.

- Implicit column (column4_pimy_std_implicit): the 100 shaking steps in one fx_newmark call with
  disp, vel and accel recorders (-time, precision 17) and a disp recorder with -dT 0.02. The
  files hold the same numbers, bit for bit, as the step-by-step run that reads the vectors back
  after every step; their time tokens are OpenSees' own; the values are within the case's GPU
  tolerance of OpenSees'; the -dT file is the every-step file's lines at the oracle's due
  indices (oracle/recorder.py).
- Explicit SSPbrick element (element_ssp_frictional_explicit): 1000 ExplicitDifference steps in
  one fx_explicit_cd call, recorded every step, bit for bit the step-by-step run's values.
"""
import os
import time as _time

import numpy as np
import pytest

from femorax_amd import core
from oracle import recorder as R
from tests import system_cases as SC
from tests.test_gpu_ssp_pimy import dev_materials, run_device_explicit
from tests.test_gpu_system import run_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def read_lines(path):
    with open(path) as f:
        return f.readlines()


def parse(lines):
    return np.array([[float(x) for x in ln.split()] for ln in lines])


def add_recorders(rec, pm, nodes, path, dT_extra=None):
    cols = pm.recorder_columns(nodes)
    idx = {}
    for resp in ("disp", "vel", "accel"):
        idx[resp] = rec.add_node(os.path.join(path, f"fx_{resp}.txt"), resp, cols, time=True, precision=17)
    if dT_extra is not None:
        idx["disp_dT"] = rec.add_node(os.path.join(path, "fx_disp_dT.txt"), "disp", cols, time=True, dT=dT_extra,
                                      precision=17)
    return idx


def test_implicit_column_recorded_by_the_library(ctx, tmp_path):
    case = "column4_pimy_std_implicit"
    m = SC.model(case)
    pm = core.Mesh(m)
    keep = []
    mats = [core.pimy_material(s["rho"], s["G"], s["K"], s["cohesion"], s["gamma_peak"], s["phi"], s["pref"], s["pdc"],
                               nsurf=s["nsurf"], b=s["b"], mass_type=s["mass_type"], keep=keep)
            for s in SC.soil_of_set(case, pm.param_group)]
    batch = core.Batch.pimy(ctx, pm, mats, rayleigh=[SC.rayleigh()] * len(mats))
    n = pm.neq
    u, v, a, du = (core.DeviceArray(ctx, n) for _ in range(4))
    for x in (u, v, a, du):
        x.zero()
    rec = core.Recorder(ctx)
    nsteps = SC.shaking_steps()
    try:
        idx = add_recorders(rec, pm, SC.recorded_nodes(SC.CASES[case]["nz"]), str(tmp_path), dT_extra=0.02)
        batch.set_stage(0)
        r = batch.newmark(u, v, a, du, 1.0, 30, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30, time=0.0)
        assert r["failed"] == 0 and r["steps"] == 30, r
        batch.set_stage(1)
        r = batch.newmark(u, v, a, du, 0.01, 100, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30, algorithm=2,
                          factor_once=True, time=r["time"])
        assert r["failed"] == 0 and r["steps"] == 100, r
        for i in idx.values():  # the `recorder` step after the stages that set up the model
            rec.activate(i)
        t, acc = SC.motion()
        series = dict(values=acc, time=t, factor=SC.GRAVITY * SC.AMP, loc=np.zeros(1, np.int32))
        t0 = _time.perf_counter()
        r = batch.newmark(u, v, a, du, SC.DT, nsteps, test=1, tol=1e-10, max_iter=25, ground_dir=0, series=series,
                          time=0.0, recorder=rec)
        ctx.sync()
        t_rec = _time.perf_counter() - t0
        assert r["failed"] == 0 and r["steps"] == nsteps, r
        rec.finish()
        assert rec.records == 3 * nsteps + len(R.due_indices(R.accumulated_times(SC.DT, nsteps), 0.02))
    finally:
        rec.close()
        batch.close()
    t0 = _time.perf_counter()
    ref = run_device(ctx, case)  # step by step, every vector read back
    t_get = _time.perf_counter() - t0
    want = SC.opensees(case)
    files = {k: read_lines(os.path.join(tmp_path, f"fx_{k}.txt")) for k in ("disp", "vel", "accel", "disp_dT")}
    for k, g, w in zip(("disp", "vel", "accel"), ref, want):
        lines = files[k]
        assert len(lines) == nsteps
        got = parse(lines)
        np.testing.assert_array_equal(got, g, err_msg=k)
        assert [ln.split(" ", 1)[0] for ln in lines] == [f"{x:.17g}" for x in w[:, 0]], k
        assert SC.relative_difference(got, w) <= SC.CASES[case]["tol_gpu"], k
        assert all(ln == R.line(row[0], row[1:], precision=17) for ln, row in zip(lines, got)), k
    due = R.due_indices(list(ref[0][:, 0]), 0.02)
    assert files["disp_dT"] == [files["disp"][i] for i in due]
    print(f"{nsteps} steps: one recorded fx_newmark call {t_rec * 1e3:.1f} ms (three recorders every step, one "
          f"every 0.02 s), step-by-step calls with read-backs (and the stages) {t_get * 1e3:.1f} ms")


def test_explicit_element_recorded_by_the_library(ctx, tmp_path):
    case = "element_ssp_frictional_explicit"
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
    dt = SC.EXPLICIT_DT
    nsteps = SC.shaking_steps(dt)
    rec = core.Recorder(ctx)
    try:
        idx = add_recorders(rec, pm, SC.recorded_nodes(SC.EXPLICIT_CASES[case]["nz"]), str(tmp_path))
        batch.set_stage(0)
        r = batch.newmark(u, v, a, du, 1.0, 30, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30, time=0.0)
        assert r["failed"] == 0 and r["steps"] == 30, r
        batch.set_stage(1)
        r = batch.newmark(u, v, a, du, 0.01, 100, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30,
                          algorithm=2, factor_once=True, time=r["time"])
        assert r["failed"] == 0 and r["steps"] == 100, r
        for i in idx.values():
            rec.activate(i)
        t, acc = SC.motion()
        series = dict(values=acc, time=t, factor=SC.GRAVITY * SC.AMP)
        vh = core.DeviceArray.from_host(ctx, v.get())
        batch.explicit_cd(dt, nsteps, u, vh, v, a, scheme=1, ground_dir=0, series=series, time=0.0, recorder=rec)
        rec.finish()
        assert rec.records == 3 * nsteps
    finally:
        rec.close()
        batch.close()
    ref = run_device_explicit(ctx, case)
    want = SC.opensees(case)
    for k, g, w in zip(("disp", "vel", "accel"), ref, want):
        lines = read_lines(os.path.join(tmp_path, f"fx_{k}.txt"))
        assert len(lines) == nsteps
        got = parse(lines)
        np.testing.assert_array_equal(got, g, err_msg=k)
        np.testing.assert_allclose(got[:, 0], w[:, 0], rtol=0, atol=1e-9)
        assert SC.relative_difference(got, w) <= SC.EXPLICIT_CASES[case]["tol_gpu"], k


def test_recorder_argument_errors(ctx, tmp_path):
    rec = core.Recorder(ctx)
    try:
        with pytest.raises(core.FxError):
            rec.add_node(os.path.join(tmp_path, "x.txt"), "disp", np.array([0, -2], np.int32))
        i = rec.add_node(os.path.join(tmp_path, "x.txt"), "disp", np.array([0, 5, -1], np.int32), precision=6)
        rec.activate(i)
        with pytest.raises(core.FxError):  # every recorder before the first activation
            rec.add_node(os.path.join(tmp_path, "y.txt"), "disp", np.array([0], np.int32))
        x = core.DeviceArray.from_host(ctx, np.arange(4, dtype=np.float64))
        with pytest.raises(core.FxError):  # equation 5 is past the 4-equation vectors
            rec.record(0.1, u=x)
        x6 = core.DeviceArray.from_host(ctx, np.arange(6, dtype=np.float64) + 0.5)
        rec.record(0.1, u=x6)
        rec.record(0.2, u=x6)
        rec.record(0.3, u=x6)
        rec.finish()
        assert read_lines(os.path.join(tmp_path, "x.txt")) == [f"{t} 0.5 5.5 0\n" for t in (0.1, 0.2, 0.3)]
        with pytest.raises(core.FxError):
            rec.record(0.4, u=x6)
    finally:
        rec.close()
    bad = core.Recorder(ctx)
    try:
        j = bad.add_node(os.path.join(tmp_path, "no_such_dir", "z.txt"), "disp", np.array([0], np.int32))
        bad.activate(j)
        with pytest.raises(core.FxError) as e:
            bad.finish()
        assert e.value.code == core.FX_EIO
    finally:
        bad.close()
