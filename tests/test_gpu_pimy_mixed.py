"""The FP32 operator of PressureIndependMultiYield batches (an FP32 copy of the stored tangents)
and the mixed-precision solves on it: FP32 CG inner solves inside FP64 iterative refinement,
alone (fx_cg_mixed) and as the Newton solver of a Newmark stage (fx_newmark_desc.mixed).

Tolerances: the FP32 operator within 2e-6 of the largest entry of the FP64 operator on the
same (FP32-rounded) input, as the elastic FP32 operator; the mixed solves stop on the FP64 true
residual at the FP64 solve's relative tolerance, so solutions and Newmark states agree with the
FP64 runs to the plastic-column tolerances of the FP64 tests (1e-8 displacements, 1e-7
velocities, accelerations and histories).
"""
import numpy as np
import pytest

from femorax_amd import core
from oracle import fxo
from tests import meshes
from tests.test_gpu_pimy_batch import DAMP, history_error, pimy_model, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def damped_batch(ctx, name, fp32, gravity=True):
    model, mats = pimy_model(name)
    pm = core.Mesh(model)
    per_set = [dict(mats[g]) for g in pm.param_group]
    for m in per_set:
        m["rayleigh"] = DAMP
        if not gravity:
            m["b"] = (0.0, 0.0, 0.0)
    keep = []
    dev = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"],
                              nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep) for m in per_set]
    batch = core.Batch.pimy(ctx, pm, dev, rayleigh=[DAMP] * len(per_set), fp32=fp32)
    batch.set_stage(1)
    return model, pm, batch, per_set


def walk(ctx, batch, n, steps):
    """Commit a few displacement states into the plastic range; the last one's tangents stored."""
    dir_ = meshes.random_vector(n, 41)
    du = core.DeviceArray(ctx, n)
    for step in range(steps):
        du.upload(dir_ * (2.5e-3 * (step + 1) * (1 if step % 3 else -0.5)))
        batch.store_tangents(du)
        if step + 1 < steps:
            batch.commit()
    return du


COEFFS = [(1.0, 0.0, 0.0, False), (1.0, 400.0, 8e4, False), (0.0, 1.0, 0.0, False), (1.0, 400.0, 8e4, True)]


@pytest.mark.parametrize("name", ["column12_tied", "distorted3"])
def test_fp32_operator_on_stored_tangents(ctx, name):
    """The FP32 operator against the FP64 operator on the same input, for the stiffness, the
    Newmark-like combination, the damping-only coefficients (committed-tangent term included) and
    the initial-stiffness form, at a plastic state; it needs stored tangents as the FP64 one."""
    model, pm, batch, per_set = damped_batch(ctx, name, fp32=True)
    n = pm.neq
    x32 = meshes.random_vector(n, 42).astype(np.float32)
    dx32, dy32 = core.DeviceArray.from_host(ctx, x32), core.DeviceArray(ctx, n, np.float32)
    dx, dy = core.DeviceArray.from_host(ctx, x32.astype(np.float64)), core.DeviceArray(ctx, n)
    with pytest.raises(core.FxError):
        batch.apply_f32(dx32, dy32)  # no stored tangents yet
    walk(ctx, batch, n, 4)
    state = batch.pimy_state(False)
    assert np.any(state[:, 26] > 0), "the walk reached the plastic range"
    for kk in COEFFS:
        k = core.coeffs(*kk)
        batch.apply(dx, dy, k)
        batch.apply_f32(dx32, dy32, k)
        assert rel(dy32.get().astype(np.float64), dy.get()) <= 2e-6, kk


def test_fp32_operator_needs_the_flag(ctx):
    model, pm, batch, per_set = damped_batch(ctx, "distorted3", fp32=False)
    n = pm.neq
    walk(ctx, batch, n, 1)
    with pytest.raises(core.FxError):
        batch.apply_f32(core.DeviceArray(ctx, n, np.float32), core.DeviceArray(ctx, n, np.float32))
    d = [core.DeviceArray(ctx, n) for _ in range(4)]
    for x in d:
        x.zero()
    with pytest.raises(core.FxError):
        batch.newmark(*d, 0.005, 1, mixed=True)


def test_mixed_cg_on_stored_tangents(ctx):
    """fx_cg_mixed on the Newmark-like stored-tangent operator of a plastic state: the same
    solution as FP64 CG at the same relative tolerance."""
    model, pm, batch, per_set = damped_batch(ctx, "column12_tied", fp32=True)
    n = pm.neq
    walk(ctx, batch, n, 4)
    k = core.coeffs(1.0, 400.0, 8e4)
    xt = core.DeviceArray.from_host(ctx, meshes.random_vector(n, 7))
    b = core.DeviceArray(ctx, n)
    batch.apply(xt, b, k)
    x64, xm = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    x64.zero()
    xm.zero()
    r64 = batch.cg(b, x64, rtol=1e-12, k=k)
    rm = batch.cg_mixed(b, xm, rtol=1e-12, inner_rtol=1e-3, k=k)
    assert r64["converged"] and rm["converged"], (r64, rm)
    assert rel(xm.get(), x64.get()) <= 1e-8
    assert rel(xm.get(), xt.get()) <= 1e-7


def test_mixed_newmark_on_a_damped_pimy_column(ctx):
    """The Newmark stage of the damped PIMY column with mixed-precision Newton solves against the
    same stage with FP64 CG: the same steps, the state to the plastic-column tolerances."""
    n = None
    states = []
    h, nsteps = 0.005, 16
    ag = 6.0 * np.sin(np.arange(1, nsteps + 1) * h * 2 * np.pi * 2)
    for fp32 in (False, True):
        model, pm, batch, per_set = damped_batch(ctx, "column12_tied", fp32=fp32, gravity=False)
        n = pm.neq
        d = [core.DeviceArray(ctx, n) for _ in range(4)]
        for x in d:
            x.zero()
        r = batch.newmark(*d, h, nsteps, ground_dir=0, ground_accel=ag, test=0, tol=1e-3, mixed=fp32, inner_rtol=1e-3)
        assert r["failed"] == 0 and r["steps"] == nsteps, r
        states.append(([x.get() for x in d[:3]], batch.pimy_state(True), r))
    (s64, c64, r64), (sm, cm, rm) = states
    assert abs(rm["newton_iterations"] - r64["newton_iterations"]) <= 2, (rm, r64)
    for g, w, tol in zip(sm, s64, (1e-8, 1e-7, 1e-7)):
        assert rel(g, w) <= tol
    assert history_error(cm, c64) <= 1e-7
    assert np.max(c64[:, 26]) > 0
