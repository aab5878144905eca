"""GPU parity of the structured fast path on rectilinear lattices (graded meshes: the plane
spacings differ along an axis; each cell takes its own spacings) against the oracle, which
computes every element from its coordinates.

Models: a box graded in z (soil layers refined toward the surface) and slightly in x, with
several materials and Rayleigh regions; the reference's tied soil column with graded element
heights (laminar ties on the lattice path). Tolerances as for the uniform lattice: one apply,
diagonal, force and the lumped vectors within 1e-12 of the largest entry, FP32 within 2e-6, the
stable step 1e-12 relative, tight CG within 1e-10.
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


def graded(coord, lo, hi, power):
    """Map the uniform planes of [lo, hi] to planes graded by power (hi fixed)."""
    t = (coord - lo) / (hi - lo)
    return lo + (hi - lo) * (1.0 - (1.0 - t) ** power)


def graded_box():
    m = S.elastic_box(12, 1.0, layers=3, seed=4, body=True)
    xyz = m["xyz"].copy()
    xyz[:, 2] = graded(xyz[:, 2], xyz[:, 2].min(), xyz[:, 2].max(), 1.6)
    xyz[:, 0] = graded(xyz[:, 0], xyz[:, 0].min(), xyz[:, 0].max(), 1.2)
    m["xyz"] = xyz
    m["region"] = m["group"].copy()
    m["rayleigh"] = [(0.4 + 0.1 * g, 4e-4, 1e-4 * (g % 2), 2e-5) for g in range(len(m["params"]))]
    return m


def graded_column():
    m = meshes.column(30)
    xyz = m["xyz"].copy()
    xyz[:, 2] = graded(xyz[:, 2], xyz[:, 2].min(), xyz[:, 2].max(), 1.8)
    m["xyz"] = xyz
    return m


MODELS = {"graded_box": graded_box, "graded_column_tied": graded_column}


def setup(ctx, name, **kw):
    model = MODELS[name]()
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, **kw)
    return model, pm, batch, fxo.OracleMesh(model), fxo.elastic_params(params)


@pytest.mark.parametrize("name", sorted(MODELS))
@pytest.mark.parametrize("depth", [4, 8])
def test_rect_apply_diagonal_match_oracle(ctx, name, depth):
    model, pm, batch, om, op = setup(ctx, name, tile_depth=depth)
    assert batch.info()["lattice"] == 1
    x = meshes.random_vector(pm.neq, 17)
    dx, dy, dd = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, pm.neq), core.DeviceArray(ctx, pm.neq)
    for kk in [(1.0, 0.0, 0.0, False), (1.0, 40.0, 1.6e3, False), (0.0, 0.0, 1.0, False), (1.0, 20.0, 4e2, True)]:
        batch.apply(dx, dy, core.coeffs(*kk))
        assert rel(dy.get(), om.apply(op, x, fxo.coeffs(*kk))) <= 1e-12, kk
        batch.diagonal(dd, core.coeffs(*kk))
        assert rel(dd.get(), om.diagonal(op, fxo.coeffs(*kk))) <= 1e-12, kk
    batch.apply(dx, dy)
    a = dy.get()
    batch.apply(dx, dy)
    np.testing.assert_array_equal(dy.get(), a)


@pytest.mark.parametrize("name", sorted(MODELS))
def test_rect_force_mass_stable_step(ctx, name):
    model, pm, batch, om, op = setup(ctx, name)
    n = pm.neq
    m, am = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    batch.lumped_mass(m, am)
    om_m, om_am = om.lumped_mass(op)
    assert rel(m.get(), om_m) <= 1e-12
    if np.any(om_am):
        assert rel(am.get(), om_am) <= 1e-12
    fb = core.DeviceArray(ctx, n)
    batch.body_load(fb)
    if np.any(om.body_load(op)):
        assert rel(fb.get(), om.body_load(op)) <= 1e-12
    u = meshes.random_vector(n, 41) * 1e-3
    v = meshes.random_vector(n, 42) * 1e-2
    du, dv, df = core.DeviceArray.from_host(ctx, u), core.DeviceArray.from_host(ctx, v), core.DeviceArray(ctx, n)
    batch.force(du, df, dv)
    assert rel(df.get(), om.force(op, u, v)) <= 1e-12
    for scheme in (0, 1):
        got, want = batch.critical_dt(scheme), om.critical_dt(op, scheme)
        assert abs(got - want) <= 1e-12 * want, (scheme, got, want)


def test_rect_fp32_and_cg(ctx):
    model, pm, batch, om, op = setup(ctx, "graded_box", fp32=True)
    x = meshes.random_vector(pm.neq, 9)
    dx32 = core.DeviceArray.from_host(ctx, x.astype(np.float32), dtype=np.float32)
    dy32 = core.DeviceArray(ctx, pm.neq, dtype=np.float32)
    batch.apply_f32(dx32, dy32)
    assert rel(dy32.get().astype(np.float64), om.apply(op, x.astype(np.float32).astype(np.float64))) <= 2e-6
    b = om.body_load(op)
    db, dxs = core.DeviceArray.from_host(ctx, b), core.DeviceArray(ctx, pm.neq)
    dxs.zero()
    r = batch.cg(db, dxs, rtol=1e-13)
    xo, ro = om.cg(op, b, rtol=1e-13, lanes=64)
    assert r["converged"] and ro["converged"]
    assert rel(dxs.get(), xo) <= 1e-10
