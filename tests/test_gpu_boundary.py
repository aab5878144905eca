"""The reference's per-batch operator boundary through the C ABI, on a model of two batches
(stdBrick<ElasticIsotropic> below, stdBrick<PressureIndependMultiYield> above, distorted
geometry), against the oracle's element functions:

- fx_run_operator with FX_OUT_SLOTS writes every element's 24 values at slot_offset + a * padded
  + e (Out::Slots, operator_kernels.hpp:138-139, assembly.hpp:28-36), for the operator, its
  diagonal and the ground-load form (unit_dir);
- InternalForce::apply_operator / operator_diagonal / mass_times_ground replayed as the reference
  runs them (internal_force.cpp:545-626) with each scatter (Gather: slots then a (batch, element,
  DOF)-ordered gather; Plain and Atomic: y zeroed, every batch adding), nodal masses added last;
- fx_batch_set_nodal_mass on a one-batch model: the operator, its diagonal, fx_mass_times_ground
  and the lumped masses with the nodal terms.
Tolerance: 1e-12 of the largest entry (element values and sums of a few of them in FP64).
"""
import numpy as np
import pytest

from femorax_amd import core
from tests.boundary_model import TwoBatchModel

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def tb():
    return TwoBatchModel()


def rel(got, want):
    return np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300)


COEFFS = [(1.0, 0.0, 0.0, False), (1.0, 250.0, 4e4, False), (1.0, 0.0, 0.0, True), (0.0, 0.0, 1.0, False)]


@pytest.mark.parametrize("kk", COEFFS)
def test_slots_hold_every_element_value(ctx, tb, kk):
    eb, pb = tb.device_batches(ctx)
    k = core.coeffs(*kk)
    x = core.DeviceArray.from_host(ctx, tb.x)
    off = [0, 24 * tb.elastic_mesh.padded]
    total = off[1] + 24 * tb.pimy_mesh.padded
    slots = core.DeviceArray(ctx, total + 7)
    for diagonal in (False, True):
        for unit_dir in ((-1, 0, 2) if not diagonal else (-1,)):
            print(f"slots: k = {kk}, diagonal = {diagonal}, unit_dir = {unit_dir}", flush=True)
            slots.upload(np.full(total + 7, 7.5))
            for which, b in enumerate((eb, pb)):
                b.run_operator(core.FX_OUT_SLOTS, x=x, slots=slots, slot_offset=off[which], diagonal=diagonal,
                               unit_dir=unit_dir, k=k)
            got = slots.get()
            for which, m in enumerate((tb.elastic_mesh, tb.pimy_mesh)):
                for e in range(m.nelem):
                    want = tb.oracle_element(which, e, k, diagonal, tb.x, unit_dir)
                    have = np.array([got[off[which] + a * m.padded + e] for a in range(24)])
                    scale = max(np.max(np.abs(want)), 1e-300)
                    if unit_dir >= 0:
                        # K r vanishes for a rigid translation up to rounding: measure against the
                        # element's own scale (its diagonal) rather than the rounding noise
                        scale = max(scale, np.max(np.abs(tb.oracle_element(which, e, k, True))))
                    assert np.max(np.abs(have - want)) <= TOL * scale, (which, e, kk, diagonal, unit_dir)
                # the padding elements' slots are not written
                pad = [got[off[which] + a * m.padded + e] for a in range(24) for e in range(m.nelem, m.padded)]
                assert all(v == 7.5 for v in pad)
            assert got[total] == 7.5
    eb.close()
    pb.close()


def apply_operator(tb, batches, scatter, k, x, y, diagonal=False, unit_dir=-1, slots=None):
    """InternalForce::apply_operator / operator_diagonal / mass_times_ground (internal_force.cpp:
    545-626) over the ABI's per-batch verb, the gather and nodal terms as the reference adds them
    (host arithmetic here stands for the reference's own kernels)."""
    ms = [tb.elastic_mesh, tb.pimy_mesh]
    off = [0, 24 * ms[0].padded]
    if scatter != core.FX_OUT_SLOTS:
        y.zero()
    for which, b in enumerate(batches):
        b.run_operator(scatter, x=x, y=y, slots=slots, slot_offset=off[which], diagonal=diagonal,
                       unit_dir=unit_dir, k=k)
    out = y.get()
    if scatter == core.FX_OUT_SLOTS:  # SlotGather::gather: (batch, element, DOF) order
        s = slots.get()
        out = np.zeros(tb.neq)
        for which, m in enumerate(ms):
            for e in range(m.nelem):
                for a in range(24):
                    q = m.eq[a * m.padded + e]
                    if q >= 0:
                        out[q] += s[off[which] + a * m.padded + e]
    if unit_dir >= 0:
        out += np.where(tb.eq_dir() == unit_dir, tb.nodal_m, 0.0)
    else:
        w = np.ones(tb.neq) if diagonal else tb.x
        if k.mass != 0.0:
            out = out + k.mass * tb.nodal_m * w
        if k.damping != 0.0:
            out = out + k.damping * tb.nodal_am * w
    return out


@pytest.mark.parametrize("scatter", [core.FX_OUT_SLOTS, core.FX_OUT_PLAIN, core.FX_OUT_ATOMIC])
def test_reference_apply_operator_sequence(ctx, tb, scatter):
    eb, pb = tb.device_batches(ctx)
    x = core.DeviceArray.from_host(ctx, tb.x)
    y = core.DeviceArray(ctx, tb.neq)
    slots = core.DeviceArray(ctx, 24 * (tb.elastic_mesh.padded + tb.pimy_mesh.padded))
    for kk in COEFFS:
        k = core.coeffs(*kk)
        for diagonal in (False, True):
            print(f"apply_operator: scatter {scatter}, k = {kk}, diagonal = {diagonal}", flush=True)
            got = apply_operator(tb, (eb, pb), scatter, k, x, y, diagonal=diagonal, slots=slots)
            want = tb.expected(k, diagonal)
            assert rel(got, want) <= TOL, (scatter, kk, diagonal, rel(got, want))
    for d in range(3):
        print(f"mass_times_ground: scatter {scatter}, direction {d}", flush=True)
        got = apply_operator(tb, (eb, pb), scatter, core.coeffs(0.0, 0.0, 1.0), None, y, unit_dir=d, slots=slots)
        want = tb.expected(core.coeffs(0.0, 0.0, 1.0), unit_dir=d)
        assert rel(got, want) <= TOL, (scatter, d, rel(got, want))
    eb.close()
    pb.close()


def test_run_operator_refusals(ctx, tb):
    """Without FX_BATCH_SLOTS an elastic batch refuses the per-element forms; bad arguments
    are refused before any launch."""
    eb, pb = tb.device_batches(ctx, slots=False)
    x = core.DeviceArray.from_host(ctx, tb.x)
    y = core.DeviceArray(ctx, tb.neq)
    slots = core.DeviceArray(ctx, 24 * tb.elastic_mesh.padded)
    with pytest.raises(core.FxError, match="FX_BATCH_SLOTS"):
        eb.run_operator(core.FX_OUT_SLOTS, x=x, slots=slots)
    with pytest.raises(core.FxError, match="FX_BATCH_SLOTS"):
        eb.run_operator(core.FX_OUT_PLAIN, y=y, unit_dir=1, k=core.coeffs(0.0, 0.0, 1.0))
    with pytest.raises(core.FxError, match="unit_dir"):
        pb.run_operator(core.FX_OUT_PLAIN, x=x, y=y, unit_dir=3)
    with pytest.raises(core.FxError, match="out must be"):
        pb.run_operator(5, x=x, y=y)
    # the fast-path adds still work without the per-element forms
    y.zero()
    eb.run_operator(core.FX_OUT_PLAIN, x=x, y=y)
    want = np.zeros(tb.neq)
    m = tb.elastic_mesh
    for e in range(m.nelem):
        ye = tb.oracle_element(0, e, core.coeffs(), False, tb.x)
        for a in range(24):
            q = m.eq[a * m.padded + e]
            if q >= 0:
                want[q] += ye[a]
    assert rel(y.get(), want) <= TOL
    eb.close()
    pb.close()


@pytest.mark.parametrize("lattice", [False, True])
def test_nodal_masses_of_a_one_batch_model(ctx, lattice):
    """fx_batch_set_nodal_mass: cM m x + cC am x in the operator (and the solvers that call it),
    cM m + cC am on its diagonal, m in fx_mass_times_ground and in the lumped masses, am v in the
    force (internal_force.cpp:366-374, 442-448, 571-572, 597-598, 619-625; batches.cpp:249-261)."""
    from femorax_amd import synthetic as S
    from oracle import fxo

    model = S.elastic_box(4, 1.0, layers=2, body=True, seed=3) if lattice else TwoBatchModel().model
    mesh = core.Mesh(model)
    params = mesh.params_for(model)
    for p in params:
        p["rayleigh"] = (0.5, 1e-3, 0.0, 0.0)
    n = mesh.neq
    rng = np.random.default_rng(11)
    m = np.where(rng.uniform(size=n) < 0.3, rng.uniform(10.0, 90.0, n), 0.0)
    am = 0.5 * m
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    b = core.Batch(ctx, mesh, params, general=not lattice, tune=False, slots=True)
    assert b.info()["lattice"] == int(lattice)
    assert b.info()["tangent_symmetry"] == core.FX_TANGENT_SPD
    b.set_nodal_mass(m, am)
    xh = np.sin(0.37 * np.arange(n))
    x = core.DeviceArray.from_host(ctx, xh)
    y = core.DeviceArray(ctx, n)
    for kk in COEFFS:
        k = core.coeffs(*kk)
        print(f"nodal masses: lattice {lattice}, k = {kk}", flush=True)
        b.apply(x, y, k)
        want = om.apply(op, xh, fxo.coeffs(*kk)) + k.mass * m * xh + k.damping * am * xh
        assert rel(y.get(), want) <= TOL, kk
        b.diagonal(y, k)
        want = om.diagonal(op, fxo.coeffs(*kk)) + k.mass * m + k.damping * am
        assert rel(y.get(), want) <= TOL, kk
    lm, la = om.lumped_mass(op)
    dm, da = core.DeviceArray(ctx, n), core.DeviceArray(ctx, n)
    b.lumped_mass(dm, da)
    assert rel(dm.get(), lm + m) <= TOL and rel(da.get(), la + am) <= TOL
    dirs = np.full(n, -1)
    for q3, q in enumerate(mesh.node_eq):
        if q >= 0:
            dirs[q] = q3 % 3
    for d in range(3):
        b.mass_times_ground(d, y)
        want = om.mass_times_ground(d, elastic=op) + np.where(dirs == d, m, 0.0)
        assert rel(y.get(), want) <= TOL, d
    # the per-element forms of this batch (lattice batches keep their connectivity for them):
    # slots gathered on the host, and the ground-load form added into y
    P = mesh.padded
    slots = core.DeviceArray(ctx, 24 * P)
    k = core.coeffs(1.0, 100.0, 3e3)
    b.run_operator(core.FX_OUT_SLOTS, x=x, slots=slots, k=k)
    s = slots.get()
    got = np.zeros(n)
    for e in range(mesh.nelem):
        for a in range(24):
            q = mesh.eq[a * P + e]
            if q >= 0:
                got[q] += s[a * P + e]
    assert rel(got, om.apply(op, xh, fxo.coeffs(1.0, 100.0, 3e3))) <= TOL
    for d in range(3):
        y.upload(np.ones(n))
        b.run_operator(core.FX_OUT_ATOMIC, y=y, unit_dir=d, k=core.coeffs(0.0, 0.0, 1.0))
        assert rel(y.get(), 1.0 + om.mass_times_ground(d, elastic=op)) <= TOL, d
    # the force: K u + C v - F_body, plus am v
    u = core.DeviceArray.from_host(ctx, 1e-3 * xh)
    v = core.DeviceArray.from_host(ctx, np.cos(0.11 * np.arange(n)))
    b.force(u, y, v)
    want = om.force(op, 1e-3 * xh, np.cos(0.11 * np.arange(n))) + am * np.cos(0.11 * np.arange(n))
    assert rel(y.get(), want) <= 1e-11
    # clearing them restores the elements-only operator
    b.set_nodal_mass(None, None)
    b.apply(x, y, core.coeffs(1.0, 0.0, 5.0))
    assert rel(y.get(), om.apply(op, xh, fxo.coeffs(1.0, 0.0, 5.0))) <= TOL
    b.close()
