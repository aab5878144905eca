"""GPU parity of the HIP hex8 operator, its diagonal and the solvers against the oracle.

Tolerances (the reference's, docs/deviations.md): one operator apply or diagonal within 1e-12 of
the largest entry (test_matrix_free_operator.cpp:129, element tolerance); FP32 operator within
2e-6 of the largest entry; CG solutions within 1e-10 relative of the oracle's at a tight rtol
(test_cg.cpp:82-143). Every call goes through the C ABI (femorax_amd/core.py).
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


def setup(ctx, model, fp32=False, general=False):
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, fp32=fp32, general=general)
    om = fxo.OracleMesh(model)
    return pm, batch, om, fxo.elastic_params(params)


CASES = {
    "c1_box16": lambda: S.config_c1(),
    "c2_layered16": lambda: S.config_c2(n=16),
    "distorted6": lambda: meshes.distorted_box(6),
    "column24_tied": lambda: meshes.column(24),
}
COEFFS = [(1.0, 0.0, 0.0, False), (1.0, 40.0, 1.6e3, False), (0.0, 0.0, 1.0, False), (1.0, 20.0, 4e2, True)]


def rel_err(got, want):
    return np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300)


STRUCTURED = {"c1_box16", "c2_layered16"}
PATHS = [(n, g) for n in sorted(CASES) for g in ((False, True) if n in STRUCTURED else (True,))]


def path_id(p):
    return f"{p[0]}-{'general' if p[1] else 'lattice'}"


@pytest.mark.parametrize("case", PATHS, ids=path_id)
def test_operator_apply_matches_oracle(ctx, case):
    name, general = case
    model = CASES[name]()
    if name == "distorted6":
        model["params"] = [dict(p, rayleigh=(0.6, 5.8e-4, 1e-4, 0.0)) for p in model["params"]]
        model["region"] = model["group"].copy()
        model["rayleigh"] = [(0.6, 5.8e-4, 1e-4, 0.0), (0.9, 3e-4, 0.0, 2e-5)]
    pm, batch, om, op = setup(ctx, model, general=general)
    assert batch.info()["lattice"] == (0 if general else 1)
    x = meshes.random_vector(pm.neq, 31)
    dx = core.DeviceArray.from_host(ctx, x)
    dy = core.DeviceArray(ctx, pm.neq)
    for kk in COEFFS:
        k = core.coeffs(*kk)
        batch.apply(dx, dy, k)
        want = om.apply(op, x, fxo.coeffs(*kk))
        got = dy.get()
        assert rel_err(got, want) <= 1e-12, (kk, rel_err(got, want))


@pytest.mark.parametrize("case", PATHS, ids=path_id)
def test_operator_diagonal_matches_oracle(ctx, case):
    name, general = case
    model = CASES[name]()
    pm, batch, om, op = setup(ctx, model, general=general)
    dd = core.DeviceArray(ctx, pm.neq)
    for kk in COEFFS:
        batch.diagonal(dd, core.coeffs(*kk))
        want = om.diagonal(op, fxo.coeffs(*kk))
        assert rel_err(dd.get(), want) <= 1e-12, kk


@pytest.mark.parametrize("general", [False, True])
def test_operator_is_deterministic(ctx, general):
    model = S.config_c2(n=24)
    pm, batch, om, op = setup(ctx, model, general=general)
    x = meshes.random_vector(pm.neq, 5)
    dx = core.DeviceArray.from_host(ctx, x)
    dy = core.DeviceArray(ctx, pm.neq)
    batch.apply(dx, dy)
    a = dy.get()
    assert rel_err(a, om.apply(op, x)) <= 1e-12
    for _ in range(3):
        batch.apply(dx, dy)
        np.testing.assert_array_equal(dy.get(), a)


def test_kernel_variants_and_tile_placement_give_the_same_bits(ctx):
    """The tuned register-budget and persistent variants and the XCD-aware tile placement change
    no bit (42^3: partial tiles on the far faces)."""
    model = S.config_c2(n=42)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    a = core.Batch(ctx, pm, params)
    b = core.Batch(ctx, pm, params, xcd=True)
    x = meshes.random_vector(pm.neq, 78)
    dx = core.DeviceArray.from_host(ctx, x)
    ya, yb = core.DeviceArray(ctx, pm.neq), core.DeviceArray(ctx, pm.neq)
    outs = []
    for v in range(8):
        a.tune(v)
        a.apply(dx, ya)
        outs.append(ya.get())
    b.apply(dx, yb)
    for o in outs[1:]:
        np.testing.assert_array_equal(outs[0], o)
    # the tile order decides the internal numbering, not the sums at a node
    np.testing.assert_array_equal(yb.get(), outs[0])
    a.tune(-1)
    assert a.info()["variant"] in range(8)


@pytest.mark.parametrize("depth", [8, 16])
def test_streamed_tile_variants_give_the_same_bits(ctx, depth):
    """8-deep tiles: the held (0-7) and streamed (8-11) kernels give the same bits; 16-deep tiles
    (always streamed) the same bits for every register budget. 42^3 cuts the last tiles in z
    (42 = 5 x 8 + 2 = 2 x 16 + 10), so the later plane fetches see missing nodes."""
    model = S.config_c2(n=42)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    a = core.Batch(ctx, pm, params, tile_depth=depth, tune=False)
    x = meshes.random_vector(pm.neq, 79)
    dx = core.DeviceArray.from_host(ctx, x)
    ya = core.DeviceArray(ctx, pm.neq)
    variants = range(12) if depth == 8 else range(4)
    # the forms without and with mass terms may run different variants: 16 (m + 1) + v
    pairs = [(2, 9), (9, 4), (3, 0)] if depth == 8 else [(1, 3), (3, 0)]
    for kk in COEFFS:
        outs = []
        for v in list(variants) + [16 * (m + 1) + v for v, m in pairs]:
            a.tune(v)
            a.apply(dx, ya, core.coeffs(*kk))
            outs.append(ya.get())
        for o in outs[1:]:
            np.testing.assert_array_equal(outs[0], o)
    a.tune(16 * (pairs[0][1] + 1) + pairs[0][0])
    assert a.info()["variant"] == pairs[0][0]
    a.tune(-1)
    assert a.info()["variant"] in variants


@pytest.mark.parametrize("shape", [(4, None), (8, None), (8, 8), (16, None)],
                         ids=["tile8x8x4", "tile8x8x8", "tile8x8x8_streamed", "tile8x8x16"])
def test_lattice_fp32_diagonal_and_mass_paths(ctx, shape):
    """Lattice path at a size with many tiles (every tile shape, held and streamed): FP32 apply,
    diagonal and the mass-term kernel against the oracle."""
    depth, variant = shape
    model = S.config_c2(n=40)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    a = core.Batch(ctx, pm, params, fp32=True, tile_depth=depth)
    if variant is not None:
        a.tune(variant)
    assert a.info()["lattice"] == 1 and a.info()["tiles"] > 16
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    x = meshes.random_vector(pm.neq, 77)
    dx = core.DeviceArray.from_host(ctx, x)
    ya = core.DeviceArray(ctx, pm.neq)
    for kk in COEFFS:
        a.apply(dx, ya, core.coeffs(*kk))
        assert rel_err(ya.get(), om.apply(op, x, fxo.coeffs(*kk))) <= 1e-12
        a.diagonal(ya, core.coeffs(*kk))
        assert rel_err(ya.get(), om.diagonal(op, fxo.coeffs(*kk))) <= 1e-12
    dx32 = core.DeviceArray.from_host(ctx, x.astype(np.float32))
    za = core.DeviceArray(ctx, pm.neq, np.float32)
    a.apply_f32(dx32, za)
    assert rel_err(za.get().astype(np.float64), om.apply(op, x.astype(np.float32).astype(np.float64))) <= 2e-6


@pytest.mark.parametrize("case", [("c1_box16", False), ("c1_box16", True), ("distorted6", True),
                                  ("column24_tied", True)], ids=path_id)
def test_operator_fp32_matches_oracle(ctx, case):
    name, general = case
    model = CASES[name]()
    pm, batch, om, op = setup(ctx, model, fp32=True, general=general)
    x = meshes.random_vector(pm.neq, 17)
    dx = core.DeviceArray.from_host(ctx, x.astype(np.float32))
    dy = core.DeviceArray(ctx, pm.neq, np.float32)
    batch.apply_f32(dx, dy)
    want = om.apply(op, x.astype(np.float32).astype(np.float64))
    assert rel_err(dy.get().astype(np.float64), want) <= 2e-6


def test_dot_matches_oracle_order(ctx):
    for n in (1, 63, 2048, 2049, 100003, 5_000_000):
        x = meshes.random_vector(min(n, 100003), n)
        x = np.resize(x, n)
        y = np.cos(np.arange(n) * 0.37)
        d = core.dot(ctx, core.DeviceArray.from_host(ctx, x), core.DeviceArray.from_host(ctx, y))
        want = fxo.dot(x, y, lanes=64)
        assert abs(d - want) <= 1e-14 * np.sum(np.abs(x * y)) + 1e-300, n


def test_dot_reduction_order_bitwise(ctx):
    """With y = 1 every product is exact, so the dot is the summation order alone: the device's
    chunk, lane and level sums (the reduction's waves included) equal the oracle's VectorOps order
    bit for bit, across one, two and many first-level chunks."""
    for n in (1, 2, 63, 64, 2047, 2048, 2049, 4097, 100003, 2048 * 2048 + 5, 6_390_144):
        x = np.resize(meshes.random_vector(min(n, 100003), n + 7), n) * (1.0 + np.arange(n) % 13)
        ones = np.ones(n)
        d = core.dot(ctx, core.DeviceArray.from_host(ctx, x), core.DeviceArray.from_host(ctx, ones))
        assert d == fxo.dot(x, ones, lanes=64), n


@pytest.mark.parametrize("name", ["c1_box16", "distorted6", "column24_tied"])
def test_cg_matches_oracle(ctx, name):
    model = CASES[name]()
    pm, batch, om, op = setup(ctx, model)
    if "nodal_load" in model:
        b = pm.to_equations(model["nodal_load"])
    else:
        b = om.body_load(op)
    if np.max(np.abs(b)) == 0:
        b = meshes.random_vector(pm.neq, 3)
    db = core.DeviceArray.from_host(ctx, b)
    dx = core.DeviceArray(ctx, pm.neq)
    dx.zero()
    r = batch.cg(db, dx, rtol=1e-8)
    xo, ro = om.cg(op, b, rtol=1e-8, lanes=64)
    assert r["converged"] and ro["converged"]
    # same stop rule, same preconditioner: the iteration counts agree up to rounding at the stop
    assert abs(r["iterations"] - ro["iterations"]) <= 2, (r["iterations"], ro["iterations"])
    assert r["true_residual"] <= 1.01e-8 * r["rhs_norm"] * 1.5
    # tight solves agree to 1e-10 relative
    dx.zero()
    r2 = batch.cg(db, dx, rtol=1e-13)
    xo2, _ = om.cg(op, b, rtol=1e-13, lanes=64)
    x = dx.get()
    assert np.max(np.abs(x - xo2)) <= 1e-10 * np.max(np.abs(xo2)), np.max(np.abs(x - xo2)) / np.max(np.abs(xo2))
    # and against a dense direct solve of the assembled operator (brute force)
    if pm.neq <= 4000:
        A = np.column_stack([om.apply(op, e) for e in np.eye(pm.neq)])
        xd = np.linalg.solve(A, b)
        assert np.max(np.abs(x - xd)) <= 1e-9 * np.max(np.abs(xd))


@pytest.mark.parametrize("name", ["c1_box16", "c2_layered16"])
def test_cg_mixed_precision_meets_rtol(ctx, name):
    model = CASES[name]()
    pm, batch, om, op = setup(ctx, model, fp32=True)
    b = pm.to_equations(model["nodal_load"])
    if name == "c2_layered16":
        b = b + om.body_load(op)
    db = core.DeviceArray.from_host(ctx, b)
    dx = core.DeviceArray(ctx, pm.neq)
    dx.zero()
    r = batch.cg_mixed(db, dx, rtol=1e-8, inner_rtol=1e-3)
    assert r["converged"], r
    assert r["true_residual"] <= 1e-8 * r["rhs_norm"]
    xo, _ = om.cg(op, b, rtol=1e-12, lanes=64)
    x = dx.get()
    # error against the all-FP64 oracle: bounded by the residual tolerance times the conditioning
    err = np.max(np.abs(x - xo)) / np.max(np.abs(xo))
    assert err <= 1e-6, err


def test_graph_replayed_cg_gives_the_launched_bits(ctx):
    """CG iterations replayed from captured hipGraphs against the same solves launched kernel by
    kernel: every replay (several solves in a row, with other work and synchronisations between
    them, a cached graph reused) gives the same bits, FP64 and mixed precision."""
    model = S.config_block(12)
    pm = core.Mesh(model)
    b = core.DeviceArray.from_host(ctx, pm.to_equations(model["nodal_load"]))
    x = core.DeviceArray(ctx, pm.neq)
    batch = core.Batch(ctx, pm, pm.params_for(model), fp32=True)
    ref = {}
    ctx.set_graphs(False)  # the reference bits: the iterations launched
    for kind in ("cg", "mixed"):
        x.zero()
        r = batch.cg(b, x, rtol=1e-10) if kind == "cg" else batch.cg_mixed(b, x, rtol=1e-10, inner_rtol=1e-3)
        ref[kind] = (x.get(), r["iterations"])
    other = core.DeviceArray(ctx, pm.neq)
    ctx.set_graphs(True)
    for rep in range(4):
        for kind in ("cg", "mixed"):
            x.zero()
            r = batch.cg(b, x, rtol=1e-10) if kind == "cg" else batch.cg_mixed(b, x, rtol=1e-10, inner_rtol=1e-3)
            np.testing.assert_array_equal(x.get(), ref[kind][0], err_msg=f"{kind} replay {rep}")
            assert r["iterations"] == ref[kind][1]
            batch.apply(x, other)  # eager work between the replays
            ctx.sync()
    ctx.set_graphs(False)


def test_tile8_shape_variants_and_partial_tiles(ctx):
    """8 x 8 x 8 tiles (two passes of four cell planes): every kernel variant gives the same bits,
    the apply matches the oracle at 42^3 (partial tiles in x, y and z) and on a box with
    nz = 5 (one pass with cells, the second all empty)."""
    for model in (S.config_c2(n=42), S.elastic_box(12, 3.0, layers=2, body=True, seed=9)):
        pm = core.Mesh(model)
        params = pm.params_for(model)
        a = core.Batch(ctx, pm, params, tile8=True, tune=False)
        assert a.info()["lattice"] == 1
        om = fxo.OracleMesh(model)
        op = fxo.elastic_params(params)
        x = meshes.random_vector(pm.neq, 91)
        dx = core.DeviceArray.from_host(ctx, x)
        ya = core.DeviceArray(ctx, pm.neq)
        outs = []
        for v in range(8):
            a.tune(v)
            a.apply(dx, ya)
            outs.append(ya.get())
        for o in outs[1:]:
            np.testing.assert_array_equal(outs[0], o)
        assert rel_err(outs[0], om.apply(op, x)) <= 1e-12
        for kk in COEFFS:
            a.apply(dx, ya, core.coeffs(*kk))
            assert rel_err(ya.get(), om.apply(op, x, fxo.coeffs(*kk))) <= 1e-12
    # a 5-deep box: the second pass of every tile has no cells
    model = S.elastic_box(5, 1.0)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    a = core.Batch(ctx, pm, params, tile8=True, tune=False)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    x = meshes.random_vector(pm.neq, 92)
    dx, ya = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, pm.neq)
    a.apply(dx, ya)
    assert rel_err(ya.get(), om.apply(op, x)) <= 1e-12


@pytest.mark.parametrize("side,tile8", [(6, False), (20, False), (20, True)])
def test_reference_benchmark_column_operator(ctx, side, tile8):
    """The reference's operator-benchmark model (synthetic_models.hpp:24-94: laminar ties,
    consistent mass, two Rayleigh regions, the top one with betaKinit): the lattice path with
    tied equation sets (every side node of a level holds the level's equations), the operator
    with the benchmark's coefficients K + 2e3 C + 1.6e7 M, K alone and the diagonal, against the
    oracle, on one tile column and on several tiles per level."""
    model = S.reference_column(side)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    a = core.Batch(ctx, pm, params, tile8=tile8)
    assert a.info()["lattice"] == 1
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    kk = S.REFERENCE_COLUMN_COEFFS
    x = meshes.random_vector(pm.neq, 93)
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, pm.neq)
    for k in (kk, (1.0, 0.0, 0.0)):
        a.apply(dx, dy, core.coeffs(*k))
        assert rel_err(dy.get(), om.apply(op, x, fxo.coeffs(*k))) <= 1e-12
    a.diagonal(dy, core.coeffs(*kk))
    assert rel_err(dy.get(), om.diagonal(op, fxo.coeffs(*kk))) <= 1e-12


@pytest.mark.parametrize("tied", [False, True])
def test_reference_numbering_apply_gives_the_solver_numbering_bits(ctx, tied):
    """fx_operator_apply on a lattice batch reads x and writes y in the reference numbering in
    place (node words in that numbering); the solver-numbering apply (the one fx_cg runs) gives
    the same values at every equation: the two results are the same multiset of bits, and match
    the oracle. FP64 and FP32, a 42^3 box (cut tiles) and a tied column (tied-set sums);
    fx_run_operator's in-place add as well."""
    from oracle import fxo
    model = meshes.column(30) if tied else S.config_c2(n=42)
    pm = core.Mesh(model)
    params = pm.params_for(model)
    batch = core.Batch(ctx, pm, params, fp32=True, tune=False)
    assert batch.info()["lattice"] == 1
    x = meshes.random_vector(pm.neq, 81)
    dx, dy, di = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, pm.neq), core.DeviceArray(ctx, pm.neq)
    om = fxo.OracleMesh(model)
    op = fxo.elastic_params(params)
    for kk in COEFFS:
        k = core.coeffs(*kk)
        batch.apply(dx, dy, k)
        want = om.apply(op, x, fxo.coeffs(*kk))
        assert np.max(np.abs(dy.get() - want)) <= 1e-12 * np.max(np.abs(want)), kk
        # the solver-numbering apply of the same x (fx_operator_time runs it on the batch's own
        # numbering; x is a permutation-invariant input only for the multiset check below)
        xs = np.full(pm.neq, 0.75)
        dxs = core.DeviceArray.from_host(ctx, xs)
        batch.time_apply(dxs, di, 1, 64, k)
        batch.apply(dxs, dy, k)
        np.testing.assert_array_equal(np.sort(dy.get()), np.sort(di.get()))
        # fx_run_operator's add (y += K x in place): onto y = x, against x + the oracle's K x
        dy.upload(x)
        batch.run_operator(core.FX_OUT_PLAIN, x=dx, y=dy, k=k)
        want = x + om.apply(op, x, fxo.coeffs(*kk))
        assert np.max(np.abs(dy.get() - want)) <= 1e-12 * np.max(np.abs(want)), kk
    x32 = core.DeviceArray.from_host(ctx, x.astype(np.float32))
    y32 = core.DeviceArray(ctx, pm.neq, np.float32)
    batch.apply_f32(x32, y32)
    want = om.apply(op, x.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(y32.get() - want)) <= 2e-6 * np.max(np.abs(want))


def test_general_path_staged_and_gathered_blocks(ctx):
    """A mesh with blocks whose local nodes fit the LDS staging and blocks that gather per element
    (disjoint elements): apply, diagonal, FP32 apply, the internal force and a repeat for bits."""
    model = meshes.partly_exploded(12)
    model["region"] = model["group"].copy()
    model["rayleigh"] = [(0.6, 5.8e-4, 1e-4, 0.0), (0.9, 3e-4, 0.0, 2e-5)]
    pm, batch, om, op = setup(ctx, model, fp32=True, general=True)
    assert batch.info()["lattice"] == 0
    x = meshes.random_vector(pm.neq, 23)
    dx, dy, dd = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, pm.neq), core.DeviceArray(ctx, pm.neq)
    for kk in COEFFS:
        batch.apply(dx, dy, core.coeffs(*kk))
        assert rel_err(dy.get(), om.apply(op, x, fxo.coeffs(*kk))) <= 1e-12, kk
        batch.diagonal(dd, core.coeffs(*kk))
        assert rel_err(dd.get(), om.diagonal(op, fxo.coeffs(*kk))) <= 1e-12, kk
    batch.apply(dx, dy)
    a = dy.get()
    batch.apply(dx, dy)
    np.testing.assert_array_equal(dy.get(), a)
    dx32 = core.DeviceArray.from_host(ctx, x.astype(np.float32), dtype=np.float32)
    dy32 = core.DeviceArray(ctx, pm.neq, dtype=np.float32)
    batch.apply_f32(dx32, dy32)
    assert rel_err(dy32.get().astype(np.float64), om.apply(op, x.astype(np.float32).astype(np.float64))) <= 2e-6
    u = meshes.random_vector(pm.neq, 41) * 1e-3
    v = meshes.random_vector(pm.neq, 42) * 1e-2
    du, dv, df = core.DeviceArray.from_host(ctx, u), core.DeviceArray.from_host(ctx, v), core.DeviceArray(ctx, pm.neq)
    batch.force(du, df, dv)
    assert rel_err(df.get(), om.force(op, u, v)) <= 1e-12


@pytest.mark.parametrize("name", ["c2_layered16", "distorted6"])
def test_caller_stream_and_stream_ordered_dropin(ctx, name):
    """FX_OPT_STREAM (the context's work on a stream the caller made) and FX_OPT_STREAM_ORDERED
    (the drop-in calls return once queued): the drop-in apply and a CG solve give the bits of the
    blocking calls on the context's own stream."""
    import ctypes as C

    core.lib()  # the library and the HIP runtime it links are loaded first
    hip = C.CDLL("libamdhip64.so.7")
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    model = CASES[name]()
    pm, batch, _, _ = setup(ctx, model)
    n = pm.neq
    x = core.DeviceArray.from_host(ctx, np.sin(0.37 * np.arange(n)))
    b = core.DeviceArray.from_host(ctx, np.cos(0.11 * np.arange(n)))
    y0, y1, s0, s1 = (core.DeviceArray(ctx, n) for _ in range(4))
    try:
        batch.apply(x, y0)
        s0.zero()
        r0 = batch.cg(b, s0, rtol=1e-10)
        ctx.set_stream(stream.value)
        ctx.set_stream_ordered(True)
        for _ in range(3):
            batch.apply(x, y1)
        assert hip.hipStreamSynchronize(stream) == 0
        np.testing.assert_array_equal(y1.get(), y0.get())
        s1.zero()
        r1 = batch.cg(b, s1, rtol=1e-10)
        assert hip.hipStreamSynchronize(stream) == 0
        assert r1["iterations"] == r0["iterations"] and r0["converged"]
        np.testing.assert_array_equal(s1.get(), s0.get())
    finally:
        ctx.set_stream_ordered(False)
        ctx.set_stream(0)
        batch.close()
        hip.hipStreamDestroy(stream)
