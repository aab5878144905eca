"""fx_cg_many: several CG + Jacobi solves advanced together on one stream (the serving loop's
admission round) give each solve the bits fx_cg gives it alone: the same iterates, counts,
residuals and x, with graph replay on and off, converged and stopped by the iteration limit.
"""
import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import synthetic as S
from tests import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def problems(ctx):
    out = []
    for model in (S.config_block(8), S.config_block(13), meshes.distorted_box(5), meshes.column(20),
                  S.config_block(17)):
        mesh = core.Mesh(model)
        batch = core.Batch(ctx, mesh, mesh.params_for(model), tune=False)
        if "nodal_load" in model:
            rhs = mesh.to_equations(model["nodal_load"])
        else:
            rhs = meshes.random_vector(mesh.neq, mesh.neq)
        out.append((batch, core.DeviceArray.from_host(ctx, rhs), core.DeviceArray(ctx, mesh.neq)))
    return out


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("max_it", [10000, 25])
def test_many_solves_give_each_solve_its_own_bits(ctx, graphs, max_it):
    ctx.set_graphs(graphs)
    try:
        ps = problems(ctx)
        alone = []
        for batch, b, x in ps:
            x.zero()
            r = batch.cg(b, x, rtol=1e-10, max_it=max_it)
            alone.append((r, x.get()))
        for _, _, x in ps:
            x.zero()
        rs = core.cg_many(ctx, [p[0] for p in ps], [p[1] for p in ps], [p[2] for p in ps], rtol=1e-10, max_it=max_it)
        for (r0, x0), r1, (_, _, x) in zip(alone, rs, ps):
            assert r1 == r0
            np.testing.assert_array_equal(x.get(), x0)
            if max_it == 25:  # the small systems converge within the limit, the others stop at it
                assert r1["converged"] or r1["iterations"] == 25
            else:
                assert r1["converged"]
        if max_it == 25:
            assert sum(not r["converged"] for r in rs) >= 2
        # a different subset and order in the mix changes nothing
        sub = [ps[3], ps[0], ps[4]]
        for _, _, x in sub:
            x.zero()
        rs = core.cg_many(ctx, [p[0] for p in sub], [p[1] for p in sub], [p[2] for p in sub], rtol=1e-10,
                          max_it=max_it)
        for i, j in enumerate((3, 0, 4)):
            assert rs[i] == alone[j][0]
            np.testing.assert_array_equal(sub[i][2].get(), alone[j][1])
    finally:
        ctx.set_graphs(False)


def test_many_refuses_a_batch_twice(ctx):
    ps = problems(ctx)[:2]
    with pytest.raises(core.FxError, match="twice"):
        core.cg_many(ctx, [ps[0][0], ps[0][0]], [ps[0][1], ps[1][1]], [ps[0][2], ps[1][2]])
