"""The serving loop (femorax_amd/serve.py): a short Poisson stream of soil columns and static
blocks served with continuous batching under a small HBM budget (so columns queue for pool
room and block batches are evicted), every request completed, and each result equal, bit for
bit, to the same request solved alone: a column in a pool of its own, a block on a fresh batch
(the stated tolerance is zero; both paths run the same kernels with per-request reductions).
"""
import numpy as np
import pytest

from femorax_amd import core
from femorax_amd import serve
from femorax_amd import synthetic as S

pytestmark = pytest.mark.gpu

DT = 0.005


@pytest.mark.parametrize("rounds", [False, True], ids=["per_request", "rounds"])
def test_each_result_is_the_same_alone_and_in_the_mix(rounds):
    reqs = serve.mixed_requests(duration=0.6, rate=60.0, nsteps=20, column_sizes=(10, 90), block_sizes=(4, 10),
                                seed=4242, nrecords=3)
    kinds = {r.kind for r in reqs}
    assert kinds == {"column", "block"} and len(reqs) >= 20
    # about 300 element slots for columns: later columns wait for room
    budget = 2 * 300 * serve.COLUMN_BYTES_PER_ELEMENT
    srv = serve.Server(hbm_budget=budget, column_share=0.5, block_workers=2, mixed=True, dt=DT, max_columns=8,
                       keep_results=True, max_record_steps=20, block_rounds=rounds)
    try:
        recs = srv.run(reqs)
    finally:
        srv.close()
    assert len(recs) == len(reqs)
    s = serve.summary(recs)
    print(s)
    assert s["failed"] == 0
    by_id = {r.rid: r for r in reqs}
    ctx = core.Context(0)
    try:
        problems = serve.BlockProblems()
        for rec in recs:
            r = by_id[rec["rid"]]
            got = rec["result"]
            if r.kind == "column":
                assert got["status"] == 1 and got["steps"] == r.nsteps
                pool = core.ColumnPool(ctx, max_elements=r.n, max_columns=1, dt=DT, max_record=10 * r.nsteps,
                                       mixed=True)
                cid = pool.add(core.pimy_materials(r.mats), r.mat_of, r.nsteps, r.record)
                while pool.run(4):
                    pass
                alone = pool.fetch(cid)
                st = pool.status(cid)
                pool.close()
                assert st["newton_iterations"] == got["newton_iterations"]
                for key in ("u", "v", "a", "history"):
                    np.testing.assert_array_equal(got[key], alone[key], err_msg=f"request {r.rid} {key}")
            else:
                assert got["converged"]
                model, mesh, rhs = problems.get(r.n)
                batch = core.Batch(ctx, mesh, mesh.params_for(model), fp32=True)
                b = core.DeviceArray.from_host(ctx, rhs)
                x = core.DeviceArray(ctx, mesh.neq)
                x.zero()
                if srv.workers[0].mixed:
                    res = batch.cg_mixed(b, x, rtol=1e-8, inner_rtol=srv.workers[0].inner_rtol)
                else:
                    res = batch.cg(b, x, rtol=1e-8)
                np.testing.assert_array_equal(got["x"], x.get(), err_msg=f"request {r.rid} x")
                assert res["iterations"] == got["iterations"]
                batch.close()
    finally:
        ctx.close()


def test_columns_wait_for_room_and_all_finish():
    """A burst of columns larger than the pool: admission in order of arrival as slots free."""
    reqs = serve.mixed_requests(duration=0.05, rate=400.0, column_frac=1.0, nsteps=6, column_sizes=(40, 60),
                                seed=77, nrecords=2)
    budget = 2 * 130 * serve.COLUMN_BYTES_PER_ELEMENT  # room for two columns at a time
    srv = serve.Server(hbm_budget=budget, column_share=0.5, block_workers=1, mixed=True, dt=DT, max_columns=4,
                       max_record_steps=6)
    try:
        recs = srv.run(reqs)
    finally:
        srv.close()
    assert len(recs) == len(reqs) >= 6
    assert all(r["result"]["status"] == 1 for r in recs)
    starts = [r["start"] for r in recs]
    assert starts == sorted(starts), "first come, first admitted"
    assert max(r["start"] - r["arrival"] for r in recs) > 0.0
