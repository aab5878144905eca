"""The device's implicit stage path against OpenSees and the oracle on the reference's system
cases (tests/system_cases.py): gravity, plastic re-equilibration by ModifiedNewton
-factoronce and 100 Newmark + Newton shaking steps of a stdBrick<PressureIndependMultiYield>
batch through fx_set_material_stage and fx_newmark, recorded after every shaking step. The
records must match OpenSees' node recorders within the reference's GPU tolerance for the case
(5e-5 element, 2e-5 column) and the oracle's records within the reference's host tolerance
(1e-5, 5e-6): the device's plastic steps may take other branches of the return map than the
oracle's at rounding (fused multiply-adds, the CG's sums), as the reference's own GPU backend
does (docs/deviations.md).
"""
import numpy as np
import pytest

from femorax_amd import core
from tests import system_cases as SC
from tests.test_oracle_system import run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def run_device(ctx, case):
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
    batch.set_stage(0)
    r = batch.newmark(u, v, a, du, 1.0, 30, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30, time=0.0)
    assert r["failed"] == 0 and r["steps"] == 30, r
    batch.set_stage(1)
    r = batch.newmark(u, v, a, du, 0.01, 100, gamma=0.6, beta=0.3025, test=0, tol=1e-7, max_iter=30, algorithm=2,
                      factor_once=True, time=r["time"])
    assert r["failed"] == 0 and r["steps"] == 100, r
    t, acc = SC.motion()
    series = dict(values=acc, time=t, factor=SC.GRAVITY * SC.AMP, loc=np.zeros(1, np.int32))
    nodes = SC.recorded_nodes(SC.CASES[case]["nz"])
    rows = [[], [], []]
    time = 0.0
    for _ in range(SC.shaking_steps()):
        r = batch.newmark(u, v, a, du, SC.DT, 1, test=1, tol=1e-10, max_iter=25, ground_dir=0, series=series,
                          time=time)
        assert r["failed"] == 0 and r["steps"] == 1, r
        time = r["time"]
        for rr, x in zip(rows, (u, v, a)):
            rr.append(np.concatenate([[time], SC.record(pm.node_eq, pm.node_rank, nodes, x.get())]))
    return tuple(np.array(rr) for rr in rows)


@pytest.mark.parametrize("case", sorted(SC.CASES))
def test_device_matches_opensees_and_oracle(ctx, case):
    got = run_device(ctx, case)
    want = SC.opensees(case)
    oracle, _ = run_oracle(case)
    for g, w, o in zip(got, want, oracle):
        assert g.shape == w.shape
        np.testing.assert_allclose(g[:, 0], w[:, 0], rtol=0, atol=1e-12)
        assert SC.relative_difference(g, w) <= SC.CASES[case]["tol_gpu"], (SC.relative_difference(g, w))
        assert SC.relative_difference(g, o) <= SC.CASES[case]["tol_host"], SC.relative_difference(g, o)
    print(case, "vs OpenSees", [f"{SC.relative_difference(g, w):.2e}" for g, w in zip(got, want)],
          "vs oracle", [f"{SC.relative_difference(g, o):.2e}" for g, o in zip(got, oracle)])
