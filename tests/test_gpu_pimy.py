"""GPU parity of the device PressureIndependMultiYield return map.

- The 34 OpenSees strain paths of the reference (tests/reference/data/nd/
  PressureIndependMultiYield_*, test_pimy_reference.cpp:97-140) replayed on the device: stress
  and tangent within 1e-12 of each file's largest value.
- 4096 points on seeded random strain walks (clay, phi = 0, and frictional, phi = 30 with
  pressure dependence), elastic stage then plastic, committed every step with discarded trials
  in between: stress, tangent and the whole trial history against the oracle within 1e-10
  (relative to each field's largest value), most values bit-identical.
- The backbone surfaces (setUpSurfaces) against the oracle's, and the reference's refusals.
"""
import ctypes as C

import numpy as np
import pytest

from femorax_amd import core
from oracle import fxo
from tests import test_oracle_pimy as T

pytestmark = pytest.mark.gpu
NS = core.PIMY_NSTATE


@pytest.fixture(scope="module")
def ctx():
    c = core.Context(0)
    yield c
    c.close()


def material_from_input(inp, nsurf, pairs, keep):
    rho, G, K, coh, gp, phi, pref, pdc = inp
    return core.pimy_material(rho, G, K, coh, gp, phi, pref, pdc, nsurf=nsurf,
                              gredu=pairs if len(pairs) else None, keep=keep)


@pytest.mark.parametrize("ref", T.pimy_files(), ids=lambda r: r.name.replace("PressureIndependMultiYield_", ""))
def test_device_pimy_replays_opensees_paths(ctx, ref):
    p, inp, pairs = T.params_from(ref)
    keep = []
    mat = material_from_input(inp, int(ref.param("noYieldSurf")), pairs, keep)
    idx = core.DeviceArray.from_host(ctx, np.zeros(1, np.int32))
    Cs, Ts = core.DeviceArray(ctx, NS), core.DeviceArray(ctx, NS)
    core.pimy_init_state(ctx, [mat], idx, Cs)
    core.pimy_init_state(ctx, [mat], idx, Ts)
    eps, sig, D = core.DeviceArray(ctx, 6), core.DeviceArray(ctx, 6), core.DeviceArray(ctx, 36)
    cst, cpre, cpre0, ceps = ref.col("stage"), ref.col("pre"), ref.col("pre0"), ref.col("eps0")
    got = []
    for row in ref.data:
        stage = [int(row[cst])]
        if row[cpre] != 0.0:
            eps.upload(row[cpre0:cpre0 + 6])
            core.pimy_points(ctx, [mat], stage, idx, Cs, Ts, eps, sig, D)
        eps.upload(row[ceps:ceps + 6])
        core.pimy_points(ctx, [mat], stage, idx, Cs, Ts, eps, sig, D)
        got.append(np.concatenate([sig.get(), D.get()]))
        Cs, Ts = Ts, Cs
    got = np.array(got)
    c = ref.col("sig0")
    want = ref.data[:, c:c + 42]
    err_s = np.max(np.abs(got[:, :6] - want[:, :6])) / np.max(np.abs(want[:, :6]))
    err_d = np.max(np.abs(got[:, 6:] - want[:, 6:])) / np.max(np.abs(want[:, 6:]))
    assert err_s <= 1e-12 and err_d <= 1e-12, (err_s, err_d)


def test_device_surfaces_match_oracle():
    L = T.lib()
    for inp in ([2.0, 6.0e4, 1.6e5, 30.0, 0.1, 0.0, 100.0, 0.0], [2.0, 6.0e4, 1.6e5, 5.0, 0.1, 30.0, 80.0, 0.5]):
        inp = np.array(inp)
        p = T.PimyParams()
        assert L.fxo_pimy_setup(inp.ctypes.data, 20, None, C.byref(p)) == 0
        size, modulus = core.pimy_surfaces(material_from_input(inp, 20, [], []))
        np.testing.assert_array_equal(size, np.array(p.size))
        np.testing.assert_array_equal(modulus, np.array(p.modulus))
    with pytest.raises(core.FxError):
        core.pimy_surfaces(core.pimy_material(2.0, 6e4, 1.6e5, 0.0, 0.1, 0.0, 100.0, 0.0))  # no cohesion, phi 0
    with pytest.raises(core.FxError):
        core.pimy_surfaces(core.pimy_material(2.0, 6e4, 1.6e5, 30.0, 0.1, 0.0, 100.0, 0.0, nsurf=25))


def test_device_pimy_random_walks_match_oracle(ctx):
    L = T.lib()
    inputs = [np.array([2.0, 6.0e4, 1.6e5, 30.0, 0.1, 0.0, 100.0, 0.0]),
              np.array([2.0, 6.0e4, 1.6e5, 5.0, 0.1, 30.0, 80.0, 0.5])]
    mats = [material_from_input(i, 20, [], []) for i in inputs]
    oparams = []
    for i in inputs:
        p = T.PimyParams()
        assert L.fxo_pimy_setup(i.ctypes.data, 20, None, C.byref(p)) == 0
        oparams.append(p)
    npts = 4096
    rng = np.random.default_rng(11)
    which = (np.arange(npts) % 2).astype(np.int32)
    dmat = core.DeviceArray.from_host(ctx, which)
    Cd, Td = core.DeviceArray(ctx, NS * npts), core.DeviceArray(ctx, NS * npts)
    core.pimy_init_state(ctx, mats, dmat, Cd)
    core.pimy_init_state(ctx, mats, dmat, Td)
    # the oracle holds the same points, point-major
    Co = np.zeros((npts, NS))
    To = np.zeros((npts, NS))
    for i in range(npts):
        L.fxo_pimy_initial_state(C.byref(oparams[which[i]]), Co[i].ctypes.data)
    np.testing.assert_array_equal(Cd.get().reshape(NS, npts).T, Co)
    deps, dsig, dD = core.DeviceArray(ctx, 6 * npts), core.DeviceArray(ctx, 6 * npts), core.DeviceArray(ctx, 36 * npts)
    eps = np.zeros((npts, 6))
    s, Dm = np.zeros(6), np.zeros(36)
    total = same = 0
    for step in range(30):
        stage = 0 if step < 5 else 1
        if step < 5:
            eps += np.array([-2e-5, -2e-5, -4e-5, 0, 0, 0])
        else:
            eps += 2e-3 * rng.uniform(-1, 1, (npts, 6)) * np.array([0.1, 0.1, 0.1, 1, 1, 1])
        for trial_only in ((True, False) if step % 7 == 3 else (False,)):
            e = eps + (1e-3 if trial_only else 0.0)
            deps.upload(np.ascontiguousarray(e.T))
            core.pimy_points(ctx, mats, [stage, stage], dmat, Cd, Td, deps, dsig, dD)
            gs = dsig.get().reshape(6, npts).T
            gD = dD.get().reshape(36, npts).T
            gT = Td.get().reshape(NS, npts).T
            os_, oD = np.zeros((npts, 6)), np.zeros((npts, 36))
            for i in range(npts):
                p = oparams[which[i]]
                p.load_stage = stage
                ei = np.ascontiguousarray(e[i])
                L.fxo_pimy_update(C.byref(p), Co[i].ctypes.data, To[i].ctypes.data, ei.ctypes.data, s.ctypes.data,
                                  Dm.ctypes.data)
                os_[i], oD[i] = s, Dm
            for got, want in ((gs, os_), (gD, oD), (gT, To)):
                scale = np.maximum(np.max(np.abs(want), axis=0), 1e-300)
                assert np.max(np.abs(got - want) / scale) <= 1e-10, step
                total += want.size
                same += int(np.sum(got == want))
        Cd, Td = Td, Cd
        Co, To = To, Co
    assert same >= 0.95 * total, same / total
