"""fx_dot_static, the dot that allocates nothing (a fixed device scratch array of the library),
against fx_dot (scratch from the context, allocated on demand) through the C ABI: the same bits
on two calls in a row, the second smaller than the first (the shared scratch keeps the first
call's partials, which the second must not read), n = 0, and the size limit."""
import ctypes as C

import numpy as np
import pytest

from femorax_amd import core

pytestmark = pytest.mark.gpu

FX_DOT_STATIC_MAX_N = 268435456


def static_dot(ctx, x, y, out):
    rc = core.lib().fx_dot_static(ctx.h, x.n, x.ptr, y.ptr, out.ptr)
    core._check(rc, ctx)
    ctx.sync()
    return out.get()[0]


def test_static_dot_matches_the_allocating_dot():
    ctx = core.Context(0)
    try:
        out = core.DeviceArray(ctx, 1)
        rng = np.random.default_rng(7)
        for n in (3_000_001, 70_000):  # the second call's partials are fewer than the first's
            print(f"fx_dot_static vs fx_dot, n = {n}", flush=True)
            xh, yh = rng.standard_normal(n), rng.standard_normal(n)
            x, y = core.DeviceArray.from_host(ctx, xh), core.DeviceArray.from_host(ctx, yh)
            want = core.dot(ctx, x, y)
            got = static_dot(ctx, x, y, out)
            assert got == want, (n, got, want)
            assert abs(got - float(np.dot(xh, yh))) <= 1e-12 * float(np.dot(np.abs(xh), np.abs(yh)))
            x.free()
            y.free()
        x = core.DeviceArray(ctx, 1)
        out.upload(np.array([5.0]))
        rc = core.lib().fx_dot_static(ctx.h, 0, x.ptr, x.ptr, out.ptr)
        core._check(rc, ctx)
        ctx.sync()
        assert out.get()[0] == 0.0
        rc = core.lib().fx_dot_static(ctx.h, FX_DOT_STATIC_MAX_N + 1, x.ptr, x.ptr, out.ptr)
        assert rc == core.FX_EINVAL
        assert b"FX_DOT_STATIC_MAX_N" in core.lib().fx_last_error(ctx.h)
    finally:
        ctx.close()
