"""Profiling driver: the operator kernels on the 128^3 layered box plus a plain copy of known size
(the calibration of the HBM traffic counters). Run under rocprofv3, e.g.
    rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv -d OUT -- python femorax_amd/profile_operator.py
The copy moves COPY_BYTES bytes (half read, half written)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from femorax_amd import core  # noqa: E402
from femorax_amd import synthetic as S  # noqa: E402

COPY_DOUBLES = 32 * 1024 * 1024  # 256 MiB read + 256 MiB written


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 128
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    model = S.config_c2(n=n)
    mesh = core.Mesh(model)
    ctx = core.Context(0)
    batch = core.Batch(ctx, mesh, mesh.params_for(model), fp32=True)
    print("batch", batch.info(), flush=True)
    x = np.sin(0.001 * np.arange(mesh.neq))
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, mesh.neq)
    dx32, dy32 = core.DeviceArray.from_host(ctx, x.astype(np.float32)), core.DeviceArray(ctx, mesh.neq, np.float32)
    a, b = core.DeviceArray(ctx, COPY_DOUBLES), core.DeviceArray(ctx, COPY_DOUBLES)
    a.zero()
    for _ in range(reps):
        core.copy(ctx, a, b)
        batch.apply(dx, dy)
        batch.apply_f32(dx32, dy32)
    ctx.sync()
    print("done", flush=True)


if __name__ == "__main__":
    main()
