"""Profiling drivers for single kernels, run under rocprofv3 (kernel trace or counters):

    python femorax_amd/profile_kernels.py general N REPS   general-geometry operator (hex8_block_kernel) on a
                                                             distorted N^3 layered box, FP64 and FP32 applies
    python femorax_amd/profile_kernels.py pimy SIDE REPS   stdBrick<PressureIndependMultiYield> force at a yielding
                                                             state (pimy_batch_kernel) on the SIDE^3 reference column
    python femorax_amd/profile_kernels.py pimy_time SIDE REPS   wall time per force call (every kernel of the call) on
                                                             the same column, elastic and yielding, fused and compacted
                                                             (pimy_time_damped: with the column's Rayleigh damping at
                                                             a velocity field, fused)
    python femorax_amd/profile_kernels.py general_time N REPS   wall time per apply (hex8_block_kernel and the
                                                             shared-node sum) on the distorted N^3 box, FP64 and FP32
    python femorax_amd/profile_kernels.py cg N REPS        CG + Jacobi to rtol 1e-8 on the C2 N^3 box (nodal load), the
                                                             iterations replayed from graphs: REPS solves after a warm one,
                                                             wall time per iteration (cg_eager: the same with every
                                                             kernel launched one by one, for kernel traces)
    python femorax_amd/profile_kernels.py ssp N REPS       SSPbrick operator on the N^3 layered box: the lattice path
                                                             (hex8_box_kernel<T, 3..5>) and the general path
    python femorax_amd/profile_kernels.py rect N REPS      stdBrick operator on the N^3 layered box graded in z
                                                             (rectilinear lattice, hex8_box_kernel<T, 6..8>)
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from femorax_amd import core  # noqa: E402
from femorax_amd import synthetic as S  # noqa: E402


def general(n, reps):
    model = S.distorted_box(n)
    mesh = core.Mesh(model)
    ctx = core.Context(0)
    batch = core.Batch(ctx, mesh, mesh.params_for(model), fp32=True)
    info = batch.info()
    assert info["lattice"] == 0, info
    print("batch", info, flush=True)
    x = np.sin(0.001 * np.arange(mesh.neq))
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, mesh.neq)
    dx32, dy32 = core.DeviceArray.from_host(ctx, x.astype(np.float32)), core.DeviceArray(ctx, mesh.neq, np.float32)
    ms64 = batch.time_apply(dx, dy, reps, 64) / reps
    ms32 = batch.time_apply(dx32, dy32, reps, 32) / reps
    print(f"general path {mesh.nelem} elements {mesh.neq} DoF: FP64 {ms64 * 1e3:.1f} us, FP32 {ms32 * 1e3:.1f} us "
          f"per apply", flush=True)
    batch.close()
    ctx.close()


def pimy(side, reps):
    model = S.reference_column(side)
    mesh = core.Mesh(model)
    ctx = core.Context(0)
    keep = []
    mats = _pimy_materials(model, keep)
    rays = [p["rayleigh"] for p in model["params"]]
    pb = core.Batch.pimy(ctx, mesh, mats, rayleigh=rays)
    pb.set_stage(1)
    z = np.linspace(0.0, 1.0, mesh.neq)
    du = core.DeviceArray.from_host(ctx, 2e-3 * np.sin(40.0 * z))
    df = core.DeviceArray(ctx, mesh.neq)
    for _ in range(reps):
        pb.force(du, df)
    ctx.sync()
    active = pb.pimy_state(False)[:, 26]
    print(f"pimy force on {mesh.nelem} elements, {int(np.sum(active > 0))} of {active.size} points past the "
          f"first surface", flush=True)
    pb.close()
    ctx.close()


def _pimy_materials(model, keep):
    mats = []
    for p in model["params"]:
        G = p["E"] / (2.0 * (1.0 + p["nu"]))
        K = p["E"] / (3.0 * (1.0 - 2.0 * p["nu"]))
        mats.append(core.pimy_material(p["rho"], G, K, 30.0, 0.1, 0.0, 100.0, 0.0, nsurf=20, b=p["b"], mass_type=0,
                                       keep=keep))
    return mats


def pimy_time(side, reps, damped=False):
    import time
    model = S.reference_column(side)
    mesh = core.Mesh(model)
    ctx = core.Context(0)
    keep = []
    mats = _pimy_materials(model, keep)
    rays = [p["rayleigh"] for p in model["params"]] if damped else None
    z = np.linspace(0.0, 1.0, mesh.neq)
    dv = core.DeviceArray.from_host(ctx, 1e-2 * np.cos(30.0 * z)) if damped else None
    for compact in ((False,) if damped else (False, True)):
        pb = core.Batch.pimy(ctx, mesh, mats, compact=compact, rayleigh=rays)
        pb.set_stage(1)
        du = core.DeviceArray.from_host(ctx, np.zeros(mesh.neq))
        df = core.DeviceArray(ctx, mesh.neq)
        for name, amp in (("elastic", 1e-7), ("yielding", 2e-3)):
            du.upload(amp * np.sin(40.0 * z))
            pb.force(du, df, dv)
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                pb.force(du, df, dv)
            ctx.sync()
            us = (time.perf_counter() - t0) / reps * 1e6
            print(f"pimy force, {mesh.nelem} elements, {'compacted' if compact else 'fused'}"
                  f"{', damped' if damped else ''}, {name}: {us:.1f} us", flush=True)
        pb.close()
    ctx.close()


def general_time(n, reps):
    model = S.distorted_box(n)
    mesh = core.Mesh(model)
    ctx = core.Context(0)
    batch = core.Batch(ctx, mesh, mesh.params_for(model), fp32=True)
    assert batch.info()["lattice"] == 0
    x = np.sin(0.001 * np.arange(mesh.neq))
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, mesh.neq)
    batch.time_apply(dx, dy, 3, 64)
    ms = min(batch.time_apply(dx, dy, reps, 64) / reps for _ in range(3))
    print(f"general apply FP64, {mesh.nelem} elements: {ms * 1e3:.1f} us", flush=True)
    dx32 = core.DeviceArray.from_host(ctx, x.astype(np.float32), dtype=np.float32)
    dy32 = core.DeviceArray(ctx, mesh.neq, dtype=np.float32)
    batch.time_apply(dx32, dy32, 3, 32)
    ms = min(batch.time_apply(dx32, dy32, reps, 32) / reps for _ in range(3))
    print(f"general apply FP32, {mesh.nelem} elements: {ms * 1e3:.1f} us", flush=True)
    batch.close()
    ctx.close()


def cg(n, reps, graphs=True):
    import time
    model = S.config_c2(n=n)
    mesh = core.Mesh(model)
    ctx = core.Context(0)
    batch = core.Batch(ctx, mesh, mesh.params_for(model))
    db = core.DeviceArray.from_host(ctx, mesh.to_equations(model["nodal_load"]))
    xs = core.DeviceArray(ctx, mesh.neq)
    ctx.set_graphs(graphs)
    for rep in range(reps + 1):
        xs.zero()
        ctx.sync()
        t0 = time.perf_counter()
        r = batch.cg(db, xs, rtol=1e-8, max_it=100000)
        ctx.sync()
        s = time.perf_counter() - t0
        print(f"CG {'warm' if rep == 0 else rep}: {r['iterations']} iterations, {s:.3f} s, "
              f"{s * 1e3 / r['iterations']:.4f} ms per iteration", flush=True)
    ctx.set_graphs(False)
    batch.close()
    ctx.close()


def ssp(n, reps):
    model = S.config_c2(n=n)
    mesh = core.Mesh(model)
    params = [dict(p, element="SSPbrick") for p in mesh.params_for(model)]
    ctx = core.Context(0)
    x = np.sin(0.001 * np.arange(mesh.neq))
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, mesh.neq)
    for general in (False, True):
        batch = core.Batch(ctx, mesh, params, element="SSPbrick", general=general)
        batch.time_apply(dx, dy, 3, 64)  # the first launches load the kernels' code
        ms = min(batch.time_apply(dx, dy, reps, 64) / reps for _ in range(3))
        print(f"SSPbrick {'general' if general else 'lattice'} path, {mesh.nelem} elements: {ms * 1e3:.1f} us per "
              f"apply", flush=True)
        batch.close()
    ctx.close()


def rect(n, reps):
    model = S.config_c2(n=n)
    z = model["xyz"][:, 2]
    lo, hi = z.min(), z.max()
    model["xyz"][:, 2] = lo + (hi - lo) * (1.0 - (1.0 - (z - lo) / (hi - lo)) ** 1.5)  # finer toward the top
    mesh = core.Mesh(model)
    ctx = core.Context(0)
    batch = core.Batch(ctx, mesh, mesh.params_for(model), fp32=True)
    info = batch.info()
    assert info["lattice"] == 1, info
    x = np.sin(0.001 * np.arange(mesh.neq))
    dx, dy = core.DeviceArray.from_host(ctx, x), core.DeviceArray(ctx, mesh.neq)
    batch.time_apply(dx, dy, 3, 64)  # the first launches load the kernels' code
    ms = min(batch.time_apply(dx, dy, reps, 64) / reps for _ in range(3))
    print(f"rectilinear lattice {mesh.nelem} elements: {ms * 1e3:.1f} us per apply", flush=True)
    batch.close()
    ctx.close()


if __name__ == "__main__":
    what = sys.argv[1]
    size = int(sys.argv[2]) if len(sys.argv) > 2 else (40 if what.startswith("pimy") else 128)
    reps = int(sys.argv[3]) if len(sys.argv) > 3 else 10
    {"general": general, "pimy": pimy, "ssp": ssp, "rect": rect, "pimy_time": pimy_time,
     "general_time": general_time, "cg": cg, "cg_eager": lambda n, r: cg(n, r, graphs=False),
     "pimy_time_damped": lambda n, r: pimy_time(n, r, damped=True)}[what](size, reps)
