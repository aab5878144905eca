"""Benchmark: the matrix-free hex8 operator apply on the 128^3 layered box (6,390,144 DoF).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu] [--solve]

A step is one apply y = K x of the FP64 operator over the whole mesh (inputs resident in HBM).
Prints one JSON line: throughput in GDoF/s, the roofline of the operator kernel (algorithmic
bytes per apply over its event-timed duration, against 8 TB/s), the FP32 operator beside it,
and the CPU baseline (the oracle's OpenMP port of the reference's CPU default path, timed on
this host's cores on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MF operator GDoF/s (% HBM peak); served solves/s + p99 latency at rtol 1e-8"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(neq, nel, nnode, w):
    """x read + y written, hex8 connectivity + parameter index (int32), node coordinates,
    node -> equation map (3 int32 per node)."""
    return 2 * w * neq + 4 * 8 * nel + 4 * nel + 3 * w * nnode + 12 * nnode


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hex8_operator_c2_128_hbm_traffic.md")


def measured_traffic(key):
    """HBM bytes per apply of the C2 128^3 operator from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (calibrated against a plain copy), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(key + " = "):
                return int(line.split("=")[1])
    except OSError:
        pass
    return None


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(model, params, x, budget_s=20.0):
    """The oracle (C restatement of the reference's CPU default: gather assembly, recomputed
    geometry... constant D) timed on the host cores; bounded to about budget_s seconds."""
    from oracle import fxo

    threads = cpu_threads()
    om = fxo.OracleMesh(model, coloring=False)
    op = fxo.elastic_params(params)
    om.apply(op, x, nthreads=threads)  # warm-up
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < 50:
        t0 = time.perf_counter()
        om.apply(op, x, nthreads=threads)
        t_total += time.perf_counter() - t0
        reps += 1
    per = t_total / reps
    return {"value": om.neq / per / 1e9, "unit": "GDoF/s", "cores": threads, "kind": "port",
            "sample": f"{reps} full FP64 operator applies of the same {om.neq}-DoF mesh "
                      f"(oracle/liboracle.so, -O3 -fopenmp -ffp-contract=off, {threads} threads), "
                      f"{per * 1e3:.1f} ms per apply"}


def bench_c1(core, S, ctx, reps=30):
    """C1: 16^3 box, E 30 MPa, 10 kPa top traction, CG + Jacobi to rtol 1e-8 from x0 = 0; each
    solve is one blocking fx_cg call (the latency a caller sees)."""
    model = S.config_c1()
    mesh = core.Mesh(model)
    batch = core.Batch(ctx, mesh, mesh.params_for(model))
    b = core.DeviceArray.from_host(ctx, mesh.to_equations(model["nodal_load"]))
    x = core.DeviceArray(ctx, mesh.neq)
    lat = []
    for i in range(reps + 3):
        x.zero()
        ctx.sync()
        t0 = time.perf_counter()
        r = batch.cg(b, x, rtol=1e-8)
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat[3:])
    return {"dof": mesh.neq, "iterations": r["iterations"], "solves_per_s": round(1.0 / lat.mean(), 2),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "note": "one blocking fx_cg per solve, rtol 1e-8, x0 = 0"}


def bench_c3(core, S, ctx, nsteps):
    """C3: 200 x 200 x 100 bricks (h 5 m), 10 layers, lumped mass, Lysmer dashpots, Ricker 5 Hz
    point force at the surface centre; CentralDifference at 0.5 of the stable step."""
    t0 = time.perf_counter()
    model = S.config_c3()
    mesh = core.Mesh(model)
    batch = core.Batch(ctx, mesh, mesh.params_for(model))
    print(f"[bench] C3 model {mesh.nelem} elements, {mesh.neq} DoF ready in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    n = mesh.neq
    dt = 0.5 * batch.critical_dt(0)
    dash = core.DeviceArray.from_host(ctx, mesh.to_equations(model["dashpots"]))
    src = model["source_node"]
    load_eq = np.array([mesh.node_eq[3 * mesh.node_rank[src] + 2]], np.int32)
    u, um, v, a = (core.DeviceArray(ctx, n) for _ in range(4))
    for d in (u, um, v, a):
        d.zero()
    warm = 20
    hist = S.ricker(np.arange(warm + nsteps) * dt, f=5.0, amplitude=1e6)[:, None]
    batch.explicit_cd(dt, warm, u, um, v, a, dash=dash, load_eq=load_eq, load_hist=hist[:warm].ravel())
    ctx.sync()
    t0 = time.perf_counter()
    batch.explicit_cd(dt, nsteps, u, um, v, a, dash=dash, load_eq=load_eq, load_hist=hist[warm:].ravel())
    ctx.sync()
    sec = time.perf_counter() - t0
    return {"elements": mesh.nelem, "dof": n, "dt": dt, "steps": nsteps, "ms_per_step": round(sec / nsteps * 1e3, 4),
            "gdof_steps_per_s": round(n * nsteps / sec / 1e9, 3),
            "note": "force on the lattice operator kernels + fused cd_solve update per step; dashpots, Ricker source"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solve", action="store_true", help="also time full plain-FP64 and mixed-precision solves")
    ap.add_argument("--c1", action="store_true", help="also time C1 (16^3 static box) solves: solves/s, latency")
    ap.add_argument("--c3", type=int, default=0, metavar="STEPS",
                    help="also time STEPS explicit CentralDifference steps of C3 (200x200x100, h 5 m)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the path would shard by "
              "element partition with a halo exchange per apply (DESIGN.md) but no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)

    from femorax_amd import core
    from femorax_amd import synthetic as S

    t0 = time.perf_counter()
    model = S.config_c2(n=args.n)
    mesh = core.Mesh(model)
    params = mesh.params_for(model)
    print(f"[bench] model {mesh.nelem} elements, {mesh.neq} DoF built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    ctx = core.Context(0)
    t0 = time.perf_counter()
    batch = core.Batch(ctx, mesh, params, fp32=True)
    info = batch.info()
    print(f"[bench] batch on device in {time.perf_counter() - t0:.1f} s: {info}", file=sys.stderr, flush=True)
    x = np.sin(0.001 * np.arange(mesh.neq))
    dx = core.DeviceArray.from_host(ctx, x)
    dy = core.DeviceArray(ctx, mesh.neq)
    dx32 = core.DeviceArray.from_host(ctx, x.astype(np.float32))
    dy32 = core.DeviceArray(ctx, mesh.neq, np.float32)

    # the operator as the solvers run it (device-resident vectors, back-to-back applies)
    batch.time_apply(dx, dy, max(1, args.warmup), 64)
    ctx.sync()
    t0 = time.perf_counter()
    ms64_total = batch.time_apply(dx, dy, args.steps, 64)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    ms64 = ms64_total / args.steps  # HIP events on the batch's stream: the roofline's denominator
    batch.time_apply(dx32, dy32, max(1, args.warmup), 32)
    ms32 = batch.time_apply(dx32, dy32, args.steps, 32) / args.steps
    # the C-ABI drop-in call (reference numbering in and out, blocking)
    for _ in range(2):
        batch.apply(dx, dy)
    t0 = time.perf_counter()
    nd = max(5, args.steps // 5)
    for _ in range(nd):
        batch.apply(dx, dy)
    ms_dropin = (time.perf_counter() - t0) / nd * 1e3
    B64 = algorithmic_bytes(mesh.neq, mesh.nelem, mesh.nnode, 8)
    B32 = algorithmic_bytes(mesh.neq, mesh.nelem, mesh.nnode, 4)
    gbs64 = B64 / (ms64 * 1e-3) / 1e9
    gbs32 = B32 / (ms32 * 1e-3) / 1e9
    value = mesh.neq * args.steps / wall / 1e9
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "GDoF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"matrix-free hex8 (stdBrick) operator apply y = K x, FP64, {args.n}^3 layered "
                               f"elastic box (8 layers, LCG seed 20261203), base fixed",
                   "elements": mesh.nelem, "dof": mesh.neq, "nodes": mesh.nnode, "lattice_path": info["lattice"],
                   "tiles": info["tiles"], "general_blocks": info["nblocks"]},
        "roofline": {"bound": "hbm", "achieved": round(gbs64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs64 / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("traffic_bytes_per_apply_f64") if args.n == 128 else None,
                     "traffic_source": "profiles/hex8_operator_c2_128_hbm_traffic.md (rocprofv3 FETCH_SIZE, "
                                       "WRITE_SIZE per apply, copy-calibrated)",
                     "kernel": "hex8_box_kernel<double, 0, *> + hex8_shared_sum_kernel<double> (one apply)",
                     "algorithmic_bytes_per_apply": B64, "kernel_ms_per_apply": round(ms64, 5)},
        "dropin_apply": {"ms": round(ms_dropin, 4), "gdof_s": round(mesh.neq / (ms_dropin * 1e-3) / 1e9, 3),
                         "note": "fx_operator_apply: reference numbering in/out (2 permutations), blocking"},
        "fp32_operator": {"gdof_s": round(mesh.neq / (ms32 * 1e-3) / 1e9, 3), "kernel_ms_per_apply": round(ms32, 5),
                          "achieved_GBs": round(gbs32, 1), "frac": round(gbs32 / HBM_PEAK_GBS, 4),
                          "algorithmic_bytes_per_apply": B32},
    }
    if args.solve:
        b = mesh.to_equations(model["nodal_load"])
        db = core.DeviceArray.from_host(ctx, b)
        xs = core.DeviceArray(ctx, mesh.neq)
        res = {}
        for name, fn in (("plain_fp64", lambda: batch.cg(db, xs, rtol=1e-8, max_it=100000)),
                         ("mixed_fp32_ir", lambda: batch.cg_mixed(db, xs, rtol=1e-8, inner_rtol=1e-4))):
            xs.zero()
            ctx.sync()
            t0 = time.perf_counter()
            r = fn()
            ctx.sync()
            r["seconds"] = round(time.perf_counter() - t0, 3)
            res[name] = r
            print(f"[bench] {name}: {r}", file=sys.stderr, flush=True)
        out["c2_static_solve"] = res
    if args.c1:
        out["c1_static_solves"] = bench_c1(core, S, ctx)
    if args.c3:
        out["c3_explicit"] = bench_c3(core, S, ctx, args.c3)
    if not args.no_cpu:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(model, params, x)
    print(json.dumps(out), flush=True)
    batch.close()
    ctx.close()


if __name__ == "__main__":
    main()
