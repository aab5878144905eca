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


def cpu_baseline(model, params, x, y_gpu=None, budget_s=20.0):
    """The oracle (C restatement of the reference's CPU operator: gather assembly, constant D,
    element geometry recomputed per apply where the reference's CPU default caches it) timed on
    the host cores, bounded to about budget_s seconds; its y also checks the GPU's y of the same
    x at full size (max |y_gpu - y_oracle| over max |y_oracle|)."""
    from oracle import fxo

    threads = cpu_threads()
    om = fxo.OracleMesh(model, coloring=False)
    op = fxo.elastic_params(params)
    y = om.apply(op, x, nthreads=threads)  # warm-up
    reps, t_total = 0, 0.0
    while t_total < budget_s and reps < 50:
        t0 = time.perf_counter()
        om.apply(op, x, nthreads=threads)
        t_total += time.perf_counter() - t0
        reps += 1
    per = t_total / reps
    out = {"value": om.neq / per / 1e9, "unit": "GDoF/s", "cores": threads, "kind": "port",
           "sample": f"{reps} full FP64 operator applies of the same {om.neq}-DoF mesh "
                     f"(oracle/liboracle.so, -O3 -fopenmp -ffp-contract=off, {threads} threads; geometry "
                     f"recomputed per apply, the reference's CPU default caches it), {per * 1e3:.1f} ms per apply"}
    if y_gpu is not None:
        out["gpu_y_max_rel_error"] = float(np.max(np.abs(y_gpu - y)) / np.max(np.abs(y)))
    return out


def bench_refcol(core, S, ctx, side, reps):
    """The reference's own operator benchmark (benchmarks/bench_operator.cpp:71-111): the side^3
    soil column of tests/common/synthetic_models.hpp (consistent mass, two Rayleigh regions,
    laminar ties), K_eff = K + 2e3 C + 1.6e7 M applied to device-resident vectors. Its published
    H200 times are beside it for context (a different GPU and build; not this line's metric)."""
    model = S.reference_column(side)
    mesh = core.Mesh(model)
    batch = core.Batch(ctx, mesh, mesh.params_for(model))
    k = core.coeffs(*S.REFERENCE_COLUMN_COEFFS)
    x = core.DeviceArray.from_host(ctx, np.full(mesh.neq, 1e-3))
    y = core.DeviceArray(ctx, mesh.neq)
    batch.time_apply(x, y, 5, 64, k=k)
    ms = batch.time_apply(x, y, reps, 64, k=k) / reps
    res = {"model": f"{side}^3 soil column (synthetic_models.hpp: consistent mass, 2 Rayleigh regions, laminar "
                    f"ties), K + 2e3 C + 1.6e7 M", "elements": mesh.nelem, "equations": mesh.neq,
           "path": "lattice" if batch.info()["lattice"] else "general", "us_per_apply": round(ms * 1e3, 2),
           "gdof_s": round(mesh.neq / (ms * 1e-3) / 1e9, 3)}
    if side == 40:
        res["reference_h200_us"] = {"gated_default": 242.3, "fastest_variant": 73.1,
                                    "source": "benchmarks/baselines/tacc-vista-gh-cuda270336.json:1909, :1789"}
    # 20 CG and GMRES(30) iterations with Jacobi on the Newmark (dt = 1e-3) operator, b = 1, x0 = 0,
    # no tolerance (bench_operator.cpp:113-185); the Jacobi setup is inside our timed call
    dt = 1e-3
    kn = core.coeffs(1.0, 0.5 / (0.25 * dt), 1.0 / (0.25 * dt * dt))
    gb = core.Batch(ctx, mesh, mesh.params_for(model), gmres=True)
    b = core.DeviceArray.from_host(ctx, np.ones(mesh.neq))
    xs = core.DeviceArray(ctx, mesh.neq)
    solves = {}
    for name, fn in (("cg20", gb.cg), ("gmres20", gb.gmres)):
        per = []
        for _ in range(6):
            xs.zero()
            ctx.sync()
            t0 = time.perf_counter()
            r = fn(b, xs, rtol=0.0, max_it=20, k=kn)
            ctx.sync()
            per.append(time.perf_counter() - t0)
        solves[name] = {"ms": round(float(np.median(per[1:])) * 1e3, 3), "iterations": r["iterations"]}
    res["solves"] = solves
    # the explicit internal force with damping at u = 1e-4, v = 1e-3 on the lumped-mass column
    # (bench_explicit.cpp:76-140)
    fmodel = S.reference_column(side, lumped=True)
    fmesh = core.Mesh(fmodel)
    fb_ = core.Batch(ctx, fmesh, fmesh.params_for(fmodel))
    du = core.DeviceArray.from_host(ctx, np.full(fmesh.neq, 1e-4))
    dv = core.DeviceArray.from_host(ctx, np.full(fmesh.neq, 1e-3))
    df = core.DeviceArray(ctx, fmesh.neq)
    for _ in range(3):
        fb_.force(du, df, dv)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        fb_.force(du, df, dv)
    ctx.sync()
    res["explicit_force_damped_us"] = round((time.perf_counter() - t0) / reps * 1e6, 2)
    # the stored-tangent operator (the path of a history material): a PressureIndependMultiYield
    # batch on the same column, tangents stored at a state, K_eff = K_t + 2e3 C + 1.6e7 M; and
    # its return-map force at an elastic and at a yielding displacement
    keep = []
    mats, rays = [], []
    for p in model["params"]:
        G = p["E"] / (2.0 * (1.0 + p["nu"]))
        K = p["E"] / (3.0 * (1.0 - 2.0 * p["nu"]))
        mats.append(core.pimy_material(p["rho"], G, K, 30.0, 0.1, 0.0, 100.0, 0.0, nsurf=20, b=p["b"], mass_type=0,
                                       keep=keep))
        rays.append(p["rayleigh"])
    pb = core.Batch.pimy(ctx, mesh, [mats[g] for g in range(len(mats))], rayleigh=rays, fp32=True)
    pb.set_stage(1)
    z = np.linspace(0.0, 1.0, mesh.neq)
    dup = core.DeviceArray.from_host(ctx, 2e-4 * np.sin(40.0 * z))
    dfp = core.DeviceArray(ctx, mesh.neq)
    pb.store_tangents(dup)
    pb.time_apply(x, y, 3, 64, k=k)
    ms_t = pb.time_apply(x, y, reps, 64, k=k) / reps
    res["pimy_stored_tangent_apply_us"] = round(ms_t * 1e3, 2)
    # the same operator in FP32 (FP32 tangent copy; the inner solves of mixed-precision CG)
    x32 = core.DeviceArray.from_host(ctx, np.sin(0.001 * np.arange(mesh.neq)).astype(np.float32))
    y32 = core.DeviceArray(ctx, mesh.neq, np.float32)
    pb.time_apply(x32, y32, 3, 32, k=k)
    res["pimy_stored_tangent_apply_fp32_us"] = round(pb.time_apply(x32, y32, reps, 32, k=k) / reps * 1e3, 2)
    for name, amp in (("elastic", 1e-7), ("yielding", 2e-3)):
        dup.upload(amp * np.sin(40.0 * z))
        pb.force(dup, dfp)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(5):
            pb.force(dup, dfp)
        ctx.sync()
        res[f"pimy_force_{name}_us"] = round((time.perf_counter() - t0) / 5 * 1e6, 1)
    res["pimy_force_note"] = ("host wall time per blocking fx_force call (5 calls after one untimed) on this batch "
                              "(two Rayleigh regions, FP32 tangent copy, material stage 1, no velocities): the "
                              "return-map kernel, the two gathers and the synchronisation; a kernel trace of this "
                              "leg measured the return-map kernel at 262-268 us (elastic: after the tangent store at a "
                              "yielding displacement, whose centre-copy bound the later updates keep) and 1,236-1,261 "
                              "us (yielding), the gathers at about 10 us each")
    active = pb.pimy_state(False)[:, 26]
    res["pimy_yielding_points_frac"] = round(float(np.mean(active > 0)), 3)
    if side == 40:
        res["reference_h200"] = {"cg20_ms": 7.676, "gmres20_ms": 8.447, "explicit_force_damped_us": 335.2,
                                 "stored_tangent_apply_us": 487.1,
                                 "source": "benchmarks/baselines/tacc-vista-gh-cuda270336.json:2089, :2134, :394, "
                                           ":1939"}
    return res


def bench_ssp(core, ctx, mesh, params, model, x, dx, dy, args, B64, check=True):
    """The SSPbrick<ElasticIsotropic> operator (one mean-strain point plus the stabilisation term) on
    the same lattice: it moves the same bytes as stdBrick's apply, so the same roofline bytes apply;
    with the oracle's SSPbrick apply on the whole model when the CPU leg runs."""
    sp = [dict(p, element="SSPbrick") for p in params]
    t0 = time.perf_counter()
    sb = core.Batch(ctx, mesh, sp, element="SSPbrick")
    setup = time.perf_counter() - t0
    sb.time_apply(dx, dy, max(1, args.warmup), 64)
    ms = sb.time_apply(dx, dy, args.steps, 64) / args.steps
    gbs = B64 / (ms * 1e-3) / 1e9
    res = {"kernel_ms_per_apply": round(ms, 5), "gdof_s": round(mesh.neq / (ms * 1e-3) / 1e9, 3),
           "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4), "setup_s": round(setup, 2),
           "kernel": "hex8_box_kernel<double, 3, 0, false, 8> + hex8_shared_sum_kernel<double>",
           "lattice_path": sb.info()["lattice"]}
    if check:
        from oracle import fxo

        sb.apply(dx, dy)
        om = fxo.OracleMesh(model, coloring=False)
        want = om.apply(fxo.elastic_params(sp), x, nthreads=cpu_threads())
        res["gpu_y_max_rel_error"] = float(np.max(np.abs(dy.get() - want)) / np.max(np.abs(want)))
    sb.close()
    return res


def bench_c1(core, S, ctx, reps=30, cpu=True):
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
    out = {"dof": mesh.neq, "iterations": r["iterations"], "solves_per_s": round(1.0 / lat.mean(), 2),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
           "note": "one blocking fx_cg per solve, rtol 1e-8, x0 = 0"}
    if cpu:
        from oracle import fxo

        threads = cpu_threads()
        om = fxo.OracleMesh(model)
        op = fxo.elastic_params(mesh.params_for(model))
        rhs = mesh.to_equations(model["nodal_load"])
        om.cg(op, rhs, rtol=1e-8, lanes=1, nthreads=threads)  # warm-up
        reps, t_total = 0, 0.0
        while t_total < 10.0 and reps < 200:
            t0 = time.perf_counter()
            om.cg(op, rhs, rtol=1e-8, lanes=1, nthreads=threads)
            t_total += time.perf_counter() - t0
            reps += 1
        out["cpu_baseline"] = {"value": round(reps / t_total, 2), "unit": "solves/s", "cores": threads, "kind": "port",
                               "sample": f"{reps} oracle CG + Jacobi solves of C1 to rtol 1e-8 (VectorOps order, "
                                         f"lanes 1), {threads} threads"}
    return out


def bench_c3(core, S, ctx, nsteps, cpu=True):
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
    out = {"elements": mesh.nelem, "dof": n, "dt": dt, "steps": nsteps, "ms_per_step": round(sec / nsteps * 1e3, 4),
           "gdof_steps_per_s": round(n * nsteps / sec / 1e9, 3),
           "note": "force on the lattice operator kernels + fused cd_solve update per step; dashpots, Ricker source"}
    # the same steps with OpenSees node recorders (disp, vel, accel; every step, -time, default
    # precision 6) of the vertical node line under the source, written by fx_recorder's writer thread
    import shutil
    import tempfile

    xyz = np.asarray(model["xyz"])
    line = np.flatnonzero((xyz[:, 0] == xyz[src, 0]) & (xyz[:, 1] == xyz[src, 1]))
    cols = mesh.recorder_columns(line)
    rsteps = min(nsteps, 500)

    def recorded(responses):
        tmp = tempfile.mkdtemp(prefix="fx_c3_rec_")
        try:
            rec = core.Recorder(ctx)
            for i in [rec.add_node(os.path.join(tmp, f"{resp}.txt"), resp, cols) for resp in responses]:
                rec.activate(i)
            ctx.sync()
            t0 = time.perf_counter()
            batch.explicit_cd(dt, rsteps, u, um, v, a, dash=dash, load_eq=load_eq,
                              load_hist=hist[warm:warm + rsteps].ravel(), recorder=rec)
            ctx.sync()
            t1 = time.perf_counter()
            rec.finish()
            t2 = time.perf_counter()
            lines = rec.records
            rec.close()
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        return {"recorders": f"{len(responses)} Node recorders ({', '.join(responses)}) of {len(line)} nodes x 3 "
                             f"DOFs, every step", "steps": rsteps, "lines": int(lines),
                "ms_per_step": round((t1 - t0) / rsteps * 1e3, 4),
                "ms_per_step_files_closed": round((t2 - t0) / rsteps * 1e3, 4),
                "overhead_vs_unrecorded": round((t1 - t0) / rsteps / (sec / nsteps) - 1.0, 4)}

    out["recorded"] = recorded(("disp", "vel", "accel"))
    out["recorded_disp_only"] = recorded(("disp",))
    out["recorded"]["note"] = ("recording v or a forms the full recorded v, a every step (the reference's cd_solve "
                               "always does; the unrecorded loop forms them in its last two steps only)")
    if cpu:
        from oracle import fxo

        threads = cpu_threads()
        om = fxo.OracleMesh(model, coloring=False)
        op = fxo.elastic_params(mesh.params_for(model))
        m, _ = om.lumped_mass(op)
        dashd = mesh.to_equations(model["dashpots"])
        uh, umh, vh, ah = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        steps = 3
        t0 = time.perf_counter()
        om.cd_run(op, m, None, dt, steps, uh, umh, vh, ah, load_eq=load_eq, load_hist=hist[:steps].ravel(), dash=dashd,
                  nthreads=threads)
        per = (time.perf_counter() - t0) / steps
        out["cpu_baseline"] = {"value": round(n / per / 1e9, 4), "unit": "GDoF*step/s", "cores": threads,
                               "kind": "port", "sample": f"{steps} oracle CentralDifference steps of the same model, "
                                                         f"{threads} threads, {per * 1e3:.0f} ms per step"}
    return out


def bench_c4(core, S, ctx, nsteps, ncol=4096, plain_columns=2, plain_steps=5, cpu=True, plain=True, compact=False,
             mixed=False, waves=0):
    """C4: ncol PressureIndependMultiYield columns 1 x 1 x N (N ~ U{20..400}, 10 layers each),
    Newmark (gamma 0.5, beta 0.25, dt 0.005 s) + Newton (NormDispIncr 1e-10) + PCG (1e-12),
    one band-limited base record at 0.3 g. All columns admitted at once into one pool and run to
    `nsteps` steps; timed from the first iteration to the last column's end."""
    cols, rec = S.config_c4(ncol=ncol, nsteps=max(nsteps, 150))
    total = sum(c["nelem"] for c in cols)
    pool = core.ColumnPool(ctx, max_elements=total + 1024, max_columns=ncol, dt=0.005, max_materials=10 * ncol,
                           max_record=ncol * 10 * nsteps + 1024, compact=compact, mixed=mixed, waves=waves,
                           tune=waves == 0)
    keep = []
    t0 = time.perf_counter()
    ids = [pool.add(core.pimy_materials(c["mats"], keep), c["mat_of"], nsteps, rec, record=True) for c in cols]
    print(f"[bench] C4 {ncol} columns, {total} elements admitted in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    ctx.sync()
    per_call = 8
    t0 = time.perf_counter()
    calls = 0
    while pool.run(per_call):
        calls += 1
    calls += 1
    sec = time.perf_counter() - t0
    st = [pool.status(i) for i in ids]
    steps = sum(s["steps"] for s in st)
    newton = sum(s["newton_iterations"] for s in st)
    krylov = sum(s["krylov_iterations"] for s in st)
    out = {"columns": ncol, "elements": total, "equations": 3 * total, "steps_per_column": nsteps,
           "failed_columns": sum(s["status"] == 2 for s in st), "seconds": round(sec, 3),
           "column_steps_per_s": round(steps / sec, 1), "pool_iterations": calls * per_call,
           "ms_per_pool_iteration": round(sec / (calls * per_call) * 1e3, 4),
           "newton_per_step": round(newton / max(steps, 1), 3), "cg_per_newton": round(krylov / max(newton, 1), 2),
           "state_bytes_full_copy": total * 2 * 150 * 8,  # upper bound: unmoved surface centres are not copied
           "return_map": "compacted (predict, compact, correct)" if compact else "fused",
           "waves_per_simd": {0: "autotuned (1, 3 or 4; timed in the first call)", 1: "compiler"}.get(waves, waves),
           "solve": "FP32 inner PCG in FP64 refinement" if mixed else "FP64 PCG",
           "note": "one material point per tied element (uniform strain), return map + per-column PCG per "
                   "Newton iteration, hipGraph replay"}
    pool.close()
    if not plain:
        return out
    # the plain per-model FP64 path (fx_newmark on one stdBrick<PIMY> batch per column), sampled
    t_plain, n_plain = 0.0, 0
    for c in cols[:plain_columns]:
        n = c["nelem"]
        model = S.soil_column(n, 1.0, params=[dict(E=1e8, nu=0.3, rho=2000.0)] * 10, layer_of=lambda k: c["mat_of"][k])
        pm = core.Mesh(model)
        batch = core.Batch.pimy(ctx, pm, core.pimy_materials([c["mats"][g] for g in pm.param_group], keep))
        batch.set_stage(1)
        d = [core.DeviceArray(ctx, pm.neq) for _ in range(4)]
        for x in d:
            x.zero()
        ctx.sync()
        t1 = time.perf_counter()
        r = batch.newmark(*d, 0.005, plain_steps, ground_dir=0, ground_accel=rec[:plain_steps])
        t_plain += time.perf_counter() - t1
        n_plain += r["steps"]
        batch.close()
    out["plain_fp64_per_model"] = {"columns": plain_columns, "steps": plain_steps,
                                   "column_steps_per_s": round(n_plain / t_plain, 2)}
    out["speedup_vs_plain"] = round(out["column_steps_per_s"] / (n_plain / t_plain), 1)
    if cpu:
        out["cpu_baseline"] = c4_cpu_baseline(core, S, cols[:8], rec)
    return out


def bench_c5(core, S, duration=60.0, rate=200.0, burst=2000, plain_columns=2, plain_blocks=4, plain_steps=50,
             log=None):
    """C5: Poisson arrivals (rate per s, `duration` s): 70% soil columns (as C4, 500 steps), 30%
    static blocks n^3 (n ~ U{8..48}, C1 loads, rtol 1e-8); HBM budget 200 GB; served with
    continuous batching in the mixed-precision form (femorax_amd/serve.py). Also: the same
    server's capacity on a burst of `burst` requests arriving at once, and the plain FP64
    per-solve path (one request at a time: fx_newmark per column model, fx_cg per block) on a
    sample of the same requests, as requests/s at its mean service time."""
    from femorax_amd import serve

    reqs = serve.mixed_requests(duration=duration, rate=rate)
    t0 = time.perf_counter()
    srv = serve.Server(hbm_budget=200e9, mixed=True, block_sizes=range(8, 49))
    setup_s = time.perf_counter() - t0
    try:
        recs = srv.run(reqs, log=log)
        served = serve.summary(recs)
        served["loop"] = srv.last_stats
        served["setup_s"] = round(setup_s, 2)  # block models made resident before the first request
        burst_reqs = serve.mixed_requests(duration=1.0, rate=float(burst), seed=S.SEED_BASE + 5)
        for r in burst_reqs:
            r.arrival = 0.0
        cap = serve.summary(srv.run(burst_reqs, log=log))
    finally:
        srv.close()
    # the plain FP64 per-solve path on a sample of the requests
    ctx = core.Context(0)
    cols = [r for r in reqs if r.kind == "column"][:plain_columns]
    blks = [r for r in reqs if r.kind == "block"][:plain_blocks]
    t_col = []
    for r in cols:
        model = S.soil_column(r.n, 1.0, params=[dict(E=1e8, nu=0.3, rho=2000.0)] * 10,
                              layer_of=lambda k, mo=r.mat_of: mo[k])
        pm = core.Mesh(model)
        batch = core.Batch.pimy(ctx, pm, core.pimy_materials([r.mats[g] for g in pm.param_group]))
        batch.set_stage(1)
        d = [core.DeviceArray(ctx, pm.neq) for _ in range(4)]
        for x in d:
            x.zero()
        ctx.sync()
        t0 = time.perf_counter()
        batch.newmark(*d, 0.005, plain_steps, ground_dir=0, ground_accel=r.record[:plain_steps])
        t_col.append((time.perf_counter() - t0) * r.nsteps / plain_steps)
        batch.close()
    t_blk = []
    problems = serve.BlockProblems()
    for r in blks:
        model, mesh, rhs = problems.get(r.n)
        batch = core.Batch(ctx, mesh, mesh.params_for(model))
        b = core.DeviceArray.from_host(ctx, rhs)
        x = core.DeviceArray(ctx, mesh.neq)
        x.zero()
        ctx.sync()
        t0 = time.perf_counter()
        batch.cg(b, x, rtol=1e-8)
        t_blk.append(time.perf_counter() - t0)
        batch.close()
    ctx.close()
    frac = sum(r.kind == "column" for r in reqs) / len(reqs)
    mean_service = frac * float(np.mean(t_col)) + (1.0 - frac) * float(np.mean(t_blk))
    plain = {"solves_per_s": round(1.0 / mean_service, 3), "column_s": round(float(np.mean(t_col)), 3),
             "block_ms": round(float(np.mean(t_blk)) * 1e3, 2),
             "sample": f"{len(cols)} columns ({plain_steps} of {cols[0].nsteps} steps, scaled), {len(blks)} blocks"}
    return {"served": served, "capacity_burst": cap, "plain_fp64_per_solve": plain,
            "speedup_served_vs_plain": round(served["solves_per_s"] / plain["solves_per_s"], 1),
            "capacity_vs_plain": round(cap["solves_per_s"] / plain["solves_per_s"], 1),
            "config": {"rate_per_s": rate, "duration_s": duration, "columns": "70%, N~U{20..400}, 500 steps",
                       "blocks": "30%, n~U{8..48}, rtol 1e-8", "hbm_budget_GB": 200, "block_workers": 2,
                       "pool": "one stream, 4 iterations per admission round",
                       "blocks_served": "one fx_cg per block request on a free worker",
                       "precision": "columns: FP32 inner PCG in FP64 refinement (FP64 return map and state); blocks: FP64 CG "
                                    "(latency-bound at these sizes, where the FP32 inner solves measured slower)"}}


def c4_cpu_baseline(core, S, cols, rec, steps=150):
    """The oracle's implicit stage (eight Gauss points per element, the reference's algorithm)
    on a sample of C4 columns, one after another, on the host cores."""
    from oracle import fxo

    cores = os.cpu_count() or 1
    sec, done, elems = 0.0, 0, 0
    for col in cols:
        n = col["nelem"]
        model = S.soil_column(n, 1.0, params=[dict(E=1e8, nu=0.3, rho=2000.0)] * 10,
                              layer_of=lambda k: col["mat_of"][k])
        om = fxo.OracleMesh(model)
        pm = core.Mesh(model)
        pe = fxo.pimy_elems([col["mats"][g] for g in pm.param_group], stage=1)
        Cs = om.pimy_initial_state(pe)
        Ts = Cs.copy()
        N = om.neq
        u, v, a, du = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N)
        gd = np.zeros(N)
        q = om.node_eq.reshape(-1, 3)[:, 0]
        gd[q[q >= 0]] = 1.0
        t0 = time.perf_counter()
        _, st = om.newmark(u, v, a, du, 0.005, steps, pimy=pe, pimy_state=[Cs, Ts], ground_dir=gd,
                           ground_accel=rec[:steps], nthreads=cores)
        sec += time.perf_counter() - t0
        done += st["steps"]
        elems += n
    return {"value": round(done / sec, 2), "unit": "column_steps/s", "cores": cores, "kind": "port",
            "sample": f"{len(cols)} columns ({elems} elements) x {steps} steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", dest="solve", action="store_false",
                    help="skip the full plain-FP64 and mixed-precision C2 solves")
    ap.add_argument("--refcol", type=int, default=40, metavar="SIDE",
                    help="time the reference's own operator benchmark model, a SIDE^3 soil column (0: skip)")
    ap.add_argument("--c1", type=int, default=1, help="time C1 (16^3 static box) solves: solves/s, latency (0: skip)")
    ap.add_argument("--c4", type=int, default=4000, metavar="STEPS",
                    help="time STEPS Newmark steps of C4 (4096 PIMY soil columns) and the plain per-model path "
                         "(0: skip)")
    ap.add_argument("--c4-columns", type=int, default=4096)
    ap.add_argument("--c5", type=float, default=60.0, metavar="SECONDS",
                    help="serve C5 (Poisson 200/s mixed columns and blocks) for SECONDS of arrivals (0: skip)")
    ap.add_argument("--ssp", type=int, default=1, help="time the SSPbrick operator on the same lattice (0: skip)")
    ap.add_argument("--c3", type=int, default=2000, metavar="STEPS",
                    help="time STEPS explicit CentralDifference steps of C3 (200x200x100, h 5 m; 0: skip)")
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
    # the input vectors first, so that the batch's autotune (which ends with the GPU busy and its
    # clocks up) runs right before the warm-up and the timed applies
    x = np.sin(0.001 * np.arange(mesh.neq))
    dx = core.DeviceArray.from_host(ctx, x)
    dy = core.DeviceArray(ctx, mesh.neq)
    dx32 = core.DeviceArray.from_host(ctx, x.astype(np.float32))
    dy32 = core.DeviceArray(ctx, mesh.neq, np.float32)
    t0 = time.perf_counter()
    batch = core.Batch(ctx, mesh, params, fp32=True)
    info = batch.info()
    print(f"[bench] batch on device in {time.perf_counter() - t0:.1f} s: {info}", file=sys.stderr, flush=True)

    # the operator as the solvers run it (device-resident vectors, back-to-back applies)
    batch.time_apply(dx, dy, max(1, args.warmup), 64)
    ctx.sync()
    t0 = time.perf_counter()
    ms64_total = batch.time_apply(dx, dy, args.steps, 64)
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    ms64 = ms64_total / args.steps  # HIP events on the batch's stream: the roofline's denominator
    # a short timed run starts before the GPU's clocks have settled (they rise over the first
    # tenths of a second of load); beside the K-step figure, the same applies after 2000
    # warm-up applies, 1000 timed (untimed for `value`)
    ms64_sustained = None
    if args.steps < 1000:
        batch.time_apply(dx, dy, 2000, 64)
        ms64_sustained = batch.time_apply(dx, dy, 1000, 64) / 1000
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
    y_dropin = dy.get()  # the drop-in apply's y (reference numbering): checked against the oracle below
    # the same calls stream-ordered (FX_OPT_STREAM_ORDERED: they return once queued, as Kokkos
    # launches do), synchronised once at the end
    ctx.set_stream_ordered(True)
    try:
        batch.apply(dx, dy)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(nd):
            batch.apply(dx, dy)
        ctx.sync()
        ms_dropin_ordered = (time.perf_counter() - t0) / nd * 1e3
    finally:
        ctx.set_stream_ordered(False)
    assert np.array_equal(dy.get(), y_dropin), "stream-ordered drop-in apply: other bits"
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
                     "algorithmic_bytes_per_apply": B64, "kernel_ms_per_apply": round(ms64, 5),
                     "sustained": None if ms64_sustained is None else {
                         "warmup": 2000, "steps": 1000, "kernel_ms_per_apply": round(ms64_sustained, 5),
                         "gdof_s": round(mesh.neq / (ms64_sustained * 1e-3) / 1e9, 3),
                         "frac": round(B64 / (ms64_sustained * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "note": "the same applies after the clocks settle; value is the K timed steps above"}},
        "dropin_apply": {"ms": round(ms_dropin, 4), "gdof_s": round(mesh.neq / (ms_dropin * 1e-3) / 1e9, 3),
                         "ms_stream_ordered": round(ms_dropin_ordered, 4),
                         "vs_internal_stream_ordered": round(ms_dropin_ordered / ms64, 3),
                         "note": "fx_operator_apply: reference numbering in and out (node words in the reference "
                                 "numbering, no permutation pass); ms: each call blocking, ms_stream_ordered: "
                                 "FX_OPT_STREAM_ORDERED, one synchronisation after the calls"},
        "fp32_operator": {"gdof_s": round(mesh.neq / (ms32 * 1e-3) / 1e9, 3), "kernel_ms_per_apply": round(ms32, 5),
                          "achieved_GBs": round(gbs32, 1), "frac": round(gbs32 / HBM_PEAK_GBS, 4),
                          "algorithmic_bytes_per_apply": B32},
    }
    if args.ssp:
        out["ssp_operator"] = bench_ssp(core, ctx, mesh, params, model, x, dx, dy, args, B64,
                                        check=not args.no_cpu)
    if args.solve:
        b = mesh.to_equations(model["nodal_load"])
        db = core.DeviceArray.from_host(ctx, b)
        xs = core.DeviceArray(ctx, mesh.neq)
        res, sols = {}, {}
        ctx.set_graphs(True)  # the CG iterations replayed from captured hipGraphs (the same kernels and bits)
        for name, fn in (("plain_fp64", lambda: batch.cg(db, xs, rtol=1e-8, max_it=100000)),
                         ("mixed_fp32_ir", lambda: batch.cg_mixed(db, xs, rtol=1e-8, inner_rtol=1e-6))):
            xs.zero()
            ctx.sync()
            t0 = time.perf_counter()
            r = fn()
            ctx.sync()
            r["seconds"] = round(time.perf_counter() - t0, 3)
            res[name] = r
            sols[name] = xs.get()
            print(f"[bench] {name}: {r}", file=sys.stderr, flush=True)
        from oracle import fxo

        threads = cpu_threads()
        om = fxo.OracleMesh(model, coloring=False)
        op = fxo.elastic_params(params)
        # each solution's residual with the FP64 oracle's operator, and the two solutions' difference
        for name in sols:
            r = b - om.apply(op, sols[name], nthreads=threads)
            res[name]["oracle_relative_residual"] = float(np.linalg.norm(r) / np.linalg.norm(b))
        res["mixed_vs_plain_max_rel_difference"] = float(np.max(np.abs(sols["mixed_fp32_ir"] - sols["plain_fp64"])) /
                                                         np.max(np.abs(sols["plain_fp64"])))
        res["mixed_speedup"] = round(res["plain_fp64"]["seconds"] / res["mixed_fp32_ir"]["seconds"], 3)
        ctx.set_graphs(False)
        for name in ("plain_fp64", "mixed_fp32_ir"):
            res[name]["ms_per_iteration"] = round(res[name]["seconds"] * 1e3 / max(1, res[name]["iterations"]), 4)
        if not args.no_cpu:  # the oracle's CG for a few iterations: time per iteration, whole solve estimated
            its = 5
            t0 = time.perf_counter()
            om.cg(op, b, rtol=1e-8, max_it=its, lanes=1, nthreads=threads)
            per = (time.perf_counter() - t0) / its
            n_it = res["plain_fp64"]["iterations"]
            res["cpu_baseline"] = {"value": round(1.0 / (per * n_it), 6), "unit": "solves/s", "cores": threads,
                                   "kind": "port", "sample": f"{its} oracle CG iterations ({per * 1e3:.0f} ms each, "
                                                             f"{threads} threads), scaled to the {n_it} iterations "
                                                             f"of the GPU solve"}
        out["c2_static_solve"] = res
    if args.refcol:
        try:
            out["reference_operator_benchmark"] = bench_refcol(core, S, ctx, args.refcol, args.steps)
        except Exception as e:  # noqa: BLE001
            out["reference_operator_benchmark"] = {"error": repr(e)}
    if args.c1:
        out["c1_static_solves"] = bench_c1(core, S, ctx, cpu=not args.no_cpu)
    if args.c3:
        out["c3_explicit"] = bench_c3(core, S, ctx, args.c3, cpu=not args.no_cpu)
    # the batched and served workloads: reported beside the operator metric; a failure there is
    # recorded in the line instead of losing the line
    if args.c4:
        try:
            out["c4_columns"] = bench_c4(core, S, ctx, args.c4, ncol=args.c4_columns, cpu=not args.no_cpu)
            # the same pool with FP32 inner PCG in FP64 refinement (same relative tolerance)
            m = bench_c4(core, S, ctx, args.c4, ncol=args.c4_columns, cpu=False, plain=False, mixed=True)
            out["c4_columns"]["mixed_precision"] = {k: m[k] for k in ("column_steps_per_s", "ms_per_pool_iteration",
                                                                     "newton_per_step", "cg_per_newton",
                                                                     "failed_columns", "solve")}
        except Exception as e:  # noqa: BLE001
            out["c4_columns"] = {"error": repr(e)}
    if args.c5:
        try:
            out["c5_serving"] = bench_c5(core, S, duration=args.c5,
                                         log=lambda s: print(s, file=sys.stderr, flush=True))
        except Exception as e:  # noqa: BLE001
            out["c5_serving"] = {"error": repr(e)}
    if not args.no_cpu:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(model, params, x, y_gpu=y_dropin)
    print(json.dumps(out), flush=True)
    batch.close()
    ctx.close()


if __name__ == "__main__":
    main()
