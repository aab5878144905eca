"""Serving loop for a mixed stream of soil-column analyses and static block solves on one GPU.

Continuous batching: requests are admitted as they arrive, whenever their HBM share has room,
while earlier requests are still on the device.
- Soil columns (PressureIndependMultiYield, Newmark + Newton + PCG) join one resident column
  pool (fx_column_*; mixed mode: FP32 inner PCG in FP64 refinement). The pool is sized from
  the column share of the HBM budget. A column is
  admitted when the pool has room (fx_column_add returns FX_EFULL otherwise) and starts its
  first step in the next iteration. Each pool iteration advances every resident column by one
  Newton iteration, each at its own step. Finished columns are read back and their slots
  freed at once.
- Static elastic blocks (n^3, CG + Jacobi to rtol 1e-8) go to worker threads. Each worker
  has its own context, so its own HIP stream: block solves overlap with the column iterations
  and with each other on the device. A free worker takes the next block request and solves it
  with fx_cg. The alternative admission rounds (block_rounds=True: a free worker takes every
  block request waiting, one per block size, and solves them together with fx_cg_many, each
  solve's next iterations launched in turn and one synchronisation per poll round for all of
  them; each solve's bits are those of its own fx_cg) measured worse on C5 at 200 requests/s,
  in one process on one box: blocks p50 / p99 16.6 / 59.5 ms in rounds against 10.2 / 23.7 ms
  one solve per request over 20 s (15.7 / 43.9 ms over 60 s), columns p50 908 against 740 ms.
  A round lasts as long as its largest block, and every block of it waits for that end. Three
  workers wait less (block wait p99 7 against 22 ms over 60 s) at the same service times. A
  worker keeps the compiled batch of each block size it
  has served resident, within the block share of the budget, evicting the least recently
  used; the block sizes the service is set up for (block_sizes) are made resident before the
  first request. Blocks solve with FP64 CG (fx_cg) by default: up to 48^3 (346k DoF) a solve is
  latency-bound, and FP32 inner CG inside FP64 refinement (fx_cg_mixed, block_mixed=True) needs
  1.5-2.6x the iterations at about the same time per iteration (measured 1.4-2.6x slower). Columns
  use the mixed PCG (mixed=True), which measures slightly faster on the pool.
Each request's result is computed by the same kernels with the same per-request reductions
alone or in the mix: a column's bits do not depend on its co-runners (the pool's solve is per
column), and a block solve always starts from x = 0 on its own batch.
"""
import collections
import math
import queue
import threading
import time

import numpy as np

from . import core
from . import synthetic as S


class Request:
    __slots__ = ("rid", "kind", "arrival", "start", "end", "n", "mats", "mat_of", "nsteps", "record", "result")

    def __init__(self, rid, kind, arrival, **kw):
        self.rid, self.kind, self.arrival = rid, kind, arrival
        self.start = self.end = None
        self.n = kw.get("n")
        self.mats = kw.get("mats")
        self.mat_of = kw.get("mat_of")
        self.nsteps = kw.get("nsteps")
        self.record = kw.get("record")
        self.result = None


def mixed_requests(duration=60.0, rate=200.0, column_frac=0.7, nsteps=500, dt=0.005, seed=S.SEED_BASE + 4,
                   nrecords=16, block_sizes=(8, 48), column_sizes=(20, 400)):
    """Poisson arrivals (LCG exponential inter-arrival times) over `duration` seconds: a
    fraction `column_frac` are soil columns (N ~ U{20..400}, 10 layers, `nsteps` steps, one of
    `nrecords` seeded base records at 0.3 g), the rest static blocks n^3 with n ~ U{8..48}."""
    rng = S.Lcg(seed)
    records = [S.ground_motion(nsteps, dt, seed + 100 + k) for k in range(nrecords)]

    def uniform01():
        return (rng.symmetric() + 1.0) * 0.5

    out, t = [], 0.0
    while True:
        t += -math.log(max(1.0 - uniform01(), 1e-300)) / rate
        if t >= duration:
            break
        rid = len(out)
        if uniform01() < column_frac:
            lo, hi = column_sizes
            n = lo + int(uniform01() * (hi - lo + 1)) % (hi - lo + 1)
            out.append(Request(rid, "column", t, n=n, mats=S.soil_layers(rng, 10), mat_of=S.column_layer_of(n, 10),
                               nsteps=nsteps, record=records[rid % nrecords]))
        else:
            lo, hi = block_sizes
            out.append(Request(rid, "block", t, n=lo + int(uniform01() * (hi - lo + 1)) % (hi - lo + 1)))
    return out


class BlockProblems:
    """Compiled block meshes and their load vectors, shared by the workers (host only)."""

    def __init__(self):
        self._lock = threading.Lock()
        self._cache = {}

    def get(self, n):
        with self._lock:
            if n not in self._cache:
                model = S.config_block(n)
                mesh = core.Mesh(model)
                self._cache[n] = (model, mesh, mesh.to_equations(model["nodal_load"]))
            return self._cache[n]


class BlockWorker(threading.Thread):
    """Solves block requests on its own context (stream), keeping batches resident by size."""

    def __init__(self, device, problems, inbox, done, budget, mixed, rtol=1e-8, keep_x=False, inner_rtol=1e-6,
                 prewarm=(), rounds=False, max_round=32, graphs=False):
        super().__init__(daemon=True)
        self.device, self.problems, self.inbox, self.done = device, problems, inbox, done
        self.budget, self.mixed, self.rtol, self.keep_x = budget, mixed, rtol, keep_x
        self.inner_rtol = inner_rtol
        self.prewarm = tuple(prewarm)
        # rounds: every block request waiting when the worker is free (one per block size) is
        # solved in one fx_cg_many call (FP64 CG); else one solve per request
        self.rounds, self.max_round, self.graphs = rounds and not mixed, max_round, graphs
        self.ready = threading.Event()
        self.error = None

    def run(self):
        ctx = None
        try:
            ctx = core.Context(self.device)
        except Exception as e:  # noqa: BLE001 - every request then reports the failure
            self.error = e
        resident = collections.OrderedDict()  # n -> (batch, b, x, bytes)
        self.used = 0
        # the block models the service is set up for: resident on the device before traffic (a
        # model seen for the first time would otherwise pay its compile and upload in its latency)
        for n in self.prewarm:
            if ctx is None:
                break
            try:
                self._resident(ctx, resident, n)
            except Exception as e:  # noqa: BLE001 - the request path reports it again
                self.error = e
        if ctx is not None and self.graphs:
            ctx.set_graphs(True)
        self.ready.set()
        stop = False
        while not stop:
            r = self.inbox.get()
            if r is None:
                break
            batch = [r]
            if self.rounds:
                # the admission round: the requests waiting now, one per block size (a second
                # request of a size waits for the next round: each size has one resident batch)
                later = []
                while len(batch) < self.max_round:
                    try:
                        q = self.inbox.get_nowait()
                    except queue.Empty:
                        break
                    if q is None:
                        stop = True
                        break
                    (later if any(o.n == q.n for o in batch) else batch).append(q)
                for q in later:
                    self.inbox.put(q)
                if stop:
                    self.inbox.put(None)
            t = time.perf_counter()
            for q in batch:
                q.start = t
            try:
                if ctx is None:
                    raise RuntimeError(f"no device context: {self.error!r}")
                if len(batch) == 1:
                    self._solve(ctx, resident, batch[0])
                else:
                    self._solve_round(ctx, resident, batch)
            except Exception as e:  # noqa: BLE001 - the requests fail, the worker goes on
                self.error = e
                for q in batch:
                    q.result = {"error": repr(e)}
            t = time.perf_counter()
            for q in batch:
                q.end = t
                self.done.put(q)
        for batch, b, x, _ in resident.values():
            batch.close()
            b.free()
            x.free()
        if ctx is not None:
            ctx.close()

    def _resident(self, ctx, resident, n):
        """The resident batch of block size n (built and uploaded on first use, least recently
        used ones evicted to stay within the worker's share of the budget)."""
        if n in resident:
            resident.move_to_end(n)
            return resident[n]
        model, mesh, rhs = self.problems.get(n)
        need = _block_bytes(mesh.neq, self.mixed)
        while resident and self.used + need > self.budget:  # admission within the block share
            _, (batch, b, x, nb) = resident.popitem(last=False)
            batch.close()
            b.free()
            x.free()
            self.used -= nb
        batch = core.Batch(ctx, mesh, mesh.params_for(model), fp32=self.mixed, tune=False)
        b = core.DeviceArray.from_host(ctx, rhs)
        x = core.DeviceArray(ctx, mesh.neq)
        resident[n] = (batch, b, x, need)
        self.used += need
        return resident[n]

    def _solve_round(self, ctx, resident, rs):
        got = [self._resident(ctx, resident, r.n) for r in rs]
        for _, _, x, _ in got:
            x.zero()
        res = core.cg_many(ctx, [g[0] for g in got], [g[1] for g in got], [g[2] for g in got], rtol=self.rtol)
        for r, g, rr in zip(rs, got, res):
            r.result = dict(rr)
            if self.keep_x:
                r.result["x"] = g[2].get()

    def _solve(self, ctx, resident, r):
        batch, b, x, _ = self._resident(ctx, resident, r.n)
        x.zero()
        if self.mixed:
            res = batch.cg_mixed(b, x, rtol=self.rtol, inner_rtol=self.inner_rtol)
        else:
            res = batch.cg(b, x, rtol=self.rtol)
        r.result = dict(res)
        if self.keep_x:
            r.result["x"] = x.get()


def _block_bytes(neq, mixed):
    """Device bytes of a resident block batch (estimate used for admission): operator data,
    solver work and the two request vectors."""
    return int(neq * (8 * 8 + (5 * 4 if mixed else 0) + 2 * 8) + neq // 3 * 16 + (1 << 20))


COLUMN_BYTES_PER_ELEMENT = 2 * core.PIMY_NSTATE * 8 + 9 * 8 + 7 * 3 * 8 + 8


class Server:
    """One GPU serving soil columns (a resident column pool) and static blocks (worker
    threads, each its own stream). hbm_budget: bytes for all resident state; column_share of
    it sizes the column pool, the rest is split between the block workers. Defaults measured on
    C5 (Poisson 200/s): the pool's solve sizes on its one stream (forked side streams share the
    process's hardware queues with the block solves and slow both), two block workers solving
    one block request at a time, four pool iterations between admissions."""

    def __init__(self, device=0, hbm_budget=200e9, column_share=0.5, block_workers=2, mixed=True, dt=0.005,
                 max_columns=4096, iterations_per_call=4, keep_results=False, max_record_steps=500,
                 one_stream=True, block_mixed=False, block_sizes=(), block_rounds=False, block_graphs=False):
        self.ctx = core.Context(device)
        self.mixed, self.iters, self.keep = mixed, iterations_per_call, keep_results
        pool_elems = int(column_share * hbm_budget / COLUMN_BYTES_PER_ELEMENT)
        pool_elems = min(pool_elems, max_columns * core.FX_COLUMN_MAX_ELEMS)
        self.pool = core.ColumnPool(self.ctx, max_elements=pool_elems, max_columns=max_columns, dt=dt,
                                    max_materials=10 * max_columns, max_record=max_columns * 10 * max_record_steps,
                                    mixed=mixed, one_stream=one_stream, tune=True)
        self.problems = BlockProblems()
        self.inbox, self.done = queue.Queue(), queue.Queue()
        per_worker = (1.0 - column_share) * hbm_budget / max(1, block_workers)
        for n in block_sizes:  # the host compile of the served block models, shared by the workers
            self.problems.get(n)
        self.workers = [BlockWorker(device, self.problems, self.inbox, self.done, per_worker, block_mixed,
                                    keep_x=keep_results, prewarm=block_sizes, rounds=block_rounds,
                                    graphs=block_graphs) for _ in range(block_workers)]
        for w in self.workers:
            w.start()
        for w in self.workers:
            w.ready.wait()

    def run(self, requests, log=None):
        """Serve `requests` (sorted by arrival, seconds from the start) in real time; returns
        per-request records (kind, n, arrival, start, end, latency, result)."""
        t0 = time.perf_counter()
        stats = collections.Counter()
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        resident = {}
        blocks_out = 0
        finished = []
        next_log = t0 + 5.0
        while pending or waiting or resident or blocks_out:
            now = time.perf_counter()
            while pending and t0 + pending[0].arrival <= now:
                r = pending.popleft()
                if r.kind == "column":
                    waiting.append(r)
                else:
                    self.inbox.put(r)
                    blocks_out += 1
            ta = time.perf_counter()
            while waiting:  # admission: the pool's slots and element room are the column share of HBM
                r = waiting[0]
                cid = self.pool.add(core.pimy_materials(r.mats), r.mat_of, r.nsteps, r.record, record=self.keep)
                if cid is None:
                    break
                waiting.popleft()
                r.start = time.perf_counter()
                resident[cid] = r
                stats["admitted"] += 1
            tb = time.perf_counter()
            stats["admit_s"] += tb - ta
            if resident:
                self.pool.run(self.iters)
                tc = time.perf_counter()
                stats["pool_s"] += tc - tb
                stats["pool_calls"] += 1
                stats["resident_sum"] += len(resident)
                for cid in self.pool.finished():
                    r = resident.pop(cid)
                    st = self.pool.status(cid)
                    r.result = {"status": st["status"], "steps": st["steps"],
                                "newton_iterations": st["newton_iterations"]}
                    if self.keep:
                        r.result.update(self.pool.fetch(cid))
                    r.end = time.perf_counter()
                    self.pool.remove(cid)
                    finished.append(r)
                stats["finish_s"] += time.perf_counter() - tc
            while True:
                try:
                    r = self.done.get_nowait()
                except queue.Empty:
                    break
                blocks_out -= 1
                finished.append(r)
            if not resident:
                if blocks_out and not waiting:
                    try:  # nothing on the pool: wait for a block or the next arrival
                        timeout = max(0.0, (t0 + pending[0].arrival) - time.perf_counter()) if pending else 0.05
                        r = self.done.get(timeout=min(timeout, 0.05))
                        blocks_out -= 1
                        finished.append(r)
                    except queue.Empty:
                        pass
                elif pending and not waiting:
                    time.sleep(max(0.0, min(0.002, t0 + pending[0].arrival - time.perf_counter())))
            if log is not None and time.perf_counter() >= next_log:
                next_log += 5.0
                log(f"[serve] t={time.perf_counter() - t0:.1f}s done {len(finished)} resident columns {len(resident)} "
                    f"waiting {len(waiting)} blocks out {blocks_out}")
        self.last_stats = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in stats.items()}
        self.last_stats["span_s"] = round(time.perf_counter() - t0, 3)
        if stats["pool_calls"]:
            self.last_stats["ms_per_pool_call"] = round(stats["pool_s"] / stats["pool_calls"] * 1e3, 3)
            self.last_stats["mean_resident"] = round(stats["resident_sum"] / stats["pool_calls"], 1)
        out = []
        for r in finished:
            out.append(dict(rid=r.rid, kind=r.kind, n=r.n, arrival=r.arrival, start=r.start - t0, end=r.end - t0,
                            latency=r.end - t0 - r.arrival, result=r.result))
        return sorted(out, key=lambda d: d["rid"])

    def close(self):
        for _ in self.workers:
            self.inbox.put(None)
        for w in self.workers:
            w.join(timeout=60)
        self.pool.close()
        self.ctx.close()


def summary(records, duration=None):
    """solves/s (completed over the span from the first arrival to the last completion),
    p50/p99 latency overall and per kind, failures."""
    lat = np.array([r["latency"] for r in records])
    span = max(r["end"] for r in records) - min(r["arrival"] for r in records)
    out = {"requests": len(records), "solves_per_s": round(len(records) / span, 2), "span_s": round(span, 3),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2), "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2)}
    for kind in ("column", "block"):
        sub = np.array([r["latency"] for r in records if r["kind"] == kind])
        if len(sub):
            out[kind] = {"count": int(len(sub)), "p50_ms": round(float(np.percentile(sub, 50)) * 1e3, 2),
                         "p99_ms": round(float(np.percentile(sub, 99)) * 1e3, 2)}
    out["failed"] = sum(1 for r in records if r["result"] is None or r["result"].get("status", 1) == 2 or
                        ("converged" in r["result"] and not r["result"]["converged"]) or "error" in r["result"])
    return out
