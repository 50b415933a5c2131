"""Serving loop: resident matrices, continuous batching of arriving solves.

Registered matrices (HLG double with their FP32 copy) and their Jacobi
preconditioners stay resident on the device.  Solve requests arrive over
time; each carries the index of a registered matrix and its right-hand side.
The loop admits queued requests in arrival order as soon as the batch engine
has a free slot and arena room for them (admission by device-memory
footprint), advances every admitted solve by one engine step per turn, and
retires solves as they converge, so requests arriving while others run join
the running batch at the next step.  The reference has no serving layer;
the per-request solve is psb_krylov's BiCGSTAB (or CG) with the DIAG
preconditioner to ||b - A x|| <= eps ||b|| (krylov/psb_dcgstab.f90,
krylov/psb_d_krylov_conv_mod.f90:279-356).

A request's right-hand side is generated on the device from (seed, stream)
with psb_hip_mv_fill_uniform (U[-1,1)), standing for a vector already
resident in device memory when the request arrives.
"""
import collections
import time

import numpy as np

from . import _native as N
from .batch import psb_hip_batch
from .vect import psb_d_hip_vect


class psb_hip_server:
    """Request loop over one batch engine (method "BICGSTAB" or "CG").  refine
    None picks restarted refinement for BiCGSTAB (fewer FP32 iterations than
    reliable updates on the convection-diffusion requests measured) and
    reliable updates for CG."""

    def __init__(self, method="BICGSTAB", max_systems=128, arena_bytes=4 << 30, check_every=4,
                 eps=1e-10, itmax=20000, inner_eps=1e-4, delta=1e-4, seed=4, refine=None,
                 use_graph=True):
        if refine is None:
            refine = "restart" if method.upper() == "BICGSTAB" else "reliable"
        self.eng = psb_hip_batch(method, max_systems=max_systems, arena_bytes=arena_bytes,
                                 check_every=check_every, delta=delta, inner_eps=inner_eps,
                                 refine=refine, use_graph=use_graph)
        self.eps, self.itmax, self.seed = eps, itmax, seed
        self.mats = []           # (engine mat id, rows, rhs staging vector, solution vector)

    def register(self, a, prec=None):
        """Register a matrix (and its DIAG preconditioner); returns its index."""
        mid = self.eng.register(a, prec)
        n = (a.a if hasattr(a, "a") else a).nr
        b = psb_d_hip_vect().all(n)
        x = psb_d_hip_vect().all(n)
        self.mats.append((mid, n, b, x))
        return len(self.mats) - 1

    def rhs_into(self, v, stream):
        """The right-hand side of request stream `stream` into device vector v."""
        N.call("psb_hip_mv_fill_uniform", v.dev, self.seed, int(stream))
        v.set_dev()

    def run(self, requests, keep_solutions=False, time_limit=120.0):
        """Serve requests = [(arrival seconds, matrix index, rhs stream)] in
        real time.  Returns one record per request: arrival, admitted, done,
        latency (done - arrival), iterations, outer, relres, status (0 when the
        time limit ended the run first) and, with keep_solutions, x."""
        order = sorted(range(len(requests)), key=lambda i: requests[i][0])
        rec = [{"arrival": float(requests[i][0]), "status": 0} for i in range(len(requests))]
        queue = collections.deque()
        live = {}                # engine slot -> (request, solution vector)
        nxt = 0
        t0 = time.perf_counter()
        while nxt < len(order) or queue or live:
            now = time.perf_counter() - t0
            if now > time_limit:
                break
            while nxt < len(order) and requests[order[nxt]][0] <= now:
                queue.append(order[nxt])
                nxt += 1
            while queue:
                r = queue[0]
                mi = requests[r][1]
                mid, n, b, x = self.mats[mi]
                if not self.eng.can_admit(mid):
                    if not live:
                        raise N.PsbError(N.OUTOFMEM, "psb_hip_server.run",
                                         f"request {r} does not fit in an empty engine")
                    break
                self.rhs_into(b, requests[r][2])
                sid = self.eng.submit(mid, b, eps=self.eps, itmax=self.itmax)
                queue.popleft()
                xo = psb_d_hip_vect().all(n) if keep_solutions else x
                live[sid] = (r, xo)
                rec[r]["admitted"] = time.perf_counter() - t0
            if not live:
                if nxt < len(order):
                    wait = requests[order[nxt]][0] - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            self.eng.step()
            done, _ = self.eng.poll()
            tnow = time.perf_counter() - t0
            for d in done:
                r, xo = live.pop(d["id"])
                self.eng.result(d["id"], xo)
                rec[r].update(done=tnow, latency=tnow - rec[r]["arrival"],
                              iterations=d["iterations"], outer=d["outer"], relres=d["relres"],
                              status=d["status"])
                if keep_solutions:
                    rec[r]["x"] = xo
        return rec

    def free(self):
        for _, _, b, x in self.mats:
            b.free()
            x.free()
        self.mats = []
        self.eng.free()


def poisson_arrivals(rate, count, nmat, seed=4, stream=0):
    """count arrival times of a Poisson process of the given rate (requests
    per second) with matrix indices uniform over nmat, from the counter-based
    host uniform stream (seed, stream) and (seed, stream + 1).  Returns
    [(t, matrix, rhs stream)] with rhs stream = 100000 * (stream + 1) + k."""
    u = np.empty(count)
    N.fn("psb_hip_fill_uniform_host")(seed, stream, count, u.ctypes.data)
    v = np.empty(count)
    N.fn("psb_hip_fill_uniform_host")(seed, stream + 1, count, v.ctypes.data)
    u01 = (u + 1.0) / 2.0                      # [0, 1)
    gaps = -np.log1p(-u01) / rate              # exponential inter-arrival times
    t = np.cumsum(gaps)
    mats = np.minimum((((v + 1.0) / 2.0) * nmat).astype(int), nmat - 1)
    return [(float(t[k]), int(mats[k]), 100000 * (stream + 1) + k) for k in range(count)]


def latency_summary(rec):
    """p50 / p99 / mean latency (ms, arrival to done) and solves per second over
    finished requests, with the latency split into queue wait (arrival to
    admission) and service (admission to done)."""
    fin = [r for r in rec if r.get("status", 0) != 0]
    if not fin:
        return {"finished": 0, "requests": len(rec)}
    lat = np.array([r["latency"] for r in fin]) * 1e3
    svc = np.array([r["done"] - r["admitted"] for r in fin]) * 1e3
    wait = np.array([r["admitted"] - r["arrival"] for r in fin]) * 1e3
    span = max(r["done"] for r in fin) - min(r["arrival"] for r in rec)
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)
    return {"requests": len(rec), "finished": len(fin),
            "converged": int(sum(r["status"] == 1 for r in fin)),
            "p50_ms": pct(lat, 50),
            "p99_ms": pct(lat, 99),
            "mean_ms": round(float(lat.mean()), 3),
            "service_p50_ms": pct(svc, 50), "service_p99_ms": pct(svc, 99),
            "queue_wait_p50_ms": pct(wait, 50), "queue_wait_p99_ms": pct(wait, 99),
            "solves_per_s": round(len(fin) / span, 2) if span > 0 else None,
            "iterations_mean": round(float(np.mean([r["iterations"] for r in fin])), 1),
            "max_relres": max(r["relres"] for r in fin)}


def serve_one_at_a_time(systems, requests, method="BICGSTAB", eps=1e-10, itmax=20000, seed=4,
                        time_limit=120.0):
    """Baseline: the same arrival stream served one request at a time, in
    arrival order, by the FP64 device-resident solver (psb_krylov_fused).
    systems = [(matrix, DIAG preconditioner, descriptor)].  Returns records
    as psb_hip_server.run."""
    from .krylov import psb_krylov_fused
    from .vect import psb_d_vect_type
    order = sorted(range(len(requests)), key=lambda i: requests[i][0])
    rec = [{"arrival": float(requests[i][0]), "status": 0} for i in range(len(requests))]
    stage = {}
    t0 = time.perf_counter()
    for r in order:
        t_arr, mi, stream = requests[r]
        now = time.perf_counter() - t0
        if now > time_limit:
            break
        if t_arr > now:
            time.sleep(t_arr - now)
        a, prec, desc = systems[mi]
        n = (a.a if hasattr(a, "a") else a).nr
        if mi not in stage:
            stage[mi] = (psb_d_hip_vect().all(n), psb_d_hip_vect().all(n))
        b, x = stage[mi]
        N.call("psb_hip_mv_fill_uniform", b.dev, seed, int(stream))
        b.set_dev()
        N.call("psb_hip_mv_setscal_d", 0.0, 1, n, 1, x.dev)
        x.set_dev()
        it, err, _ = psb_krylov_fused(method, a, prec, psb_d_vect_type(b), psb_d_vect_type(x), eps,
                                      desc, itmax=itmax, check_every=8)
        tnow = time.perf_counter() - t0
        rec[r].update(admitted=max(t_arr, now), done=tnow, latency=tnow - t_arr, iterations=it,
                      outer=0, relres=err, status=1 if err <= eps else 2)
    for b, x in stage.values():
        b.free()
        x.free()
    return rec
