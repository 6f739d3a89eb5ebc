"""Serving workload helpers: the served kernel-interpolation request stream
and a thin wrapper over the native continuous-batching server
(include/mpla_server.h, libmpla.so).

Stream (one request per arrival): Poisson arrivals, N log-uniform in
[n_min, n_max], dimension 2 or 3, points uniform in the unit cube, kernel
Gaussian (length scale 0.05) or Matern-3/2 (length scale 0.1), nugget
uniform in [0.5, 1.0], 1 to 8 right-hand sides ~ N(0, 1), tolerance 1e-8.
All draws come from the counter-based generator (mpla_amd.inputs).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import inputs, mpla


def request_stream(count, rate, seed=1235, n_min=1024, n_max=65536, tolerance=1e-8, max_rhs=8):
    """List of request dicts (host arrays) with arrival times in seconds."""
    u = inputs.uniform(8 * count, seed, 500)
    reqs = []
    t = 0.0
    for i in range(count):
        a, b, c, d, e, f = u[8 * i: 8 * i + 6]
        t += -math.log(1.0 - a) / rate
        n = int(round(math.exp(math.log(n_min) + b * (math.log(n_max) - math.log(n_min)))))
        dim = 2 if c < 0.5 else 3
        gaussian = d < 0.5
        nugget = 0.5 + 0.5 * e
        nrhs = 1 + int(f * max_rhs)
        pts = inputs.uniform(dim * n, seed, 1000 + 2 * i)
        rhs = inputs.normal(n * nrhs, seed, 1001 + 2 * i)
        reqs.append({
            "n": n, "dim": dim, "kind": mpla.MPLA_KERNEL_GAUSSIAN if gaussian else mpla.MPLA_KERNEL_MATERN32,
            "length_scale": 0.05 if gaussian else 0.1, "nugget": nugget, "nrhs": nrhs, "tolerance": tolerance,
            "arrival": t, "points": pts, "rhs": rhs,
        })
    return reqs


class Server:
    """A native server instance (one HIP stream, its own device memory pool)."""

    def __init__(self, mode=mpla.MPLA_SERVE_MIXED_BATCHED, hbm_budget_bytes=None, max_active=None,
                 iterations_per_step=None, inner_tol_floor=None, max_iterations=None):
        self.lib = mpla.lib()
        cfg = mpla.mpla_server_config()
        self.lib.mpla_server_default_config(C.byref(cfg))
        cfg.mode = mode
        if hbm_budget_bytes is not None:
            cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        if max_active is not None:
            cfg.max_active = max_active
        if iterations_per_step is not None:
            cfg.iterations_per_step = iterations_per_step
        if inner_tol_floor is not None:
            cfg.inner_tol_floor = inner_tol_floor
        if max_iterations is not None:
            cfg.max_iterations = max_iterations
        self.cfg = cfg
        self.handle = self.lib.mpla_server_create(C.byref(cfg))

    def close(self):
        if self.handle:
            self.lib.mpla_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def peak_bytes(self):
        return self.lib.mpla_server_peak_bytes(self.handle)

    def run_trace(self, reqs, arrivals=True):
        """Replays `reqs` (request dicts) in real time; returns (results, solutions)."""
        count = len(reqs)
        arr = (mpla.mpla_solve_request * count)()
        keep = []
        sols = []
        for i, r in enumerate(reqs):
            pts = np.ascontiguousarray(r["points"], dtype=np.float64)
            rhs = np.ascontiguousarray(r["rhs"], dtype=np.float64)
            sol = np.zeros(r["n"] * r["nrhs"])
            keep += [pts, rhs]
            sols.append(sol)
            q = arr[i]
            q.points, q.rhs, q.solution = pts.ctypes.data, rhs.ctypes.data, sol.ctypes.data
            q.n, q.dim, q.kind = r["n"], r["dim"], r["kind"]
            q.length_scale, q.nugget, q.nrhs, q.tolerance = r["length_scale"], r["nugget"], r["nrhs"], r["tolerance"]
            q.arrival = r["arrival"] if arrivals else 0.0
        res = (mpla.mpla_solve_result * count)()
        self.lib.mpla_server_run_trace(self.handle, arr, count, res)
        del keep
        return [res[i] for i in range(count)], [s.reshape((r["n"], r["nrhs"]), order="F") for s, r in zip(sols, reqs)]


def latency_stats(results):
    lat = [r.t_done - r.t_arrival for r in results if r.status == mpla.MPLA_REQUEST_CONVERGED]
    if not lat:
        return {"p50_ms": float("nan"), "p99_ms": float("nan")}
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(np.mean(lat) * 1e3)}


def slowest(reqs, results, count=5):
    """The `count` requests of longest latency: (latency ms, n, dim, kind,
    nrhs, iterations, refinement steps, admission wait ms)."""
    rows = []
    for r, x in zip(reqs, results):
        if x.status != mpla.MPLA_REQUEST_CONVERGED:
            continue
        rows.append({"latency_ms": (x.t_done - x.t_arrival) * 1e3, "n": r["n"], "dim": r["dim"],
                     "kind": "gaussian" if r["kind"] == mpla.MPLA_KERNEL_GAUSSIAN else "matern32",
                     "nrhs": r["nrhs"], "iterations": x.iterations, "refinement_steps": x.refinement_steps,
                     "wait_ms": (x.t_admit - x.t_arrival) * 1e3})
    rows.sort(key=lambda d: -d["latency_ms"])
    return rows[:count]
