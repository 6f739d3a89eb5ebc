#!/usr/bin/env python3
"""Benchmark of the solver hot path on one MI355X.

Workload (the headline metric's configuration): matrix-free Gaussian kernel
system (K + 1.0 I) x = b, N = 65536 points uniform in [0,1]^3, length scale
0.05, b ~ U[-1,1], x0 = 0, CG to relative residual 1e-8, entries never stored.
A step is one complete solve through the drop-in API
(mpla_kernel_conjugate_gradient_mixed: FP32 entry evaluation inside FP64
iterative refinement); the plain FP64 solve (the reference algorithm) is timed
beside it.  Inputs are generated on the device by the counter-based RNG and
stay resident in HBM; x is reset to 0 inside each timed step (as the
reference's benchmark_kernel_cg does).

Also reported: p50/p99 solve latency, the dominant kernel's roofline (its
average launch time measured with HIP events on the library's stream), dense
dgemv HBM bandwidth at N = 65536 (a 34 GB matrix), and the CPU oracle's FP64
rate on a bounded sample, timed on this host's cores.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 65536
DIM = 3
ELL = 0.05
NUGGET = 1.0
EPS = 1e-8
SEED = 1235
ITER_MAX = 20001
FLOP_PER_ENTRY = 3 * DIM + 4  # d subs, d FMA, 1 scale, 1 exp, 1 FMA accumulate
DROP_LOG2 = 15.0  # the mixed solver's default drop level at its default inner tolerance 1e-4
RELAX_LEVELS = (15.0, 10.0, 6.0)  # its relaxed inner-product screens (inner residual > 2^-5, > 2^-9, below)
# its FP64 residual screen for this solve: log2 N + log2(1/nugget) + log2(1/eps) + 12, rounded up (55)
RESIDUAL_DROP_LOG2 = float(math.ceil(math.log2(N) - math.log2(NUGGET) - math.log2(EPS) + 12.0))
PEAK_FP32_NONPACKED = 78.6e12  # one FP32 FMA per lane per cycle; the 157.3 TF figure counts packed FP32
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def sym_pairs(n, tile=512):
    """Entry evaluations of the symmetric product: every 512 x 512 tile pair
    I <= J once (diagonal tiles evaluate their full square)."""
    nt = (n + tile - 1) // tile
    return tile * tile * nt * (nt + 1) // 2


COUNTERS = os.path.join(ROOT, "profiles", "counters_kernel_matvec_gaussian_n65536.md")


def committed_traffic(name):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE x its copy-kernel calibration + WRITE_SIZE), recorded
    there as a line `traffic[<name>] = <bytes> bytes`; None when absent."""
    try:
        text = open(COUNTERS).read()
    except OSError:
        return None
    import re
    m = re.search(r"traffic\[" + re.escape(name) + r"\] = (\d+) bytes", text)
    return int(m.group(1)) if m else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def _event_ms(ext, fn, reps):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(ext)
    for _ in range(reps):
        fn()
    e1.record(ext)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_dense_cg(K, L, inst, stream_ptr, ext):
    """Dense FP64 CG: A = B B^T/N + I, N = 4096, b ~ U[-1,1], x0 = 0, relative 1e-10,
    max 2000 iterations; and 100 dgemv y = A x (the 128 MiB matrix stays in the
    256 MB Infinity Cache, so this is cache bandwidth, not HBM)."""
    import numpy as np
    import torch
    from mpla_amd import _native, mpla
    n = 4096
    Bm = torch.empty(n * n, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(Bm.data_ptr(), n * n, 1235, 10, 0, -1.0, 1.0, stream_ptr))
    torch.cuda.synchronize()
    Bm = Bm.view(n, n).t().contiguous()  # B[i, j] = element i + j*n of the generated column-major array
    A = mpla.mpla_matrix()
    L.mpla_init_matrix(C.byref(A), C.byref(inst), n, n)
    # A = B B^T / N + I, symmetrised exactly so its row- and column-major layouts coincide (input setup, untimed)
    M = Bm @ Bm.t() / n
    Ad = (0.5 * (M + M.t()) + torch.eye(n, dtype=torch.float64, device="cuda")).contiguous()
    del M
    torch.cuda.synchronize()
    assert K.mplak_copy_f64(A.data, Ad.data_ptr(), n * n, stream_ptr) == 0
    b = mpla.mpla_vector()
    x = mpla.mpla_vector()
    y = mpla.mpla_vector()
    for v in (b, x, y):
        L.mpla_init_vector(C.byref(v), C.byref(inst), n)
    _native.check(K.mplak_fill_uniform_f64(b.data, n, 1235, 11, 0, -1.0, 1.0, stream_ptr))
    core = mpla.core("mpla_generic_dgemv_core_dense")
    Ag = C.cast(C.pointer(A), C.POINTER(mpla.mpla_generic_matrix))

    def solve():
        _native.check(K.mplak_fill_f64(x.data, n, 0.0, stream_ptr))
        L.mpla_generic_conjugate_gradient_with_checkpoint_restart(C.byref(b), Ag, C.byref(x), 2001, 1e-10, core, 0,
                                                                  False, C.byref(inst))
    solve()
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        solve()
    t = (time.perf_counter() - t0) / reps
    it, rel = inst.last_iterations, inst.last_rel_res
    gemv_ms = _event_ms(ext, lambda: L.mpla_dgemv(C.byref(y), C.byref(A), C.byref(x), C.byref(inst)), 100)
    for v in (b, x, y):
        L.mpla_free_vector(C.byref(v), C.byref(inst))
    L.mpla_free_matrix(C.byref(A), C.byref(inst))
    del Ad, Bm
    torch.cuda.empty_cache()
    log(f"dense CG n=4096: {t * 1e3:.3f} ms/solve ({it} iterations, rel {rel:.2e}); dgemv {gemv_ms * 1e3:.1f} us")
    return {"solves_per_s": 1.0 / t, "ms_per_solve": t * 1e3, "iterations": it, "rel_residual": rel,
            "dgemv_us": gemv_ms * 1e3, "dgemv_gbps_cache": (8.0 * n * n + 16 * n) / (gemv_ms * 1e-3) / 1e9}


def bench_dense_large(K, L, inst, stream_ptr, ext):
    """dgemv at N = 131072 (137 GB matrix) and BiCGSTAB on A = D + R, N = 32768,
    R ~ U[-1,1]/sqrt(N), D = diag(2 + U[0,1]), relative 1e-10 (reference
    3-product iteration with the true-residual test)."""
    import torch
    from mpla_amd import _native, mpla
    out = {}
    n = 131072
    try:
        Ad = torch.empty(n * n, dtype=torch.float64, device="cuda")
    except RuntimeError as exc:  # pragma: no cover - device without 140 GB free
        out["dgemv_131072"] = f"skipped: {exc}"
        Ad = None
    if Ad is not None:
        xd = torch.empty(n, dtype=torch.float64, device="cuda")
        yd = torch.empty(n, dtype=torch.float64, device="cuda")
        _native.check(K.mplak_fill_uniform_f64(Ad.data_ptr(), n * n, 1234, 1, 0, -1.0, 1.0, stream_ptr))
        _native.check(K.mplak_fill_uniform_f64(xd.data_ptr(), n, 1234, 2, 0, -1.0, 1.0, stream_ptr))
        gwb = K.mplak_dgemv_workspace_bytes(n, n)
        gw = torch.empty(gwb // 8 + 1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        f = lambda: _native.check(K.mplak_dgemv_f64(n, n, Ad.data_ptr(), n, xd.data_ptr(), yd.data_ptr(),
                                                    gw.data_ptr(), gwb, None, stream_ptr))
        f()
        ms = _event_ms(ext, f, 10)
        gb = 8.0 * n * n + 16.0 * n
        out["dgemv_131072"] = {"ms": ms, "gbps": gb / (ms * 1e-3) / 1e9, "frac_hbm": gb / (ms * 1e-3) / PEAK_HBM}
        log(f"dgemv n={n}: {ms:.3f} ms, {gb / (ms * 1e-3) / 1e9:.1f} GB/s")
        del Ad, xd, yd, gw
        torch.cuda.empty_cache()
    n = 32768
    A = mpla.mpla_matrix()
    L.mpla_init_matrix(C.byref(A), C.byref(inst), n, n)
    _native.check(K.mplak_fill_uniform_f64(A.data, n * n, 1235, 20, 0, -1.0, 1.0, stream_ptr))
    torch.cuda.synchronize()
    # scale R by 1/sqrt(N) and add the diagonal (input setup, untimed)
    Dg = torch.empty(n, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(Dg.data_ptr(), n, 1235, 21, 0, 2.0, 3.0, stream_ptr))
    Mat = torch.empty(n * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert K.mplak_copy_f64(Mat.data_ptr(), A.data, n * n, stream_ptr) == 0
    torch.cuda.synchronize()
    Mv = Mat.view(n, n)
    Mv.mul_(1.0 / math.sqrt(n))
    Mv.diagonal().add_(Dg)
    torch.cuda.synchronize()
    assert K.mplak_copy_f64(A.data, Mat.data_ptr(), n * n, stream_ptr) == 0
    torch.cuda.synchronize()
    del Mat, Mv, Dg
    torch.cuda.empty_cache()
    b = mpla.mpla_vector()
    x = mpla.mpla_vector()
    for v in (b, x):
        L.mpla_init_vector(C.byref(v), C.byref(inst), n)
    _native.check(K.mplak_fill_uniform_f64(b.data, n, 1235, 22, 0, -1.0, 1.0, stream_ptr))
    core = mpla.core("mpla_generic_dgemv_core_dense")
    Ag = C.cast(C.pointer(A), C.POINTER(mpla.mpla_generic_matrix))

    def solve():
        _native.check(K.mplak_fill_f64(x.data, n, 0.0, stream_ptr))
        L.mpla_generic_BiCGSTAB_with_checkpoint_restart(C.byref(b), Ag, C.byref(x), 2001, 1e-10, core, 0, False,
                                                        C.byref(inst))
    solve()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        solve()
    t = (time.perf_counter() - t0) / reps
    it, rel, mv = inst.last_iterations, inst.last_rel_res, inst.last_matvecs
    gb = mv * 8.0 * n * n
    out["bicgstab_32768"] = {"ms_per_solve": t * 1e3, "solves_per_s": 1.0 / t, "iterations": it,
                             "rel_residual": rel, "products": mv, "effective_gbps": gb / t / 1e9}
    log(f"BiCGSTAB n={n}: {t * 1e3:.2f} ms/solve, {it} iterations, rel {rel:.2e}, {gb / t / 1e9:.0f} GB/s eff")
    for v in (b, x):
        L.mpla_free_vector(C.byref(v), C.byref(inst))
    L.mpla_free_matrix(C.byref(A), C.byref(inst))
    return out


def bench_block(K, L, inst, stream_ptr, args):
    """Multi-RHS Matern-3/2 system: N = 131072 points U[0,1]^3, l = 0.1, nugget
    0.5, 64 right-hand sides ~ N(0,1), block CG with the shared K*V product in
    the given precision inside FP64 refinement to 1e-8."""
    import numpy as np
    import torch
    from mpla_amd import _native, inputs, mpla
    n, R = 131072, 64
    P = torch.empty(3, n, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(P.data_ptr(), 3 * n, 1235, 30, 0, 0.0, 1.0, stream_ptr))
    table = torch.tensor([P[d].data_ptr() for d in range(3)], dtype=torch.int64, device="cuda")
    Bh = inputs.normal(n * R, 1235, 31)
    Bd = torch.as_tensor(Bh, device="cuda")
    Xd = torch.zeros(n * R, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    kd = mpla.mpla_kernel_matrix_data(table.data_ptr(), 3, 0, mpla.MPLA_KERNEL_MATERN32, mpla.MPLA_PREC_F64, 0.1, 0.5,
                                      0.0)
    A = mpla.mpla_generic_matrix()
    L.mpla_init_generic_matrix(C.byref(A), C.byref(inst), n, n)
    A.data = C.addressof(kd)
    out = {}
    names = {"bf16": mpla.MPLA_PREC_BF16, "bf16x3": mpla.MPLA_PREC_BF16X3, "f32": mpla.MPLA_PREC_F32,
             "f64": mpla.MPLA_PREC_F64}
    sols = {}
    for name in [p for p in args.block_precisions.split(",") if p]:
        Xd.zero_()
        torch.cuda.synchronize()
        rel = (C.c_double * R)()
        t0 = time.perf_counter()
        L.mpla_kernel_block_conjugate_gradient_mixed(Bd.data_ptr(), Xd.data_ptr(), R, C.byref(A), 20000, 1e-8,
                                                     names[name], rel, C.byref(inst))
        t = time.perf_counter() - t0
        sols[name] = Xd.cpu().numpy().reshape((n, R), order="F")
        out[name] = {"seconds": t, "rhs_solves_per_s": R / t, "shared_products": inst.last_iterations,
                     "max_true_rel_residual": max(rel), "converged": bool(inst.last_converged)}
        log(f"Matern block solve {name}: {t:.2f} s for {R} RHS, {inst.last_iterations} shared products, max rel {max(rel):.2e}")
    if args.block_oracle_columns and sols:
        x = mpla.mpla_vector()
        bv = mpla.mpla_vector()
        L.mpla_init_vector(C.byref(x), C.byref(inst), n)
        L.mpla_init_vector(C.byref(bv), C.byref(inst), n)
        errs = {k: [] for k in sols}
        for c in (0, 21, 42, 63):
            _native.check(K.mplak_copy_f64(bv.data, Bd.data_ptr() + 8 * c * n, n, stream_ptr))
            _native.check(K.mplak_fill_f64(x.data, n, 0.0, stream_ptr))
            L.mpla_kernel_conjugate_gradient_mixed(C.byref(bv), C.byref(A), C.byref(x), 20000, 1e-12,
                                                   mpla.MPLA_PREC_F64, C.byref(inst))
            xo = np.empty(n)
            L.mpla_copy_distributed_vector_to_cpu(xo.ctypes.data, C.byref(x), C.byref(inst))
            for k, X in sols.items():
                errs[k].append(float(np.linalg.norm(X[:, c] - xo) / np.linalg.norm(xo)))
        for k in sols:
            out[k]["rel_error_vs_fp64_columns_0_21_42_63"] = errs[k]
        L.mpla_free_vector(C.byref(x), C.byref(inst))
        L.mpla_free_vector(C.byref(bv), C.byref(inst))
    L.mpla_free_generic_matrix(C.byref(A), C.byref(inst))
    del P, Bd, Xd
    torch.cuda.empty_cache()
    return out


PUBLISHED_KERNEL_CG_S = 364.6644  # reference benchmark_kernel_cg, N = 2^17, 5 trials, P100-class GPU (2017)


def bench_reference_kernel_cg(K, L, inst, stream_ptr, ext):
    """The reference's published benchmark on the drop-in API: matrix-free
    kernel CG with the reference core (exp(-r^2), +1e-8 on coincident
    points) at N = 131072 points U[0,1]^3, b = A x_true, 5 trials of
    mpla_generic_conjugate_gradient(iter_max 50, eps 1e-15) from the same
    random x0 (restored after each trial): 250 FP64 matrix-free products, as
    src/benchmark_kernel_cg.cpp:40-140 with trials = 5."""
    import torch
    from mpla_amd import _native, mpla
    n, dim, trials = 131072, 3, 5
    P = torch.empty(dim, n, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(P.data_ptr(), dim * n, 1235, 30, 0, 0.0, 1.0, stream_ptr))
    table = torch.tensor([P[d].data_ptr() for d in range(dim)], dtype=torch.int64, device="cuda")
    kd = mpla.kernel_matrix_data(table.data_ptr(), dim, n)
    A = mpla.mpla_generic_matrix()
    L.mpla_init_generic_matrix(C.byref(A), C.byref(inst), n, n)
    A.data = C.addressof(kd)
    x, x0, b = mpla.mpla_vector(), mpla.mpla_vector(), mpla.mpla_vector()
    for v in (x, x0, b):
        L.mpla_init_vector(C.byref(v), C.byref(inst), n)
    core = mpla.core("mpla_dgemv_core_kernel_streamed_cublas")
    _native.check(K.mplak_fill_uniform_f64(x.data, n, 1235, 31, 0, 0.0, 1.0, stream_ptr))
    L.mpla_generic_dgemv(C.byref(b), C.byref(A), C.byref(x), core, C.byref(inst))
    _native.check(K.mplak_fill_uniform_f64(x0.data, n, 1235, 32, 0, 0.0, 1.0, stream_ptr))
    _native.check(K.mplak_copy_f64(x.data, x0.data, n, stream_ptr))
    torch.cuda.synchronize()
    products = 0
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(ext)
    for _ in range(trials):
        L.mpla_generic_conjugate_gradient(C.byref(b), C.byref(A), C.byref(x), 50, 1.0e-15, core, C.byref(inst))
        products += inst.last_matvecs
        _native.check(K.mplak_copy_f64(x.data, x0.data, n, stream_ptr))
    e1.record(ext)
    e1.synchronize()
    t = e0.elapsed_time(e1) * 1e-3
    for v in (x, x0, b):
        L.mpla_free_vector(C.byref(v), C.byref(inst))
    L.mpla_free_generic_matrix(C.byref(A), C.byref(inst))
    del P, table
    torch.cuda.empty_cache()
    log(f"reference benchmark_kernel_cg n={n}: {t:.3f} s for {products} FP64 products "
        f"(published: {PUBLISHED_KERNEL_CG_S} s on a P100-class GPU)")
    return {"n": n, "trials": trials, "seconds": t, "fp64_products": products, "ms_per_product": t / products * 1e3,
            "entries_per_s": float(n) * n * products / t, "published_seconds": PUBLISHED_KERNEL_CG_S,
            "published_hardware": "P100-class GPU (Piz Daint, 2017), different hardware",
            "time_ratio_published_over_this": PUBLISHED_KERNEL_CG_S / t}


def bench_served(args):
    """Served stream: Poisson arrivals of kernel-interpolation solves (N
    log-uniform in [1024, 65536], d in {2,3}, Gaussian l=0.05 or Matern-3/2
    l=0.1, nugget U[0.5,1], 1-8 RHS, tolerance 1e-8, HBM budget 200 GB).  The
    arrival rate is the plain FP64 one-at-a-time path's measured capacity
    (the first 48 requests offered at t=0), so the plain path is offered ~1x
    its capacity."""
    from mpla_amd import mpla, serving
    count = args.served_requests
    reqs = serving.request_stream(count, rate=1.0, seed=1235)
    arrivals = [r["arrival"] for r in reqs]
    # capacity of the plain path: a prefix of the stream offered at once
    head = reqs[:min(count, 48)]
    with serving.Server(mode=mpla.MPLA_SERVE_FP64_SERIAL) as srv:
        t0 = time.perf_counter()
        res_sat, _ = srv.run_trace([dict(r, arrival=0.0) for r in head])
        cap = len(head) / (time.perf_counter() - t0)
    rate = cap
    for r, a in zip(reqs, arrivals):
        r["arrival"] = a / rate  # rescale unit-rate arrivals to `rate`
    out = {"requests": count, "arrival_rate_per_s": rate, "plain_capacity_solves_per_s": cap,
           "capacity_measured_on": f"the first {len(head)} requests offered at t=0",
           "mean_n": float(sum(r["n"] for r in reqs) / count)}
    for name, mode in (("mixed_batched", mpla.MPLA_SERVE_MIXED_BATCHED), ("fp64_serial", mpla.MPLA_SERVE_FP64_SERIAL)):
        with serving.Server(mode=mode) as srv:
            t0 = time.perf_counter()
            res, _ = srv.run_trace(reqs)
            span = time.perf_counter() - t0
        ok = sum(1 for x in res if x.status == mpla.MPLA_REQUEST_CONVERGED)
        st = serving.latency_stats(res)
        out[name] = {"solves_per_s": ok / span, "makespan_s": span, "converged": ok,
                     "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st.get("mean_ms"),
                     "max_true_rel_residual": max(x.max_rel_residual for x in res),
                     "iterations": sum(x.iterations for x in res)}
        log(f"served stream {name}: {ok}/{count} converged, {ok / span:.2f} solves/s, p50 {st['p50_ms']:.1f} ms, "
            f"p99 {st['p99_ms']:.1f} ms")
    return out


def product_desc(table):
    from mpla_amd import _native
    desc = _native.mplak_kernel_desc()
    desc.points, desc.dim, desc.kind = table.data_ptr(), DIM, _native.MPLAK_KERNEL_GAUSSIAN
    desc.length_scale, desc.nugget, desc.coincident_shift = ELL, NUGGET, 0.0
    desc.row_offset, desc.row_count, desc.col_offset, desc.col_count = 0, N, 0, N
    return desc


def ordered_problem(K, table, stream_ptr, levels=(DROP_LOG2,)):
    """The solver's inner operator: the points in Hilbert order (as the mixed
    solver arranges them) and a screening plan at each given drop level."""
    import torch
    from mpla_amd import _native
    desc = product_desc(table)
    perm = torch.empty(N, dtype=torch.int64, device="cuda")
    ob = K.mplak_spatial_order_workspace_bytes(N, DIM)
    ow = torch.empty(ob // 8 + 1, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_spatial_order(table.data_ptr(), DIM, 0, N, perm.data_ptr(), ow.data_ptr(), ob, stream_ptr))
    pts = torch.empty(DIM, N, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_gather_points(table.data_ptr(), DIM, 0, perm.data_ptr(), N, pts.data_ptr(), N, stream_ptr))
    otable = torch.tensor([pts[d].data_ptr() for d in range(DIM)], dtype=torch.int64, device="cuda")
    odesc = product_desc(otable)
    pb = K.mplak_screen_plan_bytes(C.byref(odesc))
    plans = []
    for level in levels:
        plan = torch.empty(pb // 8 + 1, dtype=torch.float64, device="cuda")
        _native.check(K.mplak_screen_plan_build(C.byref(odesc), _native.MPLAK_PREC_F32, level, plan.data_ptr(), pb,
                                                stream_ptr))
        plans.append(plan)
    torch.cuda.synchronize()
    return {"desc": odesc, "plan": plans[0], "plans": plans, "levels": levels, "plan_bytes": pb,
            "keep": (pts, otable, perm)}


def measure_inner_product(K, table, xv, yv, stream_ptr, ext, reps=20):
    """Launch time (HIP events on the library stream) of the planned screened
    FP32 product, the kernel that dominates a mixed solve, at each of the
    solver's relaxed screens, launched as the solver does (one plan per
    launch, with the single-reduction CG's dot partials), and its executed
    work: the plan's count of evaluated entries (256 per kept 16 x 16 tile on
    compact target tiles, 8192 per kept 128 x 64 block on the others), 15 FLOP
    each (13 + the second accumulation) off the diagonal tiles, 13 on them."""
    import torch
    op = ordered_problem(K, table, stream_ptr, RELAX_LEVELS)
    out = [measure_level(K, op, lvl, plan, xv, yv, stream_ptr, ext, reps) for lvl, plan in zip(op["levels"],
                                                                                             op["plans"])]
    return out[0], out


def measure_level(K, op, level, plan, xv, yv, stream_ptr, ext, reps):
    import torch
    from mpla_amd import _native
    d, pb = op["desc"], op["plan_bytes"]
    nd = K.mplak_planned_dots_count(N)
    dots = torch.empty(2 * nd, dtype=torch.float64, device="cuda")
    dl = (C.c_double * 1)(level)
    pl = (C.c_void_p * 1)(plan.data_ptr())
    counts = torch.zeros(5, dtype=torch.int64, device="cuda")
    _native.check(K.mplak_screen_plan_counts(C.byref(d), plan.data_ptr(), pb, counts.data_ptr(), stream_ptr))
    torch.cuda.synchronize()
    kept_pairs, kept, diag, evals_off, evals_diag = (int(v) for v in counts.cpu())

    def launch():
        _native.check(K.mplak_kernel_matvec_planned_relaxed(C.byref(d), 1, dl, pl, pb, None, None, xv.data_ptr(),
                                                            yv.data_ptr(), dots.data_ptr(), None, stream_ptr))

    for _ in range(2):
        launch()
    # one event pair around `reps` back-to-back launches (kernel + finish), as
    # the solver issues them: the average launch duration with the short gaps
    # between the two kernels but no event packets inside the timed span
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(ext)
    for _ in range(reps):
        launch()
    e1.record(ext)
    e1.synchronize()
    t = e0.elapsed_time(e1) / reps * 1e-3
    nt = (N + 511) // 512
    flop = evals_off * (FLOP_PER_ENTRY + 2.0) + evals_diag * FLOP_PER_ENTRY
    return {"kernel": "kmv_sym_planned_kernel<GAUSSIAN, 3> + kmv_sym_finish_kernel",
            "launch_ms": t * 1e3, "drop_log2": level, "kept_pairs": kept_pairs,
            "all_pairs": nt * (nt + 1) // 2, "kept_blocks": kept, "all_blocks": 32 * nt * (nt + 1) // 2,
            "evaluated_entries": evals_off + evals_diag,
            "executed_flop": flop, "executed_tflops": flop / t / 1e12,
            "reference_equivalent_tflops": FLOP_PER_ENTRY * float(N) * N / t / 1e12}


def profile_kernels(K, table, stream_ptr):
    """Launch the measured kernels a fixed number of times, for counter runs:
    3 FP32 and 2 FP64 exact Gaussian products (N=65536), 3 planned screened
    FP32 products on the ordered points at each relaxed screen (2^-15, then
    2^-10, then 2^-6), 2 planned FP64 products at the solver's residual
    screen (2^-55), 3 dgemv (65536 x 65536) and 3 copies of 2^27 doubles
    (1 GiB read + 1 GiB written, the traffic calibration)."""
    import torch
    from mpla_amd import _native
    desc = product_desc(table)
    wb = K.mplak_kernel_matvec_workspace_bytes(C.byref(desc))
    work = torch.empty(wb // 8 + 1, dtype=torch.float64, device="cuda")
    xv = torch.empty(N, dtype=torch.float64, device="cuda")
    yv = torch.empty(N, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(xv.data_ptr(), N, SEED, 2, 0, -1.0, 1.0, stream_ptr))
    for prec, reps in ((_native.MPLAK_PREC_F32, 3), (_native.MPLAK_PREC_F64, 2)):
        for _ in range(reps):
            _native.check(K.mplak_kernel_matvec(C.byref(desc), prec, xv.data_ptr(), yv.data_ptr(), work.data_ptr(),
                                                wb, None, stream_ptr))
    torch.cuda.synchronize()
    op = ordered_problem(K, table, stream_ptr, RELAX_LEVELS)
    nd = K.mplak_planned_dots_count(N)
    dots = torch.empty(2 * nd, dtype=torch.float64, device="cuda")
    for level, plan in zip(op["levels"], op["plans"]):
        dl = (C.c_double * 1)(level)
        pl = (C.c_void_p * 1)(plan.data_ptr())
        for _ in range(3):
            _native.check(K.mplak_kernel_matvec_planned_relaxed(C.byref(op["desc"]), 1, dl, pl, op["plan_bytes"], None,
                                                                None, xv.data_ptr(), yv.data_ptr(), dots.data_ptr(),
                                                                None, stream_ptr))
    plan64 = torch.empty(op["plan_bytes"] // 8 + 1, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_screen_plan_build(C.byref(op["desc"]), _native.MPLAK_PREC_F64, RESIDUAL_DROP_LOG2,
                                            plan64.data_ptr(), op["plan_bytes"], stream_ptr))
    for _ in range(2):
        _native.check(K.mplak_kernel_matvec_planned(C.byref(op["desc"]), _native.MPLAK_PREC_F64, RESIDUAL_DROP_LOG2,
                                                    plan64.data_ptr(), op["plan_bytes"], xv.data_ptr(), yv.data_ptr(),
                                                    None, stream_ptr))
    torch.cuda.synchronize()
    del op, plan64
    log("products done")
    del work
    n = 65536
    Ad = torch.empty(n * n, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(Ad.data_ptr(), n * n, 1234, 1, 0, -1.0, 1.0, stream_ptr))
    gwb = K.mplak_dgemv_workspace_bytes(n, n)
    gw = torch.empty(gwb // 8 + 1, dtype=torch.float64, device="cuda")
    for _ in range(3):
        _native.check(K.mplak_dgemv_f64(n, n, Ad.data_ptr(), n, xv.data_ptr(), yv.data_ptr(), gw.data_ptr(), gwb,
                                        None, stream_ptr))
    torch.cuda.synchronize()
    log("dgemv done")
    cn = 1 << 27
    src = Ad[:cn]
    dst = Ad[cn:2 * cn]
    for _ in range(3):
        _native.check(K.mplak_copy_f64(dst.data_ptr(), src.data_ptr(), cn, stream_ptr))
    torch.cuda.synchronize()
    log("copies done")
    print(json.dumps({"kernels_only": True, "product_n": N, "dgemv_n": n, "copy_bytes_each_way": 8 * cn}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-matvecs", type=int, default=2)
    ap.add_argument("--skip-dgemv", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-configs", action="store_true", help="only the headline workload")
    ap.add_argument("--served-requests", type=int, default=256)
    ap.add_argument("--block-precisions", default="bf16x3,bf16")
    ap.add_argument("--block-oracle-columns", action="store_true",
                    help="also solve columns 0,21,42,63 with plain FP64 CG for the error report")
    ap.add_argument("--kernels-only", action="store_true",
                    help="launch only the measured kernels a few times (for rocprofv3 counter runs) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path runs on a 1x1 "
              "process grid and multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mpla_amd import _native, mpla

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    K = _native.mplak()
    L = mpla.lib()
    inst = mpla.mpla_instance()
    L.mpla_init_instance(C.byref(inst), mpla.MPI_COMM_WORLD)
    inst.verbosity = 0
    stream_ptr = L.mpla_instance_stream(C.byref(inst))
    ext = torch.cuda.ExternalStream(stream_ptr)

    # ---- inputs, generated in HBM ------------------------------------------
    P = torch.empty(DIM, N, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(P.data_ptr(), DIM * N, SEED, 0, 0, 0.0, 1.0, stream_ptr))
    table = torch.tensor([P[d].data_ptr() for d in range(DIM)], dtype=torch.int64, device="cuda")
    b = mpla.mpla_vector()
    x = mpla.mpla_vector()
    L.mpla_init_vector(C.byref(b), C.byref(inst), N)
    L.mpla_init_vector(C.byref(x), C.byref(inst), N)
    _native.check(K.mplak_fill_uniform_f64(b.data, N, SEED, 1, 0, -1.0, 1.0, stream_ptr))
    torch.cuda.synchronize()
    kd = mpla.mpla_kernel_matrix_data(table.data_ptr(), DIM, 0, mpla.MPLA_KERNEL_GAUSSIAN, mpla.MPLA_PREC_F64,
                                      ELL, NUGGET, 0.0)
    A = mpla.mpla_generic_matrix()
    L.mpla_init_generic_matrix(C.byref(A), C.byref(inst), N, N)
    A.data = C.addressof(kd)

    def reset_x():
        _native.check(K.mplak_fill_f64(x.data, N, 0.0, stream_ptr))

    if args.kernels_only:
        profile_kernels(K, table, stream_ptr)
        return

    def solve_mixed():
        reset_x()
        L.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), ITER_MAX, EPS,
                                               mpla.MPLA_PREC_F32, C.byref(inst))
        return inst.last_iterations, inst.last_rel_res, inst.last_converged, inst.last_matvecs

    def solve_plain():
        reset_x()
        L.mpla_generic_conjugate_gradient_with_checkpoint_restart(
            C.byref(b), C.byref(A), C.byref(x), ITER_MAX, EPS, mpla.core("mpla_dgemv_core_kernel_fused"), 0, False,
            C.byref(inst))
        return inst.last_iterations, inst.last_rel_res, inst.last_converged, inst.last_matvecs

    def read_x():
        out = np.empty(N)
        L.mpla_copy_distributed_vector_to_cpu(out.ctypes.data, C.byref(x), C.byref(inst))
        return out

    # ---- mixed precision (the product path) ---------------------------------
    log(f"warmup: {args.warmup} mixed solves")
    for _ in range(args.warmup):
        solve_mixed()
    torch.cuda.synchronize()
    lat = []
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        infos.append(solve_mixed())
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    x_mixed = read_x()
    it_mixed, rel_mixed, conv_mixed, mv_mixed = infos[-1]
    log(f"mixed: {args.steps} solves in {wall:.3f} s, inner iterations {it_mixed}, true rel residual {rel_mixed:.3e}")

    # ---- plain FP64 (reference algorithm on the GPU) ------------------------
    solve_plain()  # warm the FP64 path
    torch.cuda.synchronize()
    plat = []
    for _ in range(args.plain_steps):
        ts = time.perf_counter()
        pinfo = solve_plain()
        plat.append(time.perf_counter() - ts)
    x_plain = read_x()
    it_plain, rel_plain, conv_plain, mv_plain = pinfo
    plain_sps = args.plain_steps / sum(plat)
    log(f"plain FP64: {args.plain_steps} solves, {1 / plain_sps:.3f} s each, iterations {it_plain}")
    err_vs_fp64 = float(np.linalg.norm(x_mixed - x_plain) / np.linalg.norm(x_plain))

    # ---- dominant kernel: FP32 matrix-free product, HIP events on the stream --
    desc = product_desc(table)
    wb = K.mplak_kernel_matvec_workspace_bytes(C.byref(desc))
    work = torch.empty(wb // 8 + 1, dtype=torch.float64, device="cuda")
    yv = torch.empty(N, dtype=torch.float64, device="cuda")
    xv = torch.empty(N, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(xv.data_ptr(), N, SEED, 2, 0, -1.0, 1.0, stream_ptr))
    torch.cuda.synchronize()

    def time_kmv(prec, reps):
        for _ in range(2):
            _native.check(K.mplak_kernel_matvec(C.byref(desc), prec, xv.data_ptr(), yv.data_ptr(), work.data_ptr(),
                                                wb, None, stream_ptr))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(ext)
        for _ in range(reps):
            _native.check(K.mplak_kernel_matvec(C.byref(desc), prec, xv.data_ptr(), yv.data_ptr(), work.data_ptr(),
                                                wb, None, stream_ptr))
        e1.record(ext)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    t32 = time_kmv(_native.MPLAK_PREC_F32, 20)
    t64 = time_kmv(_native.MPLAK_PREC_F64, 5)
    entries = float(N) * N
    ach32 = FLOP_PER_ENTRY * entries / t32
    ach64 = FLOP_PER_ENTRY * entries / t64
    pairs = sym_pairs(N)
    exec32 = (FLOP_PER_ENTRY + 2) * pairs / t32
    log(f"kernel matvec: FP32 {t32 * 1e3:.3f} ms ({entries / t32:.3e} entries/s), "
        f"FP64 {t64 * 1e3:.3f} ms ({entries / t64:.3e} entries/s)")
    del work
    inner, inner_levels = measure_inner_product(K, table, xv, yv, stream_ptr, ext)
    for lv in inner_levels:
        log(f"inner product (ordered, screened at 2^-{lv['drop_log2']:g}, planned): {lv['launch_ms']:.3f} ms, "
            f"{lv['evaluated_entries']:.3e} entries, {lv['executed_tflops']:.1f} TFLOP/s executed")

    # ---- dense dgemv at N = 65536 (34 GB, far above the 256 MB MALL) ---------
    dgemv = None
    if not args.skip_dgemv:
        n = 65536
        Ad = torch.empty(n * n, dtype=torch.float64, device="cuda")
        xd = torch.empty(n, dtype=torch.float64, device="cuda")
        yd = torch.empty(n, dtype=torch.float64, device="cuda")
        _native.check(K.mplak_fill_uniform_f64(Ad.data_ptr(), n * n, 1234, 1, 0, -1.0, 1.0, stream_ptr))
        _native.check(K.mplak_fill_uniform_f64(xd.data_ptr(), n, 1234, 2, 0, -1.0, 1.0, stream_ptr))
        gwb = K.mplak_dgemv_workspace_bytes(n, n)
        gw = torch.empty(gwb // 8 + 1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()

        def gemv():
            _native.check(K.mplak_dgemv_f64(n, n, Ad.data_ptr(), n, xd.data_ptr(), yd.data_ptr(), gw.data_ptr(),
                                            gwb, None, stream_ptr))

        gemv()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record(ext)
        for _ in range(reps):
            gemv()
        e1.record(ext)
        e1.synchronize()
        tg = e0.elapsed_time(e1) / reps * 1e-3
        gbytes = 8.0 * n * n + 16.0 * n
        dgemv = {"n": n, "ms": tg * 1e3, "gbps": gbytes / tg / 1e9, "frac_hbm": gbytes / tg / PEAK_HBM,
                 "algorithmic_bytes": gbytes}
        log(f"dgemv n={n}: {tg * 1e3:.3f} ms, {gbytes / tg / 1e9:.1f} GB/s")
        del Ad, xd, yd, gw
        torch.cuda.empty_cache()

    # ---- CPU oracle on a bounded sample --------------------------------------
    cpu = None
    if not args.skip_cpu:
        from oracle import binding as oracle
        pts = P.cpu().numpy()
        k = oracle.Kernel(pts, oracle.GAUSSIAN, ELL, nugget=NUGGET)
        xh = xv.cpu().numpy()
        tc = time.perf_counter()
        for _ in range(args.cpu_matvecs):
            k.matvec(xh, mode=0)
        tcpu = (time.perf_counter() - tc) / args.cpu_matvecs
        solve_cpu = tcpu * (1 + it_plain)
        cpu = {"value": 1.0 / solve_cpu, "unit": "solves/s", "cores": oracle.threads(), "kind": "port",
               "sample": f"{args.cpu_matvecs} FP64 matrix-free products of the N={N} Gaussian operator "
                         f"({tcpu:.2f} s each, OpenMP), extrapolated to the {1 + it_plain} products of one plain "
                         f"FP64 CG solve; entries/s {entries / tcpu:.3e}"}
        log(f"cpu oracle: {tcpu:.2f} s per product on {oracle.threads()} threads")

    configs = {}
    if not args.skip_configs:
        configs["dense_cg_n4096"] = bench_dense_cg(K, L, inst, stream_ptr, ext)
        configs["dense_dgemv_bicgstab"] = bench_dense_large(K, L, inst, stream_ptr, ext)
        configs["matern_block_n131072"] = bench_block(K, L, inst, stream_ptr, args)
        configs["served_stream"] = bench_served(args)
        configs["reference_benchmark_kernel_cg_n131072"] = bench_reference_kernel_cg(K, L, inst, stream_ptr, ext)

    value = args.steps / wall
    out = {
        "metric": "kernel-CG solves/sec + p50/p99 latency @1e-8 rel residual; dgemv HBM GB/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-entries+f64-refinement",
        "data": "synthetic (counter-based RNG, seed 1235), resident in HBM",
        "config": {"workload": "matrix-free Gaussian kernel CG: N=65536 points U[0,1]^3, l=0.05, (K+1.0I)x=b, "
                               "b~U[-1,1], x0=0, rel residual 1e-8; mixed FP32 entries + FP64 refinement",
                   "n": N, "dim": DIM, "length_scale": ELL, "nugget": NUGGET, "tolerance": EPS},
        "latency_ms": {"p50": percentile(lat, 50) * 1e3, "p99": percentile(lat, 99) * 1e3},
        "mixed": {"inner_iterations": it_mixed, "true_rel_residual": rel_mixed, "converged": bool(conv_mixed),
                  "products": mv_mixed, "rel_error_vs_fp64_solution": err_vs_fp64},
        "plain_fp64": {"solves_per_s": plain_sps, "iterations": it_plain, "rel_residual": rel_plain,
                       "converged": bool(conv_plain),
                       "latency_ms": {"p50": percentile(plat, 50) * 1e3, "p99": percentile(plat, 99) * 1e3}},
        "speedup_mixed_vs_plain": value / plain_sps,
        "roofline": {"bound": "valu", "achieved": inner["executed_tflops"], "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": inner["executed_tflops"] * 1e12 / PEAK_FP32,
                     "traffic": committed_traffic("planned_inner_product_n65536"),
                     "kernel": inner["kernel"], "launch_ms": inner["launch_ms"],
                     "basis": "FLOP of the evaluated entries (the plan's count: 256 per kept 16x16 tile, 8192 per "
                              "kept 128x64 block of non-compact tiles; 15 FLOP each off the diagonal tiles, 13 on "
                              "them) per launch / launch time (HIP events, library stream), at the solver's base "
                              "screen 2^-15; the relaxed screens 2^-10 and 2^-6 are in inner_product_levels",
                     "frac_of_nonpacked_fp32": inner["executed_tflops"] * 1e12 / PEAK_FP32_NONPACKED,
                     "inner_product": inner, "inner_product_levels": inner_levels},
        "roofline_exact_fp32": {"bound": "valu", "achieved": ach32 / 1e12, "peak": PEAK_FP32 / 1e12,
                                "unit": "TFLOP/s", "frac": ach32 / PEAK_FP32,
                                "kernel": "kmv_sym_kernel<float, GAUSSIAN, 3> + kmv_sym_finish_kernel",
                                "launch_ms": t32 * 1e3, "entries_per_s": entries / t32,
                                "basis": "algorithmic: N^2 entries x 13 FLOP per exact product; the symmetric "
                                         "kernel evaluates each pair once (N(N+1)/2 evaluations, 15 FLOP each)",
                                "executed_tflops": exec32 / 1e12, "executed_frac": exec32 / PEAK_FP32},
        "roofline_fp64_kernel": {"bound": "valu", "achieved": ach64 / 1e12, "peak": PEAK_FP64 / 1e12,
                                 "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64, "launch_ms": t64 * 1e3,
                                 "entries_per_s": entries / t64},
        "dgemv": dgemv,
        "cpu_baseline": cpu,
        "configs": configs,
    }
    if dgemv is not None:
        out["roofline_dgemv"] = {"bound": "hbm", "achieved": dgemv["gbps"], "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                                 "frac": dgemv["frac_hbm"], "traffic": committed_traffic("dgemv_n65536")}
    L.mpla_free_vector(C.byref(b), C.byref(inst))
    L.mpla_free_vector(C.byref(x), C.byref(inst))
    L.mpla_free_generic_matrix(C.byref(A), C.byref(inst))
    L.mpla_destroy_instance(C.byref(inst))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
