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
PEAK_BF16_MFMA = 2.5e15  # dense
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
    cols = (0, 21, 42, 63)
    sols = {}
    # the shared K*V product alone, per precision (HIP events on the library stream)
    desc = _native.mplak_kernel_desc()
    desc.points, desc.dim, desc.kind = table.data_ptr(), 3, _native.MPLAK_KERNEL_MATERN32
    desc.length_scale, desc.nugget, desc.coincident_shift = 0.1, 0.5, 0.0
    desc.row_offset, desc.row_count, desc.col_offset, desc.col_count = 0, n, 0, n
    Yd = torch.empty(n * R, dtype=torch.float64, device="cuda")
    ext = torch.cuda.ExternalStream(stream_ptr)
    kv = {}
    for name in [p for p in args.block_precisions.split(",") if p and p != "f64"]:
        prec = names[name]
        wb = K.mplak_kernel_mrhs_workspace_bytes(C.byref(desc), prec, R)
        w = torch.empty(wb // 8 + 1, dtype=torch.float64, device="cuda")
        f = lambda: _native.check(K.mplak_kernel_mrhs(C.byref(desc), prec, R, Bd.data_ptr(), n, Yd.data_ptr(), n,
                                                      w.data_ptr(), wb, None, stream_ptr))
        f()
        ms = _event_ms(ext, f, 5)
        gen = 16.0 * n * n
        con = 2.0 * n * n * R * (3 if name == "bf16x3" else 1)
        mfma_peak = PEAK_BF16_MFMA if name.startswith("bf16") else PEAK_FP32
        kv[name] = {"kernel": "mrhs_mfma_kernel", "ms": ms, "entry_tflops": gen / (ms * 1e-3) / 1e12,
                    "entry_frac_valu": gen / (ms * 1e-3) / PEAK_FP32,
                    "contraction_tflops": con / (ms * 1e-3) / 1e12,
                    "contraction_frac_mfma": con / (ms * 1e-3) / mfma_peak,
                    "mfma_peak_tflops": mfma_peak / 1e12}
        log(f"K*V {name}: {ms:.2f} ms, entries {gen / (ms * 1e-3) / 1e12:.1f} TFLOP/s, "
            f"contraction {con / (ms * 1e-3) / 1e12:.0f} TFLOP/s")
        del w
    del Yd
    out["kv_product"] = kv
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
    if not args.no_block_oracle_columns and sols:
        # the FP64 reference of the four columns: plain batched FP64 CG (one
        # shared FP64 product of the four columns per iteration) to 1e-12
        B4 = torch.stack([Bd[c * n:(c + 1) * n] for c in cols]).reshape(-1).contiguous()
        X4 = torch.zeros(len(cols) * n, dtype=torch.float64, device="cuda")
        rel4 = (C.c_double * len(cols))()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.mpla_kernel_block_conjugate_gradient_mixed(B4.data_ptr(), X4.data_ptr(), len(cols), C.byref(A), 20000,
                                                     1e-12, mpla.MPLA_PREC_F64, rel4, C.byref(inst))
        t_or = time.perf_counter() - t0
        XO = X4.cpu().numpy().reshape((n, len(cols)), order="F")
        errs = {k: [float(np.linalg.norm(X[:, c] - XO[:, q]) / np.linalg.norm(XO[:, q])) for q, c in enumerate(cols)]
                for k, X in sols.items()}
        for k in sols:
            out[k]["rel_error_vs_fp64_columns_0_21_42_63"] = errs[k]
        out["fp64_columns"] = {"columns": list(cols), "tolerance": 1e-12, "seconds": t_or,
                               "iterations": inst.last_iterations, "max_rel_residual": max(rel4),
                               "solver": "plain batched FP64 CG (mpla_kernel_block_conjugate_gradient_mixed, "
                                         "MPLA_PREC_F64)"}
        log(f"Matern block: FP64 columns {cols} in {t_or:.1f} s ({inst.last_iterations} iterations); errors " +
            ", ".join(f"{k} {max(v):.2e}" for k, v in errs.items()))
        del B4, X4
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
    modes = (("mixed_batched", mpla.MPLA_SERVE_MIXED_BATCHED), ("fp64_serial", mpla.MPLA_SERVE_FP64_SERIAL))
    for name, mode in modes:
        with serving.Server(mode=mode) as srv:
            t0 = time.perf_counter()
            res, _ = srv.run_trace(reqs)
            span = time.perf_counter() - t0
        ok = sum(1 for x in res if x.status == mpla.MPLA_REQUEST_CONVERGED)
        st = serving.latency_stats(res)
        out[name] = {"solves_per_s": ok / span, "makespan_s": span, "converged": ok,
                     "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st.get("mean_ms"),
                     "max_true_rel_residual": max(x.max_rel_residual for x in res),
                     "iterations": sum(x.iterations for x in res),
                     "slowest": serving.slowest(reqs, res, 5)}
        if mode == mpla.MPLA_SERVE_FP64_SERIAL and not args.skip_cpu:
            out["cpu"] = cpu_served(reqs, res)
        log(f"served stream {name}: {ok}/{count} converged, {ok / span:.2f} solves/s, p50 {st['p50_ms']:.1f} ms, "
            f"p99 {st['p99_ms']:.1f} ms")
    # capacity: the first `saturated` requests all offered at t = 0 (closed system, no pacing)
    sat = reqs[:min(count, args.served_saturated)]
    out["saturated"] = {"requests": len(sat), "offered": "all at t=0"}
    for name, mode in modes:
        with serving.Server(mode=mode) as srv:
            t0 = time.perf_counter()
            res, _ = srv.run_trace(sat, arrivals=False)
            span = time.perf_counter() - t0
        ok = sum(1 for x in res if x.status == mpla.MPLA_REQUEST_CONVERGED)
        st = serving.latency_stats(res)
        out["saturated"][name] = {"capacity_solves_per_s": ok / span, "makespan_s": span, "converged": ok,
                                  "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                                  "max_true_rel_residual": max(x.max_rel_residual for x in res)}
        log(f"served saturated {name}: {ok}/{len(sat)} in {span:.2f} s = {ok / span:.2f} solves/s, "
            f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms")
    a, b = out["saturated"]["mixed_batched"], out["saturated"]["fp64_serial"]
    out["saturated"]["capacity_ratio_mixed_over_fp64"] = a["capacity_solves_per_s"] / b["capacity_solves_per_s"]
    return out


def cpu_dense(args):
    """CPU legs of the dense configurations with the oracle (OpenMP, this
    host's cores): the N = 4096 CG solve and 100 dgemv, dgemv at N = 65536
    (the headline dgemv size, 34 GB on the host) and BiCGSTAB at N = 32768 in
    full, on the same generated inputs as the GPU legs."""
    import numpy as np
    from oracle import binding as oracle
    out = {}
    n = 4096
    B = oracle.uniform(n * n, 1235, 10, -1.0, 1.0).reshape((n, n), order="F")
    M = B @ B.T / n
    A = np.asfortranarray(0.5 * (M + M.T) + np.eye(n))
    del M, B
    b = oracle.uniform(n, 1235, 11, -1.0, 1.0)
    op = oracle.dense_op(A, mode=0)
    t0 = time.perf_counter()
    xs, info = oracle.cg_relative(op, b, np.zeros(n), 2001, 1e-10)
    tcg = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(100):
        oracle.dgemv(A, xs, mode=0)
    tg = (time.perf_counter() - t0) / 100
    out["dense_cg_n4096"] = {"solves_per_s": 1.0 / tcg, "ms_per_solve": tcg * 1e3, "iterations": info.iterations,
                             "dgemv_ms": tg * 1e3, "dgemv_gbps": (8.0 * n * n + 16 * n) / tg / 1e9,
                             "cores": oracle.threads(), "kind": "port",
                             "sample": "full CG solve (relative 1e-10) and 100 dgemv, A = B B^T/N + I built with "
                                       "numpy (same B as the GPU leg)"}
    log(f"cpu dense CG n=4096: {tcg * 1e3:.1f} ms ({info.iterations} iterations), dgemv {tg * 1e3:.2f} ms")
    del A, op
    n = 65536
    A = oracle.uniform(n * n, 1234, 1, -1.0, 1.0).reshape((n, n), order="F")
    xg = oracle.uniform(n, 1234, 2, -1.0, 1.0)
    oracle.dgemv(A, xg, mode=0)
    t0 = time.perf_counter()
    for _ in range(3):
        oracle.dgemv(A, xg, mode=0)
    tg = (time.perf_counter() - t0) / 3
    out["dgemv_65536"] = {"ms": tg * 1e3, "gbps": (8.0 * n * n + 16 * n) / tg / 1e9, "cores": oracle.threads(),
                          "kind": "port", "sample": "3 dgemv on the 34 GB matrix of the GPU leg (host RAM)"}
    log(f"cpu dgemv n={n}: {tg * 1e3:.0f} ms, {(8.0 * n * n) / tg / 1e9:.1f} GB/s")
    del A
    n = 32768
    A = oracle.uniform(n * n, 1235, 20, -1.0, 1.0).reshape((n, n), order="F")
    A *= 1.0 / math.sqrt(n)
    A[np.arange(n), np.arange(n)] += oracle.uniform(n, 1235, 21, 2.0, 3.0)
    b = oracle.uniform(n, 1235, 22, -1.0, 1.0)
    t0 = time.perf_counter()
    _, info = oracle.bicgstab(oracle.dense_op(A, mode=0), b, np.zeros(n), 2001, 1e-10)
    tb = time.perf_counter() - t0
    out["bicgstab_32768"] = {"ms_per_solve": tb * 1e3, "solves_per_s": 1.0 / tb, "iterations": info.iterations,
                             "products": info.matvecs, "cores": oracle.threads(), "kind": "port",
                             "sample": "full BiCGSTAB solve (3 products per iteration, relative 1e-10)"}
    log(f"cpu BiCGSTAB n={n}: {tb:.2f} s ({info.iterations} iterations)")
    del A
    return out


def cpu_block(shared_products, seconds_budget=None):
    """CPU K*V of the Matern-3/2 N = 131072 system: one right-hand side's
    product with the oracle, extrapolated (labelled) to the 64 columns of one
    shared product and to the products of one block solve."""
    import numpy as np
    from oracle import binding as oracle
    n = 131072
    pts = oracle.uniform(3 * n, 1235, 30).reshape(3, n)
    k = oracle.Kernel(pts, oracle.MATERN32, 0.1, nugget=0.5)
    v = oracle.normal(n, 1235, 31)
    t0 = time.perf_counter()
    k.matvec(v, mode=0)
    t1 = time.perf_counter() - t0
    out = {"one_column_product_s": t1, "kv_product_s_extrapolated": 64 * t1, "cores": oracle.threads(),
           "kind": "port", "sample": "1 FP64 matrix-free product of one column (OpenMP), extrapolated x64 columns"}
    if shared_products:
        out["block_solve_s_extrapolated"] = 64 * t1 * shared_products
        out["extrapolated_to_shared_products"] = shared_products
    log(f"cpu Matern K*V: {t1:.2f} s per column, {64 * t1:.0f} s per 64-column product (extrapolated)")
    return out


def cpu_served(reqs, results, budget_s=25.0):
    """The served stream's plain FP64 path on the CPU: the requests of a
    prefix of the trace solved one at a time (oracle CG per right-hand side,
    relative 1e-8), as many as fit a time budget predicted from the GPU
    iteration counts; service times only (arrivals ignored)."""
    import numpy as np
    from oracle import binding as oracle
    from mpla_amd import mpla
    rate = None
    served, times = 0, []
    predicted = 0.0
    for r, res in zip(reqs, results):
        if rate is not None:
            cost = (res.iterations + 2) * float(r["n"]) ** 2 / rate
            if predicted + cost > budget_s:
                break
            predicted += cost
        pts = r["points"].reshape(r["dim"], r["n"])
        kind = oracle.GAUSSIAN if r["kind"] == mpla.MPLA_KERNEL_GAUSSIAN else oracle.MATERN32
        k = oracle.Kernel(pts, kind, r["length_scale"], nugget=r["nugget"])
        Bm = r["rhs"].reshape((r["n"], r["nrhs"]), order="F")
        t0 = time.perf_counter()
        products = 0
        for c in range(r["nrhs"]):
            _, info = oracle.cg_relative(k.op(mode=0), np.ascontiguousarray(Bm[:, c]), np.zeros(r["n"]), 20000,
                                         r["tolerance"])
            products += info.matvecs
        t = time.perf_counter() - t0
        times.append(t)
        served += 1
        rate = products * float(r["n"]) ** 2 / t
    tot = sum(times)
    out = {"requests": served, "solves_per_s": served / tot if tot else None, "seconds": tot,
           "service_ms_p50": percentile(times, 50) * 1e3, "service_ms_p99": percentile(times, 99) * 1e3,
           "cores": oracle.threads(), "kind": "port",
           "sample": f"the first {served} requests of the trace, one at a time (oracle FP64 CG per right-hand side)"}
    log(f"cpu served prefix: {served} requests in {tot:.1f} s")
    return out


def bench_hmatrix(K, L, inst, stream_ptr, ext):
    """The H-matrix plugin at N = 131072 beside the exact product on the same
    points: the reference H-matrix driver's setting (src/h_matrix_test.cu:
    Halton points in [0,1]^2, exp(-r^2), c_leaf 512, k 20, eta 1.5) through
    mpla_init_hmglib with the Gaussian system assembler (virtual entries), and
    through the closed-form kernel path in 2-D and 3-D."""
    import numpy as np
    import torch
    from mpla_amd import _native, mpla
    n = 131072
    out = {"n": n, "c_leaf": 512, "k": 20, "eta": 1.5, "kernel": "exp(-r^2)", "points": "Halton"}
    xv = torch.empty(n, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_fill_uniform_f64(xv.data_ptr(), n, SEED, 40, 0, -1.0, 1.0, stream_ptr))
    yv = torch.empty(n, dtype=torch.float64, device="cuda")
    for dim, virtual in ((2, True), (2, False), (3, False)):
        name = f"d{dim}_" + ("assembler" if virtual else "closed_form")
        ph = np.empty((dim, n))
        assert L.mpla_halton_points(ph.ctypes.data, dim, n, 0, n) == 0  # host memory
        pts = torch.as_tensor(ph, device="cuda")
        torch.cuda.synchronize()
        table = torch.tensor([pts[d].data_ptr() for d in range(dim)], dtype=torch.int64, device="cuda")
        ptrs = (C.c_void_p * dim)(*[pts[d].data_ptr() for d in range(dim)])
        coords = (C.c_void_p * 2)(C.cast(ptrs, C.c_void_p), C.cast(ptrs, C.c_void_p))
        A = mpla.mpla_generic_matrix()
        L.mpla_init_generic_matrix(C.byref(A), C.byref(inst), n, n)
        data = mpla.h_matrix_data()
        A.data = C.addressof(data)
        assem_pp = None
        t0 = time.perf_counter()
        if virtual:
            assem_pp = torch.zeros(1, dtype=torch.int64, device="cuda")
            L.create_gaussian_kernel_system_assembler_object(assem_pp.data_ptr(), table.data_ptr(), n, dim)
            torch.cuda.synchronize()
            ids = torch.arange(n, dtype=torch.int32, device="cuda")
            idp = (C.c_void_p * 2)(ids.data_ptr(), ids.data_ptr())
            counts = (C.c_int * 2)(n, n)
            t0 = time.perf_counter()
            L.mpla_init_hmglib(C.byref(A), counts, coords, idp, int(assem_pp.cpu()[0]), 1.5, dim, 32, 512, 20,
                               b"\0", b"\0", 1 << 24, 1 << 24, C.byref(inst))
        else:
            L.mpla_init_hmatrix_kernel(C.byref(A), coords, dim, mpla.MPLA_KERNEL_GAUSSIAN, 1.0 / math.sqrt(2.0), 0.0,
                                       1.5, 512, 20, 0.0, C.byref(inst))
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        info = mpla.mpla_hmatrix_info()
        L.mpla_hmatrix_get_info(C.byref(A), C.byref(info))
        xm, ym = mpla.mpla_vector(), mpla.mpla_vector()
        L.mpla_init_vector(C.byref(xm), C.byref(inst), n)
        L.mpla_init_vector(C.byref(ym), C.byref(inst), n)
        _native.check(K.mplak_copy_f64(xm.data, xv.data_ptr(), n, stream_ptr))
        hcore, fcore = mpla.core("mpla_dgemv_core_hmglib_h_matrix"), mpla.core("mpla_dgemv_core_hmglib_full_matrix")
        fh = lambda: L.mpla_generic_dgemv(C.byref(ym), C.byref(A), C.byref(xm), hcore, C.byref(inst))
        fh()
        t_h = _event_ms(ext, fh, 10)
        yh = torch.empty(n, dtype=torch.float64, device="cuda")
        _native.check(K.mplak_copy_f64(yh.data_ptr(), ym.data, n, stream_ptr))
        ff = lambda: L.mpla_generic_dgemv(C.byref(ym), C.byref(A), C.byref(xm), fcore, C.byref(inst))
        t_f = _event_ms(ext, ff, 1)
        # the exact symmetric FP64 product on the same points (the fastest exact product of this build)
        d = _native.mplak_kernel_desc()
        d.points, d.dim, d.kind = table.data_ptr(), dim, _native.MPLAK_KERNEL_GAUSSIAN
        d.length_scale, d.nugget, d.coincident_shift = 1.0 / math.sqrt(2.0), 0.0, 0.0
        d.row_offset, d.row_count, d.col_offset, d.col_count = 0, n, 0, n
        wb = K.mplak_kernel_matvec_workspace_bytes(C.byref(d))
        w = torch.empty(wb // 8 + 1, dtype=torch.float64, device="cuda")
        fe = lambda: _native.check(K.mplak_kernel_matvec(C.byref(d), _native.MPLAK_PREC_F64, xv.data_ptr(),
                                                         yv.data_ptr(), w.data_ptr(), wb, None, stream_ptr))
        fe()
        t_e = _event_ms(ext, fe, 3)
        torch.cuda.synchronize()
        err = float(torch.linalg.norm(yh - yv) / torch.linalg.norm(yv))
        near_flop = float(info.near_entries) * (3 * dim + 4)
        far_bytes = 8.0 * float(info.stored_terms)
        out[name] = {"setup_s": setup, "h_product_ms": t_h, "full_core_ms": t_f, "exact_product_ms": t_e,
                     "speedup_vs_exact": t_e / t_h, "rel_error_vs_exact": err,
                     "near_entries": info.near_entries, "far_entries": info.far_entries,
                     "near_fraction": info.near_entries / float(n) / n, "far_blocks": info.far_blocks,
                     "near_blocks": info.near_blocks, "mean_rank": info.mean_rank, "max_rank": info.max_rank,
                     "stored_terms": info.stored_terms, "device_bytes": info.device_bytes,
                     "near_tflops": near_flop / (t_h * 1e-3) / 1e12,
                     "near_frac_fp64": near_flop / (t_h * 1e-3) / PEAK_FP64,
                     "far_factor_gbps": far_bytes / (t_h * 1e-3) / 1e9}
        log(f"H-matrix {name}: setup {setup:.2f} s, product {t_h:.3f} ms (full {t_f:.2f} ms, exact {t_e:.3f} ms), "
            f"rel error {err:.2e}, near {info.near_entries / float(n) / n:.3f}, mean rank {info.mean_rank:.1f}")
        L.mpla_destroy_hmglib(C.byref(A), C.byref(inst))
        L.mpla_free_vector(C.byref(xm), C.byref(inst))
        L.mpla_free_vector(C.byref(ym), C.byref(inst))
        L.mpla_free_generic_matrix(C.byref(A), C.byref(inst))
        if assem_pp is not None:
            L.destroy_gaussian_kernel_system_assembler_object(assem_pp.data_ptr())
        del w, pts, table, yh
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
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
    ap.add_argument("--served-saturated", type=int, default=96,
                    help="requests of the stream offered at t=0 to measure each server mode's capacity")
    ap.add_argument("--block-precisions", default="bf16x3,bf16,f32")
    ap.add_argument("--no-block-oracle-columns", action="store_true",
                    help="skip solving columns 0,21,42,63 with plain FP64 CG for the error report")
    ap.add_argument("--only", default="",
                    help="run only these configurations (comma-separated: dense, dense_large, block, served, "
                         "reference_kernel_cg, hmatrix, cpu_dense) and print their results")
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
    if args.only:
        # one or more configurations alone (for profiler runs); no headline line
        runs = {"dense": lambda: bench_dense_cg(K, L, inst, stream_ptr, ext),
                "dense_large": lambda: bench_dense_large(K, L, inst, stream_ptr, ext),
                "block": lambda: bench_block(K, L, inst, stream_ptr, args),
                "served": lambda: bench_served(args),
                "reference_kernel_cg": lambda: bench_reference_kernel_cg(K, L, inst, stream_ptr, ext),
                "hmatrix": lambda: bench_hmatrix(K, L, inst, stream_ptr, ext),
                "cpu_dense": lambda: cpu_dense(args)}
        res = {name: runs[name]() for name in args.only.split(",")}
        print(json.dumps({"only": res}), flush=True)
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

    def solve_plain_ordered():
        reset_x()
        L.mpla_kernel_conjugate_gradient_mixed(C.byref(b), C.byref(A), C.byref(x), ITER_MAX, EPS,
                                               mpla.MPLA_PREC_F64, C.byref(inst))
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
    # outside the timed region: the mixed solution's residual with the exact
    # FP64 product (every entry, caller order)
    b_host = np.empty(N)
    L.mpla_copy_distributed_vector_to_cpu(b_host.ctypes.data, C.byref(b), C.byref(inst))
    desc_x = product_desc(table)
    wbx = K.mplak_kernel_matvec_workspace_bytes(C.byref(desc_x))
    wx = torch.empty(wbx // 8 + 1, dtype=torch.float64, device="cuda")
    kx = torch.empty(N, dtype=torch.float64, device="cuda")
    _native.check(K.mplak_kernel_matvec(C.byref(desc_x), _native.MPLAK_PREC_F64, x.data, kx.data_ptr(),
                                        wx.data_ptr(), wbx, None, stream_ptr))
    torch.cuda.synchronize()
    exact_rel = float(np.linalg.norm(b_host - kx.cpu().numpy()) / np.linalg.norm(b_host))
    del wx, kx
    log(f"mixed: exact FP64 residual of the solution {exact_rel:.3e}")
    assert exact_rel < EPS, f"mixed solution misses the tolerance on the exact operator: {exact_rel}"

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
    # plain FP64 CG on Hilbert-ordered points with the FP64-exact screen
    solve_plain_ordered()
    torch.cuda.synchronize()
    olat = []
    for _ in range(args.plain_steps):
        ts = time.perf_counter()
        oinfo = solve_plain_ordered()
        olat.append(time.perf_counter() - ts)
    x_ord = read_x()
    ord_sps = args.plain_steps / sum(olat)
    log(f"plain FP64 (ordered, FP64-exact screen): {1 / ord_sps:.3f} s each, iterations {oinfo[0]}")
    best_plain = max(plain_sps, ord_sps)

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
        configs["hmatrix_n131072"] = bench_hmatrix(K, L, inst, stream_ptr, ext)
        if not args.skip_cpu:
            dense = cpu_dense(args)
            configs["dense_cg_n4096"]["cpu"] = dense["dense_cg_n4096"]
            configs["dense_dgemv_bicgstab"]["cpu"] = {"dgemv_65536": dense["dgemv_65536"],
                                                      "bicgstab_32768": dense["bicgstab_32768"]}
            blk = configs["matern_block_n131072"]
            sp = next((blk[p]["shared_products"] for p in ("bf16x3", "f32", "bf16") if p in blk), 0)
            blk["cpu"] = cpu_block(sp)
        if cpu is not None:
            configs["gaussian_n65536"] = {"cpu": cpu}

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
                  "products": mv_mixed, "rel_error_vs_fp64_solution": err_vs_fp64,
                  "exact_rel_residual": exact_rel},
        "plain_fp64": {"solves_per_s": plain_sps, "iterations": it_plain, "rel_residual": rel_plain,
                       "converged": bool(conv_plain), "operator": "every entry (drop-in fused core)",
                       "latency_ms": {"p50": percentile(plat, 50) * 1e3, "p99": percentile(plat, 99) * 1e3}},
        "plain_fp64_screened": {"solves_per_s": ord_sps, "iterations": oinfo[0], "rel_residual": oinfo[1],
                                "converged": bool(oinfo[2]),
                                "operator": "Hilbert-ordered points, FP64 product screened at 2^-(53+log2 N) "
                                            "(mpla_kernel_conjugate_gradient_mixed with MPLA_PREC_F64)",
                                "rel_diff_vs_all_entries_solution":
                                    float(np.linalg.norm(x_ord - x_plain) / np.linalg.norm(x_plain)),
                                "latency_ms": {"p50": percentile(olat, 50) * 1e3, "p99": percentile(olat, 99) * 1e3}},
        "speedup_mixed_vs_plain": value / best_plain,
        "speedup_basis": "the faster of plain_fp64 and plain_fp64_screened",
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
    kv = configs.get("matern_block_n131072", {}).get("kv_product", {})
    if "bf16" in kv:
        k16 = kv["bf16"]
        out["roofline_mrhs"] = {
            "bound": "valu", "kernel": "mrhs_mfma_kernel (BF16 MFMA, FP32 entries), Matern N=131072, R=64",
            "achieved": k16["entry_tflops"], "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": k16["entry_frac_valu"], "launch_ms": k16["ms"],
            "contraction_tflops": k16["contraction_tflops"], "contraction_frac_mfma": k16["contraction_frac_mfma"],
            "basis": "entry generation 16 FLOP x N^2 per K*V (the bound: each entry is used by 64 columns on the "
                     "matrix cores); contraction 2 N^2 R against the 2.5 PF dense BF16 MFMA peak",
            "by_precision": kv}
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
