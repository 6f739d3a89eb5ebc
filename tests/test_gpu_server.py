"""The continuous-batching solve server (no reference counterpart).

Bars: every request's true FP64 relative residual (recomputed on the host
with the oracle's product) below its tolerance; the solution within 1e-6 of
the FP64 oracle solution; each request's result bitwise identical whether it
ran alone or in a mix with other requests (a stronger statement than the
tolerance the mixed-precision path states, 1e-6 relative).
"""
import numpy as np
import pytest

from mpla_amd import mpla, serving
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu


def small_stream(count, seed, rate=400.0):
    return serving.request_stream(count, rate, seed=seed, n_min=500, n_max=5000)


def host_check(oracle, r, X):
    pts = r["points"].reshape(r["dim"], r["n"])
    k = oracle.Kernel(pts, 0 if r["kind"] == mpla.MPLA_KERNEL_GAUSSIAN else 1, r["length_scale"], nugget=r["nugget"])
    B = r["rhs"].reshape((r["n"], r["nrhs"]), order="F")
    worst = 0.0
    for c in range(r["nrhs"]):
        worst = max(worst, np.linalg.norm(B[:, c] - k.matvec(X[:, c])) / np.linalg.norm(B[:, c]))
    return k, B, worst


def test_server_single_request_vs_oracle(oracle):
    require_gpu()
    reqs = small_stream(1, seed=3)
    r = reqs[0]
    with serving.Server() as srv:
        res, sols = srv.run_trace(reqs)
    assert res[0].status == mpla.MPLA_REQUEST_CONVERGED
    k, B, worst = host_check(oracle, r, sols[0])
    assert worst < r["tolerance"] * 1.01
    assert res[0].max_rel_residual < r["tolerance"]
    xo, _ = oracle.cg_relative(k.op(), B[:, 0], np.zeros(r["n"]), 20000, 1e-12)
    assert np.linalg.norm(sols[0][:, 0] - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("pairs", [None, "1"])
@pytest.mark.parametrize("mode", [mpla.MPLA_SERVE_MIXED_BATCHED, mpla.MPLA_SERVE_FP64_SERIAL])
def test_server_mix_equals_alone(oracle, monkeypatch, mode, pairs):
    """Requests that overlap on the GPU give the same bits as each request
    alone, with the products split by size (slabs at these n) and with every
    product on the tile-pair split (MPLAK_RAGGED_PAIRS=1)."""
    require_gpu()
    if pairs is not None:
        monkeypatch.setenv("MPLAK_RAGGED_PAIRS", pairs)
    reqs = small_stream(10, seed=11, rate=2000.0)
    with serving.Server(mode=mode) as srv:
        res_mix, sol_mix = srv.run_trace(reqs)
    assert all(x.status == mpla.MPLA_REQUEST_CONVERGED for x in res_mix)
    if mode == mpla.MPLA_SERVE_MIXED_BATCHED:
        # the requests really were on the device together
        spans = sorted((x.t_admit, x.t_done) for x in res_mix)
        assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for i, r in enumerate(reqs):
        alone = dict(r, arrival=0.0)
        with serving.Server(mode=mode) as srv:
            res1, sol1 = srv.run_trace([alone])
        assert res1[0].status == mpla.MPLA_REQUEST_CONVERGED
        assert res1[0].iterations == res_mix[i].iterations
        assert np.array_equal(sol1[0], sol_mix[i]), f"request {i} differs alone vs in the mix"
        _, _, worst = host_check(oracle, r, sol_mix[i])
        assert worst < r["tolerance"] * 1.01


def test_server_budget_queues_requests(oracle):
    require_gpu()
    reqs = small_stream(6, seed=5, rate=5000.0)
    lib = mpla.lib()
    sizes = sorted(lib.mpla_server_request_bytes(r["n"], r["dim"], r["nrhs"]) for r in reqs)
    budget = sizes[-1] + sizes[-2]
    with serving.Server(hbm_budget_bytes=budget) as srv:
        res, sols = srv.run_trace(reqs)
        peak = srv.peak_bytes()
    assert all(x.status == mpla.MPLA_REQUEST_CONVERGED for x in res)
    assert peak <= budget
    assert max(x.t_admit - x.t_arrival for x in res) > 0  # someone waited for memory
    for r, X in zip(reqs, sols):
        _, _, worst = host_check(oracle, r, X)
        assert worst < r["tolerance"] * 1.01


def test_server_rejects_requests_beyond_limits():
    require_gpu()
    r = small_stream(1, seed=2)[0]
    too_many = dict(r, nrhs=9, rhs=np.zeros(r["n"] * 9))
    with serving.Server(hbm_budget_bytes=10 << 20) as srv:
        res, _ = srv.run_trace([too_many, dict(r, n=200000, points=np.zeros(r["dim"] * 200000),
                                               rhs=np.zeros(200000 * r["nrhs"]))])
    assert [x.status for x in res] == [mpla.MPLA_REQUEST_REJECTED, mpla.MPLA_REQUEST_REJECTED]


def test_mixed_and_fp64_serial_agree(oracle):
    require_gpu()
    reqs = small_stream(4, seed=17, rate=3000.0)
    with serving.Server(mode=mpla.MPLA_SERVE_MIXED_BATCHED) as srv:
        _, mixed = srv.run_trace(reqs)
    with serving.Server(mode=mpla.MPLA_SERVE_FP64_SERIAL) as srv:
        _, plain = srv.run_trace(reqs)
    for a, b in zip(mixed, plain):
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)


def test_server_screened_ordered_vs_exact(oracle, monkeypatch):
    """Mixed mode orders each request's points and screens its products by
    default; MPLA_IR_DROP_LOG2=0 keeps every entry in the request's order.
    Both meet every request's tolerance and agree to the solution bar."""
    require_gpu()
    reqs = serving.request_stream(6, 3000.0, seed=23, n_min=2000, n_max=20000)
    with serving.Server() as srv:
        res_s, sol_s = srv.run_trace(reqs)
    monkeypatch.setenv("MPLA_IR_DROP_LOG2", "0")
    with serving.Server() as srv:
        res_e, sol_e = srv.run_trace(reqs)
    for r, a, b, ra, rb in zip(reqs, sol_s, sol_e, res_s, res_e):
        assert ra.status == rb.status == mpla.MPLA_REQUEST_CONVERGED
        _, _, worst = host_check(oracle, r, a)
        assert worst < r["tolerance"] * 1.01
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)


def test_server_zero_rhs_column_retires(oracle):
    """A right-hand side that is all zero is solved exactly by x = 0 at the
    first residual; the request retires with its other columns converged."""
    require_gpu()
    r = serving.request_stream(1, 100.0, seed=29, n_min=3000, n_max=3000)[0]
    r = dict(r, nrhs=2)
    B = np.zeros((r["n"], 2), order="F")
    B[:, 0] = serving.request_stream(1, 100.0, seed=31, n_min=3000, n_max=3000)[0]["rhs"][: r["n"]]
    r["rhs"] = B.reshape(-1, order="F").copy()
    for mode in (mpla.MPLA_SERVE_MIXED_BATCHED, mpla.MPLA_SERVE_FP64_SERIAL):
        with serving.Server(mode=mode) as srv:
            res, sols = srv.run_trace([r])
        assert res[0].status == mpla.MPLA_REQUEST_CONVERGED
        assert np.isfinite(res[0].max_rel_residual) and res[0].max_rel_residual < r["tolerance"]
        assert np.all(sols[0][:, 1] == 0.0)
        pts = r["points"].reshape(r["dim"], r["n"])
        k = oracle.Kernel(pts, 0 if r["kind"] == mpla.MPLA_KERNEL_GAUSSIAN else 1, r["length_scale"],
                          nugget=r["nugget"])
        assert np.linalg.norm(B[:, 0] - k.matvec(sols[0][:, 0])) / np.linalg.norm(B[:, 0]) < r["tolerance"] * 1.01


@pytest.mark.parametrize("nugget", [1e-3, 1e-5])
def test_server_small_nugget_mixed(oracle, nugget):
    """Mixed-mode serving at typical interpolation nuggets: every request
    converges to its tolerance on the oracle's residual."""
    require_gpu()
    reqs = [dict(r, nugget=nugget) for r in serving.request_stream(3, 3000.0, seed=37, n_min=4000, n_max=8192)]
    with serving.Server() as srv:
        res, sols = srv.run_trace(reqs)
    for r, x, X in zip(reqs, res, sols):
        assert x.status == mpla.MPLA_REQUEST_CONVERGED
        _, _, worst = host_check(oracle, r, X)
        assert worst < r["tolerance"] * 1.01
        assert abs(worst - x.max_rel_residual) <= 1e-2 * worst + 1e-15
