/*
 * mpla_server.h -- continuous-batching solve server for kernel-interpolation
 * systems (K + nugget I) X = B on one GPU.  No reference counterpart.
 *
 * Requests of different sizes arrive at any time.  A scheduling step admits
 * queued requests while slots and the HBM budget allow (first come, first
 * served), keeps each admitted request's points, right-hand sides, iterate
 * and Krylov vectors resident on the device, and advances every active
 * request with ragged launches that cover all of them at once (each request's
 * nrhs right-hand sides share its products).  Finished requests are retired
 * and their memory reused.  A request's arithmetic does not depend on what
 * else is in the batch, so its result is the same alone and in any mix.
 *
 * Modes:
 *   MPLA_SERVE_MIXED_BATCHED  FP32 entry evaluation in the inner CG, FP64
 *                             residuals / dot products / updates (iterative
 *                             refinement), all active requests batched;
 *   MPLA_SERVE_FP64_SERIAL    plain FP64 CG, one request on the GPU at a time.
 * Every result meets its tolerance on the FP64 relative residual.  In mixed
 * mode the FP64 products drop only entries whose sum moves a residual by at
 * most 2^-12 of the request's tolerance (at most the entries below
 * 2^-(53 + log2 n), which change no product beyond FP64 rounding).
 */
#ifndef mpla_server_h__
#define mpla_server_h__

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPLA_SERVE_MIXED_BATCHED 0
#define MPLA_SERVE_FP64_SERIAL 1

#define MPLA_REQUEST_PENDING 0
#define MPLA_REQUEST_CONVERGED 1
#define MPLA_REQUEST_ITER_LIMIT 2 /* stopped unconverged: iteration limit, 40 refinement steps or a non-finite residual */
#define MPLA_REQUEST_REJECTED (-1) /* cannot fit the budget or exceeds the limits */
#define MPLA_REQUEST_UNKNOWN (-2)

#define MPLA_SERVER_MAX_RHS 8
#define MPLA_SERVER_MAX_DIM 4

typedef struct mpla_server mpla_server;

typedef struct mpla_server_config
{
	size_t hbm_budget_bytes;  /* device bytes all active requests may hold together */
	int max_active;           /* requests resident at once (slots) */
	int mode;                 /* MPLA_SERVE_* */
	int iterations_per_step;  /* inner CG iterations per scheduling step (rounded up to even) */
	double inner_tol_floor;   /* smallest relative tolerance asked of an FP32 inner solve */
	int max_iterations;       /* per-request cap on CG iterations */
} mpla_server_config;

typedef struct mpla_solve_request
{
	const double* points;   /* host, SoA: coordinate d of point i at points[d * n + i] */
	const double* rhs;      /* host, n x nrhs column-major */
	double* solution;       /* host, n x nrhs column-major, written when the request finishes (may be NULL) */
	int64_t n;
	int dim;                /* 1 .. MPLA_SERVER_MAX_DIM */
	int kind;               /* MPLA_KERNEL_GAUSSIAN / MPLA_KERNEL_MATERN32 */
	double length_scale;
	double nugget;
	int nrhs;               /* 1 .. MPLA_SERVER_MAX_RHS */
	double tolerance;       /* on ||b - A x|| / ||b|| per right-hand side */
	double arrival;         /* seconds after the start of mpla_server_run_trace */
} mpla_solve_request;

typedef struct mpla_solve_result
{
	int status;             /* MPLA_REQUEST_* */
	int iterations;         /* CG iterations (products) the request used */
	int refinement_steps;   /* FP64 residual evaluations */
	double max_rel_residual;/* largest true FP64 relative residual over its right-hand sides */
	double t_arrival;       /* seconds on the server clock */
	double t_admit;
	double t_done;
} mpla_solve_result;

/* Defaults: 200 GB budget, 64 slots, mixed batched, 4 iterations per step,
 * inner tolerance floor 1e-5, 20000 iterations. */
void mpla_server_default_config(mpla_server_config* cfg);

mpla_server* mpla_server_create(const mpla_server_config* cfg);
void mpla_server_destroy(mpla_server* server);

/* Device bytes a request holds while it is active (host only). */
size_t mpla_server_request_bytes(int64_t n, int dim, int nrhs);

/* Queue a request (its host arrays must stay valid until it finishes).
 * Returns the request id (>= 0) or MPLA_REQUEST_REJECTED.  Thread-safe. */
int mpla_server_submit(mpla_server* server, const mpla_solve_request* request);

/* One scheduling step.  Returns the number of requests still pending or active. */
int mpla_server_step(mpla_server* server);

/* Result of request `id` (status MPLA_REQUEST_PENDING while it runs). */
int mpla_server_result(mpla_server* server, int id, mpla_solve_result* result);

/* Replays a trace in real time: request i is submitted once its `arrival`
 * time has passed, steps run until every request has finished.  results[i]
 * receives request i's outcome.  Returns the number of requests that did not
 * converge. */
int mpla_server_run_trace(mpla_server* server, const mpla_solve_request* requests, int count,
                          mpla_solve_result* results);

/* Seconds on the server clock (the clock of t_arrival / t_admit / t_done). */
double mpla_server_clock(mpla_server* server);

/* Largest number of device bytes held by active requests so far. */
size_t mpla_server_peak_bytes(mpla_server* server);

#ifdef __cplusplus
}
#endif

#endif
