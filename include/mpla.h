/*
 * mpla.h -- drop-in replacement of MPLA's public API (src/mpla.h of the
 * reference) on one MI355X.
 *
 * Same struct names, the fields callers read, and the same function names and
 * signatures as the reference header (src/mpla.h:25-126), rebuilt on the
 * gfx950 kernel library (mplak.h).  Differences a caller can observe:
 *   - the process grid is 1 x 1 (one process, one GPU); the block
 *     distribution arithmetic is kept and computes the 1 x 1 case;
 *   - functions have C linkage (so FFI layers can bind them directly);
 *   - there is no cuBLAS handle: user plugin cores that called cuBLAS on
 *     instance->cublas_handle must call the exported cores or mplak.h instead;
 *   - work is enqueued on the instance's HIP stream and never needs a host
 *     synchronisation, except where a value is returned to the host
 *     (mpla_ddot, file I/O, solver exit);
 *   - solvers print only their final summary line unless MPLA_VERBOSE >= 2.
 * Errors keep the reference convention: print and exit(EXIT_FAILURE).
 *
 * Functions below the "Extensions" line have no reference counterpart.
 */
#ifndef mpla_h__
#define mpla_h__

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

/* Single-process communicator.  When <mpi.h> has been included first its
 * MPI_Comm is used; otherwise a one-rank stand-in keeps reference drivers
 * (MPI_Init / MPI_COMM_WORLD / MPI_Finalize) compiling unchanged. */
#if !defined(MPI_VERSION) && !defined(MPLA_NO_MPI_STANDIN)
typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
static inline int MPI_Init(int* argc, char*** argv) { (void)argc; (void)argv; return 0; }
static inline int MPI_Finalize(void) { return 0; }
#endif

#ifdef __cplusplus
extern "C" {
#endif

struct mpla_instance
{
	union { MPI_Comm comm; uint64_t comm_storage; }; /* fixed 8-byte slot */
	int proc_count;
	int proc_rows;
	int proc_cols;
	int cur_proc_coord[2];
	int cur_proc_row;
	int cur_proc_col;
	int cur_proc_rank;
	bool is_parent;
	/* runtime state of this build (no counterpart in the reference) */
	void* stream;              /* hipStream_t all work of this instance runs on */
	void* runtime;             /* opaque: workspace arena, solver scratch */
	const int32_t* skip_flag;  /* device flag the exported cores honour (set by solvers) */
	int last_iterations;       /* loop counter k of the last solve, as the reference prints it */
	int last_converged;        /* 1 if the last solve met its stopping test */
	double last_rel_res;       /* stopping quantity at exit of the last solve */
	int last_matvecs;          /* operator applications of the last solve */
	int verbosity;             /* 0 silent, 1 final line (default), 2 per-iteration lines */
};

struct mpla_matrix
{
	double* data;
	int mat_row_count, mat_col_count;
	int cur_proc_row_count;
	int cur_proc_col_count;
	int cur_proc_row_offset;
	int cur_proc_col_offset;
	int** proc_row_count;
	int** proc_col_count;
	int** proc_row_offset;
	int** proc_col_offset;
};

struct mpla_generic_matrix
{
	void* data;
	int mat_row_count, mat_col_count;
	int cur_proc_row_count;
	int cur_proc_col_count;
	int cur_proc_row_offset;
	int cur_proc_col_offset;
	int** proc_row_count;
	int** proc_col_count;
	int** proc_row_offset;
	int** proc_col_offset;
};

struct mpla_vector
{
	double* data;
	int vec_row_count, vec_col_count;
	int cur_proc_row_count;
	int cur_proc_row_offset;
	int** proc_row_count;
	int** proc_row_offset;
};

typedef void (*mpla_dgemv_core_t)(struct mpla_vector*, struct mpla_generic_matrix*,
                                  struct mpla_vector*, struct mpla_instance*);

/* src/mpla.cpp:26-29 */
void info(void);

/* src/mpla.cpp:31-63: grid from the communicator size (1 x 1 here), rank 0 is parent */
void mpla_init_instance(struct mpla_instance* instance, MPI_Comm comm);

/* src/mpla.cpp:66-99: rows-only grid {P, 1} */
void mpla_init_instance_block_rows(struct mpla_instance* instance, MPI_Comm comm);

/* src/mpla.cpp:102-105 */
void mpla_destroy_instance(struct mpla_instance* instance);

/* src/mpla.cpp:108-215: balanced 2-D blocks, local block column-major, lda = local rows */
void mpla_init_matrix(struct mpla_matrix* matrix, struct mpla_instance* instance, int mat_row_count, int mat_col_count);

/* src/mpla.cpp:217-321: as above; data is owned by the caller (plugin payload) */
void mpla_init_generic_matrix(struct mpla_generic_matrix* matrix, struct mpla_instance* instance, int mat_row_count, int mat_col_count);

/* src/mpla.cpp:323-373: row-blocked over process rows, replicated over columns */
void mpla_init_vector(struct mpla_vector* vector, struct mpla_instance* instance, int vec_row_count);

/* src/mpla.cpp:375-445: blocked over process columns */
void mpla_init_vector_for_block_rows(struct mpla_vector* vector, struct mpla_instance* instance, int vec_row_count);

/* src/mpla.cpp:449-516: all-gather over the process column, keep the column block */
void mpla_redistribute_vector_for_dgesv(struct mpla_vector* b_redist, struct mpla_vector* b, struct mpla_matrix* A, struct mpla_instance* instance);

/* src/mpla.cpp:605-652: int32 length + raw float64, written by rank 0 */
void mpla_save_vector(struct mpla_vector* x, char* filename, struct mpla_instance* instance);

/* src/mpla.cpp:686-713: exits when the stored length differs */
void mpla_load_vector(struct mpla_vector* x, char* filename, struct mpla_instance* instance);

/* src/mpla.cpp:562-603 */
void mpla_free_vector(struct mpla_vector* x, struct mpla_instance* instance);
void mpla_free_matrix(struct mpla_matrix* A, struct mpla_instance* instance);
void mpla_free_generic_matrix(struct mpla_generic_matrix* A, struct mpla_instance* instance);

/* src/mpla.cpp:716-751: b = A x (stored dense, HIP dgemv) */
void mpla_dgemv(struct mpla_vector* b, struct mpla_matrix* A, struct mpla_vector* x, struct mpla_instance* instance);

/* src/mpla.cpp:753-770: *xy = <x, y> returned on the host */
void mpla_ddot(double* xy, struct mpla_vector* x, struct mpla_vector* y, struct mpla_instance* instance);

/* src/mpla.cpp:772-804: b = A x through the plugin core */
void mpla_generic_dgemv(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x, void (*mpla_dgemv_core)(struct mpla_vector*, struct mpla_generic_matrix*, struct mpla_vector*, struct mpla_instance*), struct mpla_instance* instance);

/* src/mpla.cpp:806-810 */
void mpla_daxpy(struct mpla_vector* y, double alpha, struct mpla_vector* x, struct mpla_instance* instance);

/* src/mpla.cpp:812-815 */
void mpla_vector_set_zero(struct mpla_vector* x, struct mpla_instance* instance);

/* src/mpla.cpp:817-885: CG, stops on sqrt(<r,r>/N) < epsilon */
void mpla_generic_conjugate_gradient(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x, int iter_max, double epsilon, void (*mpla_dgemv_core)(struct mpla_vector*, struct mpla_generic_matrix*, struct mpla_vector*, struct mpla_instance*), struct mpla_instance* instance);

/* src/mpla.cpp:887-1001: CG, stops on ||r||/||b|| < epsilon; x,r,d,z checkpoint files.
 * iterations_until_checkpoint <= 0 never checkpoints (the reference divides by it). */
void mpla_generic_conjugate_gradient_with_checkpoint_restart(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x, int iter_max, double epsilon, void (*mpla_dgemv_core)(struct mpla_vector*, struct mpla_generic_matrix*, struct mpla_vector*, struct mpla_instance*), int iterations_until_checkpoint, bool restart, struct mpla_instance* instance);

/* src/mpla.cpp:1003-1183: BiCGSTAB with the true-residual stopping test */
void mpla_generic_BiCGSTAB_with_checkpoint_restart(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x, int iter_max, double epsilon, void (*mpla_dgemv_core)(struct mpla_vector*, struct mpla_generic_matrix*, struct mpla_vector*, struct mpla_instance*), int iterations_until_checkpoint, bool restart, struct mpla_instance* instance);

/* src/mpla.cpp:654-683: gather the full vector into host memory */
void mpla_copy_distributed_vector_to_cpu(double* x_cpu, struct mpla_vector* x, struct mpla_instance* instance);

/* src/mpla.cpp:1187-1193: print the pending HIP error and exit */
void checkCUDAError(const char* msg);

/* ---------------------------------------------------------------- Extensions */

/* Stored-dense plugin core: b = A x with A->data a column-major local block
 * (lda = A->cur_proc_row_count).  The reference drivers define this core
 * themselves on cuBLAS (src/test_cg.cpp:38-43); mpla_generic_dgemv_core_cublas
 * is exported as a weak alias so those drivers link unchanged after deleting
 * their cuBLAS body (or keep their own definition, which takes precedence). */
void mpla_generic_dgemv_core_dense(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);
void mpla_generic_dgemv_core_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);

/* Declare a user core safe for HIP graph capture: stream-ordered on
 * instance->stream, no host synchronisation or allocation.  Exported cores
 * are registered already.  Solvers capture chunks of iterations with such
 * cores and run other cores eagerly, one host check per iteration. */
void mpla_register_graph_safe_core(mpla_dgemv_core_t core);

/* Block distribution arithmetic of mpla_init_* for an arbitrary grid (host
 * only, usable without a GPU): counts[p], offsets[p] for p < parts.  Returns
 * 0, or -1 when floor(total/parts) == 0. */
int mpla_block_distribution(int total, int parts, int* counts, int* offsets);

/* Device pointer of the instance's HIP stream (hipStream_t). */
void* mpla_instance_stream(struct mpla_instance* instance);

#ifdef __cplusplus
}
#endif

#endif
