/*
 * hmglib_adapter.h -- drop-in for the reference's H-matrix plugin
 * (src/hmglib_adapter.h:21-29, src/hmglib_adapter.cu:6-132), which binds the
 * MPLA solvers to the hmglib H-matrix library.  hmglib is not part of this
 * build: its setup and product are rebuilt in libmpla on gfx950, so the
 * struct h_matrix_data below is this library's own.  It keeps the fields the
 * reference adapter sets; callers declare one, point mpla_generic_matrix::data
 * at it and pass the matrix to the functions below.
 *
 * The operator (one process, the 1 x 1 grid):
 *   - each point set is ordered by a cluster tree: a box is bisected along
 *     its longest edge at the median point (cardinality-balanced), down to
 *     leaves of at most c_leaf points;
 *   - a block of row cluster s and column cluster t is admissible when
 *     min(diam B_s, diam B_t) <= eta dist(B_s, B_t) for their bounding boxes
 *     (dist > 0); the block tree is refined until blocks are admissible or
 *     both clusters are leaves;
 *   - admissible blocks are stored as U V^T from adaptive cross
 *     approximation with partial pivoting (at most k terms, fewer once a
 *     term falls below epsilon times the block's Frobenius-norm estimate,
 *     epsilon = 0: k terms), computed once at init;
 *   - inadmissible leaf blocks are evaluated on the fly in every product.
 * The product is deterministic (fixed summation orders).
 */
#ifndef hmglib_adapter_h__
#define hmglib_adapter_h__

#include <stdint.h>

#include "mpla.h"

#ifdef __cplusplus
extern "C" {
#endif

struct system_assembler;

struct h_matrix_data
{
	int point_count[2];           /* local row / column points */
	int dim;
	int bits;                     /* accepted (hmglib's Morton-code bits); not used */
	double eta;                   /* admissibility parameter */
	int max_level;                /* accepted; the trees stop at c_leaf */
	int c_leaf;                   /* leaf size */
	int k;                        /* largest rank of an admissible block (<= 32) */
	double epsilon;               /* ACA tolerance; 0: k terms per block */
	char root_level_set_1;        /* accepted (process-grid levels); 0 on one process */
	char root_level_set_2;
	int max_batched_dense_size;   /* accepted; the near field is never stored */
	int max_batched_aca_size;     /* accepted; every block is compressed at init */
	struct system_assembler* assem; /* device pointer, or NULL for closed-form kernel entries */
	double** coords_d[2];         /* host arrays of dim device pointers: local row / column coordinates */
	unsigned int* point_ids_d[2]; /* device: the point id of each local row / column */
	void* impl;                   /* this library's operator state */
};

/* Entries K(i, j) = assem->get_matrix_entry(row id i, column id j) with the
 * ids of all_point_ids; coordinates from device arrays all_coords[s][d]
 * (global, the local block is taken at the matrix's offsets).  Replaces
 * src/hmglib_adapter.cu:6-57. */
void mpla_init_hmglib(struct mpla_generic_matrix* A, int global_point_count[2], double** all_coords[2], unsigned int* all_point_ids[2], struct system_assembler* assem, double eta, int dim, int bits, int c_leaf, int k, char root_level_set_1, char root_level_set_2, int max_batched_dense_size, int max_batched_aca_size, struct mpla_instance* instance);

/* As mpla_init_hmglib with host coordinates and point ids equal to the
 * global indices.  Replaces src/hmglib_adapter.cu:59-100. */
void mpla_init_hmglib_coords_from_host(struct mpla_generic_matrix* A, int global_point_count[2], double** all_coords[2], struct system_assembler* assem, double eta, int dim, int bits, int c_leaf, int k, int max_batched_dense_size, int max_batched_aca_size, struct mpla_instance* instance);

/* Frees what init allocated (the struct itself is the caller's).
 * Replaces src/hmglib_adapter.cu:104-110. */
void mpla_destroy_hmglib(struct mpla_generic_matrix* A, struct mpla_instance* instance);

/* Plugin cores: b = H x_redist (the compressed operator) and b = K x_redist
 * (every entry).  Replace src/hmglib_adapter.cu:113-132. */
void mpla_dgemv_core_hmglib_h_matrix(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);

void mpla_dgemv_core_hmglib_full_matrix(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);

/* b[i] = assem->get_rhs_entry(id of row point i), or the Gaussian erf
 * right-hand side of the row points for closed-form entries: hmglib's
 * helper the reference's H-matrix driver calls (src/h_matrix_test.cu:225). */
void set_gaussian_kernel_rhs(double* b, struct h_matrix_data* data);

/* ---- No reference counterpart ------------------------------------------ */

/* Closed-form kernel entries instead of a virtual assembler: k(|p_i - p_j|)
 * of MPLA_KERNEL_GAUSSIAN / MPLA_KERNEL_MATERN32 (mpla_kernel.h) with the
 * given length scale, plus nugget where the global row equals the global
 * column.  all_coords as for mpla_init_hmglib (device coordinates); dim 1..4;
 * epsilon: the ACA tolerance (0: k terms). */
void mpla_init_hmatrix_kernel(struct mpla_generic_matrix* A, double** all_coords[2], int dim, int kernel, double length_scale, double nugget, double eta, int c_leaf, int k, double epsilon, struct mpla_instance* instance);

struct mpla_hmatrix_info
{
	int64_t row_leaves, col_leaves;
	int64_t near_blocks, far_blocks;
	int64_t near_entries;         /* entries of the inadmissible blocks (evaluated per product) */
	int64_t far_entries;          /* entries of the admissible blocks */
	int64_t stored_terms;         /* sum over admissible blocks of (rows + columns) x rank */
	int32_t max_rank, k;
	double mean_rank;
	int64_t device_bytes;         /* device memory the operator holds */
	int64_t near_pairs;           /* > 0: symmetric near field, this many chunk pairs each
	                                 evaluated once for both sides (0: per row) */
};

void mpla_hmatrix_get_info(struct mpla_generic_matrix* A, struct mpla_hmatrix_info* info);

/* perm[k] = local index of the k-th point in cluster-tree order of the row
 * (which = 0) or column (which = 1) set; perm is a host array. */
void mpla_hmatrix_get_permutation(struct mpla_generic_matrix* A, int which, int64_t* perm);

/* Host only (no GPU): the cluster-tree order and block partition of n
 * points (coordinates coords[d * n + i], host) for one point set used as
 * rows and columns, as the init functions build it.  perm (n entries)
 * receives the order; counts[0..3] = leaves, near blocks, far blocks,
 * near entries.  Returns 0, or 1 for invalid arguments. */
int mpla_hmatrix_partition(const double* coords, int64_t n, int dim, double eta, int c_leaf, int64_t* perm, int64_t* counts);

#ifdef __cplusplus
}
#endif

#endif
