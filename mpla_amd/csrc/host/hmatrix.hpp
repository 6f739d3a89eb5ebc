// Host side of the H-matrix plugin (hmglib_adapter.h): cluster trees, the
// block partition and the device launches of the entries behind a virtual
// system_assembler (hmatrix_virtual.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "../../../include/mplak.h"

struct system_assembler;

namespace mpla_hm {

// Cluster of a tree over one point set: points perm[lo, hi) in tree order.
struct Cluster {
  int64_t lo, hi;
  int32_t child[2];  // -1 for a leaf
  double bmin[4], bmax[4];
};

struct Tree {
  std::vector<int64_t> perm;     // tree order -> local point index
  std::vector<Cluster> nodes;    // nodes[0] is the root
  std::vector<int32_t> leaves;   // leaf node ids in tree order
};

// Cardinality-balanced bisection: a cluster with more than c_leaf points is
// split along the longest edge of its bounding box (the lowest axis on ties)
// into the floor(n/2) points with the smallest (coordinate, index) keys and
// the rest; each leaf lists its points in increasing index.
Tree build_tree(const double* coords, int64_t n, int dim, int c_leaf);

struct Partition {
  std::vector<std::pair<int32_t, int32_t>> near, far;  // (row node, column node) pairs, recursion order
};

// Block partition of rows x columns: admissible when min(diam) <= eta dist
// with dist > 0; otherwise both leaves -> near, else refine the non-leaf
// side(s).
Partition partition(const Tree& rows, const Tree& cols, int dim, double eta);

// Entries behind a virtual assembler (relocatable device code).
int virtual_aca(system_assembler* assem, const int64_t* row_ids, const int64_t* col_ids,
                const mplak_hm_block* blocks, int64_t nblocks, int kmax, double eps, double* U, double* V,
                int32_t* rank, void* stream);
int virtual_apply(system_assembler* assem, const int64_t* row_ids, const int64_t* col_ids,
                  const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* near_ptr, const int64_t* near_cols,
                  const int64_t* far_ptr, const int32_t* far_blk, const mplak_hm_block* blocks, const int32_t* rank,
                  int kmax, const double* U, const double* t, const double* xs, double* ys, void* stream);
int virtual_rhs(system_assembler* assem, const int64_t* ids, int64_t n, double* b, void* stream);
// stored near field (layout of mplak_hm_apply_stored): one workgroup per near range
int virtual_near_fill(system_assembler* assem, const int64_t* row_ids, const int64_t* col_ids,
                      const int64_t* range_leaf_lo, const int64_t* near_cols, const int64_t* near_off,
                      int64_t nranges, double* stored, void* stream);

}  // namespace mpla_hm
