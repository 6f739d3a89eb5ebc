// H-matrix plugin (include/hmglib_adapter.h): the reference binds its solvers
// to hmglib through src/hmglib_adapter.cu:6-132; here the operator is rebuilt
// on this library.  The host orders both point sets by a cluster tree and
// builds the block partition (O(n log n) work at init); the device
// compresses every admissible block once by adaptive cross approximation and
// applies near and far field in each product (mplak_hm_* kernels, or the
// virtual-assembler kernels of hmatrix_virtual.hip).

#include "hmatrix.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "../../../include/hmglib_adapter.h"
#include "../../../include/mpla_kernel.h"
#include "runtime.hpp"

using namespace mpla_rt;

namespace mpla_hm {

namespace {

void box_of(const double* coords, int64_t n, int dim, const int64_t* idx, int64_t count, double* bmin, double* bmax) {
  for (int d = 0; d < 4; ++d) {
    bmin[d] = 0.0;
    bmax[d] = 0.0;
  }
  for (int d = 0; d < dim; ++d) {
    double lo = INFINITY, hi = -INFINITY;
    const double* c = coords + (size_t)d * n;
    for (int64_t k = 0; k < count; ++k) {
      lo = std::min(lo, c[idx[k]]);
      hi = std::max(hi, c[idx[k]]);
    }
    bmin[d] = lo;
    bmax[d] = hi;
  }
}

double diam(const Cluster& c, int dim) {
  double s = 0.0;
  for (int d = 0; d < dim; ++d) s += (c.bmax[d] - c.bmin[d]) * (c.bmax[d] - c.bmin[d]);
  return std::sqrt(s);
}

double dist(const Cluster& a, const Cluster& b, int dim) {
  double s = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double g = std::max(0.0, std::max(a.bmin[d] - b.bmax[d], b.bmin[d] - a.bmax[d]));
    s += g * g;
  }
  return std::sqrt(s);
}

}  // namespace

Tree build_tree(const double* coords, int64_t n, int dim, int c_leaf) {
  Tree t;
  t.perm.resize(n);
  std::iota(t.perm.begin(), t.perm.end(), int64_t(0));
  if (n == 0) return t;
  t.nodes.push_back(Cluster{0, n, {-1, -1}, {}, {}});
  std::vector<int32_t> stack{0};
  while (!stack.empty()) {
    const int32_t id = stack.back();
    stack.pop_back();
    const int64_t lo = t.nodes[id].lo, hi = t.nodes[id].hi;
    Cluster c = t.nodes[id];
    box_of(coords, n, dim, t.perm.data() + lo, hi - lo, c.bmin, c.bmax);
    if (hi - lo <= c_leaf) {
      std::sort(t.perm.begin() + lo, t.perm.begin() + hi);
      t.nodes[id] = c;
      continue;
    }
    int axis = 0;
    for (int d = 1; d < dim; ++d)
      if (c.bmax[d] - c.bmin[d] > c.bmax[axis] - c.bmin[axis]) axis = d;
    const double* ca = coords + (size_t)axis * n;
    const int64_t mid = lo + (hi - lo) / 2;
    std::nth_element(t.perm.begin() + lo, t.perm.begin() + mid, t.perm.begin() + hi, [ca](int64_t a, int64_t b) {
      return ca[a] < ca[b] || (ca[a] == ca[b] && a < b);
    });
    const int32_t c0 = (int32_t)t.nodes.size();
    t.nodes.push_back(Cluster{lo, mid, {-1, -1}, {}, {}});
    t.nodes.push_back(Cluster{mid, hi, {-1, -1}, {}, {}});
    c.child[0] = c0;
    c.child[1] = c0 + 1;
    t.nodes[id] = c;
    // the second child is popped last, so children are visited first-to-last
    stack.push_back(c0 + 1);
    stack.push_back(c0);
  }
  // leaves in tree order (increasing lo)
  for (int32_t i = 0; i < (int32_t)t.nodes.size(); ++i)
    if (t.nodes[i].child[0] < 0) t.leaves.push_back(i);
  std::sort(t.leaves.begin(), t.leaves.end(),
            [&t](int32_t a, int32_t b) { return t.nodes[a].lo < t.nodes[b].lo; });
  return t;
}

Partition partition(const Tree& rows, const Tree& cols, int dim, double eta) {
  Partition p;
  if (rows.nodes.empty() || cols.nodes.empty()) return p;
  std::vector<std::pair<int32_t, int32_t>> stack{{0, 0}};
  while (!stack.empty()) {
    const auto st = stack.back();
    stack.pop_back();
    const Cluster& a = rows.nodes[st.first];
    const Cluster& b = cols.nodes[st.second];
    const double ds = dist(a, b, dim);
    if (ds > 0.0 && std::min(diam(a, dim), diam(b, dim)) <= eta * ds) {
      p.far.push_back(st);
      continue;
    }
    const bool la = a.child[0] < 0, lb = b.child[0] < 0;
    if (la && lb) {
      p.near.push_back(st);
      continue;
    }
    // push in reverse so that blocks are visited in (row child, column child) order
    std::vector<std::pair<int32_t, int32_t>> kids;
    if (la) {
      kids = {{st.first, b.child[0]}, {st.first, b.child[1]}};
    } else if (lb) {
      kids = {{a.child[0], st.second}, {a.child[1], st.second}};
    } else {
      kids = {{a.child[0], b.child[0]}, {a.child[0], b.child[1]}, {a.child[1], b.child[0]}, {a.child[1], b.child[1]}};
    }
    for (auto it = kids.rbegin(); it != kids.rend(); ++it) stack.push_back(*it);
  }
  return p;
}

}  // namespace mpla_hm

namespace {

using namespace mpla_hm;

constexpr int kMaxDim = 4;
constexpr int64_t kItemColumns = 4096;  // columns per far-field coefficient item
constexpr int64_t kChunkRows = 128;     // rows per apply item (two per lane of a 64-lane wave)
constexpr int64_t kNearChunk = 512;     // rows per chunk of the symmetric near field (mplak_hm_pair)

struct HMatrix {
  int dim = 0;
  bool closed_form = false;
  int kind = 0;
  double length_scale = 1.0, nugget = 0.0;
  system_assembler* assem = nullptr;
  int kmax = 0;
  double eps = 0.0;
  int64_t nr = 0, nc = 0, row_off = 0, col_off = 0;
  Tree rt, ct;
  Partition part;
  mpla_hmatrix_info info{};
  std::vector<void*> allocs;
  // device
  int64_t *perm_r = nullptr, *perm_c = nullptr;   // sorted -> local index
  int64_t *gid_r = nullptr, *gid_c = nullptr;     // sorted -> point id (virtual entries)
  int64_t* ids_r = nullptr;                       // local row -> point id (right-hand side)
  double *pts_r = nullptr, *pts_c = nullptr;      // sorted scaled coordinates (closed form)
  mplak_hm_block* blocks = nullptr;
  int32_t* rank = nullptr;
  double *U = nullptr, *V = nullptr, *t = nullptr, *fpart = nullptr;
  mplak_hm_item* items = nullptr;
  int64_t* first_item = nullptr;
  mplak_hm_chunk* chunks = nullptr;
  int64_t *near_ptr = nullptr, *near_cols = nullptr, *far_ptr = nullptr;
  int32_t* far_blk = nullptr;
  int64_t *full_ptr = nullptr, *full_cols = nullptr, *nofar_ptr = nullptr;
  double *xs = nullptr, *ys = nullptr;
  // symmetric near field (closed form, the same points on both sides):
  // pairs of near chunks, their partial rows, and per near chunk the list
  // of partial blocks (2 pair + side) that make up its rows
  mplak_hm_pair* pairs = nullptr;
  int64_t npairs = 0;
  double* pair_part = nullptr;
  int64_t *chunk_near = nullptr, *near_start = nullptr, *ent_ptr = nullptr, *ent = nullptr;
  // stored near field (virtual entries): entries evaluated once at init
  double* near_st = nullptr;
  int64_t *near_off = nullptr, *leaf_range = nullptr;
  int64_t nblocks = 0, nitems = 0, nchunks = 0;
  double** coords_host[2] = {nullptr, nullptr};  // the h_matrix_data coords_d arrays
};

template <typename T>
T* dalloc(HMatrix* h, size_t count) {
  void* p = nullptr;
  const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
  hip_ck(hipMalloc(&p, bytes), "H-matrix hipMalloc");
  h->allocs.push_back(p);
  h->info.device_bytes += (int64_t)bytes;
  return static_cast<T*>(p);
}

template <typename T>
T* upload(HMatrix* h, const std::vector<T>& v, hipStream_t s) {
  T* d = dalloc<T>(h, v.size());
  // pageable source: wait for the copy before the caller reuses it
  if (!v.empty()) hip_ck(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "H-matrix upload");
  hip_ck(hipStreamSynchronize(s), "H-matrix upload");
  return d;
}

HMatrix* impl_of(mpla_generic_matrix* A, const char* what) {
  h_matrix_data* data = static_cast<h_matrix_data*>(A->data);
  if (data == nullptr || data->impl == nullptr) fail(what, "H-matrix not initialised (A->data)");
  return static_cast<HMatrix*>(data->impl);
}

mplak_hm_points points_of(const HMatrix* h) {
  mplak_hm_points p;
  p.rows = h->pts_r;
  p.cols = h->pts_c;
  p.row_ld = h->nr;
  p.col_ld = h->nc;
  p.dim = h->dim;
  p.kind = h->kind;
  return p;
}

// Host coordinates of one local point set: coords[d * n + i].
std::vector<double> host_coords(double* const* dev_coords, int dim, int64_t n, hipStream_t s) {
  std::vector<double> c((size_t)dim * n);
  for (int d = 0; d < dim; ++d)
    if (n > 0)
      hip_ck(hipMemcpyAsync(c.data() + (size_t)d * n, dev_coords[d], sizeof(double) * n, hipMemcpyDeviceToHost, s),
             "H-matrix coordinates");
  hip_ck(hipStreamSynchronize(s), "H-matrix coordinates");
  return c;
}

// Symmetric near field: when rows and columns are the same points (the two
// trees, and so the leaves, coincide) and the near set is symmetric, the
// inadmissible blocks become pairs of near chunks (512-row pieces of the
// leaves) A <= B, each evaluated once for the rows of both
// (mplak_hm_near_pairs).  MPLA_HM_SYMMETRIC=0 keeps the per-row form.
void build_pairs(HMatrix* h, const std::vector<int32_t>& leaf_of, const std::vector<mplak_hm_chunk>& chunks,
                 hipStream_t s) {
  const char* v = getenv("MPLA_HM_SYMMETRIC");
  if (v != nullptr && *v != '\0' && atoi(v) == 0) return;
  const Tree& rt = h->rt;
  const Tree& ct = h->ct;
  if (rt.leaves.size() != ct.leaves.size()) return;
  const int64_t nleaf = (int64_t)rt.leaves.size();
  for (int64_t l = 0; l < nleaf; ++l)
    if (rt.nodes[rt.leaves[l]].lo != ct.nodes[ct.leaves[l]].lo || rt.nodes[rt.leaves[l]].hi != ct.nodes[ct.leaves[l]].hi)
      return;
  std::vector<int32_t> cleaf_of(ct.nodes.size(), -1);
  for (size_t i = 0; i < ct.leaves.size(); ++i) cleaf_of[ct.leaves[i]] = (int32_t)i;
  std::vector<std::pair<int32_t, int32_t>> near;
  for (const auto& nb : h->part.near) near.push_back({leaf_of[nb.first], cleaf_of[nb.second]});
  std::sort(near.begin(), near.end());
  for (const auto& ab : near)
    if (!std::binary_search(near.begin(), near.end(), std::make_pair(ab.second, ab.first))) return;
  std::vector<int64_t> first_chunk(nleaf + 1), near_start, near_len;
  for (int64_t l = 0; l < nleaf; ++l) {
    first_chunk[l] = (int64_t)near_start.size();
    const Cluster& c = rt.nodes[rt.leaves[l]];
    for (int64_t r0 = c.lo; r0 < c.hi; r0 += kNearChunk) {
      near_start.push_back(r0);
      near_len.push_back(std::min(kNearChunk, c.hi - r0));
    }
  }
  first_chunk[nleaf] = (int64_t)near_start.size();
  std::vector<mplak_hm_pair> pairs;
  std::vector<std::vector<int64_t>> ents(near_start.size());
  for (const auto& ab : near) {
    if (ab.first > ab.second) continue;
    for (int64_t ca = first_chunk[ab.first]; ca < first_chunk[ab.first + 1]; ++ca)
      for (int64_t cb = first_chunk[ab.second]; cb < first_chunk[ab.second + 1]; ++cb) {
        if (ab.first == ab.second && cb < ca) continue;
        const int64_t k = (int64_t)pairs.size();
        pairs.push_back(mplak_hm_pair{near_start[ca], near_len[ca], near_start[cb], near_len[cb]});
        ents[ca].push_back(2 * k);
        if (cb != ca) ents[cb].push_back(2 * k + 1);
      }
  }
  if (pairs.empty() || pairs.size() > 0x7fffffffu) return;
  std::vector<int64_t> ent_ptr{0}, ent;
  for (const auto& e : ents) {
    ent.insert(ent.end(), e.begin(), e.end());
    ent_ptr.push_back((int64_t)ent.size());
  }
  std::vector<int64_t> chunk_near;
  for (const auto& c : chunks) {
    const int64_t lo = rt.nodes[rt.leaves[c.leaf]].lo;
    chunk_near.push_back(first_chunk[c.leaf] + (c.r0 - lo) / kNearChunk);
  }
  h->pairs = upload(h, pairs, s);
  h->npairs = (int64_t)pairs.size();
  h->near_start = upload(h, near_start, s);
  h->ent_ptr = upload(h, ent_ptr, s);
  h->ent = upload(h, ent, s);
  h->chunk_near = upload(h, chunk_near, s);
  h->pair_part = dalloc<double>(h, 2 * pairs.size() * (size_t)kNearChunk);
  h->info.near_pairs = h->npairs;
}

// Everything after the coordinates and ids are in h_matrix_data: trees,
// partition, device lists, compression.
void setup(mpla_generic_matrix* A, h_matrix_data* data, HMatrix* h, const std::vector<int64_t>& ids_r_host,
           const std::vector<int64_t>& ids_c_host, mpla_instance* inst) {
  const char* what = "H-matrix setup";
  hipStream_t s = runtime(inst)->stream;
  const int dim = h->dim;
  const std::vector<double> cr = host_coords(data->coords_d[0], dim, h->nr, s);
  const std::vector<double> cc = host_coords(data->coords_d[1], dim, h->nc, s);
  h->rt = build_tree(cr.data(), h->nr, dim, data->c_leaf);
  h->ct = build_tree(cc.data(), h->nc, dim, data->c_leaf);
  h->part = partition(h->rt, h->ct, dim, data->eta);
  const Tree& rt = h->rt;
  const Tree& ct = h->ct;

  // leaf index of every row node that is a leaf; far blocks reach the leaves below their row cluster
  std::vector<int32_t> leaf_of(rt.nodes.size(), -1);
  for (size_t i = 0; i < rt.leaves.size(); ++i) leaf_of[rt.leaves[i]] = (int32_t)i;
  const int64_t nleaf = (int64_t)rt.leaves.size();
  std::vector<std::vector<std::pair<int64_t, int64_t>>> near_of(nleaf);
  std::vector<std::vector<int32_t>> far_of(nleaf);
  h->info = mpla_hmatrix_info{};
  for (const auto& nb : h->part.near) {
    const Cluster& a = rt.nodes[nb.first];
    const Cluster& b = ct.nodes[nb.second];
    near_of[leaf_of[nb.first]].push_back({b.lo, b.hi - b.lo});
    h->info.near_entries += (a.hi - a.lo) * (b.hi - b.lo);
  }
  std::vector<mplak_hm_block> blocks;
  std::vector<mplak_hm_item> items;
  std::vector<int64_t> first_item{0};
  int64_t uoff = 0, voff = 0;
  for (const auto& fb : h->part.far) {
    const Cluster& a = rt.nodes[fb.first];
    const Cluster& b = ct.nodes[fb.second];
    mplak_hm_block B;
    B.r0 = a.lo;
    B.m = a.hi - a.lo;
    B.c0 = b.lo;
    B.n = b.hi - b.lo;
    B.uoff = uoff;
    B.voff = voff;
    uoff += B.m * h->kmax;
    voff += B.n * h->kmax;
    const int32_t id = (int32_t)blocks.size();
    blocks.push_back(B);
    for (int64_t j0 = 0; j0 < B.n; j0 += kItemColumns)
      items.push_back(mplak_hm_item{id, 0, j0, std::min(kItemColumns, B.n - j0)});
    first_item.push_back((int64_t)items.size());
    h->info.far_entries += B.m * B.n;
    // every leaf under the row cluster
    std::vector<int32_t> st{fb.first};
    while (!st.empty()) {
      const int32_t c = st.back();
      st.pop_back();
      if (rt.nodes[c].child[0] < 0) {
        far_of[leaf_of[c]].push_back(id);
      } else {
        st.push_back(rt.nodes[c].child[1]);
        st.push_back(rt.nodes[c].child[0]);
      }
    }
  }
  if (blocks.size() > 0x7fffffffu) fail(what, "too many admissible blocks");
  // CSR lists per leaf, row chunks of up to 128 rows
  std::vector<int64_t> near_ptr{0}, near_cols, far_ptr{0}, full_ptr{0}, full_cols, nofar_ptr(nleaf + 1, 0);
  std::vector<int32_t> far_blk;
  std::vector<mplak_hm_chunk> chunks;
  for (int64_t l = 0; l < nleaf; ++l) {
    for (const auto& nc : near_of[l]) {
      near_cols.push_back(nc.first);
      near_cols.push_back(nc.second);
    }
    near_ptr.push_back((int64_t)near_cols.size() / 2);
    for (int32_t b : far_of[l]) far_blk.push_back(b);
    far_ptr.push_back((int64_t)far_blk.size());
    for (int32_t cl : ct.leaves) {
      full_cols.push_back(ct.nodes[cl].lo);
      full_cols.push_back(ct.nodes[cl].hi - ct.nodes[cl].lo);
    }
    full_ptr.push_back((int64_t)full_cols.size() / 2);
    const Cluster& c = rt.nodes[rt.leaves[l]];
    for (int64_t r0 = c.lo; r0 < c.hi; r0 += kChunkRows)
      chunks.push_back(mplak_hm_chunk{(int32_t)l, (int32_t)std::min<int64_t>(kChunkRows, c.hi - r0), r0});
  }
  h->info.row_leaves = nleaf;
  h->info.col_leaves = (int64_t)ct.leaves.size();
  h->info.near_blocks = (int64_t)h->part.near.size();
  h->info.far_blocks = (int64_t)blocks.size();
  h->info.k = h->kmax;

  // device memory: refuse what the device cannot hold before allocating
  const double factor_bytes = 8.0 * (double)(uoff + voff);
  size_t free_b = 0, total_b = 0;
  hip_ck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  if (factor_bytes > 0.9 * (double)free_b) {
    char msg[160];
    snprintf(msg, sizeof msg, "the ACA factors need %.3g GB, %.3g GB free: lower k or raise c_leaf",
             factor_bytes / 1e9, (double)free_b / 1e9);
    fail(what, msg);
  }
  h->perm_r = upload(h, rt.perm, s);
  h->perm_c = upload(h, ct.perm, s);
  std::vector<int64_t> g(h->nr);
  for (int64_t k = 0; k < h->nr; ++k) g[k] = ids_r_host[rt.perm[k]];
  h->gid_r = upload(h, g, s);
  g.assign(h->nc, 0);
  for (int64_t k = 0; k < h->nc; ++k) g[k] = ids_c_host[ct.perm[k]];
  h->gid_c = upload(h, g, s);
  h->ids_r = upload(h, ids_r_host, s);
  h->blocks = upload(h, blocks, s);
  h->items = upload(h, items, s);
  h->first_item = upload(h, first_item, s);
  h->chunks = upload(h, chunks, s);
  h->near_ptr = upload(h, near_ptr, s);
  h->near_cols = upload(h, near_cols, s);
  h->far_ptr = upload(h, far_ptr, s);
  h->far_blk = upload(h, far_blk, s);
  h->full_ptr = upload(h, full_ptr, s);
  h->full_cols = upload(h, full_cols, s);
  h->nofar_ptr = upload(h, nofar_ptr, s);
  h->nblocks = (int64_t)blocks.size();
  h->nitems = (int64_t)items.size();
  h->nchunks = (int64_t)chunks.size();
  if (h->closed_form && h->nr == h->nc && cr == cc) build_pairs(h, leaf_of, chunks, s);
  h->rank = dalloc<int32_t>(h, blocks.size());
  h->U = dalloc<double>(h, (size_t)uoff);
  h->V = dalloc<double>(h, (size_t)voff);
  h->t = dalloc<double>(h, blocks.size() * (size_t)h->kmax);
  h->fpart = dalloc<double>(h, items.size() * (size_t)h->kmax);
  h->xs = dalloc<double>(h, (size_t)h->nc);
  h->ys = dalloc<double>(h, (size_t)h->nr);
  if (h->closed_form) {
    h->pts_r = dalloc<double>(h, (size_t)dim * h->nr);
    h->pts_c = dalloc<double>(h, (size_t)dim * h->nc);
    double** tab = dalloc<double*>(h, 2 * kMaxDim);
    double* host_tab[2 * kMaxDim] = {};
    for (int d = 0; d < dim; ++d) {
      host_tab[d] = data->coords_d[0][d];
      host_tab[kMaxDim + d] = data->coords_d[1][d];
    }
    hip_ck(hipMemcpyAsync(tab, host_tab, sizeof(host_tab), hipMemcpyHostToDevice, s), "H-matrix upload");
    hip_ck(hipStreamSynchronize(s), "H-matrix upload");
    if (h->nr > 0)
      ck(mplak_hm_gather_scaled(tab, dim, 0, h->perm_r, h->nr, h->kind, h->length_scale, h->pts_r, h->nr, s), what);
    if (h->nc > 0)
      ck(mplak_hm_gather_scaled(tab + kMaxDim, dim, 0, h->perm_c, h->nc, h->kind, h->length_scale, h->pts_c, h->nc,
                                s),
         what);
    const mplak_hm_points p = points_of(h);
    ck(mplak_hm_aca(&p, h->blocks, h->nblocks, h->kmax, h->eps, h->U, h->V, h->rank, s), what);
  } else {
    ck(virtual_aca(h->assem, h->gid_r, h->gid_c, h->blocks, h->nblocks, h->kmax, h->eps, h->U, h->V, h->rank, s),
       what);
  }
  if (!h->closed_form) {
    // Virtual entries cost a device call each: evaluate the inadmissible
    // blocks once and keep them (column-major per near range, leading
    // dimension the leaf's row count), when they take at most half of the
    // free device memory; the near field of a product is then HBM-bound.
    hip_ck(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (8.0 * (double)h->info.near_entries <= 0.5 * (double)free_b) {
      std::vector<int64_t> range_leaf_lo, near_off, leaf_range;
      int64_t off = 0;
      for (int64_t l = 0; l < nleaf; ++l) {
        const Cluster& c = rt.nodes[rt.leaves[l]];
        leaf_range.push_back(c.lo);
        leaf_range.push_back(c.hi - c.lo);
        for (const auto& nc : near_of[l]) {
          range_leaf_lo.push_back(c.lo);
          range_leaf_lo.push_back(c.hi - c.lo);
          near_off.push_back(off);
          off += (c.hi - c.lo) * nc.second;
        }
      }
      int64_t* d_rll = upload(h, range_leaf_lo, s);
      h->near_off = upload(h, near_off, s);
      h->leaf_range = upload(h, leaf_range, s);
      h->near_st = dalloc<double>(h, (size_t)off);
      ck(virtual_near_fill(h->assem, h->gid_r, h->gid_c, d_rll, h->near_cols, h->near_off, (int64_t)near_off.size(),
                           h->near_st, s),
         what);
    }
  }
  std::vector<int32_t> ranks(blocks.size());
  if (!ranks.empty())
    hip_ck(hipMemcpyAsync(ranks.data(), h->rank, sizeof(int32_t) * ranks.size(), hipMemcpyDeviceToHost, s),
           "H-matrix ranks");
  hip_ck(hipStreamSynchronize(s), what);
  double rank_sum = 0.0;
  for (size_t b = 0; b < ranks.size(); ++b) {
    h->info.max_rank = std::max(h->info.max_rank, ranks[b]);
    rank_sum += ranks[b];
    h->info.stored_terms += (blocks[b].m + blocks[b].n) * ranks[b];
  }
  h->info.mean_rank = ranks.empty() ? 0.0 : rank_sum / (double)ranks.size();
  (void)A;
}

// h_matrix_data fields and the local coordinate / id copies (src/hmglib_adapter.cu:13-56)
HMatrix* start(mpla_generic_matrix* A, int dim, int c_leaf, int k, double eta, const char* what) {
  h_matrix_data* data = static_cast<h_matrix_data*>(A->data);
  if (data == nullptr) fail(what, "A->data must point to a struct h_matrix_data");
  if (dim < 1 || dim > kMaxDim) fail(what, "dim must be 1..4");
  if (c_leaf < 1) fail(what, "c_leaf must be >= 1");
  if (k < 1 || k > MPLAK_HM_MAX_RANK) fail(what, "k must be 1..32");
  if (!(eta > 0.0)) fail(what, "eta must be > 0");
  HMatrix* h = new HMatrix();
  h->dim = dim;
  h->kmax = k;
  h->nr = A->cur_proc_row_count;
  h->nc = A->cur_proc_col_count;
  h->row_off = A->cur_proc_row_offset;
  h->col_off = A->cur_proc_col_offset;
  memset(data, 0, sizeof(*data));
  data->point_count[0] = (int)h->nr;
  data->point_count[1] = (int)h->nc;
  data->dim = dim;
  data->eta = eta;
  data->max_level = 50;
  data->c_leaf = c_leaf;
  data->k = k;
  data->impl = h;
  for (int s2 = 0; s2 < 2; ++s2) {
    const int64_t n = s2 == 0 ? h->nr : h->nc;
    double** c = new double*[dim];
    for (int d = 0; d < dim; ++d) c[d] = dalloc<double>(h, (size_t)n);
    h->coords_host[s2] = c;
    data->coords_d[s2] = c;
    data->point_ids_d[s2] = reinterpret_cast<unsigned int*>(dalloc<unsigned int>(h, (size_t)n));
  }
  return h;
}

void copy_coords(h_matrix_data* data, HMatrix* h, double** all_coords[2], hipMemcpyKind kind, hipStream_t s) {
  for (int d = 0; d < h->dim; ++d) {
    if (h->nr > 0)
      hip_ck(hipMemcpyAsync(data->coords_d[0][d], all_coords[0][d] + h->row_off, sizeof(double) * h->nr, kind, s),
             "H-matrix coordinates");
    if (h->nc > 0)
      hip_ck(hipMemcpyAsync(data->coords_d[1][d], all_coords[1][d] + h->col_off, sizeof(double) * h->nc, kind, s),
             "H-matrix coordinates");
  }
}

std::vector<int64_t> index_ids(int64_t n, int64_t offset) {
  std::vector<int64_t> v(n);
  for (int64_t i = 0; i < n; ++i) v[i] = offset + i;
  return v;
}

void upload_ids(h_matrix_data* data, int which, const std::vector<int64_t>& ids, hipStream_t s) {
  std::vector<unsigned int> u(ids.size());
  for (size_t i = 0; i < ids.size(); ++i) u[i] = (unsigned int)ids[i];
  if (!u.empty())
    hip_ck(hipMemcpyAsync(data->point_ids_d[which], u.data(), sizeof(unsigned int) * u.size(), hipMemcpyHostToDevice, s),
           "H-matrix point ids");
  hip_ck(hipStreamSynchronize(s), "H-matrix point ids");
}

// y (rows, caller order) = the sorted operator applied to x (columns, caller order)
void apply(HMatrix* h, bool full, const double* x, double* y, mpla_instance* inst) {
  const char* what = full ? "mpla_dgemv_core_hmglib_full_matrix" : "mpla_dgemv_core_hmglib_h_matrix";
  hipStream_t s = runtime(inst)->stream;
  if (h->nr == 0) return;
  if (h->nc > 0) ck(mplak_gather_f64(h->xs, x, h->perm_c, h->nc, s), what);
  const bool far = !full && h->nblocks > 0;
  if (far) {
    ck(mplak_hm_far_partials(h->blocks, h->items, h->nitems, h->rank, h->kmax, h->V, h->xs, h->fpart, s), what);
    ck(mplak_hm_far_sum(h->first_item, h->nblocks, h->rank, h->kmax, h->fpart, h->t, s), what);
  }
  const int64_t* np = full ? h->full_ptr : h->near_ptr;
  const int64_t* nc = full ? h->full_cols : h->near_cols;
  const int64_t* fp = far ? h->far_ptr : h->nofar_ptr;
  if (h->closed_form && !full && h->npairs > 0) {
    const mplak_hm_points p = points_of(h);
    ck(mplak_hm_near_pairs(&p, h->pairs, h->npairs, h->xs, h->pair_part, s), what);
    ck(mplak_hm_apply_pairs(h->chunks, h->nchunks, h->chunk_near, h->near_start, h->ent_ptr, h->ent, h->pair_part,
                            fp, h->far_blk, h->blocks, h->rank, h->kmax, h->U, h->t, h->ys, s),
       what);
  } else if (h->closed_form) {
    const mplak_hm_points p = points_of(h);
    ck(mplak_hm_apply(&p, h->chunks, h->nchunks, np, nc, fp, h->far_blk, h->blocks, h->rank, h->kmax, h->U, h->t,
                      h->xs, h->ys, s),
       what);
  } else if (!full && h->near_st != nullptr) {
    ck(mplak_hm_apply_stored(h->chunks, h->nchunks, h->leaf_range, h->near_ptr, h->near_cols, h->near_off, h->near_st,
                             fp, h->far_blk, h->blocks, h->rank, h->kmax, h->U, h->t, h->xs, h->ys, s),
       what);
  } else {
    ck(virtual_apply(h->assem, h->gid_r, h->gid_c, h->chunks, h->nchunks, np, nc, fp, h->far_blk, h->blocks, h->rank,
                     h->kmax, h->U, h->t, h->xs, h->ys, s),
       what);
  }
  ck(mplak_hm_scatter_nugget(h->ys, h->perm_r, h->nr, h->closed_form ? h->nugget : 0.0, x, h->row_off - h->col_off,
                             h->nc, y, s),
     what);
}

}  // namespace

extern "C" {

void mpla_init_hmglib(struct mpla_generic_matrix* A, int global_point_count[2], double** all_coords[2],
                      unsigned int* all_point_ids[2], struct system_assembler* assem, double eta, int dim, int bits,
                      int c_leaf, int k, char root_level_set_1, char root_level_set_2, int max_batched_dense_size,
                      int max_batched_aca_size, struct mpla_instance* instance) {
  const char* what = "mpla_init_hmglib";
  if (assem == nullptr) fail(what, "assem must be a device system_assembler");
  (void)global_point_count;
  HMatrix* h = start(A, dim, c_leaf, k, eta, what);
  h_matrix_data* data = static_cast<h_matrix_data*>(A->data);
  data->bits = bits;
  data->root_level_set_1 = root_level_set_1;
  data->root_level_set_2 = root_level_set_2;
  data->max_batched_dense_size = max_batched_dense_size;
  data->max_batched_aca_size = max_batched_aca_size;
  data->assem = assem;
  h->assem = assem;
  hipStream_t s = runtime(instance)->stream;
  copy_coords(data, h, all_coords, hipMemcpyDeviceToDevice, s);
  std::vector<int64_t> ids[2];
  for (int w = 0; w < 2; ++w) {
    const int64_t n = w == 0 ? h->nr : h->nc;
    const int64_t off = w == 0 ? h->row_off : h->col_off;
    std::vector<unsigned int> u(n);
    if (n > 0)
      hip_ck(hipMemcpyAsync(u.data(), all_point_ids[w] + off, sizeof(unsigned int) * n, hipMemcpyDeviceToHost, s),
             "H-matrix point ids");
    hip_ck(hipStreamSynchronize(s), "H-matrix point ids");
    ids[w].assign(u.begin(), u.end());
    upload_ids(data, w, ids[w], s);
  }
  setup(A, data, h, ids[0], ids[1], instance);
}

void mpla_init_hmglib_coords_from_host(struct mpla_generic_matrix* A, int global_point_count[2],
                                       double** all_coords[2], struct system_assembler* assem, double eta, int dim,
                                       int bits, int c_leaf, int k, int max_batched_dense_size,
                                       int max_batched_aca_size, struct mpla_instance* instance) {
  const char* what = "mpla_init_hmglib_coords_from_host";
  if (assem == nullptr) fail(what, "assem must be a device system_assembler");
  (void)global_point_count;
  HMatrix* h = start(A, dim, c_leaf, k, eta, what);
  h_matrix_data* data = static_cast<h_matrix_data*>(A->data);
  data->bits = bits;
  data->max_batched_dense_size = max_batched_dense_size;
  data->max_batched_aca_size = max_batched_aca_size;
  data->assem = assem;
  h->assem = assem;
  hipStream_t s = runtime(instance)->stream;
  copy_coords(data, h, all_coords, hipMemcpyHostToDevice, s);
  const std::vector<int64_t> ir = index_ids(h->nr, h->row_off), ic = index_ids(h->nc, h->col_off);
  upload_ids(data, 0, ir, s);
  upload_ids(data, 1, ic, s);
  setup(A, data, h, ir, ic, instance);
}

void mpla_init_hmatrix_kernel(struct mpla_generic_matrix* A, double** all_coords[2], int dim, int kernel,
                              double length_scale, double nugget, double eta, int c_leaf, int k, double epsilon,
                              struct mpla_instance* instance) {
  const char* what = "mpla_init_hmatrix_kernel";
  if (kernel != MPLA_KERNEL_GAUSSIAN && kernel != MPLA_KERNEL_MATERN32) fail(what, "unknown kernel");
  if (!(length_scale > 0.0)) fail(what, "length_scale must be > 0");
  if (!(epsilon >= 0.0)) fail(what, "epsilon must be >= 0");
  HMatrix* h = start(A, dim, c_leaf, k, eta, what);
  h_matrix_data* data = static_cast<h_matrix_data*>(A->data);
  data->epsilon = epsilon;
  h->closed_form = true;
  h->kind = kernel == MPLA_KERNEL_GAUSSIAN ? MPLAK_KERNEL_GAUSSIAN : MPLAK_KERNEL_MATERN32;
  h->length_scale = length_scale;
  h->nugget = nugget;
  h->eps = epsilon;
  hipStream_t s = runtime(instance)->stream;
  copy_coords(data, h, all_coords, hipMemcpyDeviceToDevice, s);
  const std::vector<int64_t> ir = index_ids(h->nr, h->row_off), ic = index_ids(h->nc, h->col_off);
  upload_ids(data, 0, ir, s);
  upload_ids(data, 1, ic, s);
  setup(A, data, h, ir, ic, instance);
}

void mpla_destroy_hmglib(struct mpla_generic_matrix* A, struct mpla_instance* instance) {
  h_matrix_data* data = static_cast<h_matrix_data*>(A->data);
  if (data == nullptr || data->impl == nullptr) return;
  HMatrix* h = static_cast<HMatrix*>(data->impl);
  hip_ck(hipStreamSynchronize(runtime(instance)->stream), "mpla_destroy_hmglib");
  for (void* p : h->allocs) (void)hipFree(p);
  delete[] h->coords_host[0];
  delete[] h->coords_host[1];
  delete h;
  data->impl = nullptr;
  data->coords_d[0] = data->coords_d[1] = nullptr;
  data->point_ids_d[0] = data->point_ids_d[1] = nullptr;
}

void mpla_dgemv_core_hmglib_h_matrix(struct mpla_vector* b, struct mpla_generic_matrix* A,
                                     struct mpla_vector* x_redist, struct mpla_instance* instance) {
  apply(impl_of(A, "mpla_dgemv_core_hmglib_h_matrix"), false, x_redist->data, b->data, instance);
}

void mpla_dgemv_core_hmglib_full_matrix(struct mpla_vector* b, struct mpla_generic_matrix* A,
                                        struct mpla_vector* x_redist, struct mpla_instance* instance) {
  apply(impl_of(A, "mpla_dgemv_core_hmglib_full_matrix"), true, x_redist->data, b->data, instance);
}

void set_gaussian_kernel_rhs(double* b, struct h_matrix_data* data) {
  const char* what = "set_gaussian_kernel_rhs";
  if (data == nullptr || data->impl == nullptr) fail(what, "H-matrix not initialised");
  HMatrix* h = static_cast<HMatrix*>(data->impl);
  if (h->closed_form) {
    // the integral of exp(-|p - y|^2) over the unit cube at each row point
    double** tab = nullptr;
    hip_ck(hipMalloc(reinterpret_cast<void**>(&tab), sizeof(double*) * kMaxDim), what);
    hip_ck(hipMemcpy(tab, data->coords_d[0], sizeof(double*) * h->dim, hipMemcpyHostToDevice), what);
    ck(mplak_gaussian_kernel_rhs(tab, h->dim, 0, h->nr, b, nullptr), what);
    hip_ck(hipDeviceSynchronize(), what);
    (void)hipFree(tab);
  } else {
    ck(virtual_rhs(h->assem, h->ids_r, h->nr, b, nullptr), what);
    hip_ck(hipDeviceSynchronize(), what);
  }
}

void mpla_hmatrix_get_info(struct mpla_generic_matrix* A, struct mpla_hmatrix_info* info) {
  *info = impl_of(A, "mpla_hmatrix_get_info")->info;
}

void mpla_hmatrix_get_permutation(struct mpla_generic_matrix* A, int which, int64_t* perm) {
  HMatrix* h = impl_of(A, "mpla_hmatrix_get_permutation");
  const std::vector<int64_t>& p = which == 0 ? h->rt.perm : h->ct.perm;
  std::copy(p.begin(), p.end(), perm);
}

int mpla_hmatrix_partition(const double* coords, int64_t n, int dim, double eta, int c_leaf, int64_t* perm,
                           int64_t* counts) {
  if (coords == nullptr || n < 0 || dim < 1 || dim > kMaxDim || !(eta > 0.0) || c_leaf < 1) return 1;
  const Tree t = build_tree(coords, n, dim, c_leaf);
  const Partition p = partition(t, t, dim, eta);
  if (perm) std::copy(t.perm.begin(), t.perm.end(), perm);
  if (counts) {
    int64_t near_entries = 0;
    for (const auto& nb : p.near)
      near_entries += (t.nodes[nb.first].hi - t.nodes[nb.first].lo) * (t.nodes[nb.second].hi - t.nodes[nb.second].lo);
    counts[0] = (int64_t)t.leaves.size();
    counts[1] = (int64_t)p.near.size();
    counts[2] = (int64_t)p.far.size();
    counts[3] = near_entries;
  }
  return 0;
}

}  // extern "C"
