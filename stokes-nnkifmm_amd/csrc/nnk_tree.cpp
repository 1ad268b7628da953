// Octree construction and interaction lists.
//
// Semantics follow the reference exactly (Classes/OcTree.m):
//  * points = unique([pot; tgt],'rows','stable'); masks by ismember   (:80-82)
//  * root cube [m m m M M M] from the min/max over all coordinates      (:97-98)
//  * capacity = min(nodeCapacity, nPts-1)                              (:105)
//  * Divide visits nodes in order, errors when level+1 >= maxDepth,
//    splits when a node holds more than capacity points, recursing into
//    the 8 new children before the next sibling (DFS creation order)  (:165-178)
//  * centre = (lo+hi)/2, child k = 1 + 4[x>c] + 2[y>c] + [z>c]          (:191-221)
//  * closed-box adjacency                       (Tools/Tree/nodeintersection.m)
//  * U, V, W lists and the W coarsening, X = find(Wpadded == i)         (:265-366)
// De-duplication, subdivision and lists run on the GPU (nnk_tree_gpu.hip:
// three stable radix sorts of the rows, path keys and one radix sort,
// level-by-level node creation, preorder ids, per-node tree walks for the
// lists); this file holds the checks, the per-leaf point lists and the
// NEAREST attachment.
#include "nnk_tree.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace nnk {

int build_tree(const double *pot, int64_t npot, const double *tgt, int64_t ntgt, int capacity,
               int max_depth, Tree &t, std::string &err, const double *cube) {
  t = Tree();
  using clk = std::chrono::steady_clock;
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  const auto t0 = clk::now();
  if (npot < 0 || ntgt < 0 || npot + ntgt < 2) {
    err = "OcTree needs at least two points";
    return NNK_EARG;
  }
  for (int64_t i = 0; i < 3 * npot; i++)
    if (pot[i] != pot[i]) { err = "OcTree: NaN in potential points"; return NNK_EARG; }
  for (int64_t i = 0; i < 3 * ntgt; i++)
    if (tgt[i] != tgt[i]) { err = "OcTree: NaN in target points"; return NNK_EARG; }
  dedup_gpu(pot, npot, tgt, ntgt, t);
  const int64_t np = t.npts;
  t.pot_ord.assign(np, -1);
  t.tgt_ord.assign(np, -1);
  for (int64_t i = 0; i < np; i++) {
    if (t.is_pot[i]) t.pot_ord[i] = t.npot++;
    if (t.is_tgt[i]) t.tgt_ord[i] = t.ntgt++;
  }
  if (np < 2) {
    err = "OcTree needs at least two distinct points";
    return NNK_EARG;
  }
  double m = t.points[0], M = t.points[0];
  for (int64_t i = 0; i < 3 * np; i++) {
    m = std::min(m, t.points[i]);
    M = std::max(M, t.points[i]);
  }
  if (cube) {
    if (!(cube[0] <= m && M <= cube[1])) {
      err = "OcTree: the root cube does not contain every point";
      return NNK_EARG;
    }
    m = cube[0];
    M = cube[1];
  }
  t.max_depth = max_depth;
  t.capacity = (int32_t)std::min<int64_t>(capacity, np - 1);
  const auto t1 = clk::now();
  int rcode = subdivide_gpu(t, m, M, err);
  if (rcode) return rcode;
  const auto t2 = clk::now();
  t.max_level = 0;
  for (int n = 0; n < t.count; n++) t.max_level = std::max(t.max_level, t.level[n]);
  t.pot_leaf.assign(t.count + 1, 0);
  t.tgt_leaf.assign(t.count + 1, 0);
  for (int64_t i = 0; i < np; i++) {
    if (t.is_pot[i]) t.pot_leaf[t.point_index[i]] = 1;
    if (t.is_tgt[i]) t.tgt_leaf[t.point_index[i]] = 1;
  }
  t.pt_ptr.assign(t.count + 2, 0);
  for (int64_t i = 0; i < np; i++) t.pt_ptr[t.point_index[i] + 1]++;
  for (int n = 1; n <= t.count; n++) t.pt_ptr[n + 1] += t.pt_ptr[n];
  t.pt_rows.resize(np);
  {
    std::vector<int64_t> fill(t.count + 2, 0);
    for (int64_t i = 0; i < np; i++) {
      int n = t.point_index[i];
      t.pt_rows[t.pt_ptr[n] + fill[n]++] = i;
    }
  }
  const auto t3 = clk::now();
  build_lists_gpu(t);
  const auto t4 = clk::now();
  t.build_ms[0] = ms(t0, t1);
  t.build_ms[1] = ms(t1, t2);
  t.build_ms[2] = ms(t2, t3);
  t.build_ms[3] = ms(t3, t4);
  return NNK_OK;
}

int validate_nn_csc(int64_t ncol, int64_t nfine, const double *fine, const int64_t *colptr, const int64_t *row,
                    const double *val, std::string &err) {
  if (nfine < 0) { err = "negative fine point count"; return NNK_EARG; }
  if (!colptr) { err = "null NN colptr"; return NNK_EARG; }
  if (nfine > 0 && !fine) { err = "null fine points"; return NNK_EARG; }
  if (colptr[0] != 0) { err = "NN colptr[0] must be 0"; return NNK_EARG; }
  for (int64_t c = 0; c < ncol; c++)
    if (colptr[c + 1] < colptr[c]) { err = "NN colptr not monotone"; return NNK_EARG; }
  const int64_t nnz = colptr[ncol];
  if (nnz > 0 && (!row || !val)) { err = "null NN row indices or values"; return NNK_EARG; }
  for (int64_t e = 0; e < nnz; e++)
    if (row[e] < 0 || row[e] >= nfine) { err = "NN row index out of range"; return NNK_EARG; }
  return NNK_OK;
}

int set_nearest(Tree &t, const double *fine, int64_t nfine, const int64_t *colptr,
                const int64_t *row, const double *val, std::string &err) {
  if (t.npot != t.ntgt) {
    err = "NEAREST mode needs potential points == target points";
    return NNK_EARG;
  }
  if (int rc = validate_nn_csc(t.npot, nfine, fine, colptr, row, val, err)) return rc;
  const int64_t nnz = colptr[t.npot];
  t.nearest = true;
  t.nfine = nfine;
  t.fine.assign(fine, fine + 3 * nfine);
  t.nn_colptr.assign(colptr, colptr + t.npot + 1);
  t.nn_row.assign(row, row + nnz);
  t.nn_val.assign(val, val + nnz);
  return NNK_OK;
}

}  // namespace nnk
