// Octree construction and interaction lists.
//
// Semantics follow the reference exactly (Classes/OcTree.m):
//  * points = unique([pot; tgt],'rows','stable'); masks by ismember   (:80-82)
//    (on the GPU: three stable radix sorts by z, y, x, runs of equal rows)
//  * root cube [m m m M M M] from the min/max over all coordinates      (:97-98)
//  * capacity = min(nodeCapacity, nPts-1)                              (:105)
//  * Divide visits nodes in order, errors when level+1 >= maxDepth,
//    splits when a node holds more than capacity points, recursing into
//    the 8 new children before the next sibling (DFS creation order)  (:165-178)
//  * centre = (lo+hi)/2, child k = 1 + 4[x>c] + 2[y>c] + [z>c]          (:191-221)
// The de-duplication and the subdivision run on the GPU (nnk_tree_gpu.hip:
// path keys, one radix sort, level-by-level node creation, preorder ids);
// the lists run here.
//  * closed-box adjacency                       (Tools/Tree/nodeintersection.m)
//  * U, V, W lists and the W coarsening, X = find(Wpadded == i)         (:265-366)
// The lists are found by descending from the root through the nodes whose
// closed box meets the query box (a child box lies inside its parent box,
// because child corners are copies of the parent's lo/centre/hi), instead
// of the reference's scan over all nodes.  The W coarsening is computed from
// a per-node "complete subtree" flag; the result, including its order
// (sorted, with the last merged node appended), is the reference's.
#include "nnk_tree.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace nnk {

namespace {

inline bool isect(const double *c1, const double *c2) {
  return !(c2[0] > c1[3] || c2[3] < c1[0] || c2[1] > c1[4] || c2[4] < c1[1] ||
           c2[2] > c1[5] || c2[5] < c1[2]);
}

// Visit every node whose closed box meets q and whose level <= maxlev.
template <class F>
void descend(const Tree &t, const double *q, int maxlev, F &&visit) {
  int stack[8 * 64];
  int sp = 0;
  stack[sp++] = 1;
  while (sp) {
    int n = stack[--sp];
    if (!isect(q, t.box(n))) continue;
    visit(n);
    if (!t.is_leaf(n) && t.level[n - 1] < maxlev)
      for (int k = 7; k >= 0; k--) stack[sp++] = t.children[8 * (n - 1) + k];
  }
}

void build_lists(Tree &t) {
  const int nc = t.count;
  // complete(x): x is a potential leaf, or a non-leaf whose 8 children are
  // complete -- exactly the nodes the W coarsening ends up merging.
  std::vector<uint8_t> complete(nc + 1, 0);
  {
    std::vector<int> order(nc);
    for (int i = 0; i < nc; i++) order[i] = i + 1;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return t.level[a - 1] > t.level[b - 1]; });
    for (int n : order) {
      if (t.is_leaf(n)) complete[n] = t.pot_leaf[n];
      else {
        bool all = true;
        for (int k = 0; k < 8; k++) all = all && complete[t.children[8 * (n - 1) + k]];
        complete[n] = all;
      }
    }
  }
  std::vector<std::vector<int32_t>> L[3];
  for (int q = 0; q < 3; q++) L[q].assign(nc + 1, {});
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 2; i <= nc; i++) {
    const double *bi = t.box(i);
    const int li = t.level[i - 1];
    const bool tl = t.tgt_leaf[i];
    // U: potential leaves adjacent to i, sorted (OcTree.m:323-327)
    if (tl) {
      auto &U = L[0][i];
      descend(t, bi, 1 << 30, [&](int n) {
        if (t.is_leaf(n) && t.pot_leaf[n]) U.push_back(n);
      });
      std::sort(U.begin(), U.end());
    }
    // V: children of the non-leaf same-level neighbours of the parent that
    // are not adjacent to i (OcTree.m:330-339)
    {
      auto &V = L[1][i];
      int P = t.parent[i - 1];
      int lp = t.level[P - 1];
      descend(t, t.box(P), lp, [&](int n) {
        if (n == P || t.level[n - 1] != lp || t.is_leaf(n)) return;
        for (int k = 0; k < 8; k++) {
          int c = t.children[8 * (n - 1) + k];
          if (!isect(bi, t.box(c))) V.push_back(c);
        }
      });
      std::sort(V.begin(), V.end());
    }
    // W: potential leaves under the non-leaf same-level neighbours of i
    // that are not adjacent to i, coarsened (OcTree.m:343-360)
    if (tl) {
      auto &W = L[2][i];
      std::vector<int> st;
      descend(t, bi, li, [&](int n) {
        if (t.level[n - 1] == li && !t.is_leaf(n)) st.push_back(n);
      });
      if (!st.empty()) {
        std::vector<int> stack;
        for (int c : st)
          for (int k = 0; k < 8; k++) stack.push_back(t.children[8 * (c - 1) + k]);
        int last = 0;  // last merged node: smallest level, then largest id
        while (!stack.empty()) {
          int y = stack.back();
          stack.pop_back();
          bool leaf = t.is_leaf(y);
          if (isect(bi, t.box(y))) {
            if (!leaf)
              for (int k = 0; k < 8; k++) stack.push_back(t.children[8 * (y - 1) + k]);
          } else if (complete[y]) {
            W.push_back(y);
            if (!leaf) {
              if (!last || t.level[y - 1] < t.level[last - 1] ||
                  (t.level[y - 1] == t.level[last - 1] && y > last))
                last = y;
            }
          } else if (!leaf) {
            for (int k = 0; k < 8; k++) stack.push_back(t.children[8 * (y - 1) + k]);
          }
        }
        std::sort(W.begin(), W.end());
        if (last) {
          W.erase(std::find(W.begin(), W.end(), last));
          W.push_back(last);
        }
      }
    }
  }
  // X(i) = { j : i in W(j) }, in the order of find() on the padded W
  // matrix: position in W(j) first, then j (OcTree.m:279-288).
  std::vector<std::vector<int32_t>> X(nc + 1);
  size_t maxw = 0;
  for (int n = 1; n <= nc; n++) maxw = std::max(maxw, L[2][n].size());
  {
    std::vector<int32_t> withw;
    for (int n = 1; n <= nc; n++)
      if (!L[2][n].empty()) withw.push_back(n);
    for (size_t k = 0; k < maxw; k++)
      for (int j : withw)
        if (L[2][j].size() > k) X[L[2][j][k]].push_back(j);
  }
  std::vector<std::vector<int32_t>> *all[4] = {&L[0], &L[1], &L[2], &X};
  for (int q = 0; q < 4; q++) {
    auto &src = *all[q];
    t.lptr[q].assign(nc + 2, 0);
    for (int n = 1; n <= nc; n++) t.lptr[q][n + 1] = t.lptr[q][n] + (int64_t)src[n].size();
    t.lidx[q].resize(t.lptr[q][nc + 1]);
    for (int n = 1; n <= nc; n++)
      std::copy(src[n].begin(), src[n].end(), t.lidx[q].begin() + t.lptr[q][n]);
  }
}

}  // namespace

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
  build_lists(t);
  const auto t4 = clk::now();
  t.build_ms[0] = ms(t0, t1);
  t.build_ms[1] = ms(t1, t2);
  t.build_ms[2] = ms(t2, t3);
  t.build_ms[3] = ms(t3, t4);
  return NNK_OK;
}

int set_nearest(Tree &t, const double *fine, int64_t nfine, const int64_t *colptr,
                const int64_t *row, const double *val, std::string &err) {
  if (nfine < 0) { err = "negative fine point count"; return NNK_EARG; }
  if (t.npot != t.ntgt) {
    err = "NEAREST mode needs potential points == target points";
    return NNK_EARG;
  }
  int64_t nnz = colptr[t.npot];
  for (int64_t c = 0; c < t.npot; c++)
    if (colptr[c + 1] < colptr[c]) { err = "NN colptr not monotone"; return NNK_EARG; }
  for (int64_t e = 0; e < nnz; e++)
    if (row[e] < 0 || row[e] >= nfine) { err = "NN row index out of range"; return NNK_EARG; }
  t.nearest = true;
  t.nfine = nfine;
  t.fine.assign(fine, fine + 3 * nfine);
  t.nn_colptr.assign(colptr, colptr + t.npot + 1);
  t.nn_row.assign(row, row + nnz);
  t.nn_val.assign(val, val + nnz);
  return NNK_OK;
}

}  // namespace nnk
