// Octree subdivision on the GPU with the reference's node numbering
// (Classes/OcTree.m:155-222).
//
// A point's path through the tree depends only on the root cube: at each
// level its box is split at the centre (lo+hi)/2 and it goes to child
// 1 + 4[x>cx] + 2[y>cy] + [z>cz] (:191-221), the child box being copies of
// the parent's lo/centre/hi.  So every point's path key (3 bits per level,
// most significant first) is computed independently, with the same
// floating-point operations as the reference, and one stable radix sort
// groups the points of every node of every level into a contiguous range
// (points of a node stay in ascending row order).
//
// The nodes are then created level by level (breadth first): a node splits
// when it holds more than `capacity` points, and its 8 children (all of
// them, also empty ones, :191-221) find their point ranges by binary search
// in the sorted keys.  The reference numbers nodes in the order Divide
// creates them: a splitting node allocates 8 consecutive ids, and Divide
// recurses into each child before the next sibling (:165-178).  A node's
// children therefore get ids 2 + 8 r .. 9 + 8 r, r = the number of splitting
// nodes before it in preorder; r is found with one bottom-up pass (number of
// splitting nodes per subtree) and one top-down pass.
//
// The interaction lists are found per node by walks over the tree (no
// stack: siblings have consecutive ids), counted, scanned, filled and
// sorted with segmented sorts; X is the inverse of W by one radix sort.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "nnk_common.hpp"
#include "nnk_tree.hpp"

namespace nnk {

namespace {

constexpr int kKeyLevels = 21;   // 63 bits of path key

__host__ __device__ inline int key_shift(int level) { return 3 * (kKeyLevels - level); }

__global__ void k_tree_keys(const double *__restrict__ pts, int64_t np, double m, double M, int nlev,
                            uint64_t *__restrict__ key, int32_t *__restrict__ row) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double lo[3] = {m, m, m}, hi[3] = {M, M, M};
  const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  uint64_t k = 0;
  for (int l = 1; l <= nlev; l++) {
    int d = 0;
    for (int a = 0; a < 3; a++) {
      const double c = (lo[a] + hi[a]) / 2.0;
      const bool up = p[a] > c;
      d = 2 * d + (up ? 1 : 0);
      if (up) lo[a] = c;
      else hi[a] = c;
    }
    k |= (uint64_t)d << key_shift(l);
  }
  key[i] = k;
  row[i] = (int32_t)i;
}

__global__ void k_split_flags(int64_t n, const int64_t *__restrict__ beg, const int64_t *__restrict__ end,
                              int64_t cap, int32_t *__restrict__ split) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) split[i] = (end[i] - beg[i] > cap) ? 1 : 0;
}

// first index in [lo, hi) with key >= x
__device__ inline int64_t lower_bound(const uint64_t *__restrict__ key, int64_t lo, int64_t hi, uint64_t x) {
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (key[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// children (level l+1) of the splitting nodes of level l
__global__ void k_make_children(int64_t n, int l, const int32_t *__restrict__ split,
                                const int32_t *__restrict__ scan, const int64_t *__restrict__ beg,
                                const int64_t *__restrict__ end, const uint64_t *__restrict__ pfx,
                                const double *__restrict__ corners, const uint64_t *__restrict__ key,
                                int64_t *__restrict__ cbeg, int64_t *__restrict__ cend,
                                uint64_t *__restrict__ cpfx, double *__restrict__ ccorners) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !split[i]) return;
  const int sh = key_shift(l + 1);
  const double *c = corners + 6 * i;
  const double co[9] = {c[0], c[1], c[2], (c[0] + c[3]) / 2.0, (c[1] + c[4]) / 2.0, (c[2] + c[5]) / 2.0,
                        c[3], c[4], c[5]};
  // lo/hi rows of the 8 children in the 3x3 table [lo; centre; hi]
  const int rows[8][6] = {{0, 1, 2, 3, 4, 5}, {0, 1, 5, 3, 4, 8}, {0, 4, 2, 3, 7, 5}, {0, 4, 5, 3, 7, 8},
                          {3, 1, 2, 6, 4, 5}, {3, 1, 5, 6, 4, 8}, {3, 4, 2, 6, 7, 5}, {3, 4, 5, 6, 7, 8}};
  int64_t b = beg[i];
  for (int k = 0; k < 8; k++) {
    const int64_t ch = 8 * (int64_t)scan[i] + k;
    const uint64_t p = pfx[i] | ((uint64_t)k << sh);
    const int64_t e = k == 7 ? end[i] : lower_bound(key, b, end[i], pfx[i] | ((uint64_t)(k + 1) << sh));
    cbeg[ch] = b;
    cend[ch] = e;
    cpfx[ch] = p;
    for (int a = 0; a < 6; a++) ccorners[6 * ch + a] = co[rows[k][a]];
    b = e;
  }
}

// number of splitting nodes in each subtree (children level done first)
__global__ void k_subtree_splits(int64_t n, const int32_t *__restrict__ split, const int32_t *__restrict__ scan,
                                 const int64_t *__restrict__ sub_next, int64_t *__restrict__ sub) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t s = 0;
  if (split[i]) {
    s = 1;
    for (int k = 0; k < 8; k++) s += sub_next[8 * (int64_t)scan[i] + k];
  }
  sub[i] = s;
}

// preorder rank among splitting nodes and ids of the children
__global__ void k_child_ids(int64_t n, const int32_t *__restrict__ split, const int32_t *__restrict__ scan,
                            const int64_t *__restrict__ rank, const int32_t *__restrict__ id,
                            const int64_t *__restrict__ sub_next, int64_t *__restrict__ crank,
                            int32_t *__restrict__ cid, int32_t *__restrict__ cparent) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !split[i]) return;
  int64_t acc = rank[i] + 1;
  for (int k = 0; k < 8; k++) {
    const int64_t c = 8 * (int64_t)scan[i] + k;
    crank[c] = acc;
    cid[c] = (int32_t)(2 + 8 * rank[i] + k);
    cparent[c] = id[i];
    acc += sub_next[c];
  }
}

// node arrays in id order, and the leaf id of every point
__global__ void k_emit_nodes(int64_t n, int l, const int32_t *__restrict__ split, const int32_t *__restrict__ scan,
                             const int32_t *__restrict__ id, const int32_t *__restrict__ parent,
                             const double *__restrict__ corners, const int64_t *__restrict__ beg,
                             const int64_t *__restrict__ end, const int32_t *__restrict__ cid,
                             const int32_t *__restrict__ rows, double *__restrict__ o_corners,
                             int32_t *__restrict__ o_level, int32_t *__restrict__ o_parent,
                             int32_t *__restrict__ o_children, int32_t *__restrict__ o_pindex) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t d = id[i] - 1;
  for (int a = 0; a < 6; a++) o_corners[6 * d + a] = corners[6 * i + a];
  o_level[d] = l;
  o_parent[d] = parent[i];
  for (int k = 0; k < 8; k++) o_children[8 * d + k] = split[i] ? cid[8 * (int64_t)scan[i] + k] : 0;
  if (!split[i])
    for (int64_t p = beg[i]; p < end[i]; p++) o_pindex[rows[p]] = id[i];
}

// ---- unique([pot; tgt], 'rows', 'stable') and the ismember masks ----

// order-preserving integer image of a double (-0 and 0 map alike)
__device__ inline uint64_t sortable(double v) {
  v = v == 0.0 ? 0.0 : v;
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

__global__ void k_row_key(int64_t n, const double *__restrict__ R, const int32_t *__restrict__ idx, int coord,
                          uint64_t *__restrict__ key, int32_t *__restrict__ val) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int32_t i = idx ? idx[j] : (int32_t)j;
  key[j] = sortable(R[3 * (int64_t)i + coord]);
  val[j] = i;
}

// heads of runs of equal rows in sorted order
__global__ void k_row_heads(int64_t n, const double *__restrict__ R, const int32_t *__restrict__ ord,
                            int32_t *__restrict__ head) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int h = 1;
  if (j > 0) {
    const double *a = R + 3 * (int64_t)ord[j], *b = R + 3 * (int64_t)ord[j - 1];
    h = !(a[0] == b[0] && a[1] == b[1] && a[2] == b[2]);
  }
  head[j] = h;
}

// rep[i]: first (smallest) row index equal to row i; gscan = inclusive scan of heads
__global__ void k_row_rep(int64_t n, const int32_t *__restrict__ ord, const int32_t *__restrict__ head,
                          const int32_t *__restrict__ gscan, int32_t *__restrict__ leader, int pass,
                          int32_t *__restrict__ rep) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int32_t g = gscan[j] - 1;
  if (pass == 0) {
    if (head[j]) leader[g] = ord[j];
  } else {
    rep[ord[j]] = leader[g];
  }
}

__global__ void k_row_keep(int64_t n, const int32_t *__restrict__ rep, int32_t *__restrict__ keep) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = rep[i] == (int32_t)i ? 1 : 0;
}

__global__ void k_row_emit(int64_t n, int64_t npot, const double *__restrict__ R, const int32_t *__restrict__ rep,
                           const int32_t *__restrict__ keep, const int32_t *__restrict__ uidx,
                           double *__restrict__ pts, uint8_t *__restrict__ is_pot, uint8_t *__restrict__ is_tgt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (keep[i])
    for (int d = 0; d < 3; d++) pts[3 * (int64_t)uidx[i] + d] = R[3 * i + d];
  const int32_t u = uidx[rep[i]];
  if (i < npot) is_pot[u] = 1;
  else is_tgt[u] = 1;
}

// ---- interaction lists (OcTree.m:265-366), one thread per node ----

struct TreeView {
  const double *corners;      // 6 per node [id-1]
  const int32_t *level, *parent, *children;
  const uint8_t *pot_leaf, *tgt_leaf, *complete;   // [id]
};

__device__ inline bool isect(const double *c1, const double *c2) {
  return !(c2[0] > c1[3] || c2[3] < c1[0] || c2[1] > c1[4] || c2[4] < c1[1] || c2[2] > c1[5] || c2[5] < c1[2]);
}
__device__ inline bool is_leaf(const TreeView &T, int n) { return T.children[8 * (n - 1)] == 0; }
__device__ inline const double *box(const TreeView &T, int n) { return T.corners + 6 * (int64_t)(n - 1); }

// Preorder walk of the nodes strictly below `top`, children in order, with
// no stack: siblings have consecutive ids, so the walk moves to the next
// sibling or climbs through the parent.  enter(n) returns whether to visit
// n's children.
template <class F>
__device__ void walk_below(const TreeView &T, int top, F &&enter) {
  if (is_leaf(T, top)) return;
  int n = T.children[8 * (top - 1)];
  for (;;) {
    if (enter(n) && !is_leaf(T, n)) {
      n = T.children[8 * (n - 1)];
      continue;
    }
    for (;;) {
      const int p = T.parent[n - 1];
      if (n - T.children[8 * (p - 1)] < 7) {
        n++;
        break;
      }
      n = p;
      if (n == top) return;
    }
  }
}

// every node (root included) whose closed box meets q, descending below
// nodes of level < maxlev (the host reference walk of the lists)
template <class F>
__device__ void descend(const TreeView &T, const double *q, int maxlev, F &&visit) {
  if (!isect(q, box(T, 1))) return;
  visit(1);
  if (T.level[0] >= maxlev) return;
  walk_below(T, 1, [&](int n) {
    if (!isect(q, box(T, n))) return false;
    visit(n);
    return T.level[n - 1] < maxlev;
  });
}

// U, V, W of node i (FILL = false: sizes only).  U: potential leaves
// adjacent to i (:323-327); V: children, not adjacent to i, of the non-leaf
// same-level neighbours of the parent (:330-339); W: below the non-leaf
// same-level neighbours of i, the nodes not adjacent to i whose parent is
// adjacent and whose subtree is complete, else their children (:343-360);
// `last` is the merged non-leaf of smallest level, then largest id.
template <bool FILL>
__global__ void k_lists(TreeView T, int nc, int32_t *__restrict__ cnt, const int64_t *__restrict__ ptrU,
                        const int64_t *__restrict__ ptrV, const int64_t *__restrict__ ptrW,
                        int32_t *__restrict__ idxU, int32_t *__restrict__ idxV, int32_t *__restrict__ idxW,
                        int32_t *__restrict__ lastW) {
  const int i = 2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > nc) return;
  const double *bi = box(T, i);
  const int li = T.level[i - 1];
  const bool tl = T.tgt_leaf[i];
  if (tl) {
    int64_t o = FILL ? ptrU[i] : 0;
    int c = 0;
    descend(T, bi, 1 << 30, [&](int n) {
      if (is_leaf(T, n) && T.pot_leaf[n]) {
        if (FILL) idxU[o + c] = n;
        c++;
      }
    });
    if (!FILL) cnt[0 * (nc + 2) + i + 1] = c;
  }
  {
    const int P = T.parent[i - 1];
    const int lp = T.level[P - 1];
    int64_t o = FILL ? ptrV[i] : 0;
    int c = 0;
    descend(T, box(T, P), lp, [&](int n) {
      if (n == P || T.level[n - 1] != lp || is_leaf(T, n)) return;
      for (int k = 0; k < 8; k++) {
        const int ch = T.children[8 * (n - 1) + k];
        if (!isect(bi, box(T, ch))) {
          if (FILL) idxV[o + c] = ch;
          c++;
        }
      }
    });
    if (!FILL) cnt[1 * (nc + 2) + i + 1] = c;
  }
  if (tl) {
    int64_t o = FILL ? ptrW[i] : 0;
    int c = 0, last = 0;
    descend(T, bi, li, [&](int st) {
      if (T.level[st - 1] != li || is_leaf(T, st)) return;
      walk_below(T, st, [&](int y) {
        const bool leaf = is_leaf(T, y);
        if (isect(bi, box(T, y))) return !leaf;
        if (T.complete[y]) {
          if (FILL) idxW[o + c] = y;
          c++;
          if (!leaf && (!last || T.level[y - 1] < T.level[last - 1] ||
                        (T.level[y - 1] == T.level[last - 1] && y > last)))
            last = y;
          return false;
        }
        return !leaf;
      });
    });
    if (!FILL) cnt[2 * (nc + 2) + i + 1] = c;
    else lastW[i] = last;
  }
}

// W(i) sorted, then its last merged node moved to the end
__global__ void k_w_last(int nc, const int64_t *__restrict__ ptrW, int32_t *__restrict__ idxW,
                         const int32_t *__restrict__ lastW) {
  const int i = 2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > nc || !lastW[i]) return;
  const int64_t b = ptrW[i], e = ptrW[i + 1];
  int64_t p = b;
  while (p < e && idxW[p] != lastW[i]) p++;
  if (p == e) return;
  for (; p + 1 < e; p++) idxW[p] = idxW[p + 1];
  idxW[e - 1] = lastW[i];
}

// X(y) = { j : y in W(j) }, ordered by the position k of y in W(j), then j
// (find on the padded W matrix, :279-288): keys (y, k), values j, in j order
__global__ void k_x_pairs(int nc, const int64_t *__restrict__ ptrW, const int32_t *__restrict__ idxW,
                          uint64_t *__restrict__ key, int32_t *__restrict__ val, int32_t *__restrict__ cntX) {
  const int j = 2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j > nc) return;
  for (int64_t e = ptrW[j]; e < ptrW[j + 1]; e++) {
    const int32_t y = idxW[e];
    key[e] = ((uint64_t)(uint32_t)y << 32) | (uint64_t)(e - ptrW[j]);
    val[e] = j;
    atomicAdd(&cntX[y + 1], 1);
  }
}

struct Level {
  int64_t n = 0, nsplit = 0;
  DBuf beg, end, pfx, corners, split, scan, sub, rank, id, parent;
};

inline unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void build_lists_gpu(Tree &t) {
  const int nc = t.count;
  for (int q = 0; q < 4; q++) {
    t.lptr[q].assign(nc + 2, 0);
    t.lidx[q].clear();
  }
  if (nc < 2) return;
  // complete(x): a potential leaf, or a non-leaf whose 8 children are
  // complete -- exactly the nodes the W coarsening ends up merging
  std::vector<uint8_t> complete(nc + 1, 0);
  {
    std::vector<std::vector<int>> bylev(t.max_level + 1);
    for (int n = 1; n <= nc; n++) bylev[t.level[n - 1]].push_back(n);
    for (int l = t.max_level; l >= 0; l--)
      for (int n : bylev[l]) {
        if (t.is_leaf(n)) complete[n] = t.pot_leaf[n];
        else {
          bool all = true;
          for (int k = 0; k < 8; k++) all = all && complete[t.children[8 * (n - 1) + k]];
          complete[n] = all;
        }
      }
  }
  DBuf dc, dl, dp, dch, dpl, dtl, dcomp, cnt, ptr, tmp;
  dc.upload(t.corners);
  dl.upload(t.level);
  dp.upload(t.parent);
  dch.upload(t.children);
  dpl.upload(t.pot_leaf);
  dtl.upload(t.tgt_leaf);
  dcomp.upload(complete);
  TreeView T{dc.as<double>(), dl.as<int32_t>(), dp.as<int32_t>(), dch.as<int32_t>(),
              dpl.as<uint8_t>(), dtl.as<uint8_t>(), dcomp.as<uint8_t>()};
  const int64_t stride = nc + 2;
  cnt.alloc(sizeof(int32_t) * 3 * stride);
  NNK_HIP(hipMemset(cnt.p, 0, sizeof(int32_t) * 3 * stride));
  const unsigned nb = (unsigned)((nc - 1 + 127) / 128);
  hipLaunchKernelGGL(k_lists<false>, dim3(nb), dim3(128), 0, 0, T, nc, cnt.as<int32_t>(), nullptr, nullptr,
                     nullptr, nullptr, nullptr, nullptr, nullptr);
  NNK_HIP(hipGetLastError());
  // ptr[q][n] = sum of sizes of nodes < n (CSR over ids, count+2 entries)
  ptr.alloc(sizeof(int64_t) * 4 * stride);
  size_t tb = 0;
  NNK_HIP(rocprim::inclusive_scan(nullptr, tb, cnt.as<int32_t>(), ptr.as<int64_t>(), (size_t)stride, rocprim::plus<int64_t>()));
  tmp.alloc(tb);
  int64_t tot[4];
  for (int q = 0; q < 3; q++) {
    NNK_HIP(rocprim::inclusive_scan(tmp.p, tb, cnt.as<int32_t>() + q * stride,
                                             ptr.as<int64_t>() + q * stride, (size_t)stride, rocprim::plus<int64_t>()));
    NNK_HIP(hipMemcpy(&tot[q], ptr.as<int64_t>() + q * stride + nc + 1, 8, hipMemcpyDeviceToHost));
  }
  DBuf idx[4], lastW, sorted;
  for (int q = 0; q < 3; q++) idx[q].alloc(sizeof(int32_t) * std::max<int64_t>(tot[q], 1));
  lastW.alloc(sizeof(int32_t) * (nc + 1));
  NNK_HIP(hipMemset(lastW.p, 0, sizeof(int32_t) * (nc + 1)));
  const int64_t *pU = ptr.as<int64_t>(), *pV = pU + stride, *pW = pV + stride;
  hipLaunchKernelGGL(k_lists<true>, dim3(nb), dim3(128), 0, 0, T, nc, nullptr, pU, pV, pW, idx[0].as<int32_t>(),
                     idx[1].as<int32_t>(), idx[2].as<int32_t>(), lastW.as<int32_t>());
  NNK_HIP(hipGetLastError());
  // sort every list (segments = nodes)
  for (int q = 0; q < 3; q++) {
    if (!tot[q]) continue;
    const int64_t *pq = pU + q * stride;
    sorted.alloc(sizeof(int32_t) * tot[q]);
    size_t sb = 0;
    NNK_HIP(rocprim::segmented_radix_sort_keys(nullptr, sb, idx[q].as<int32_t>(), sorted.as<int32_t>(),
                                               (unsigned)tot[q], (unsigned)(nc + 1), pq, pq + 1));
    DBuf st;
    st.alloc(sb);
    NNK_HIP(rocprim::segmented_radix_sort_keys(st.p, sb, idx[q].as<int32_t>(), sorted.as<int32_t>(),
                                               (unsigned)tot[q], (unsigned)(nc + 1), pq, pq + 1));
    std::swap(idx[q], sorted);
  }
  if (tot[2]) {
    hipLaunchKernelGGL(k_w_last, dim3(nb), dim3(128), 0, 0, nc, pW, idx[2].as<int32_t>(), lastW.as<int32_t>());
    NNK_HIP(hipGetLastError());
  }
  // X from W
  DBuf cntX, ptrX;
  cntX.alloc(sizeof(int32_t) * stride);
  NNK_HIP(hipMemset(cntX.p, 0, sizeof(int32_t) * stride));
  ptrX.alloc(sizeof(int64_t) * stride);
  tot[3] = tot[2];
  idx[3].alloc(sizeof(int32_t) * std::max<int64_t>(tot[3], 1));
  if (tot[2]) {
    DBuf k0, k1, v0;
    k0.alloc(sizeof(uint64_t) * tot[2]);
    k1.alloc(sizeof(uint64_t) * tot[2]);
    v0.alloc(sizeof(int32_t) * tot[2]);
    hipLaunchKernelGGL(k_x_pairs, dim3(nb), dim3(128), 0, 0, nc, pW, idx[2].as<int32_t>(), k0.as<uint64_t>(),
                       v0.as<int32_t>(), cntX.as<int32_t>());
    NNK_HIP(hipGetLastError());
    size_t sb = 0;
    NNK_HIP(rocprim::radix_sort_pairs(nullptr, sb, k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<int32_t>(),
                                               idx[3].as<int32_t>(), (int)tot[2], 0, 64));
    DBuf st;
    st.alloc(sb);
    NNK_HIP(rocprim::radix_sort_pairs(st.p, sb, k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<int32_t>(),
                                               idx[3].as<int32_t>(), (int)tot[2], 0, 64));
  }
  NNK_HIP(rocprim::inclusive_scan(tmp.p, tb, cntX.as<int32_t>(), ptrX.as<int64_t>(), (size_t)stride, rocprim::plus<int64_t>()));
  for (int q = 0; q < 4; q++) {
    const int64_t *pq = q < 3 ? pU + q * stride : ptrX.as<int64_t>();
    NNK_HIP(hipMemcpy(t.lptr[q].data(), pq, sizeof(int64_t) * stride, hipMemcpyDeviceToHost));
    t.lidx[q].resize(tot[q]);
    if (tot[q]) NNK_HIP(hipMemcpy(t.lidx[q].data(), idx[q].p, sizeof(int32_t) * tot[q], hipMemcpyDeviceToHost));
  }
}

int dedup_gpu(const double *pot, int64_t npot, const double *tgt, int64_t ntgt, Tree &t) {
  const int64_t n = npot + ntgt;
  if (n > INT32_MAX) throw Error(NNK_EARG, "OcTree: too many points");
  DBuf R, ka, kb, va, vb, tmp, head, gscan, leader, rep, keep, uidx, pts, ip, it;
  R.alloc(sizeof(double) * 3 * n);
  if (npot) NNK_HIP(hipMemcpy(R.p, pot, sizeof(double) * 3 * npot, hipMemcpyHostToDevice));
  if (ntgt)
    NNK_HIP(hipMemcpy(R.as<double>() + 3 * npot, tgt, sizeof(double) * 3 * ntgt, hipMemcpyHostToDevice));
  ka.alloc(sizeof(uint64_t) * n);
  kb.alloc(sizeof(uint64_t) * n);
  va.alloc(sizeof(int32_t) * n);
  vb.alloc(sizeof(int32_t) * n);
  size_t tb = 0;
  NNK_HIP(rocprim::radix_sort_pairs(nullptr, tb, ka.as<uint64_t>(), kb.as<uint64_t>(), va.as<int32_t>(),
                                             vb.as<int32_t>(), (int)n, 0, 64));
  tmp.alloc(tb);
  // lexicographic (x, y, z) order, equal rows in index order: stable sorts
  // by z, then y, then x
  const int32_t *ord = nullptr;
  for (int coord = 2; coord >= 0; coord--) {
    hipLaunchKernelGGL(k_row_key, dim3(blocks(n)), dim3(256), 0, 0, n, R.as<double>(), ord, coord,
                       ka.as<uint64_t>(), va.as<int32_t>());
    NNK_HIP(hipGetLastError());
    NNK_HIP(rocprim::radix_sort_pairs(tmp.p, tb, ka.as<uint64_t>(), kb.as<uint64_t>(), va.as<int32_t>(),
                                               vb.as<int32_t>(), (int)n, 0, 64));
    std::swap(va, vb);   // va: current order
    ord = va.as<int32_t>();
  }
  head.alloc(sizeof(int32_t) * n);
  gscan.alloc(sizeof(int32_t) * n);
  hipLaunchKernelGGL(k_row_heads, dim3(blocks(n)), dim3(256), 0, 0, n, R.as<double>(), ord, head.as<int32_t>());
  NNK_HIP(hipGetLastError());
  size_t sb = 0;
  NNK_HIP(rocprim::inclusive_scan(nullptr, sb, head.as<int32_t>(), gscan.as<int32_t>(), (size_t)n, rocprim::plus<int32_t>()));
  DBuf stmp;
  stmp.alloc(sb);
  NNK_HIP(rocprim::inclusive_scan(stmp.p, sb, head.as<int32_t>(), gscan.as<int32_t>(), (size_t)n, rocprim::plus<int32_t>()));
  leader.alloc(sizeof(int32_t) * n);
  rep.alloc(sizeof(int32_t) * n);
  for (int pass = 0; pass < 2; pass++) {
    hipLaunchKernelGGL(k_row_rep, dim3(blocks(n)), dim3(256), 0, 0, n, ord, head.as<int32_t>(), gscan.as<int32_t>(),
                       leader.as<int32_t>(), pass, rep.as<int32_t>());
    NNK_HIP(hipGetLastError());
  }
  keep.alloc(sizeof(int32_t) * n);
  uidx.alloc(sizeof(int32_t) * n);
  hipLaunchKernelGGL(k_row_keep, dim3(blocks(n)), dim3(256), 0, 0, n, rep.as<int32_t>(), keep.as<int32_t>());
  NNK_HIP(hipGetLastError());
  NNK_HIP(rocprim::exclusive_scan(stmp.p, sb, keep.as<int32_t>(), uidx.as<int32_t>(), 0, (size_t)n, rocprim::plus<int32_t>()));
  int32_t last[2];
  NNK_HIP(hipMemcpy(&last[0], keep.as<int32_t>() + n - 1, 4, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(&last[1], uidx.as<int32_t>() + n - 1, 4, hipMemcpyDeviceToHost));
  const int64_t np = (int64_t)last[0] + last[1];
  pts.alloc(sizeof(double) * 3 * np);
  ip.alloc(np);
  it.alloc(np);
  NNK_HIP(hipMemset(ip.p, 0, np));
  NNK_HIP(hipMemset(it.p, 0, np));
  hipLaunchKernelGGL(k_row_emit, dim3(blocks(n)), dim3(256), 0, 0, n, npot, R.as<double>(), rep.as<int32_t>(),
                     keep.as<int32_t>(), uidx.as<int32_t>(), pts.as<double>(), ip.as<uint8_t>(), it.as<uint8_t>());
  NNK_HIP(hipGetLastError());
  t.npts = np;
  t.points.resize(3 * np);
  t.is_pot.resize(np);
  t.is_tgt.resize(np);
  NNK_HIP(hipMemcpy(t.points.data(), pts.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(t.is_pot.data(), ip.p, np, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(t.is_tgt.data(), it.p, np, hipMemcpyDeviceToHost));
  return NNK_OK;
}

int subdivide_gpu(Tree &t, double m, double M, std::string &err) {
  const int64_t np = t.npts;
  // levels a path key can describe; Divide errors before a node is
  // created at level max_depth - 1 (OcTree.m:166-168)
  const int nlev = std::max(0, std::min(kKeyLevels, t.max_depth - 1));
  DBuf dpts, key0, key1, row0, row1, tmp;
  dpts.alloc(sizeof(double) * 3 * np);
  NNK_HIP(hipMemcpy(dpts.p, t.points.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice));
  key0.alloc(sizeof(uint64_t) * np);
  key1.alloc(sizeof(uint64_t) * np);
  row0.alloc(sizeof(int32_t) * np);
  row1.alloc(sizeof(int32_t) * np);
  hipLaunchKernelGGL(k_tree_keys, dim3(blocks(np)), dim3(256), 0, 0, dpts.as<double>(), np, m, M, nlev,
                     key0.as<uint64_t>(), row0.as<int32_t>());
  NNK_HIP(hipGetLastError());
  {
    size_t tb = 0;
    NNK_HIP(rocprim::radix_sort_pairs(nullptr, tb, key0.as<uint64_t>(), key1.as<uint64_t>(),
                                               row0.as<int32_t>(), row1.as<int32_t>(), (int)np, 0, 63));
    tmp.alloc(tb);
    NNK_HIP(rocprim::radix_sort_pairs(tmp.p, tb, key0.as<uint64_t>(), key1.as<uint64_t>(),
                                               row0.as<int32_t>(), row1.as<int32_t>(), (int)np, 0, 63));
  }
  const uint64_t *key = key1.as<uint64_t>();
  // ---- levels, breadth first ----
  std::vector<Level> L(1);
  {
    Level &r = L[0];
    r.n = 1;
    const int64_t b = 0;
    const uint64_t z = 0;
    const double rc[6] = {m, m, m, M, M, M};
    r.beg.upload(std::vector<int64_t>{b});
    r.end.upload(std::vector<int64_t>{np});
    r.pfx.upload(std::vector<uint64_t>{z});
    r.corners.upload(std::vector<double>(rc, rc + 6));
  }
  if (t.max_depth <= 1) {
    err = "Max level reached: Increase Max Depth";
    return NNK_EMAXDEPTH;
  }
  for (int l = 0;; l++) {
    Level &P = L[l];
    P.split.alloc(sizeof(int32_t) * P.n);
    P.scan.alloc(sizeof(int32_t) * P.n);
    hipLaunchKernelGGL(k_split_flags, dim3(blocks(P.n)), dim3(256), 0, 0, P.n, P.beg.as<int64_t>(),
                       P.end.as<int64_t>(), (int64_t)t.capacity, P.split.as<int32_t>());
    NNK_HIP(hipGetLastError());
    size_t tb = 0;
    NNK_HIP(rocprim::exclusive_scan(nullptr, tb, P.split.as<int32_t>(), P.scan.as<int32_t>(), 0, (size_t)P.n, rocprim::plus<int32_t>()));
    DBuf stmp;
    stmp.alloc(tb);
    NNK_HIP(rocprim::exclusive_scan(stmp.p, tb, P.split.as<int32_t>(), P.scan.as<int32_t>(), 0, (size_t)P.n, rocprim::plus<int32_t>()));
    int32_t last[2];
    NNK_HIP(hipMemcpy(&last[0], P.split.as<int32_t>() + P.n - 1, 4, hipMemcpyDeviceToHost));
    NNK_HIP(hipMemcpy(&last[1], P.scan.as<int32_t>() + P.n - 1, 4, hipMemcpyDeviceToHost));
    P.nsplit = (int64_t)last[0] + last[1];
    if (P.nsplit == 0) break;
    // the children are at level l+1, and Divide is called on each of them
    if (l + 2 >= t.max_depth) {
      err = "Max level reached: Increase Max Depth";
      return NNK_EMAXDEPTH;
    }
    if (l + 1 > nlev) {
      err = "OcTree: more than 21 levels are not supported";
      return NNK_EMAXDEPTH;
    }
    L.emplace_back();
    Level &C = L[l + 1];
    Level &Pp = L[l];   // L may have moved
    C.n = 8 * Pp.nsplit;
    if (C.n + 1 > INT32_MAX) { err = "OcTree: too many nodes"; return NNK_EARG; }
    C.beg.alloc(sizeof(int64_t) * C.n);
    C.end.alloc(sizeof(int64_t) * C.n);
    C.pfx.alloc(sizeof(uint64_t) * C.n);
    C.corners.alloc(sizeof(double) * 6 * C.n);
    hipLaunchKernelGGL(k_make_children, dim3(blocks(Pp.n)), dim3(256), 0, 0, Pp.n, l, Pp.split.as<int32_t>(),
                       Pp.scan.as<int32_t>(), Pp.beg.as<int64_t>(), Pp.end.as<int64_t>(), Pp.pfx.as<uint64_t>(),
                       Pp.corners.as<double>(), key, C.beg.as<int64_t>(), C.end.as<int64_t>(),
                       C.pfx.as<uint64_t>(), C.corners.as<double>());
    NNK_HIP(hipGetLastError());
  }
  const int nl = (int)L.size();
  // ---- subtree split counts (bottom up), ranks and ids (top down) ----
  for (int l = nl - 1; l >= 0; l--) {
    Level &P = L[l];
    P.sub.alloc(sizeof(int64_t) * P.n);
    const int64_t *next = l + 1 < nl ? L[l + 1].sub.as<int64_t>() : nullptr;
    if (!next) {  // deepest level: nothing splits
      NNK_HIP(hipMemset(P.sub.p, 0, sizeof(int64_t) * P.n));
    } else {
      hipLaunchKernelGGL(k_subtree_splits, dim3(blocks(P.n)), dim3(256), 0, 0, P.n, P.split.as<int32_t>(),
                         P.scan.as<int32_t>(), next, P.sub.as<int64_t>());
      NNK_HIP(hipGetLastError());
    }
  }
  int64_t total = 1;
  for (int l = 0; l < nl; l++) total += 8 * L[l].nsplit;
  {
    Level &r = L[0];
    r.rank.upload(std::vector<int64_t>{0});
    r.id.upload(std::vector<int32_t>{1});
    r.parent.upload(std::vector<int32_t>{0});
  }
  for (int l = 0; l + 1 < nl; l++) {
    Level &P = L[l], &C = L[l + 1];
    C.rank.alloc(sizeof(int64_t) * C.n);
    C.id.alloc(sizeof(int32_t) * C.n);
    C.parent.alloc(sizeof(int32_t) * C.n);
    hipLaunchKernelGGL(k_child_ids, dim3(blocks(P.n)), dim3(256), 0, 0, P.n, P.split.as<int32_t>(),
                       P.scan.as<int32_t>(), P.rank.as<int64_t>(), P.id.as<int32_t>(), C.sub.as<int64_t>(),
                       C.rank.as<int64_t>(), C.id.as<int32_t>(), C.parent.as<int32_t>());
    NNK_HIP(hipGetLastError());
  }
  // ---- node arrays in id order, leaf id per point ----
  DBuf oc, olev, opar, ochi, opi;
  oc.alloc(sizeof(double) * 6 * total);
  olev.alloc(sizeof(int32_t) * total);
  opar.alloc(sizeof(int32_t) * total);
  ochi.alloc(sizeof(int32_t) * 8 * total);
  opi.alloc(sizeof(int32_t) * np);
  for (int l = 0; l < nl; l++) {
    Level &P = L[l];
    const int32_t *cid = l + 1 < nl ? L[l + 1].id.as<int32_t>() : nullptr;
    hipLaunchKernelGGL(k_emit_nodes, dim3(blocks(P.n)), dim3(256), 0, 0, P.n, l, P.split.as<int32_t>(),
                       P.scan.as<int32_t>(), P.id.as<int32_t>(), P.parent.as<int32_t>(), P.corners.as<double>(),
                       P.beg.as<int64_t>(), P.end.as<int64_t>(), cid, row1.as<int32_t>(), oc.as<double>(),
                       olev.as<int32_t>(), opar.as<int32_t>(), ochi.as<int32_t>(), opi.as<int32_t>());
    NNK_HIP(hipGetLastError());
  }
  t.count = (int32_t)total;
  t.corners.resize(6 * total);
  t.level.resize(total);
  t.parent.resize(total);
  t.children.resize(8 * total);
  t.point_index.resize(np);
  NNK_HIP(hipMemcpy(t.corners.data(), oc.p, sizeof(double) * 6 * total, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(t.level.data(), olev.p, sizeof(int32_t) * total, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(t.parent.data(), opar.p, sizeof(int32_t) * total, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(t.children.data(), ochi.p, sizeof(int32_t) * 8 * total, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(t.point_index.data(), opi.p, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
  return NNK_OK;
}

}  // namespace nnk
