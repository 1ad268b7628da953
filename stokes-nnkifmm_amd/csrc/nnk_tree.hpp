// Adaptive octree and U/V/W/X interaction lists with the semantics of the
// reference's Classes/OcTree.m (ids, corners, strict '>' split, closed-box
// adjacency, list order), built in O(N log N) instead of O(nodes^2).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace nnk {

struct Tree {
  // points = unique([pot; tgt], 'rows', 'stable')   (OcTree.m:80)
  int64_t npts = 0, npot = 0, ntgt = 0;
  std::vector<double> points;          // npts x 3
  std::vector<uint8_t> is_pot, is_tgt; // ismember masks (OcTree.m:81-82)
  std::vector<int64_t> pot_ord, tgt_ord; // ordinal among potentials/targets, -1 if not
  std::vector<int32_t> point_index;    // 1-based leaf id per point
  int32_t capacity = 0, max_depth = 21, max_level = 0;
  int32_t count = 0;                   // nodeCount
  std::vector<double> corners;         // 6 per node, [id-1]
  std::vector<int32_t> level, parent;  // [id-1]
  std::vector<int32_t> children;       // 8 per node, 0 for a leaf
  std::vector<uint8_t> pot_leaf, tgt_leaf; // [id] (index 0 unused)
  // Points per node (CSR over ids; only leaves hold points), point order.
  std::vector<int64_t> pt_ptr, pt_rows;
  // Lists U,V,W,X: CSR over ids 1..count, ptr has count+2 entries,
  // list of node n is idx[ptr[n] .. ptr[n+1]).
  std::vector<int64_t> lptr[4];
  std::vector<int32_t> lidx[4];
  // NEAREST mode (Tools/Nearest): fine points and NN (Q x npot, CSC).
  bool nearest = false;
  int64_t nfine = 0;
  std::vector<double> fine;
  std::vector<int64_t> nn_colptr, nn_row;
  std::vector<double> nn_val;
  // build time in ms: 0 de-duplication, 1 subdivision, 2 point lists,
  // 3 interaction lists
  double build_ms[4] = {0, 0, 0, 0};

  bool is_leaf(int n) const { return children[8 * (n - 1)] == 0; }
  const double *box(int n) const { return &corners[6 * (n - 1)]; }
};

// Build; returns 0 or an error code with a message in err.  cube: null for
// the bounding cube of the points (OcTree.m), or a fixed root
// [cube[0] x3, cube[1] x3] containing every point.
int build_tree(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
               int capacity, int max_depth, Tree &t, std::string &err, const double *cube = nullptr);
// GPU subdivision of the root cube [m m m M M M] over t.points with
// t.capacity and t.max_depth: fills count, corners, level, parent, children
// and point_index (nnk_tree_gpu.hip)
int subdivide_gpu(Tree &t, double m, double M, std::string &err);
// GPU unique([pot; tgt], 'rows', 'stable') (-0 equals 0) and the ismember
// masks: fills npts, points, is_pot, is_tgt (nnk_tree_gpu.hip)
int dedup_gpu(const double *pot, int64_t npot, const double *tgt, int64_t ntgt, Tree &t);
// GPU U/V/W/X lists of a subdivided tree: fills lptr, lidx (nnk_tree_gpu.hip)
void build_lists_gpu(Tree &t);
int set_nearest(Tree &t, const double *fine, int64_t nfine, const int64_t *colptr,
                const int64_t *row, const double *val, std::string &err);
// Checks a nearest map (nfine x ncol, CSC with 0-based rows) before any of
// it is read on the GPU: colptr[0] == 0 and monotone, rows in [0, nfine),
// no null array that would be read.  Returns NNK_OK or NNK_EARG with err.
int validate_nn_csc(int64_t ncol, int64_t nfine, const double *fine, const int64_t *colptr, const int64_t *row,
                    const double *val, std::string &err);

enum { NNK_OK = 0, NNK_EARG = 1, NNK_EMAXDEPTH = 2, NNK_EHIP = 3, NNK_ENOMEM = 4, NNK_ESTATE = 5 };

}  // namespace nnk
