// Device-resident KIFMM operator for one tree (the reference's KIFMM object,
// Classes/KIFMM.m, specialised to one corona resolution).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <vector>

#include "nnk_common.hpp"
#include "nnk_graph.hpp"
#include "nnk_tree.hpp"

namespace nnk {

struct RedArgs;

// rows per workgroup tile and k depth per chunk of the FP64 gather GEMM;
// columns per tile of its M2L launches (GBN1) and of the others (GBN0:
// expansions, M2M, L2L, check-to-equivalent solves)
constexpr int GBM = 160, GBK = 16, GBN1 = 64, GBN0 = 32;

struct GemmJob {
  const double *const *A;  // nb operator pointers, each At[Kpad][Mpad] (k-major)
  const int32_t *gidx;     // nb x N gather slots (-1 = none), or null = column index
  const int32_t *cidx;     // N output slots, or null = column index
  int nb, N;
  // optional, per column tile ct: the operator blocks with a source in the
  // tile, in increasing order, tlist[tptr[ct] .. tptr[ct+1])
  const int32_t *tptr = nullptr, *tlist = nullptr;
};

// the column-tile active lists of a gather table (nb x N, -1 = none)
void tile_lists(const std::vector<int32_t> &gidx, int nb, int N, int BN, std::vector<int32_t> &ptr,
                std::vector<int32_t> &list);

// Refuses (NNK_ESTATE) a gather job whose tables a kernel could index out
// of bounds: the gather table must be nb x N with entries in [-1, count),
// the tile pointers must cover ceil(N / BN) tiles monotonically, and every
// tile-list entry must be a block in [0, nb).
void check_gather_job(const std::vector<int32_t> &gidx, int nb, int N, int BN, const std::vector<int32_t> &tptr,
                      const std::vector<int32_t> &tlist, int64_t count, const char *what);

// The 316 V-list offsets as slots (ox+3)*49 + (oy+3)*7 + (oz+3) of the 7^3
// offset table, in the order the M2L visits them (every layout: class,
// merged and per-offset jobs): grouped by the parity of (ox, oy, oz), and
// within a group along a boustrophedon path (z fastest) on which
// consecutive offsets differ by 2 in one coordinate.  A column tile's
// targets are the children of one class of consecutive parents (a compact
// box, longest in z), so the sources it gathers for consecutive offsets
// are mostly the same nodes, which the L2 still holds.
const std::vector<int> &m2l_slot_order();

// The V-list offsets a child of class a (bits x, y, z) can see, as slots,
// in the order of that class's gather-table blocks (189 of them):
// m2l_slot_order restricted to the class.
std::vector<int> class_offset_slots(int a);

// Host plan of one level's split-precision M2L: per child class a job over
// its 189 offsets (gather table gv[a], 189 x |cls[a]|), or one merged job
// over all 316 offsets and every node of the level, whichever walks fewer
// (column tile, operator block) pairs.  The merged table (316 x N) is built
// only for levels of at most 8 BN nodes: with n nodes per class the class
// jobs walk ~8 ceil(n/BN) 189 pairs and the merged job ~ceil(8n/BN) 316,
// so larger levels never gain from it.  Every table chosen is checked with
// check_gather_job (count = node count).
struct M2lLevelPlan {
  bool merged = false;
  int64_t class_pairs = 0, merged_pairs = -1;   // -1: merged table not built
  std::vector<int32_t> ctp[8], ctl[8];          // class jobs' tile lists
  std::vector<int> msl;                         // merged job: the 316 slots in block order
  std::vector<int32_t> mcols, mg, mtp, mtl;     // merged job: columns, gather table, tile lists
};
void plan_m2l_level(const std::vector<int32_t> *cls, const std::vector<int32_t> *gv, int64_t count, int BN,
                    M2lLevelPlan &out);

// The per-offset M2L plan of one level (sparse adaptive levels): the
// V-list (target, source) pairs grouped by offset slot, as jobs of at most
// BN pairs of one slot, each product into its own column of a pair buffer,
// then a fixed-order sum per target.  Chosen (use) where its padded work
// (BN per job tile, plus rt / M per pair for the pair buffer's round trip)
// is below base_work (or, with base_work < 0, below the class tiles'
// work: BN-column tiles times their active blocks) and the buffer
// (npairs x M doubles) fits cap_elems.  Jobs are laid out for the 8 XCDs:
// the slot-ordered tiles are split into 8 contiguous runs and tile k of
// run x is job 8k + x (empty jobs: n = 0, slot = -1).  node/ptr/idx: per
// target its pairs, ordered by slot.  Tables are checked as for the class
// jobs (count = node count).
struct M2lPairPlan {
  bool use = false;
  int64_t class_work = 0, pair_work = 0, npairs = 0;
  std::vector<int32_t> job_slot, job_p0, job_n;
  std::vector<int32_t> src;             // per pair: source slot
  std::vector<int32_t> node, ptr, idx;
};
void plan_m2l_pairs_host(const std::vector<int32_t> *cls, const std::vector<int32_t> *gv, int64_t count, int BN,
                         int64_t M, int64_t cap_elems, int64_t rt, int64_t base_work, M2lPairPlan &out);
// pair buffer limit of the per-offset M2L (doubles): 2 GiB
constexpr int64_t kPairBufMaxElems = (int64_t)1 << 28;
// the pair buffer's round trip (16 M bytes per pair) relative to a pair
// column's product, times M: ~4 TB/s for the round trip against ~45 TF/s
// for 2 M^2 FP64 flops, or ~700 TF/s for the 6 M^2 flops of two-part
// 16-bit products
constexpr int64_t kPairRtFP64 = 90, kPairRt16 = 470;

struct GemmBatch {
  DBuf jobs;                 // GemmJob[n]
  std::vector<DBuf> owned;   // index/pointer arrays referenced by the jobs
  int njobs = 0, maxN = 0;
};

struct LevelPlan {
  int level = 0;
  // upward: nodes whose equivalent density is computed at this level
  int n_up = 0;
  DBuf up_nodes;              // slots (id-1)
  GemmBatch m2m;              // children -> parent check surface
  int m2m_split = 1;          // split-K factor of the M2M launch (one or two children per part)
  DBuf m2m_slot;              // column -> parent slot, for its split reduction
  GemmBatch up_solve1, up_solve2;
  // downward
  int n_down = 0;
  DBuf down_nodes;
  GemmBatch m2l;              // 8 classes x 189 translations (reads up densities)
  int m2l_split = 1;          // split-K factor of the M2L launch (both forms)
  DBuf m2l_slot;              // job * maxN + column -> output slot (-1 none), for the split reduction
  GemmBatch l2l;              // 8 classes, parent down density (accumulates)
  // Either FP64 form, where the class tiles would mostly multiply empty
  // columns (sparse adaptive levels): the M2L as jobs of at most 64 V-list
  // pairs of one offset, each product into its own column of the pair
  // buffer, then a fixed-order sum per target (by offset slot).  Built
  // where the planner's work estimate favours it; m2l_pair_on selects it.
  int m2l_npairs = 0;         // pairs of the per-offset plan (0: none built)
  bool m2l_pair_on = false;
  GemmBatch m2lp;             // one operator, <= 64 pairs per job
  int m2lp_nodes = 0;         // targets of the per-target sums (the level's down nodes)
  DBuf m2lp_node, m2lp_ptr, m2lp_idx;  // target slot; its pairs [ptr[j], ptr[j+1]) of idx
  DBuf mx_m2lp;               // mixed forms: the same jobs as GemmJobBf16[mx_m2lp_jobs]
  int mx_m2lp_jobs = 0;
  int mx_pair_depth = 0;      // register-ring depth (2, 3) or 4 (DMA-staged) of their launch
  GemmBatch dn_solve1, dn_solve2;
  int n_x = 0;
  DBuf x_nodes;               // nodes with a non-empty X list
  // mixed form
  GemmBatch mx_m2m, mx_l2l, mx_xP;
  int mx_m2m_split = 1;       // as m2m_split, for the mixed form's M2M
  DBuf mx_m2m_slot;
  DBuf mx_m2l;                // GemmJobBf16[8]
  int mx_split = 1, mx_depth = 3;  // split-K factor and register-ring depth of the M2L launch
                                   // (depth 4: the DMA-staged kernel; 5: the same with
                                   // 64 x 48 wave tiles)
  int mx_maxnb = 0;                // most operator blocks of one M2L job
  bool mx_tn3 = false;             // the class jobs also exist with 192-column tiles (mx_m2l3)
  DBuf mx_m2l3;                    // GemmJobBf16[mx_m2l_jobs], tile lists over 192 columns
  DBuf mx_slot;
  std::vector<DBuf> mx_owned;
  int mx_m2l_jobs = 0, mx_m2l_maxN = 0;
  // FP16 form: the same split-FP16 GEMM for the down expansion of the
  // target leaves, over this level's down coefficients (dn_rows)
  DBuf hx_dn, dn_rows;
  int hx_dn_n = 0, hx_dn_maxN = 0, n_dn_rows = 0;
};

// host copies of a level's node lists (slots = id-1), shared by both forms
struct HostLevel {
  std::vector<int32_t> up_leaves, up_nonleaves, m2m_g;   // m2m_g: 8 x nonleaves
  std::vector<int32_t> dn, tleaves, wnodes, xnodes;       // xnodes: ids
  std::vector<int32_t> cls[8], gv[8], gp[8];              // gv: 189 x |cls|
};

// compressed (truncated-SVD) operators of one level for the mixed form;
// read-only once built (shared between Fmm objects through the cache)
struct MixedOps {
  DBuf Pup, Vr, Pdn, Ur;      // FP64 At layout: R x K, K x R, R x K, K x R
  DBuf VrT, UrT;              // R rows of length K (precompute inputs)
  DBuf Mc, Lc;                // 8 each, FP64 At (R x R); absent at level 2
  DBuf Cf;                    // 316 M2L operators, NP 16-bit parts in MFMA fragment order
  DBuf Uf;                    // FP16 form: per operator row, the unscale factor
  // FP16 form: the down expansion U_R (K x R) as FP16 parts in fragment
  // order, with its row unscale factors.  (The up expansion S^-1 U^T, M2M
  // and L2L stay FP64: their S^-1 amplifies the rounding of 22-bit inputs
  // beyond the 1e-6 target.)
  DBuf UrF, UrU;
  // FP64-M2L truncated form: the 316 R x R operators in FP64 At layout
  // (KpadR x MpadR each, rows R..KpadR-1 zero)
  DBuf Cd;
  std::vector<const double *> dop;    // per offset slot (343)
  std::vector<const uint16_t *> cop;  // per offset slot (343)
  std::vector<const float *> uop;
  size_t bytes() const;
};

// operators of one level (all boxes of a level share them); read-only once
// built (shared between Fmm objects through the cache)
struct LevelOps {
  DBuf Wup, Vup, Wdn, Udn;    // check-to-equivalent factors, At layout
  DBuf M2M, L2L;              // 8 operators each (this level's child class)
  DBuf T;                     // 316 translation operators (absent in the mixed forms)
  std::vector<double *> tslot;  // host copy: slot -> device pointer or null
  std::vector<double> svdU, svdS, svdV;  // host SVD of K(up -> down), row-major
  size_t bytes() const;
};

// Process-wide cache of per-level operators.  The operators of a level
// depend only on (device, eps, mu, corona resolution, shells, box width,
// level) and, for the mixed forms, on (precision, SVD cut, common rank), so
// trees whose root cube is unchanged (time stepping, repeated solves) reuse
// them instead of repeating the SVDs and kernel-matrix formation.  Live
// entries are always shared; up to `limit` bytes of entries whose last
// owner is gone are kept (least recently used first out).
struct OpKey {
  int dev, kind, res, shells, level, rank;  // kind: 0 FP64 (with T), 1 no T, 2+p mixed precision p
  double eps, mu, w, tau;
  bool operator<(const OpKey &o) const;
};
class OpCache {
 public:
  static OpCache &get();
  std::shared_ptr<const void> find(const OpKey &k);
  void put(const OpKey &k, std::shared_ptr<const void> v, size_t bytes);
  void set_limit(int64_t bytes);
  // hits, misses, bytes retained (beyond live owners), entries retained
  void stats(int64_t *out4);
  void clear();

 private:
  void trim();
  struct Ent {
    std::weak_ptr<const void> live;
    std::shared_ptr<const void> kept;
    size_t bytes = 0;
    uint64_t tick = 0;
  };
  std::map<OpKey, Ent> map_;
  int64_t limit_ = 0, hits_ = 0, misses_ = 0;
  uint64_t tick_ = 0;
};

class Fmm {
 public:
  // precision 0: FP64 form (reference parity); 1: mixed form (truncated SVD
  // at relative tolerance tau, two-part split-BF16 M2L, FP32-math direct
  // sums); 2: as 1 with three-part split-BF16 M2L; 3: as 1 with two-part
  // split-FP16 M2L on power-of-two scaled operands; 4: truncated SVD with
  // every operation in FP64 (R x R M2L on the FP64 matrix cores, FP64
  // direct sums)
  Fmm(std::shared_ptr<const Tree> t, double eps, double mu, int res, int shells, int precision = 0,
      double tau = 1e-10);
  ~Fmm();
  Fmm(const Fmm &) = delete;
  Fmm &operator=(const Fmm &) = delete;
  // f: npot x 3 (potential order), u: ntgt x 3 (target order); device ptrs.
  void apply(const double *d_f, double *d_u, hipStream_t s);
  // the same matvec replayed from a HIP graph captured on first use (one
  // launch per call instead of one per kernel); bit-identical to apply
  void apply_graph(const double *d_f, double *d_u, hipStream_t s);
  // autotune the M2L launch shapes and the leaf kernel's workgroup size on
  // the first apply (off by default)
  void set_autotune(bool on) { tune_ = on; }
  // per level: M2L split-K factor and (mixed forms) register-ring depth
  void launch_config(int level, int *split, int *depth) const;
  // fixes the M2L launch of one level (split-K factor, register-ring depth
  // or 4 for the DMA-staged FP16 kernel); throws if the level has no M2L
  // or the form cannot run it.  A captured graph is dropped.
  void set_launch_config(int level, int split, int depth);
  // threads per target leaf of the leaf kernel (L2T + P2P + M2T)
  int leaf_threads() const { return leaf_threads_; }
  // leaf-local product (reducedPass.m) on device pointers
  void reduced_apply(const double *d_f, double *d_u, hipStream_t s);
  // leaf-block solve (ReducedSolve.m): f = blockdiag(K_leaf)^-1 v, with the
  // per-leaf LU factors formed on first use; needs potentials == targets
  void reduced_solve(const double *d_v, double *d_f, hipStream_t s);
  int64_t npot() const { return npot_; }
  int64_t ntgt() const { return ntgt_; }
  int nq() const { return nq_; }
  const Tree &tree() const { return *tp_; }
  // copy an internal per-node array (0 check, 1 up, 2 down density) to host
  void debug_copy(int which, double *host) const;
  int K() const { return K_; }
  int rank() const { return R_; }
  // Phase timing with HIP events on the apply stream (no host sync inside
  // apply).  Phases: 0 strengths, 1 S2M, 2 upward GEMMs, 3 M2L, 4 L2L,
  // 5 P2L, 6 downward solves, 7 leaf L2T+P2P+M2T.
  static constexpr int kPhases = 8;
  void set_profiling(bool on);
  // totals in ms and launch counts since the last call; resets
  void phase_times(double *ms, int64_t *count);
  // work counts of one matvec: 0 M2L pairs, 1 P2P pairs, 2 S2M pairs,
  // 3 L2T pairs, 4 M2T pairs, 5 P2L pairs, 6 up-solve nodes,
  // 7 down-solve nodes, 8 M2M parents, 9 L2L nodes, 10 sources, 11 targets
  static constexpr int kStats = 12;
  void stats(int64_t *out) const;

 private:
  void plan();
  // K(up -> down) of the canonical box of a level, and the level's
  // operators from that matrix's SVD (A = U diag(S) V^T, row-major)
  std::vector<double> up_to_down(int level) const;
  std::shared_ptr<const LevelOps> build_ops(int level, bool with_T, std::vector<double> U, std::vector<double> S,
                                            std::vector<double> V);
  // the 316 translation operators of a level into T (tslot: 343 pointers)
  void form_T(int level, DBuf &T, std::vector<double *> &tslot);
  OpKey key(int kind, int level) const;
  void gemm(const GemmBatch &b, const double *X, int ldx, double *C, int ldc, int beta, hipStream_t s,
            int tag = 0, int nsplit = 1, const DBuf *slot = nullptr);
  // split-K factor for a launch of `tiles` workgroups over blocks of nb
  // operators, and the reduction of its partial tiles into C
  static int split_for(int64_t tiles, int nb);
  // split-K factor of an M2M launch (8 child blocks) of `tiles` workgroups
  // whose partial tiles take `part_elems` doubles per split
  static int split_m2m(int64_t tiles, size_t part_elems);
  void split_reduce(const DBuf &slot, int ncols, int M, int ldc, int nsplit, double *C, hipStream_t s);
  void reserve_part(size_t elems);
  void gemm_dims(const GemmBatch &b, const double *X, int ldx, double *C, int ldc, int beta,
                 hipStream_t s, int M, int K, int Mpad, int Kpad, int nsplit = 1, const DBuf *slot = nullptr,
                 int tag = 0);
  // forms 1-3: M2L on 16-bit parts and FP32-math direct sums; form 4 keeps
  // both in FP64
  bool m16() const { return prec_ >= 1 && prec_ <= 3; }
  bool fp32_math() const { return prec_ >= 1 && prec_ <= 3; }
  // leading dimension of the M2L output (FP64: check potentials; truncated
  // forms: down coefficients)
  int m2l_ld() const { return prec_ >= 1 ? RP_ : K_; }
  // form 4: the M2L of level l, FP64 gather GEMM over the R x R operators
  void launch_m2l_t(int l, int split, hipStream_t s);
  // Either FP64 form: the per-offset pair plan of level P (cls[a]: output
  // slots of child class a, gv[a]: its nb x N gather table, ops[a]: its nb
  // operators in class block order, M: product rows), built and selected
  // where its padded work plus the pair buffer's round trip is below the
  // class tiles'; the launch of the plan
  void plan_m2l_pairs(LevelPlan &P, const std::vector<int32_t> *cls, const std::vector<int32_t> *gv,
                      const std::vector<const double *> *ops, int M);
  void launch_m2l_pairs(const LevelPlan &P, const double *X, int ldx, double *C, int ldc, int M, int Mpad,
                        int Kpad, hipStream_t s);
  // the mixed forms' per-offset pair plan of a level (base_work: the padded
  // work of the class or merged jobs it replaces) and its launch by the
  // 16-bit M2L kernel of register-ring depth 2/3 or the DMA-staged one (4)
  void plan_m2l_pairs16(LevelPlan &P, const HostLevel &H, const MixedOps &X, int64_t base_work);
  void launch_m2l_pairs16(const LevelPlan &P, int depth, hipStream_t s);
  // the per-target sums of a per-offset plan (Y: M rows per pair)
  void pair_sum(const LevelPlan &P, int M, double *C, int ldc, hipStream_t s);
  void build_mixed();
  void plan_mixed();
  void apply_mixed(const double *d_f, double *d_u, hipStream_t s);
  void apply_impl(const double *d_f, double *d_u, hipStream_t s);
  void launch_m2l_mixed(int l, int depth, int split, hipStream_t s);
  bool dma_ok(const LevelPlan &P, int split, int depth = 4) const;
  // time the M2L launch candidates of every level on the current densities
  // (split-K factor s/2, s, 2s; the kernel variants of the form) and keep
  // the fastest, re-timing the near-ties
  void autotune(hipStream_t s);
  void autotune_leaf(hipStream_t s);
  int leaf_threads_ = 256;          // threads per target leaf in the leaf kernel
  void launch_surf(const int32_t *nodes, int n, const int64_t *lptr, const int32_t *lidx, int down,
                   int accumulate, double *chk, hipStream_t s);
  void launch_leaf(int far, double *d_u, hipStream_t s);
  void strengths(const double *d_f, hipStream_t s);
  void mark(int phase, bool begin, hipStream_t s);

  std::shared_ptr<const Tree> tp_;
  const Tree &tree_;
  double eps_, mu_;
  int res_, shells_, nq_, K_, Kpad_, Mpad_;
  int64_t npot_, ntgt_, nsrc_ = 0;
  int maxlev_ = 0;
  double w0_ = 0;
  // geometry and point data
  DBuf corners_, surf_idx_;
  DBuf src_ptr_, src_pos_, src_q_, src_from_, ctr_ptr_, ctr_col_, ctr_w_;
  DBuf tgt_ptr_, tgt_pos_, tgt_out_;
  DBuf uptr_, uidx_, wptr_, widx_, xptr_, xidx_;
  DBuf pleaves_, tleaves_;
  std::vector<int32_t> src_ptr_h_;  // host copy of src_ptr_ (launch order by work)
  int n_pleaves_ = 0, n_tleaves_ = 0;
  int max_leaf_tgt_ = 0, ncu_ = 256;  // most targets in a leaf; compute units of the device
  // work buffers
  DBuf chk_, uppot_, downpot_, tmp_;
  std::vector<std::shared_ptr<const LevelOps>> ops_;
  std::vector<LevelPlan> plans_;
  std::vector<HostLevel> hl_;
  int prec_ = 0;
  double tau_ = 1e-10;
  int R_ = 0, RP_ = 0, RPm_ = 0, MpadR_ = 0, KpadR_ = 0;
  std::vector<std::shared_ptr<const MixedOps>> mops_;
  DBuf chat_, dhat_, chat_x_, upall_;
  int np_ = 2, KCm_ = 0, RT_ = 0;   // mixed M2L: 16-bit parts, 32-deep k chunks, 16-row tiles
  bool f16_ = false;                // FP16 parts (with scaling) instead of BF16
  DBuf amax_, xun_;                 // FP16 form: max |coefficient|, coefficient unscale factor
  // FP16 form: split down coefficients ([node][KC][2][32], zero row at nc)
  // and their unscale factor per level
  DBuf dhat_x_, dun_lv_;
  int RTk_ = 0;                     // K rows in 16-row tiles (whole 128-row blocks)
  void launch_h16(const DBuf &jobs, int njobs, int maxN, int M, int KC, const uint16_t *X, double *C, int ldc,
                  const double *xun, int beta, hipStream_t s);
  void split_h(const DBuf &rows, int nrows, const double *X, int ldx, int len, int KC, uint16_t *out,
               double *xun, hipStream_t s);
  DBuf part_;                       // split-K partial tiles
  size_t part_elems_ = 0;
  DBuf pairbuf_;                    // per-offset M2L: one product column per V-list pair
  size_t pairbuf_elems_ = 0;
  int n_upall_ = 0;
  // mixed forms: the per-node expansions of every level, one job per level
  // in one launch each (their nodes do not depend on other levels' results
  // at the point where they run): up expansion of the leaves with sources,
  // equivalent densities of the W-list nodes, down expansion of the target
  // leaves (FP64 down expansion only)
  GemmBatch mx_upP_all_, mx_wV_all_, mx_downU_all_;
  // the X-list owners of every level, for one P2L launch (both forms); the
  // FP64 form puts the P2L check potentials in xchk_ and adds them level by
  // level after M2L and L2L
  DBuf x_all_, xchk_;
  int n_x_all_ = 0;
  void reduced_factor();
  bool red_ready_ = false;
  DBuf red_blk_, red_off_, red_piv_, red_perm_, red_poff_, red_ord2loc_;
  struct RedArgs red_args() const;
  int red_maxn_ = 0;
  int64_t stats_[kStats] = {0};
  bool prof_ = false;
  bool tune_ = false, tuned_ = false;
  GraphOp gop_;
  struct Mark { int phase; hipEvent_t a, b; };
  std::vector<Mark> marks_;
  hipEvent_t open_ = nullptr;
};

}  // namespace nnk
