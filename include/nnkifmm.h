/*
 * nnkifmm.h -- C ABI of the MI355X KIFMM library (libnnkifmm.so).
 *
 * The reference (Walkersb101/Stokes-NNKIFMM) is MATLAB; a MEX gateway binds
 * these entry points so that the reference's classes and functions keep
 * their signatures.  Each declaration names the reference
 * interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *  - Every function returns 0 on success or a positive status code; the
 *    message of the last failure on the calling thread is nnk_last_error().
 *    No C++ exception crosses this boundary.
 *  - Point arrays are n x 3 row-major doubles ([x1 y1 z1 x2 ...]), which is
 *    MATLAB's (3,n) column-major layout, i.e. the reference's "format 1"
 *    GMRES vector (Classes/KIFMM.m:84-85, 98-99).
 *  - Node ids are 1-based as in Classes/OcTree.m.
 *  - Host-pointer entry points copy in and out; *_device entry points take
 *    device pointers and a hipStream_t (passed as void*, 0 = null stream).
 *  - Handles are not re-entrant: one call at a time per handle.
 */
#ifndef NNKIFMM_H
#define NNKIFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nnk_tree nnk_tree;
typedef struct nnk_fmm nnk_fmm;
typedef struct nnk_mobility nnk_mobility;
typedef struct nnk_batch nnk_batch;

enum {
  NNK_OK = 0,
  NNK_EARG = 1,      /* invalid argument                                   */
  NNK_EMAXDEPTH = 2, /* OcTree.m:168-170 'Max level reached'               */
  NNK_EHIP = 3,      /* HIP runtime failure                                */
  NNK_ENOMEM = 4,
  NNK_ESTATE = 5     /* inconsistent state / unsupported configuration     */
};

enum { NNK_LIST_U = 0, NNK_LIST_V = 1, NNK_LIST_W = 2, NNK_LIST_X = 3 };
/* NNK_FP64: plain FP64 form, reproduces the reference's velocities.
 * NNK_MIXED: truncated-SVD pseudo-inverses (FP64), M2L on the BF16 matrix
 * cores with operands split in two BF16 parts (3 products, ~16 bits) and
 * FP32 accumulation promoted to FP64 per translation, FP32-math direct sums
 * with FP64 accumulation.
 * NNK_MIXED3: as NNK_MIXED with three BF16 parts (6 products, ~24 bits).
 * NNK_MIXED_H: as NNK_MIXED with two FP16 parts (3 products, ~22 bits) on
 * operands scaled by powers of two (per operator row; per matvec for the
 * coefficients).
 * NNK_FP64_T: the truncated-SVD pseudo-inverses and R x R translations of
 * the mixed forms with every operation in FP64 (M2L on the FP64 matrix
 * cores, FP64 direct sums): the FP64 comparator at the mixed forms'
 * accuracy. */
enum { NNK_FP64 = 0, NNK_MIXED = 1, NNK_MIXED3 = 2, NNK_MIXED_H = 3, NNK_FP64_T = 4 };

/* Library version (major*10000 + minor*100 + patch). */
int32_t nnk_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *nnk_last_error(void);

/* Replaces: OcTree(potPoints, targetPoints, 'nodeCapacity', c, 'maxDepth', d)
 *           Classes/OcTree.m:77-113 (unique/ismember, root cube, Divide,
 *           geninteractionlists). */
int32_t nnk_tree_create(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
                        int32_t node_capacity, int32_t max_depth, nnk_tree **out);

/* No counterpart in the reference: nnk_tree_create with the root cube fixed
 *           to [cube[0] cube[0] cube[0] cube[1] cube[1] cube[1]] instead of
 *           the points' bounding cube (OcTree.m:97-98); every point must lie
 *           in it.  Subdivision is unchanged.  For time stepping: with a fixed
 *           root the level widths, and so the per-level operators, stay the
 *           same from one right-hand side to the next (nnk_op_cache_limit). */
int32_t nnk_tree_create_in_cube(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
                                int32_t node_capacity, int32_t max_depth, const double *cube,
                                nnk_tree **out);

/* Replaces: OcTree(..., 'nearest', 1, 'finePoints', F, 'NN', NN)
 *           Classes/OcTree.m:115-128.  NN is Q x npot in CSC form (MATLAB
 *           sparse jc/ir/pr with 0-based row indices). */
int32_t nnk_tree_set_nearest(nnk_tree *t, const double *fine, int64_t nfine,
                             const int64_t *nn_colptr, const int64_t *nn_rowidx,
                             const double *nn_val);

/* No counterpart in the reference (its OcTree is timed with tic/toc): the
 *           tree's build time in ms by phase: 0 de-duplication of the points,
 *           1 subdivision (GPU), 2 per-leaf point lists, 3 U/V/W/X lists. */
int32_t nnk_tree_build_times(const nnk_tree *t, double *ms4);

/* Replaces: properties points/nodeCount/nodeCapacity (Classes/OcTree.m:59-74). */
int32_t nnk_tree_sizes(const nnk_tree *t, int64_t *npts, int64_t *npot, int64_t *ntgt,
                       int32_t *node_count, int32_t *node_capacity);

/* Replaces: properties points, potentials, targets, pointIndex,
 *           nodeCorners (count x 6 row-major), nodeLevel, nodeParents,
 *           nodeChildren (count x 8 row-major)   Classes/OcTree.m:59-74. */
int32_t nnk_tree_export(const nnk_tree *t, double *points, uint8_t *is_pot, uint8_t *is_tgt,
                        int32_t *point_index, double *corners, int32_t *level,
                        int32_t *parent, int32_t *children);

/* Replaces: property interactions{node, kind}  Classes/OcTree.m:270-288.
 * ptr has node_count+1 entries; the list of node n (1-based) is
 * idx[ptr[n-1] .. ptr[n]). */
int32_t nnk_tree_list_size(const nnk_tree *t, int32_t kind, int64_t *total);
int32_t nnk_tree_list_export(const nnk_tree *t, int32_t kind, int64_t *ptr, int32_t *idx);

void nnk_tree_destroy(nnk_tree *t);

/* Replaces: KIFMM(tree, [eps mu], 'coronaRes', r, 'coronaShells', s)
 *           Classes/KIFMM.m:46-61.  Builds the device-resident operator:
 *           per-level check-to-equivalent factors and translation
 *           operators, point data in leaf order.  precision is NNK_FP64,
 *           NNK_MIXED, NNK_MIXED3 or NNK_MIXED_H; svd_tol (mixed forms only, <= 0 for the
 *           default 1e-8)
 *           is the relative singular-value cut of the pseudo-inverses.  The
 *           tree must stay alive while the handle exists. */
int32_t nnk_fmm_create(const nnk_tree *t, double eps, double mu, int32_t corona_res,
                       int32_t corona_shells, int32_t precision, double svd_tol, nnk_fmm **out);

/* Replaces: KIFMM.computeVel  Classes/KIFMM.m:63-105 (upwardPass +
 *           downwardPass).  f: npot x 3 host array in points(potentials)
 *           order; u: ntgt x 3 host array in points(targets) order. */
int32_t nnk_fmm_apply(nnk_fmm *h, const double *f, double *u);
int32_t nnk_fmm_apply_device(nnk_fmm *h, const double *d_f, double *d_u, void *stream);

/* Replaces: KIFMM.reducedComputeVel  Classes/KIFMM.m:157-198
 *           (Tools/Passes/reducedPass.m). */
int32_t nnk_fmm_reduced_apply(nnk_fmm *h, const double *f, double *u);
int32_t nnk_fmm_reduced_apply_device(nnk_fmm *h, const double *d_f, double *d_u, void *stream);

/* Replaces: KIFMM.reducedSolve  Classes/KIFMM.m:200-242 (Mobility/ReducedSolve.m):
 *           per target leaf, K(leaf sources -> leaf targets) \ v, with the NN
 *           map folded in NEAREST mode.  Needs potentials == targets.  v, f:
 *           N x 3.  The leaf blocks are LU-factored and inverted once,
 *           on first use; each solve is one block-diagonal product (same
 *           solution to rounding). */
int32_t nnk_fmm_reduced_solve(nnk_fmm *h, const double *v, double *f);
int32_t nnk_fmm_reduced_solve_device(nnk_fmm *h, const double *d_v, double *d_f, void *stream);

void nnk_fmm_destroy(nnk_fmm *h);

/* Replaces: NN = nearestMatrix(coarse, fine, blockSize, threads)
 *           Tools/Nearest/nearestMatrix.m:17-41.  NN is q x n in CSC form:
 *           colptr (n+1), rowidx (0-based), val; rows ascending in a column.
 *           If cap < nnz the call fails with NNK_EARG and sets *nnz. */
int32_t nnk_nearest_matrix(const double *coarse, int64_t n, const double *fine, int64_t q,
                           int64_t *colptr, int64_t *rowidx, double *val, int64_t cap,
                           int64_t *nnz);

/* Replaces: the operators of mobilityProblem (Mobility/mobilityProblem.m:35-57):
 *           B, BT from genMatrixComponets (Mobility/genMatrixComponets.m:334-337)
 *           as CSR (B: nb6 x 3N, BT: 3N x nb6, nb6 = 6*nsw), A1 = B*B' and
 *           A2 = BT'*BT dense row-major (nb6 x nb6); fmm is the computeVel
 *           operator and fmm_sparse its coronaRes-4 twin (computeVelSparse,
 *           KIFMM.m:107-155; may be null if no preconditioner is used).  Both
 *           handles must outlive the mobility handle.  Vectors are
 *           [f (3N, all x, all y, all z); U (3 nsw); Omega (3 nsw)].
 *           nb6 = 0 (no bodies; B, BT empty, A1 and A2 unused) gives the
 *           resistance operator S alone: nnk_mobility_solve then is
 *           gmres(@(x) fmm.computeVel(x), vel, [], tol, maxit)
 *           (FlowField.m:17, DirectMethodComparison.m:28,
 *           Rigid_Resistance.m:24), preconditioned by the leaf-block solve
 *           if precon != 0. */
int32_t nnk_mobility_create(nnk_fmm *fmm, nnk_fmm *fmm_sparse, int64_t nb6,
                            const int64_t *B_rowptr, const int32_t *B_col, const double *B_val,
                            const int64_t *BT_rowptr, const int32_t *BT_col, const double *BT_val,
                            const double *A1, const double *A2, nnk_mobility **out);

/* Replaces: mobilityMatrixMulti(f, swimmers, B, BT)  Mobility/MobilityMatrixMulti.m:17-36. */
int32_t nnk_mobility_apply(nnk_mobility *h, const double *x, double *y);
int32_t nnk_mobility_apply_device(nnk_mobility *h, const double *d_x, double *d_y, void *stream);

/* Replaces: mobilityPrecon(f, swimmers, B, BT, A1, A2, tol, maxIter)
 *           Mobility/MobilityPrecon.m:17-39 (tol and maxIter are unused there). */
int32_t nnk_mobility_precon(nnk_mobility *h, const double *x, double *y);
int32_t nnk_mobility_precon_device(nnk_mobility *h, const double *d_x, double *d_y, void *stream);

/* Replaces: x = gmres(afunc, input, [], tol, maxIter) and, with precon != 0,
 *           afunc = A(P(x)) followed by x = P(x)  (Mobility/mobilityProblem.m:47-58).
 *           Unrestarted GMRES from x0 = 0.  rhs, x: host vectors of length
 *           3N + nb6.  resvec (optional) receives up to maxit+1 residual norms;
 *           *iters is the iteration count, *relres the true relative residual
 *           of the (preconditioned) system, *flag 0 if relres <= tol else 1. */
int32_t nnk_mobility_solve(nnk_mobility *h, const double *rhs, double tol, int32_t maxit,
                           int32_t precon, double *x, int32_t *flag, double *relres,
                           int32_t *iters, double *resvec);

/* No counterpart in the reference (MATLAB launches nothing): route every
 *           later nnk_fmm_apply / nnk_fmm_apply_device call (on = 1) through
 *           a HIP graph of the whole matvec, captured on the first call and
 *           replayed with one launch; results are bit-identical to on = 0. */
int32_t nnk_fmm_set_graph(nnk_fmm *h, int32_t on);

/* No counterpart in the reference: with on = 1 the first later apply
 *           times the M2L launch candidates of every level on its own data
 *           (split-K factor s/2, s, 2s of the default s; for the FP64
 *           forms also the per-offset pair jobs, depth 1, where planned;
 *           for the mixed forms also the register-ring depth 2 or 3 and
 *           the DMA-staged kernels, depths 4-8, and where planned the
 *           per-offset pair jobs, 12-14, 17) and keeps the fastest
 *           (candidates within 5 % of the fastest are timed three more
 *           times, interleaved, and keep their best time; within 0.5 %
 *           the earlier candidate in that order wins).  The
 *           choice changes summation order only (results stay within
 *           rounding of each other and deterministic afterwards).
 *           nnk_fmm_launch_config reports the per-level choice (split,
 *           depth; 0 where a level has no M2L) for levels 0..nlev-1. */
int32_t nnk_fmm_set_autotune(nnk_fmm *h, int32_t on);
int32_t nnk_fmm_launch_config(const nnk_fmm *h, int32_t *split, int32_t *depth, int32_t nlev);
/* No counterpart in the reference: fixes the M2L launch of levels
 *           0..nlev-1 (entries with split 0 are left as they are), e.g. to
 *           replay a configuration that nnk_fmm_launch_config reported.
 *           split: split-K factor 1..64 within the reserved part buffer;
 *           depth: 0 for the FP64 forms ("fp64", "fp64t") with their
 *           class tiles, or 1 for their per-offset pair jobs (split 1;
 *           levels where the planner built them: sparse adaptive levels,
 *           whose class tiles would mostly multiply empty columns); 2 or 3
 *           (register-ring depth) for
 *           the mixed forms, 2 only for "mixed3"; 4 selects the DMA-staged
 *           kernel (any mixed form; for "mixedh" at most 320 operator blocks
 *           per split-K part; 5 and 6 its 64 x 48 wave-tile and four-stage
 *           variants, class jobs of at most 192 blocks; 7 depth 4 with the
 *           LDS fragment reads of each chunk overlapping the products of the
 *           previous one, two-part forms, depth 4's limits; 8 the same with
 *           four LDS stages, depth 6's limits); 12-14 and 17 run the per-offset
 *           pair jobs (split 1, levels where the planner built them) through
 *           the kernel of depth 2-4 or 7.  Mixed-form results
 *           are bit-identical across the depths of one split; the FP64
 *           forms' two layouts agree to rounding.  NNK_EARG for a level without M2L or a value the
 *           form cannot run; a captured graph is recaptured on next use. */
int32_t nnk_fmm_set_launch_config(nnk_fmm *h, const int32_t *split, const int32_t *depth, int32_t nlev);
/* No counterpart in the reference: the workgroup size of the leaf kernel
 *           (L2T + P2P + M2T per target leaf), 128, 256 or 512 threads; with
 *           nnk_fmm_set_autotune the first apply times the three and keeps
 *           the fastest (summation order of the P2P sums changes, not their
 *           accuracy).  256 without autotuning. */
int32_t nnk_fmm_leaf_threads(const nnk_fmm *h, int32_t *threads);

/* Replaces nothing; serves the time-stepping callers of mobilityProblem
 *           (ode45 in Data Generation/Squiremers.m:47-54,
 *           SquiremersGrav.m:55-67), which rebuild the KIFMM every right-hand
 *           side (mobilityProblem.m:26-35).  The per-level operators (SVD
 *           factors of the check-to-equivalent solves, M2M, L2L, M2L, and the
 *           compressed mixed-form operators) depend only on (device, eps, mu,
 *           coronaRes, shells, box width, level[, precision, svd_tol, rank]),
 *           so handles whose root cube is unchanged share them.  Handles alive
 *           at the same time always share; nnk_op_cache_limit keeps up to
 *           `bytes` of operators whose last handle is gone (least recently
 *           used dropped first; 0, the default, keeps none).  Shared
 *           operators are built by the same deterministic code, so results
 *           are bit-identical with or without reuse.  nnk_op_cache_stats
 *           returns hits, misses, bytes kept and entries kept. */
int32_t nnk_op_cache_limit(int64_t bytes);
int32_t nnk_op_cache_stats(int64_t *hits, int64_t *misses, int64_t *bytes, int64_t *entries);

/* No counterpart in the reference: the same for the GMRES operator (A or
 *           A P) of nnk_mobility_solve / nnk_mobility_solve_refined, one
 *           graph per operator form, captured on first use. */
int32_t nnk_mobility_set_graph(nnk_mobility *h, int32_t on);

/* No counterpart in the reference (which solves in FP64 only): the same
 *           solve by iterative refinement.  Outer residuals use the FP64
 *           operator of nnk_mobility_create; each correction is a GMRES solve
 *           on `inner` (a KIFMM handle on the same tree, usually a mixed
 *           form) to relative tolerance inner_tol, preconditioned (precon
 *           != 0) with `inner_sparse` as the coronaRes-4 operator if given
 *           (else the FP64 one).  Stops when the true FP64 relative residual
 *           of the unpreconditioned system is <= tol or maxit inner
 *           iterations are used.  *iters = inner iterations, *outer =
 *           refinement steps. */
int32_t nnk_mobility_solve_refined(nnk_mobility *h, nnk_fmm *inner, nnk_fmm *inner_sparse, const double *rhs,
                                   double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                   int32_t *flag, double *relres, int32_t *iters, int32_t *outer);

/* No counterpart in the reference (its slot is the gmres call of
 *           Mobility/mobilityProblem.m:53,57): the same solve by flexible
 *           GMRES (Saad 1993).  The outer Arnoldi process runs in FP64 on the
 *           FP64 operator of nnk_mobility_create; its k-th basis vector is
 *           preconditioned by a GMRES solve on `inner` to inner_tol (right-
 *           preconditioned as in nnk_mobility_solve_refined when precon !=
 *           0), and x is assembled from those preconditioned vectors.
 *           Arguments and outputs as nnk_mobility_solve_refined; *outer =
 *           outer Arnoldi steps (at most 50, no restart). */
int32_t nnk_mobility_solve_fgmres(nnk_mobility *h, nnk_fmm *inner, nnk_fmm *inner_sparse, const double *rhs,
                                  double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                  int32_t *flag, double *relres, int32_t *iters, int32_t *outer);

void nnk_mobility_destroy(nnk_mobility *h);

/* Replaces: blockcomputation(x, x0, q, blockSize, [eps mu])
 *           Tools/Other/blockcomputation.m:17-28 (dense K(x->x0) q, here
 *           without forming K).  src: ns x 3, q: ns x 3, tgt: nt x 3,
 *           u: nt x 3. */
int32_t nnk_direct_apply(const double *src, int64_t ns, const double *q, const double *tgt,
                         int64_t nt, double eps, double mu, double *u);
int32_t nnk_direct_apply_device(const double *d_src, int64_t ns, const double *d_q,
                                const double *d_tgt, int64_t nt, double eps, double mu,
                                double *d_u, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nnk_direct_apply_device with the same kernels and the
 * same summation order (bit-identical results), but its per-source-part
 * partial sums go to a static device array of 16 MiB (2^21 doubles) that
 * the library holds on every GPU it runs on, instead of a stream-ordered
 * allocation.  It needs 24 * S * nt bytes, S = source parts (S = 1, none,
 * for nt > 130,816 targets; S * nt < 655,360 otherwise), so every size
 * fits; a size that did not would be refused with NNK_EARG.  No hipMalloc,
 * hipFree or allocating library call is made.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race. */
int32_t nnk_direct_apply_device_noalloc(const double *d_src, int64_t ns, const double *d_q,
                                        const double *d_tgt, int64_t nt, double eps, double mu,
                                        double *d_u, void *stream);

/* No counterpart in the reference (which times with tic/toc): per-phase
 * HIP-event timing of nnk_fmm_apply* on the apply stream.  Phases:
 * 0 strengths, 1 S2M, 2 upward GEMMs (M2M + solves), 3 M2L, 4 L2L, 5 P2L,
 * 6 downward solves, 7 leaf L2T+P2P+M2T.  nnk_fmm_phase_times returns the
 * totals (ms) and launch counts recorded since the previous call. */
int32_t nnk_fmm_set_profiling(nnk_fmm *h, int32_t on);
int32_t nnk_fmm_phase_times(nnk_fmm *h, double *ms, int64_t *count, int32_t nphases);

/* No counterpart in the reference: work counts of one matvec, for roofline
 * accounting.  0 M2L pairs, 1 P2P pairs, 2 S2M pairs, 3 L2T pairs,
 * 4 M2T pairs, 5 P2L pairs, 6 up-solve nodes, 7 down-solve nodes, 8 M2M
 * parents, 9 L2L nodes, 10 sources, 11 targets; out[nstats]; also returns
 * the density length 3*Nq in out[12] if nstats > 12 and the pseudo-inverse
 * rank R of the mixed form in out[13] if nstats > 13. */
int32_t nnk_fmm_stats(const nnk_fmm *h, int64_t *out, int32_t nstats);

/* Batched mobility problems and their adjoints.  No counterpart in the
 * reference, which solves one problem per mobilityProblem call
 * (Mobility/mobilityProblem.m:24-72): each problem here is that call's
 * system for its own bodies (and optional fixed boundary), with S the
 * direct regularized-Stokeslet sum from the problem's quadrature points
 * (nearest map) onto its force points -- what the KIFMM reduces to when the
 * tree is the root and 8 leaves.  Problem p owns force points
 * [pt_off[p], pt_off[p+1]), quadrature points [fine_off[p], fine_off[p+1])
 * (world coordinates, n x 3) and bodies [body_off[p], body_off[p+1]).  The
 * nearest map is CSR over quadrature points (nn_ptr: Q+1 entries, nn_col:
 * global force-point index in the same problem).  Per force point: weight
 * (NN column sum), arm (NN-weighted rotated quadrature position) and rot
 * (rotated, untranslated position), as genMatrixComponets forms AF, AM and
 * AOm (Mobility/genMatrixComponets.m:34-76), and pt_body (body index local
 * to the problem, -1 for boundary points; a body's points are contiguous).
 * eps: per problem.  precision NNK_FP64 or NNK_MIXED (FP32 kernel
 * evaluation, FP64 accumulation).  Vectors concatenate the problems, each
 * [f (3 N_p: all x, all y, all z); U (3 nsw_p: all x, all y, all z);
 * Omega (3 nsw_p)] as in mobilityProblem; nnk_batch_size gives the total. */
int32_t nnk_batch_create(int32_t nprob, const int64_t *pt_off, const double *pts,
                         const int64_t *fine_off, const double *fine, const int64_t *nn_ptr,
                         const int32_t *nn_col, const double *nn_val, const double *weight,
                         const double *arm, const double *rot, const int32_t *pt_body,
                         const int64_t *body_off, const double *eps, double mu, int32_t precision,
                         nnk_batch **out);
int32_t nnk_batch_size(const nnk_batch *h, int64_t *n);
/* y = A x for every problem (mobilityMatrixMulti, MobilityMatrixMulti.m:17-36,
 * per problem); the _transpose form applies A^T. */
int32_t nnk_batch_apply(nnk_batch *h, const double *x, double *y);
int32_t nnk_batch_apply_transpose(nnk_batch *h, const double *x, double *y);
int32_t nnk_batch_apply_device(nnk_batch *h, const double *d_x, double *d_y, void *stream);
/* on = 1: replay the batched operators from HIP graphs captured on first use. */
int32_t nnk_batch_set_graph(nnk_batch *h, int32_t on);
/* on = 1: the first operator application times the batched P2P at 1, 2
 * and 4 wavefronts per workgroup (sources of a 64-target tile split over
 * the wavefronts) and keeps the fastest; nnk_batch_p2p_waves reports the
 * choice (4 without autotuning).  Summation order changes, not accuracy. */
int32_t nnk_batch_set_autotune(nnk_batch *h, int32_t on);
int32_t nnk_batch_p2p_waves(const nnk_batch *h, int32_t *waves);
/* Unrestarted GMRES from x = 0 on every problem at once, on A (transpose =
 * 0) or A^T (transpose = 1): lockstep iterations, a problem stops when its
 * residual estimate reaches tol (tol / 2 with FP32-math operators, whose
 * estimate drifts from the true residual).  flag (0 when the recomputed
 * relative residual is <= tol, as MATLAB's gmres), relres (that residual),
 * iters: nprob entries each. */
int32_t nnk_batch_solve(nnk_batch *h, const double *rhs, double tol, int32_t maxit, int32_t transpose,
                        double *x, int32_t *flag, double *relres, int32_t *iters);
/* Adjoint gradient: with A z = b and A^T lambda = dJ/dz, writes
 * dJ/dtheta = -d(lambda^T A z)/dtheta for the force-point positions (gx,
 * 3 per point), the quadrature-point positions (gy, 3 per point) and eps
 * (geps, one per problem), holding b and the nearest map fixed. */
int32_t nnk_batch_gradient(nnk_batch *h, const double *z, const double *lambda, double *gx, double *gy,
                           double *geps);
void nnk_batch_destroy(nnk_batch *h);

#ifdef __cplusplus
}
#endif
#endif /* NNKIFMM_H */
