/*
 * nnk_gateway.h -- the marshalling layer between MATLAB/Octave (MEX) and
 * libnnkifmm's C ABI (include/nnkifmm.h), on plain pointers.
 *
 * nnkifmm_mex.c only unpacks mxArrays and calls these functions, so the
 * whole of what the MATLAB side hands the library -- and in which layout --
 * is exercised without MATLAB by a C program (tests/native/gateway_test.c).
 *
 * MATLAB conventions handled here:
 *  - point sets arrive as (3,n) column-major double arrays (the shims pass
 *    points.'), which is the ABI's n x 3 row-major layout;
 *  - sparse matrices arrive as MATLAB CSC: jc (ncol+1) and ir (nnz) of type
 *    mwIndex (size_t in the 64-bit API), 0-based rows, values pr;
 *    B and BT reach nnkgw_mobility_create as the CSC of B.' and BT.', which
 *    is the CSR of B and BT;
 *  - dense A1, A2 arrive as A1.', A2.' (column-major of the transpose =
 *    row-major);
 *  - exports fill MATLAB column-major double arrays ((npts,3) points,
 *    (count,6) corners, (count,8) children), 1-based ids;
 *  - handles are ids into a registry (never raw pointers), checked for kind
 *    and liveness on every call.  A released tree or KIFMM handle stays
 *    alive until the handles built on it are released too.
 */
#ifndef NNK_GATEWAY_H
#define NNK_GATEWAY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t nnkgw_id;   /* 0 = no handle */
typedef size_t nnkgw_index;  /* MATLAB mwIndex (64-bit API) */

enum { NNKGW_TREE = 1, NNKGW_FMM = 2, NNKGW_MOBILITY = 3 };
enum { NNKGW_APPLY = 0, NNKGW_REDUCED = 1, NNKGW_RSOLVE = 2 };

/* message of the last failure (gateway or library) */
const char *nnkgw_last_error(void);

/* OcTree(pot, tgt, 'nodeCapacity', cap, 'maxDepth', depth[, 'rootCube', cube]):
 * pot3, tgt3 are (3,n) column-major; cube = [lo hi] or NULL. */
int32_t nnkgw_tree_create(const double *pot3, int64_t npot, const double *tgt3, int64_t ntgt, int32_t cap,
                          int32_t depth, const double *cube, nnkgw_id *out);
/* 'nearest', 1, 'finePoints', fine, 'NN', NN: fine3 (3,q); NN q x ncol CSC */
int32_t nnkgw_tree_set_nearest(nnkgw_id tree, const double *fine3, int64_t q, int64_t ncol,
                               const nnkgw_index *jc, const nnkgw_index *ir, const double *pr);
int32_t nnkgw_tree_sizes(nnkgw_id tree, int64_t *npts, int64_t *npot, int64_t *ntgt, int32_t *count,
                         int32_t *cap);
/* properties in MATLAB layout (all double): points (npts,3), potentials,
 * targets, pointIndex (npts), nodeCorners (count,6), nodeLevel, nodeParents
 * (count), nodeChildren (count,8).  Any output may be NULL. */
int32_t nnkgw_tree_export(nnkgw_id tree, double *points, double *potentials, double *targets,
                          double *point_index, double *corners, double *level, double *parent,
                          double *children);
/* interactions(:, kind+1): total entries, then ptr (count+1, 0-based
 * offsets) and idx (1-based node ids) as doubles; ptr/idx may be NULL */
int32_t nnkgw_tree_list(nnkgw_id tree, int32_t kind, int64_t *total, double *ptr, double *idx);

/* nearestMatrix(coarse, fine): first call with jc == NULL returns nnz;
 * second call fills jc (n+1), ir (nnz), pr (nnz) */
int32_t nnkgw_nearest_matrix(const double *coarse3, int64_t n, const double *fine3, int64_t q, int64_t *nnz,
                             nnkgw_index *jc, nnkgw_index *ir, double *pr);

/* KIFMM(tree, [eps mu], 'coronaRes', res, 'coronaShells', shells,
 * 'precision', p, 'svdTol', tol) */
int32_t nnkgw_fmm_create(nnkgw_id tree, double eps, double mu, int32_t res, int32_t shells, int32_t precision,
                         double svd_tol, nnkgw_id *out);
/* computeVel (NNKGW_APPLY), reducedComputeVel (NNKGW_REDUCED),
 * reducedSolve (NNKGW_RSOLVE) on (3,n) column-major vectors */
int32_t nnkgw_fmm_apply(nnkgw_id fmm, int32_t op, const double *in3, double *out3);

/* the mobility operator: fmm4 may be 0 (no preconditioner); Bt_* is the
 * CSC of B.' (n3 x nb6), BTt_* the CSC of BT.' (nb6 x n3), A1t = A1.',
 * A2t = A2.' (nb6 x nb6) */
int32_t nnkgw_mobility_create(nnkgw_id fmm, nnkgw_id fmm4, int64_t nb6, int64_t n3, const nnkgw_index *Bt_jc,
                              const nnkgw_index *Bt_ir, const double *Bt_pr, const nnkgw_index *BTt_jc,
                              const nnkgw_index *BTt_ir, const double *BTt_pr, const double *A1t,
                              const double *A2t, nnkgw_id *out);
/* mobilityMatrixMulti (precon = 0) or mobilityPrecon (precon = 1) */
int32_t nnkgw_mobility_apply(nnkgw_id mob, int32_t precon, const double *x, double *y);
/* gmres(afunc, rhs, [], tol, maxit) of mobilityProblem; resvec may be NULL */
int32_t nnkgw_mobility_solve(nnkgw_id mob, const double *rhs, double tol, int32_t maxit, int32_t precon,
                             double *x, int32_t *flag, double *relres, int32_t *iters, double *resvec);
/* the same by iterative refinement with inner (mixed-form) operators */
int32_t nnkgw_mobility_solve_refined(nnkgw_id mob, nnkgw_id inner, nnkgw_id inner4, const double *rhs,
                                     double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                     int32_t *flag, double *relres, int32_t *iters, int32_t *outer);
int32_t nnkgw_mobility_solve_fgmres(nnkgw_id mob, nnkgw_id inner, nnkgw_id inner4, const double *rhs,
                                    double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                    int32_t *flag, double *relres, int32_t *iters, int32_t *outer);
int32_t nnkgw_mobility_size(nnkgw_id mob, int64_t *n);

/* handle-class delete(): release one handle (its object is destroyed once
 * no handle built on it is alive) */
int32_t nnkgw_free(nnkgw_id id);
/* mexAtExit: destroy every live object (mobility, then KIFMM, then trees) */
void nnkgw_free_all(void);
/* live registry entries of a kind (0 = all), for tests */
int64_t nnkgw_live(int32_t kind);

/* layout helpers (also used by the exports): (rows,cols) row-major ->
 * column-major, and mwIndex -> int64 / int32 with range checks */
void nnkgw_rowmajor_to_colmajor(const double *in, int64_t rows, int64_t cols, double *out);
int32_t nnkgw_index_to_i64(const nnkgw_index *in, int64_t n, int64_t *out);
int32_t nnkgw_index_to_i32(const nnkgw_index *in, int64_t n, int32_t *out);

#ifdef __cplusplus
}
#endif
#endif /* NNK_GATEWAY_H */
