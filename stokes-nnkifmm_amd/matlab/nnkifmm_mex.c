/*
 * nnkifmm_mex.c -- MEX gateway of libnnkifmm for MATLAB / GNU Octave.
 *
 * One mexFunction dispatching on a command string.  It only unpacks
 * mxArrays and calls the marshalling layer nnk_gateway.c (which the C test
 * tests/native/gateway_test.c drives the same way); handles travel as uint64
 * registry ids wrapped by nnkHandle.m, whose delete() calls 'free'.
 *
 * Build (MATLAB):  mex nnkifmm_mex.c nnk_gateway.c -I<repo>/include \
 *                      -L<repo>/stokes-nnkifmm_amd/lib -lnnkifmm
 * Build (Octave):  mkoctfile --mex nnkifmm_mex.c nnk_gateway.c -I<repo>/include \
 *                      -L<repo>/stokes-nnkifmm_amd/lib -lnnkifmm
 *
 * Commands (points and vectors in (3,n) column-major, i.e. points.'):
 *   id = nnkifmm_mex('tree', pot3, tgt3, cap, depth[, cube])
 *        nnkifmm_mex('set_nearest', tree, fine3, NN)           NN sparse q x npot
 *   [points, potentials, targets, pointIndex, nodeCorners, nodeLevel,
 *    nodeParents, nodeChildren] = nnkifmm_mex('tree_export', tree)
 *   [ptr, idx] = nnkifmm_mex('tree_list', tree, kind)           kind 0..3 = U V W X
 *   NN = nnkifmm_mex('nearest_matrix', coarse3, fine3)
 *   id = nnkifmm_mex('fmm', tree, [eps mu], coronaRes, coronaShells[, precision, svdTol])
 *   u3 = nnkifmm_mex('apply' | 'reduced' | 'rsolve', fmm, f3)
 *   id = nnkifmm_mex('mobility', fmm, fmm4 or [], B.', BT.', A1.', A2.')
 *   y  = nnkifmm_mex('mobility_apply' | 'precon', mob, x)
 *   [x, flag, relres, iter, resvec] = nnkifmm_mex('solve', mob, rhs, tol, maxit, precon)
 *   [x, flag, relres, iter, outer] = nnkifmm_mex('solve_refined' | 'solve_fgmres', mob, inner,
 *                                                inner4 or [], rhs, tol, maxit, innerTol, precon)
 *        nnkifmm_mex('opcache', bytes)
 *        nnkifmm_mex('free', id)
 */
#include <stdint.h>
#include <string.h>

#include "mex.h"
#include "nnk_gateway.h"
#include "nnkifmm.h"

static int g_atexit = 0;

static void ok(int32_t rc) {
  if (rc) mexErrMsgIdAndTxt("nnkifmm:fail", "%s", nnkgw_last_error());
}

static void need(int nrhs, int n, const char *cmd) {
  if (nrhs < n) mexErrMsgIdAndTxt("nnkifmm:args", "nnkifmm_mex('%s'): too few arguments", cmd);
}

static nnkgw_id id_of(const mxArray *a) {
  if (mxIsEmpty(a)) return 0;
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt("nnkifmm:args", "bad handle");
  return *(const uint64_t *)mxGetData(a);
}

static mxArray *mk_id(nnkgw_id id) {
  mxArray *a = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *(uint64_t *)mxGetData(a) = id;
  return a;
}

/* a real double matrix with 3 rows (the points.' the shims pass) */
static const double *pts3(const mxArray *a, int64_t *n) {
  if (!mxIsDouble(a) || mxIsComplex(a) || mxIsSparse(a) || (!mxIsEmpty(a) && mxGetM(a) != 3))
    mexErrMsgIdAndTxt("nnkifmm:args", "points must be a real 3 x n double matrix");
  *n = (int64_t)mxGetN(a);
  return mxGetPr(a);
}

static void sparse_in(const mxArray *a, const char *what) {
  if (!mxIsSparse(a) || mxIsComplex(a) || !mxIsDouble(a))
    mexErrMsgIdAndTxt("nnkifmm:args", "%s must be a real sparse double matrix", what);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  char cmd[32];
  if (!g_atexit) {
    mexAtExit(nnkgw_free_all);
    g_atexit = 1;
  }
  if (nrhs < 1 || mxGetString(prhs[0], cmd, sizeof cmd)) mexErrMsgIdAndTxt("nnkifmm:args", "missing command");
  int64_t n1 = 0, n2 = 0;

  if (!strcmp(cmd, "tree")) {
    need(nrhs, 5, cmd);
    const double *p = pts3(prhs[1], &n1), *t = pts3(prhs[2], &n2);
    const double *cube = nrhs > 5 && !mxIsEmpty(prhs[5]) ? mxGetPr(prhs[5]) : NULL;
    nnkgw_id id;
    ok(nnkgw_tree_create(p, n1, t, n2, (int32_t)mxGetScalar(prhs[3]), (int32_t)mxGetScalar(prhs[4]), cube,
                         &id));
    plhs[0] = mk_id(id);

  } else if (!strcmp(cmd, "set_nearest")) {
    need(nrhs, 4, cmd);
    const double *f = pts3(prhs[2], &n1);
    sparse_in(prhs[3], "NN");
    if ((int64_t)mxGetM(prhs[3]) != n1) mexErrMsgIdAndTxt("nnkifmm:args", "NN must have one row per fine point");
    ok(nnkgw_tree_set_nearest(id_of(prhs[1]), f, n1, (int64_t)mxGetN(prhs[3]), mxGetJc(prhs[3]),
                              mxGetIr(prhs[3]), mxGetPr(prhs[3])));

  } else if (!strcmp(cmd, "tree_export")) {
    need(nrhs, 2, cmd);
    nnkgw_id id = id_of(prhs[1]);
    int64_t np;
    int32_t nc;
    ok(nnkgw_tree_sizes(id, &np, NULL, NULL, &nc, NULL));
    mxArray *o[8] = {mxCreateDoubleMatrix(np, 3, mxREAL), mxCreateDoubleMatrix(np, 1, mxREAL),
                     mxCreateDoubleMatrix(np, 1, mxREAL), mxCreateDoubleMatrix(np, 1, mxREAL),
                     mxCreateDoubleMatrix(nc, 6, mxREAL), mxCreateDoubleMatrix(nc, 1, mxREAL),
                     mxCreateDoubleMatrix(nc, 1, mxREAL), mxCreateDoubleMatrix(nc, 8, mxREAL)};
    ok(nnkgw_tree_export(id, mxGetPr(o[0]), mxGetPr(o[1]), mxGetPr(o[2]), mxGetPr(o[3]), mxGetPr(o[4]),
                         mxGetPr(o[5]), mxGetPr(o[6]), mxGetPr(o[7])));
    for (int k = 0; k < 8; k++) {
      if (k < (nlhs > 0 ? nlhs : 1)) plhs[k] = o[k];
      else mxDestroyArray(o[k]);
    }

  } else if (!strcmp(cmd, "tree_list")) {
    need(nrhs, 3, cmd);
    nnkgw_id id = id_of(prhs[1]);
    int32_t kind = (int32_t)mxGetScalar(prhs[2]), nc;
    int64_t total;
    ok(nnkgw_tree_sizes(id, NULL, NULL, NULL, &nc, NULL));
    ok(nnkgw_tree_list(id, kind, &total, NULL, NULL));
    plhs[0] = mxCreateDoubleMatrix(nc + 1, 1, mxREAL);
    mxArray *idx = mxCreateDoubleMatrix(total, 1, mxREAL);
    ok(nnkgw_tree_list(id, kind, &total, mxGetPr(plhs[0]), mxGetPr(idx)));
    if (nlhs > 1) plhs[1] = idx;
    else mxDestroyArray(idx);

  } else if (!strcmp(cmd, "nearest_matrix")) {
    need(nrhs, 3, cmd);
    const double *c = pts3(prhs[1], &n1), *f = pts3(prhs[2], &n2);
    int64_t nnz = 0;
    ok(nnkgw_nearest_matrix(c, n1, f, n2, &nnz, NULL, NULL, NULL));
    plhs[0] = mxCreateSparse(n2, n1, nnz > 0 ? nnz : 1, mxREAL);
    ok(nnkgw_nearest_matrix(c, n1, f, n2, &nnz, mxGetJc(plhs[0]), mxGetIr(plhs[0]), mxGetPr(plhs[0])));

  } else if (!strcmp(cmd, "fmm")) {
    need(nrhs, 5, cmd);
    const double *kp = mxGetPr(prhs[2]);
    if (mxGetNumberOfElements(prhs[2]) < 2) mexErrMsgIdAndTxt("nnkifmm:args", "kernelPar must be [eps mu]");
    int32_t prec = nrhs > 5 ? (int32_t)mxGetScalar(prhs[5]) : NNK_FP64;
    double tol = nrhs > 6 ? mxGetScalar(prhs[6]) : 0.0;
    nnkgw_id id;
    ok(nnkgw_fmm_create(id_of(prhs[1]), kp[0], kp[1], (int32_t)mxGetScalar(prhs[3]),
                        (int32_t)mxGetScalar(prhs[4]), prec, tol, &id));
    plhs[0] = mk_id(id);

  } else if (!strcmp(cmd, "apply") || !strcmp(cmd, "reduced") || !strcmp(cmd, "rsolve")) {
    need(nrhs, 3, cmd);
    const double *f = pts3(prhs[2], &n1);
    (void)n1;
    plhs[0] = mxCreateDoubleMatrix(3, mxGetN(prhs[2]), mxREAL);
    ok(nnkgw_fmm_apply(id_of(prhs[1]), !strcmp(cmd, "apply") ? NNKGW_APPLY
                                       : !strcmp(cmd, "reduced") ? NNKGW_REDUCED : NNKGW_RSOLVE,
                       f, mxGetPr(plhs[0])));

  } else if (!strcmp(cmd, "mobility")) {
    need(nrhs, 7, cmd);
    const mxArray *Bt = prhs[3], *BTt = prhs[4];
    sparse_in(Bt, "B.'");
    sparse_in(BTt, "BT.'");
    const int64_t nb6 = (int64_t)mxGetN(Bt), n3 = (int64_t)mxGetN(BTt);
    if ((int64_t)mxGetM(Bt) != n3 || (int64_t)mxGetM(BTt) != nb6 || (int64_t)mxGetM(prhs[5]) != nb6 ||
        (int64_t)mxGetN(prhs[5]) != nb6 || (int64_t)mxGetM(prhs[6]) != nb6 || (int64_t)mxGetN(prhs[6]) != nb6)
      mexErrMsgIdAndTxt("nnkifmm:args", "B.' (3N x 6nsw), BT.' (6nsw x 3N), A1.', A2.' (6nsw x 6nsw) expected");
    nnkgw_id id;
    ok(nnkgw_mobility_create(id_of(prhs[1]), id_of(prhs[2]), nb6, n3, mxGetJc(Bt), mxGetIr(Bt), mxGetPr(Bt),
                             mxGetJc(BTt), mxGetIr(BTt), mxGetPr(BTt), mxGetPr(prhs[5]), mxGetPr(prhs[6]), &id));
    plhs[0] = mk_id(id);

  } else if (!strcmp(cmd, "mobility_apply") || !strcmp(cmd, "precon")) {
    need(nrhs, 3, cmd);
    nnkgw_id id = id_of(prhs[1]);
    int64_t n;
    ok(nnkgw_mobility_size(id, &n));
    if ((int64_t)mxGetNumberOfElements(prhs[2]) != n) mexErrMsgIdAndTxt("nnkifmm:args", "vector length");
    plhs[0] = mxCreateDoubleMatrix(n, 1, mxREAL);
    ok(nnkgw_mobility_apply(id, cmd[0] == 'p', mxGetPr(prhs[2]), mxGetPr(plhs[0])));

  } else if (!strcmp(cmd, "solve")) {
    need(nrhs, 6, cmd);
    nnkgw_id id = id_of(prhs[1]);
    int64_t n;
    ok(nnkgw_mobility_size(id, &n));
    if ((int64_t)mxGetNumberOfElements(prhs[2]) != n) mexErrMsgIdAndTxt("nnkifmm:args", "vector length");
    int32_t maxit = (int32_t)mxGetScalar(prhs[4]), flag = 1, it = 0;
    double relres = 0;
    plhs[0] = mxCreateDoubleMatrix(n, 1, mxREAL);
    mxArray *rv = mxCreateDoubleMatrix(maxit + 1, 1, mxREAL);
    ok(nnkgw_mobility_solve(id, mxGetPr(prhs[2]), mxGetScalar(prhs[3]), maxit, (int32_t)mxGetScalar(prhs[5]),
                            mxGetPr(plhs[0]), &flag, &relres, &it, mxGetPr(rv)));
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(flag);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(relres);
    if (nlhs > 3) plhs[3] = mxCreateDoubleScalar(it);
    if (nlhs > 4) {
      mxSetM(rv, (mwSize)it + 1);
      plhs[4] = rv;
    } else {
      mxDestroyArray(rv);
    }

  } else if (!strcmp(cmd, "solve_refined") || !strcmp(cmd, "solve_fgmres")) {
    need(nrhs, 9, cmd);
    nnkgw_id id = id_of(prhs[1]);
    int64_t n;
    ok(nnkgw_mobility_size(id, &n));
    if ((int64_t)mxGetNumberOfElements(prhs[4]) != n) mexErrMsgIdAndTxt("nnkifmm:args", "vector length");
    int32_t flag = 1, it = 0, outer = 0;
    double relres = 0;
    plhs[0] = mxCreateDoubleMatrix(n, 1, mxREAL);
    ok((!strcmp(cmd, "solve_fgmres") ? nnkgw_mobility_solve_fgmres : nnkgw_mobility_solve_refined)(
        id, id_of(prhs[2]), id_of(prhs[3]), mxGetPr(prhs[4]), mxGetScalar(prhs[5]), (int32_t)mxGetScalar(prhs[6]),
        mxGetScalar(prhs[7]), (int32_t)mxGetScalar(prhs[8]), mxGetPr(plhs[0]), &flag, &relres, &it, &outer));
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(flag);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(relres);
    if (nlhs > 3) plhs[3] = mxCreateDoubleScalar(it);
    if (nlhs > 4) plhs[4] = mxCreateDoubleScalar(outer);

  } else if (!strcmp(cmd, "opcache")) {
    need(nrhs, 2, cmd);
    if (nnk_op_cache_limit((int64_t)mxGetScalar(prhs[1]))) mexErrMsgIdAndTxt("nnkifmm:fail", "%s", nnk_last_error());

  } else if (!strcmp(cmd, "free")) {
    need(nrhs, 2, cmd);
    ok(nnkgw_free(id_of(prhs[1])));

  } else {
    mexErrMsgIdAndTxt("nnkifmm:args", "unknown command '%s'", cmd);
  }
}
