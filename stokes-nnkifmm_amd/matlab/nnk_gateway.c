/*
 * nnk_gateway.c -- MATLAB-side marshalling for libnnkifmm (see
 * nnk_gateway.h).  Plain C99; built with gcc into libnnk_gateway.so and
 * linked by the MEX file nnkifmm_mex.c.  Single-threaded by design: the
 * MATLAB client thread is the only caller (parfor workers are separate
 * processes), so the handle registry takes no lock.
 */
#include "nnk_gateway.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nnkifmm.h"

/* ------------------------------------------------------------------ errors */

static char g_err[512];

const char *nnkgw_last_error(void) { return g_err; }

static int32_t fail(int32_t rc, const char *msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return rc;
}

/* pass a library status through, keeping its message */
static int32_t lib(int32_t rc) {
  if (rc) snprintf(g_err, sizeof g_err, "%s", nnk_last_error());
  else g_err[0] = 0;
  return rc;
}

/* ---------------------------------------------------------------- registry */

typedef struct {
  int32_t kind;      /* 0 = free slot */
  void *obj;
  int32_t released;  /* the MATLAB handle is gone */
  int32_t refs;      /* live objects built on this one */
  nnkgw_id dep[2];   /* objects this one was built on */
  int64_t size;      /* mobility: vector length 3N + nb6 */
} entry;

static entry *g_tab = NULL;
static int64_t g_n = 0, g_cap = 0;

static int32_t reg(int32_t kind, void *obj, nnkgw_id d0, nnkgw_id d1, nnkgw_id *out) {
  if (g_n == g_cap) {
    int64_t nc = g_cap ? 2 * g_cap : 64;
    entry *t = (entry *)realloc(g_tab, sizeof(entry) * (size_t)nc);
    if (!t) return fail(NNK_ENOMEM, "nnkifmm: out of host memory");
    g_tab = t;
    g_cap = nc;
  }
  entry *e = &g_tab[g_n++];
  e->kind = kind;
  e->obj = obj;
  e->released = 0;
  e->refs = 0;
  e->dep[0] = d0;
  e->dep[1] = d1;
  e->size = 0;
  if (d0) g_tab[d0 - 1].refs++;
  if (d1) g_tab[d1 - 1].refs++;
  *out = (nnkgw_id)g_n;
  return NNK_OK;
}

static void *get(nnkgw_id id, int32_t kind) {
  if (id == 0 || (int64_t)id > g_n) return NULL;
  entry *e = &g_tab[id - 1];
  if (e->kind != kind || e->released) return NULL;
  return e->obj;
}

static void destroy(entry *e) {
  if (e->kind == NNKGW_MOBILITY) nnk_mobility_destroy((nnk_mobility *)e->obj);
  else if (e->kind == NNKGW_FMM) nnk_fmm_destroy((nnk_fmm *)e->obj);
  else if (e->kind == NNKGW_TREE) nnk_tree_destroy((nnk_tree *)e->obj);
  e->kind = 0;
  e->obj = NULL;
}

/* destroy a released entry nothing depends on any more, then its bases */
static void collect(nnkgw_id id) {
  entry *e = &g_tab[id - 1];
  if (!e->kind || !e->released || e->refs > 0) return;
  nnkgw_id d[2] = {e->dep[0], e->dep[1]};
  destroy(e);
  for (int k = 0; k < 2; k++)
    if (d[k]) {
      g_tab[d[k] - 1].refs--;
      collect(d[k]);
    }
}

int32_t nnkgw_free(nnkgw_id id) {
  if (id == 0 || (int64_t)id > g_n || !g_tab[id - 1].kind || g_tab[id - 1].released)
    return fail(NNK_EARG, "nnkifmm: free of an unknown or released handle");
  g_tab[id - 1].released = 1;
  collect(id);
  g_err[0] = 0;
  return NNK_OK;
}

void nnkgw_free_all(void) {
  for (int32_t kind = NNKGW_MOBILITY; kind >= NNKGW_TREE; kind--)
    for (int64_t i = 0; i < g_n; i++)
      if (g_tab[i].kind == kind) destroy(&g_tab[i]);
  free(g_tab);
  g_tab = NULL;
  g_n = g_cap = 0;
}

int64_t nnkgw_live(int32_t kind) {
  int64_t c = 0;
  for (int64_t i = 0; i < g_n; i++)
    if (g_tab[i].kind && (kind == 0 || g_tab[i].kind == kind)) c++;
  return c;
}

/* ----------------------------------------------------------------- layouts */

void nnkgw_rowmajor_to_colmajor(const double *in, int64_t rows, int64_t cols, double *out) {
  for (int64_t r = 0; r < rows; r++)
    for (int64_t c = 0; c < cols; c++) out[c * rows + r] = in[r * cols + c];
}

int32_t nnkgw_index_to_i64(const nnkgw_index *in, int64_t n, int64_t *out) {
  for (int64_t i = 0; i < n; i++) {
    if (in[i] > (nnkgw_index)INT64_MAX) return fail(NNK_EARG, "nnkifmm: sparse index out of range");
    out[i] = (int64_t)in[i];
  }
  return NNK_OK;
}

int32_t nnkgw_index_to_i32(const nnkgw_index *in, int64_t n, int32_t *out) {
  for (int64_t i = 0; i < n; i++) {
    if (in[i] > (nnkgw_index)INT32_MAX) return fail(NNK_EARG, "nnkifmm: sparse index out of range");
    out[i] = (int32_t)in[i];
  }
  return NNK_OK;
}

/* -------------------------------------------------------------------- tree */

int32_t nnkgw_tree_create(const double *pot3, int64_t npot, const double *tgt3, int64_t ntgt, int32_t cap,
                          int32_t depth, const double *cube, nnkgw_id *out) {
  if (!out) return fail(NNK_EARG, "nnkifmm: tree: null output");
  *out = 0;
  nnk_tree *t = NULL;
  int32_t rc = cube ? nnk_tree_create_in_cube(pot3, npot, tgt3, ntgt, cap, depth, cube, &t)
                    : nnk_tree_create(pot3, npot, tgt3, ntgt, cap, depth, &t);
  if (rc) return lib(rc);
  if ((rc = reg(NNKGW_TREE, t, 0, 0, out))) nnk_tree_destroy(t);
  return rc ? rc : lib(NNK_OK);
}

int32_t nnkgw_tree_set_nearest(nnkgw_id tree, const double *fine3, int64_t q, int64_t ncol,
                               const nnkgw_index *jc, const nnkgw_index *ir, const double *pr) {
  nnk_tree *t = (nnk_tree *)get(tree, NNKGW_TREE);
  if (!t) return fail(NNK_EARG, "nnkifmm: nearest: not a live tree handle");
  int64_t npot = 0;
  int32_t rc = nnk_tree_sizes(t, NULL, &npot, NULL, NULL, NULL);
  if (rc) return lib(rc);
  if (ncol != npot) return fail(NNK_EARG, "nnkifmm: nearest: NN must have one column per potential point");
  if (!jc) return fail(NNK_EARG, "nnkifmm: nearest: null NN");
  const int64_t nnz = (int64_t)jc[ncol];
  int64_t *cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncol + 1));
  int64_t *ri = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
  if (!cp || !ri) rc = fail(NNK_ENOMEM, "nnkifmm: out of host memory");
  if (!rc) rc = nnkgw_index_to_i64(jc, ncol + 1, cp);
  if (!rc && nnz > 0) rc = ir ? nnkgw_index_to_i64(ir, nnz, ri) : fail(NNK_EARG, "nnkifmm: nearest: null NN rows");
  if (!rc) rc = lib(nnk_tree_set_nearest(t, fine3, q, cp, nnz > 0 ? ri : NULL, nnz > 0 ? pr : NULL));
  free(cp);
  free(ri);
  return rc;
}

int32_t nnkgw_tree_sizes(nnkgw_id tree, int64_t *npts, int64_t *npot, int64_t *ntgt, int32_t *count,
                         int32_t *cap) {
  nnk_tree *t = (nnk_tree *)get(tree, NNKGW_TREE);
  if (!t) return fail(NNK_EARG, "nnkifmm: sizes: not a live tree handle");
  return lib(nnk_tree_sizes(t, npts, npot, ntgt, count, cap));
}

int32_t nnkgw_tree_export(nnkgw_id tree, double *points, double *potentials, double *targets,
                          double *point_index, double *corners, double *level, double *parent,
                          double *children) {
  nnk_tree *t = (nnk_tree *)get(tree, NNKGW_TREE);
  if (!t) return fail(NNK_EARG, "nnkifmm: export: not a live tree handle");
  int64_t np = 0;
  int32_t nc = 0;
  int32_t rc = nnk_tree_sizes(t, &np, NULL, NULL, &nc, NULL);
  if (rc) return lib(rc);
  double *P = (double *)malloc(sizeof(double) * (size_t)(3 * np + 1));
  uint8_t *ip = (uint8_t *)malloc((size_t)np + 1), *it = (uint8_t *)malloc((size_t)np + 1);
  int32_t *pi = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np + 1));
  double *C = (double *)malloc(sizeof(double) * (size_t)(6 * nc + 1));
  int32_t *lv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nc + 1));
  int32_t *pa = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nc + 1));
  int32_t *ch = (int32_t *)malloc(sizeof(int32_t) * (size_t)(8 * nc + 1));
  if (!P || !ip || !it || !pi || !C || !lv || !pa || !ch) rc = fail(NNK_ENOMEM, "nnkifmm: out of host memory");
  if (!rc) rc = lib(nnk_tree_export(t, P, ip, it, pi, C, lv, pa, ch));
  if (!rc) {
    if (points) nnkgw_rowmajor_to_colmajor(P, np, 3, points);
    for (int64_t i = 0; i < np; i++) {
      if (potentials) potentials[i] = ip[i];
      if (targets) targets[i] = it[i];
      if (point_index) point_index[i] = pi[i];
    }
    if (corners) nnkgw_rowmajor_to_colmajor(C, nc, 6, corners);
    for (int32_t n = 0; n < nc; n++) {
      if (level) level[n] = lv[n];
      if (parent) parent[n] = pa[n];
      if (children)
        for (int k = 0; k < 8; k++) children[(int64_t)k * nc + n] = ch[8 * (int64_t)n + k];
    }
  }
  free(P); free(ip); free(it); free(pi); free(C); free(lv); free(pa); free(ch);
  return rc;
}

int32_t nnkgw_tree_list(nnkgw_id tree, int32_t kind, int64_t *total, double *ptr, double *idx) {
  nnk_tree *t = (nnk_tree *)get(tree, NNKGW_TREE);
  if (!t) return fail(NNK_EARG, "nnkifmm: list: not a live tree handle");
  if (!total) return fail(NNK_EARG, "nnkifmm: list: null output");
  int32_t nc = 0;
  int32_t rc = nnk_tree_sizes(t, NULL, NULL, NULL, &nc, NULL);
  if (!rc) rc = nnk_tree_list_size(t, kind, total);
  if (rc) return lib(rc);
  if (!ptr && !idx) return lib(NNK_OK);
  int64_t *p = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nc + 1));
  int32_t *x = (int32_t *)malloc(sizeof(int32_t) * (size_t)(*total + 1));
  if (!p || !x) rc = fail(NNK_ENOMEM, "nnkifmm: out of host memory");
  if (!rc) rc = lib(nnk_tree_list_export(t, kind, p, x));
  if (!rc) {
    if (ptr)
      for (int32_t n = 0; n <= nc; n++) ptr[n] = (double)p[n];
    if (idx)
      for (int64_t e = 0; e < *total; e++) idx[e] = x[e];
  }
  free(p);
  free(x);
  return rc;
}

int32_t nnkgw_nearest_matrix(const double *coarse3, int64_t n, const double *fine3, int64_t q, int64_t *nnz,
                             nnkgw_index *jc, nnkgw_index *ir, double *pr) {
  if (!nnz || n <= 0) return fail(NNK_EARG, "nnkifmm: nearestMatrix: bad argument");
  int64_t *cp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  if (!cp) return fail(NNK_ENOMEM, "nnkifmm: out of host memory");
  int32_t rc;
  if (!jc) {  /* size query: a call with no room reports nnz */
    int64_t r1 = 0;
    double v1 = 0;
    rc = nnk_nearest_matrix(coarse3, n, fine3, q, cp, &r1, &v1, 0, nnz);
    if (rc == NNK_EARG && *nnz > 0) rc = NNK_OK;
    free(cp);
    return lib(rc);
  }
  const int64_t cap = *nnz;
  int64_t *ri = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap > 0 ? cap : 1));
  if (!ri) { free(cp); return fail(NNK_ENOMEM, "nnkifmm: out of host memory"); }
  rc = lib(nnk_nearest_matrix(coarse3, n, fine3, q, cp, ri, pr, cap, nnz));
  if (!rc) {
    for (int64_t j = 0; j <= n; j++) jc[j] = (nnkgw_index)cp[j];
    for (int64_t e = 0; e < *nnz; e++) ir[e] = (nnkgw_index)ri[e];
  }
  free(cp);
  free(ri);
  return rc;
}

/* ------------------------------------------------------------------- KIFMM */

int32_t nnkgw_fmm_create(nnkgw_id tree, double eps, double mu, int32_t res, int32_t shells, int32_t precision,
                         double svd_tol, nnkgw_id *out) {
  if (!out) return fail(NNK_EARG, "nnkifmm: KIFMM: null output");
  *out = 0;
  nnk_tree *t = (nnk_tree *)get(tree, NNKGW_TREE);
  if (!t) return fail(NNK_EARG, "nnkifmm: KIFMM: not a live tree handle");
  nnk_fmm *f = NULL;
  int32_t rc = nnk_fmm_create(t, eps, mu, res, shells, precision, svd_tol, &f);
  if (rc) return lib(rc);
  if ((rc = reg(NNKGW_FMM, f, tree, 0, out))) nnk_fmm_destroy(f);
  return rc ? rc : lib(NNK_OK);
}

int32_t nnkgw_fmm_apply(nnkgw_id fmm, int32_t op, const double *in3, double *out3) {
  nnk_fmm *f = (nnk_fmm *)get(fmm, NNKGW_FMM);
  if (!f) return fail(NNK_EARG, "nnkifmm: apply: not a live KIFMM handle");
  switch (op) {
    case NNKGW_APPLY: return lib(nnk_fmm_apply(f, in3, out3));
    case NNKGW_REDUCED: return lib(nnk_fmm_reduced_apply(f, in3, out3));
    case NNKGW_RSOLVE: return lib(nnk_fmm_reduced_solve(f, in3, out3));
    default: return fail(NNK_EARG, "nnkifmm: apply: unknown operation");
  }
}

/* ---------------------------------------------------------------- mobility */

int32_t nnkgw_mobility_create(nnkgw_id fmm, nnkgw_id fmm4, int64_t nb6, int64_t n3, const nnkgw_index *Bt_jc,
                              const nnkgw_index *Bt_ir, const double *Bt_pr, const nnkgw_index *BTt_jc,
                              const nnkgw_index *BTt_ir, const double *BTt_pr, const double *A1t,
                              const double *A2t, nnkgw_id *out) {
  if (!out) return fail(NNK_EARG, "nnkifmm: mobility: null output");
  *out = 0;
  nnk_fmm *f = (nnk_fmm *)get(fmm, NNKGW_FMM);
  nnk_fmm *f4 = fmm4 ? (nnk_fmm *)get(fmm4, NNKGW_FMM) : NULL;
  if (!f || (fmm4 && !f4)) return fail(NNK_EARG, "nnkifmm: mobility: not a live KIFMM handle");
  if (nb6 < 0 || n3 < 0 || !Bt_jc || !BTt_jc) return fail(NNK_EARG, "nnkifmm: mobility: bad B / BT");
  const int64_t nzB = (int64_t)Bt_jc[nb6], nzT = (int64_t)BTt_jc[n3];
  int64_t *Brp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nb6 + 1));
  int64_t *Trp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n3 + 1));
  int32_t *Bci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nzB + 1));
  int32_t *Tci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nzT + 1));
  int32_t rc = (!Brp || !Trp || !Bci || !Tci) ? fail(NNK_ENOMEM, "nnkifmm: out of host memory") : NNK_OK;
  /* CSC of B.' = CSR of B; CSC of BT.' = CSR of BT */
  if (!rc) rc = nnkgw_index_to_i64(Bt_jc, nb6 + 1, Brp);
  if (!rc) rc = nnkgw_index_to_i64(BTt_jc, n3 + 1, Trp);
  if (!rc && nzB > 0) rc = Bt_ir ? nnkgw_index_to_i32(Bt_ir, nzB, Bci) : fail(NNK_EARG, "nnkifmm: mobility: null B");
  if (!rc && nzT > 0) rc = BTt_ir ? nnkgw_index_to_i32(BTt_ir, nzT, Tci) : fail(NNK_EARG, "nnkifmm: mobility: null BT");
  nnk_mobility *m = NULL;
  /* A1.', A2.' in MATLAB's column-major order are A1, A2 row-major */
  if (!rc) rc = lib(nnk_mobility_create(f, f4, nb6, Brp, Bci, Bt_pr, Trp, Tci, BTt_pr, A1t, A2t, &m));
  free(Brp); free(Trp); free(Bci); free(Tci);
  if (rc) return rc;
  if ((rc = reg(NNKGW_MOBILITY, m, fmm, fmm4, out))) {
    nnk_mobility_destroy(m);
    return rc;
  }
  g_tab[*out - 1].size = n3 + nb6;
  return lib(NNK_OK);
}

int32_t nnkgw_mobility_apply(nnkgw_id mob, int32_t precon, const double *x, double *y) {
  nnk_mobility *m = (nnk_mobility *)get(mob, NNKGW_MOBILITY);
  if (!m) return fail(NNK_EARG, "nnkifmm: mobility apply: not a live mobility handle");
  return lib(precon ? nnk_mobility_precon(m, x, y) : nnk_mobility_apply(m, x, y));
}

int32_t nnkgw_mobility_solve(nnkgw_id mob, const double *rhs, double tol, int32_t maxit, int32_t precon,
                             double *x, int32_t *flag, double *relres, int32_t *iters, double *resvec) {
  nnk_mobility *m = (nnk_mobility *)get(mob, NNKGW_MOBILITY);
  if (!m) return fail(NNK_EARG, "nnkifmm: solve: not a live mobility handle");
  return lib(nnk_mobility_solve(m, rhs, tol, maxit, precon, x, flag, relres, iters, resvec));
}

int32_t nnkgw_mobility_solve_refined(nnkgw_id mob, nnkgw_id inner, nnkgw_id inner4, const double *rhs,
                                     double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                     int32_t *flag, double *relres, int32_t *iters, int32_t *outer) {
  nnk_mobility *m = (nnk_mobility *)get(mob, NNKGW_MOBILITY);
  nnk_fmm *fi = (nnk_fmm *)get(inner, NNKGW_FMM);
  nnk_fmm *f4 = inner4 ? (nnk_fmm *)get(inner4, NNKGW_FMM) : NULL;
  if (!m || !fi || (inner4 && !f4)) return fail(NNK_EARG, "nnkifmm: refined solve: not a live handle");
  return lib(nnk_mobility_solve_refined(m, fi, f4, rhs, tol, maxit, inner_tol, precon, x, flag, relres, iters,
                                        outer));
}

int32_t nnkgw_mobility_solve_fgmres(nnkgw_id mob, nnkgw_id inner, nnkgw_id inner4, const double *rhs,
                                    double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                    int32_t *flag, double *relres, int32_t *iters, int32_t *outer) {
  nnk_mobility *m = (nnk_mobility *)get(mob, NNKGW_MOBILITY);
  nnk_fmm *fi = (nnk_fmm *)get(inner, NNKGW_FMM);
  nnk_fmm *f4 = inner4 ? (nnk_fmm *)get(inner4, NNKGW_FMM) : NULL;
  if (!m || !fi || (inner4 && !f4)) return fail(NNK_EARG, "nnkifmm: flexible GMRES: not a live handle");
  return lib(nnk_mobility_solve_fgmres(m, fi, f4, rhs, tol, maxit, inner_tol, precon, x, flag, relres, iters,
                                       outer));
}

int32_t nnkgw_mobility_size(nnkgw_id mob, int64_t *n) {
  if (!get(mob, NNKGW_MOBILITY) || !n) return fail(NNK_EARG, "nnkifmm: size: not a live mobility handle");
  *n = g_tab[mob - 1].size;
  return lib(NNK_OK);
}
