/*
 * gateway_test.c -- drives libnnkifmm through the MATLAB marshalling layer
 * (stokes-nnkifmm_amd/matlab/nnk_gateway.c) exactly as nnkifmm_mex.c does:
 * (3,n) column-major point sets, MATLAB CSC sparse matrices with mwIndex
 * (size_t) indices, B and BT handed over as the CSC of B.' and BT.', A1.'
 * and A2.' as column-major arrays, handles as registry ids.
 *
 * usage: nnk_gateway_test <dir>
 * Reads from <dir> (raw little-endian files written by the test):
 *   sizes.i64   n q nnzNN nb6 nnzBt nnzBTt cap
 *   pot3.f64 (3 x n)  fine3.f64 (3 x q)  nn_jc.idx nn_ir.idx nn_pr.f64
 *   f3.f64 (3 x n)  bt_jc.idx bt_ir.idx bt_pr.f64  btt_jc.idx btt_ir.idx
 *   btt_pr.f64  a1t.f64  a2t.f64  x.f64 (3n + nb6)  params.f64 (eps mu tol)
 * Writes: points.f64 (n,3 column-major), corners.f64, children.f64,
 *   level.f64, pointindex.f64, list<k>_ptr.f64 / list<k>_idx.f64,
 *   u_apply.f64, u_reduced.f64, y_apply.f64, y_precon.f64, x_solve.f64,
 *   solve.f64 (flag relres iters), and prints one line per step.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nnk_gateway.h"

static const char *g_dir;

static void *load(const char *name, size_t elem, size_t n) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", g_dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  void *p = malloc(elem * (n ? n : 1));
  if (n && fread(p, elem, n, f) != n) { fprintf(stderr, "short read %s\n", path); exit(2); }
  fclose(f);
  return p;
}

static void save(const char *name, const double *p, size_t n) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", g_dir, name);
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, sizeof(double), n, f) != n) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
  fclose(f);
}

#define CHECK(call)                                                            \
  do {                                                                         \
    int32_t rc_ = (call);                                                      \
    if (rc_) {                                                                 \
      fprintf(stderr, "FAILED %s: %d %s\n", #call, rc_, nnkgw_last_error());   \
      return 1;                                                                \
    }                                                                          \
  } while (0)

#define EXPECT(cond)                                        \
  do {                                                      \
    if (!(cond)) {                                          \
      fprintf(stderr, "EXPECTATION FAILED: %s\n", #cond);   \
      return 1;                                             \
    }                                                       \
  } while (0)

int main(int argc, char **argv) {
  if (argc < 2) { fprintf(stderr, "usage: %s <dir>\n", argv[0]); return 2; }
  g_dir = argv[1];
  int64_t *sz = (int64_t *)load("sizes.i64", 8, 7);
  const int64_t n = sz[0], q = sz[1], nnzNN = sz[2], nb6 = sz[3], nzB = sz[4], nzT = sz[5];
  const int32_t cap = (int32_t)sz[6];
  const int64_t n3 = 3 * n, nv = n3 + nb6;
  double *pot3 = (double *)load("pot3.f64", 8, (size_t)n3);
  double *fine3 = (double *)load("fine3.f64", 8, (size_t)(3 * q));
  nnkgw_index *jc = (nnkgw_index *)load("nn_jc.idx", sizeof(nnkgw_index), (size_t)n + 1);
  nnkgw_index *ir = (nnkgw_index *)load("nn_ir.idx", sizeof(nnkgw_index), (size_t)nnzNN);
  double *pr = (double *)load("nn_pr.f64", 8, (size_t)nnzNN);
  double *f3 = (double *)load("f3.f64", 8, (size_t)n3);
  nnkgw_index *bj = (nnkgw_index *)load("bt_jc.idx", sizeof(nnkgw_index), (size_t)nb6 + 1);
  nnkgw_index *bi = (nnkgw_index *)load("bt_ir.idx", sizeof(nnkgw_index), (size_t)nzB);
  double *bp = (double *)load("bt_pr.f64", 8, (size_t)nzB);
  nnkgw_index *tj = (nnkgw_index *)load("btt_jc.idx", sizeof(nnkgw_index), (size_t)n3 + 1);
  nnkgw_index *ti = (nnkgw_index *)load("btt_ir.idx", sizeof(nnkgw_index), (size_t)nzT);
  double *tp = (double *)load("btt_pr.f64", 8, (size_t)nzT);
  double *a1t = (double *)load("a1t.f64", 8, (size_t)(nb6 * nb6));
  double *a2t = (double *)load("a2t.f64", 8, (size_t)(nb6 * nb6));
  double *x = (double *)load("x.f64", 8, (size_t)nv);
  double *par = (double *)load("params.f64", 8, 3);

  /* OcTree(pot, pot, 'nodeCapacity', cap, 'nearest', 1, 'finePoints', fine, 'NN', NN) */
  nnkgw_id tree = 0, fmm = 0, fmm4 = 0, mob = 0;
  CHECK(nnkgw_tree_create(pot3, n, pot3, n, cap, 21, NULL, &tree));
  CHECK(nnkgw_tree_set_nearest(tree, fine3, q, n, jc, ir, pr));
  int64_t npts = 0;
  int32_t count = 0;
  CHECK(nnkgw_tree_sizes(tree, &npts, NULL, NULL, &count, NULL));
  printf("tree: %lld points, %d nodes\n", (long long)npts, count);
  fflush(stdout);
  double *P = malloc(sizeof(double) * (size_t)(3 * npts)), *C = malloc(sizeof(double) * (size_t)(6 * count));
  double *CH = malloc(sizeof(double) * (size_t)(8 * count)), *LV = malloc(sizeof(double) * (size_t)count);
  double *PI = malloc(sizeof(double) * (size_t)npts);
  CHECK(nnkgw_tree_export(tree, P, NULL, NULL, PI, C, LV, NULL, CH));
  save("points.f64", P, (size_t)(3 * npts));
  save("corners.f64", C, (size_t)(6 * count));
  save("children.f64", CH, (size_t)(8 * count));
  save("level.f64", LV, (size_t)count);
  save("pointindex.f64", PI, (size_t)npts);
  for (int kind = 0; kind < 4; kind++) {
    int64_t total = 0;
    CHECK(nnkgw_tree_list(tree, kind, &total, NULL, NULL));
    double *lp = malloc(sizeof(double) * (size_t)(count + 1)), *li = malloc(sizeof(double) * (size_t)(total + 1));
    CHECK(nnkgw_tree_list(tree, kind, &total, lp, li));
    char nm[64];
    snprintf(nm, sizeof nm, "list%d_ptr.f64", kind);
    save(nm, lp, (size_t)count + 1);
    snprintf(nm, sizeof nm, "list%d_idx.f64", kind);
    save(nm, li, (size_t)total);
    free(lp);
    free(li);
  }

  /* KIFMM(tree, [eps mu]) and its coronaRes-4 twin */
  CHECK(nnkgw_fmm_create(tree, par[0], par[1], 6, 1, 0, 0.0, &fmm));
  CHECK(nnkgw_fmm_create(tree, par[0], par[1], 4, 1, 0, 0.0, &fmm4));
  double *u = malloc(sizeof(double) * (size_t)n3);
  CHECK(nnkgw_fmm_apply(fmm, NNKGW_APPLY, f3, u));
  save("u_apply.f64", u, (size_t)n3);
  CHECK(nnkgw_fmm_apply(fmm, NNKGW_REDUCED, f3, u));
  save("u_reduced.f64", u, (size_t)n3);
  printf("apply: done\n");
  fflush(stdout);

  /* mobility operator from B.', BT.', A1.', A2.' */
  CHECK(nnkgw_mobility_create(fmm, fmm4, nb6, n3, bj, bi, bp, tj, ti, tp, a1t, a2t, &mob));
  int64_t len = 0;
  CHECK(nnkgw_mobility_size(mob, &len));
  EXPECT(len == nv);
  double *y = malloc(sizeof(double) * (size_t)nv);
  CHECK(nnkgw_mobility_apply(mob, 0, x, y));
  save("y_apply.f64", y, (size_t)nv);
  CHECK(nnkgw_mobility_apply(mob, 1, x, y));
  save("y_precon.f64", y, (size_t)nv);
  int32_t flag = -1, iters = 0;
  double relres = -1.0;
  CHECK(nnkgw_mobility_solve(mob, x, par[2], 200, 1, y, &flag, &relres, &iters, NULL));
  save("x_solve.f64", y, (size_t)nv);
  double info[3] = {(double)flag, relres, (double)iters};
  save("solve.f64", info, 3);
  printf("solve: flag %d relres %.3e iters %d\n", flag, relres, iters);
  fflush(stdout);

  /* handle lifetimes: releasing the tree and KIFMM handles first keeps the
     objects alive (the mobility operator uses them); a released id is
     refused; releasing the mobility handle then frees everything */
  EXPECT(nnkgw_live(0) == 4);
  CHECK(nnkgw_free(tree));
  CHECK(nnkgw_free(fmm));
  EXPECT(nnkgw_live(0) == 4);
  EXPECT(nnkgw_fmm_apply(fmm, NNKGW_APPLY, f3, u) != 0);
  EXPECT(nnkgw_free(tree) != 0);
  CHECK(nnkgw_mobility_apply(mob, 0, x, y));   /* still usable */
  CHECK(nnkgw_free(mob));
  EXPECT(nnkgw_live(NNKGW_MOBILITY) == 0 && nnkgw_live(NNKGW_TREE) == 1 && nnkgw_live(NNKGW_FMM) == 1);
  CHECK(nnkgw_free(fmm4));
  EXPECT(nnkgw_live(0) == 0);
  EXPECT(nnkgw_tree_sizes(tree, &npts, NULL, NULL, &count, NULL) != 0);
  /* mexAtExit path: free_all with live objects of every kind */
  CHECK(nnkgw_tree_create(pot3, n, pot3, n, cap, 21, NULL, &tree));
  CHECK(nnkgw_fmm_create(tree, par[0], par[1], 6, 1, 0, 0.0, &fmm));
  EXPECT(nnkgw_live(0) == 2);
  nnkgw_free_all();
  EXPECT(nnkgw_live(0) == 0);
  printf("handles: done\nGATEWAY_OK\n");
  return 0;
}
