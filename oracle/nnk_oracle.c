/*
 * nnk_oracle.c -- CPU restatement of the reference's KIFMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library and the timed CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * It restates, in FP64 C with OpenMP (standing in for MATLAB parfor), the
 * algorithm of the MATLAB reference as written:
 *   kernel                 Tools/kernel.m:15-49
 *   blockcomputation       Tools/Other/blockcomputation.m:17-28
 *   genupsurf/gendownsurf  Tools/KIFMM/genupsurf.m:30-53, gendownsurf.m:41-59
 *   OcTree build           Classes/OcTree.m:77-222
 *   interaction lists      Classes/OcTree.m:265-366, Tools/Tree/{isleaf,nodeintersection,getChildren}.m
 *   nearestMatrix          Tools/Nearest/nearestMatrix.m:17-41
 *   upwardPass / Up        Tools/Passes/upwardPass.m:18-99, Up.m:30-34,
 *                          Tools/Nearest/UpNearest.m:33-37
 *   downwardPass           Tools/Passes/downwardPass.m:23-305,
 *                          DownNonleaf.m:41-47, Downleaf.m:36-43 (+Nearest)
 *   reducedPass            Tools/Passes/reducedPass.m:21-80
 *   ReducedSolve           Mobility/ReducedSolve.m:88-133
 * Every per-node surface, dense kernel block and pivoted LU solve is formed
 * per node and per call, exactly as the reference does.
 *
 * MATLAB's mldivide on a dense square non-symmetric matrix is LU with
 * partial pivoting (LAPACK getrf/getrs); that is what lu_solve() does.  On
 * the symmetric reduced-solve blocks mldivide would try Cholesky first;
 * the oracle uses the same LU there (a rounding-level difference; parity of
 * that choice is unpinned since MATLAB cannot run here).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_OK 0
#define OR_EMAXDEPTH 1
#define OR_EARG 2
#define OR_ENOMEM 3

static char g_err[512];
const char *or_last_error(void) { return g_err; }

static const double OR_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------ */
/* kernel.m: S(x0, x) = [(r^2 + 2 eps^2) I + r r^T] / (8 pi mu (r^2+eps^2)^{3/2}),
 * r = x0 - x.  3x3 block for one (target, source) pair, row-major.       */
static inline void stokeslet_block(const double *x0, const double *x,
                                   double eps, double mu, double S[9]) {
  double rx = x0[0] - x[0], ry = x0[1] - x[1], rz = x0[2] - x[2];
  double r2 = rx * rx + ry * ry + rz * rz;
  double e2 = eps * eps;
  double invs = (1.0 / pow(r2 + e2, 1.5)) * (1.0 / (8.0 * OR_PI * mu));
  double iso = r2 + 2.0 * e2;
  S[0] = (rx * rx + iso) * invs; S[1] = (rx * ry) * invs; S[2] = (rx * rz) * invs;
  S[3] = (ry * rx) * invs; S[4] = (ry * ry + iso) * invs; S[5] = (ry * rz) * invs;
  S[6] = (rz * rx) * invs; S[7] = (rz * ry) * invs; S[8] = (rz * rz + iso) * invs;
}

/* Dense (3m x 3n) matrix, row-major, rows = targets x0, cols = sources x. */
int or_kernel(const double *x, int64_t n, const double *x0, int64_t m,
              double eps, double mu, double *S) {
  for (int64_t i = 0; i < m; i++)
    for (int64_t j = 0; j < n; j++) {
      double B[9];
      stokeslet_block(x0 + 3 * i, x + 3 * j, eps, mu, B);
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          S[(3 * i + a) * (3 * n) + 3 * j + b] = B[3 * a + b];
    }
  return OR_OK;
}

/* blockcomputation.m: y(3m) (+)= K(x -> x0) q(3n); rows in order, columns
 * summed in source order. */
static void gemv_kernel(const double *x, int64_t n, const double *x0, int64_t m,
                        const double *q, double eps, double mu, double *y,
                        int accumulate) {
  for (int64_t i = 0; i < m; i++) {
    double acc[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; j++) {
      double B[9];
      stokeslet_block(x0 + 3 * i, x + 3 * j, eps, mu, B);
      for (int a = 0; a < 3; a++)
        acc[a] += B[3 * a] * q[3 * j] + B[3 * a + 1] * q[3 * j + 1] +
                  B[3 * a + 2] * q[3 * j + 2];
    }
    for (int a = 0; a < 3; a++) {
      if (accumulate) y[3 * i + a] += acc[a];
      else y[3 * i + a] = acc[a];
    }
  }
}

int or_direct(const double *src, int64_t n, const double *q, const double *tgt,
              int64_t m, double eps, double mu, double *u) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++)
    gemv_kernel(src, n, tgt + 3 * i, 1, q, eps, mu, u + 3 * i, 0);
  return OR_OK;
}

/* ------------------------------------------------------------------ */
/* LU with partial pivoting (Golub & Van Loan alg. 3.4.1), in place on a
 * row-major n x n matrix, then forward/back substitution for one rhs.   */
static int lu_solve(double *A, int n, double *b) {
  for (int k = 0; k < n; k++) {
    int p = k;
    double amax = fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; i++) {
      double v = fabs(A[(size_t)i * n + k]);
      if (v > amax) { amax = v; p = i; }
    }
    if (p != k) {
      for (int j = 0; j < n; j++) {
        double t = A[(size_t)k * n + j];
        A[(size_t)k * n + j] = A[(size_t)p * n + j];
        A[(size_t)p * n + j] = t;
      }
      double t = b[k]; b[k] = b[p]; b[p] = t;
    }
    double piv = A[(size_t)k * n + k];
    if (piv == 0.0) continue; /* singular column: MATLAB warns and goes on */
    for (int i = k + 1; i < n; i++) {
      double l = A[(size_t)i * n + k] / piv;
      A[(size_t)i * n + k] = l;
      if (l != 0.0) {
        double *ri = A + (size_t)i * n, *rk = A + (size_t)k * n;
        for (int j = k + 1; j < n; j++) ri[j] -= l * rk[j];
      }
    }
  }
  for (int i = 0; i < n; i++) { /* L y = P b (unit lower) */
    double s = b[i];
    for (int j = 0; j < i; j++) s -= A[(size_t)i * n + j] * b[j];
    b[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) { /* U x = y */
    double s = b[i];
    for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * b[j];
    b[i] = s / A[(size_t)i * n + i];
  }
  return 0;
}

int or_lu_solve(double *A, int n, double *b) { return lu_solve(A, n, b); }

/* ------------------------------------------------------------------ */
/* MATLAB linspace(d1,d2,n): d1 + ((0:n-1).*(d2-d1))./(n-1), endpoints exact */
static void linspace_m(double d1, double d2, int n, double *y) {
  int n1 = n - 1;
  for (int k = 0; k < n; k++) y[k] = d1 + ((double)k * (d2 - d1)) / (double)n1;
  y[0] = d1;
  y[n1] = d2;
}

int or_surface_count(int res, int shells) {
  int inner = res - 2 * shells;
  if (inner < 0) inner = 0;
  return res * res * res - inner * inner * inner;
}

/* genupsurf.m / gendownsurf.m: meshgrid(x,y,z) is (ny,nx,nz) with X varying
 * along dim 2; the kept points are taken column-major: y fastest, then x,
 * then z.  down != 0 widens the box by its width on every side.          */
int or_surface(const double *corners, int res, int shells, int down,
               double *out) {
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) { lo[d] = corners[d]; hi[d] = corners[3 + d]; }
  if (down) {
    for (int d = 0; d < 3; d++) {
      double w = hi[d] - lo[d];
      lo[d] = lo[d] - w;
      hi[d] = hi[d] + w;
    }
  }
  double xs[64], ys[64], zs[64];
  if (res > 64 || res < 2) { snprintf(g_err, sizeof g_err, "bad res"); return OR_EARG; }
  linspace_m(lo[0], hi[0], res, xs);
  linspace_m(lo[1], hi[1], res, ys);
  linspace_m(lo[2], hi[2], res, zs);
  int c = 0;
  for (int iz = 0; iz < res; iz++)
    for (int ix = 0; ix < res; ix++)
      for (int iy = 0; iy < res; iy++) {
        int inside = iy >= shells && iy < res - shells && ix >= shells &&
                     ix < res - shells && iz >= shells && iz < res - shells;
        if (inside) continue;
        out[3 * c] = xs[ix]; out[3 * c + 1] = ys[iy]; out[3 * c + 2] = zs[iz];
        c++;
      }
  return c;
}

/* ------------------------------------------------------------------ */
/* Octree (Classes/OcTree.m).  Node ids are 1-based as in the reference;
 * arrays are stored 0-based at [id-1].                                   */
typedef struct {
  int32_t n, cap;
  int32_t *ptr, *idx; /* CSR over nodes, 1-based ids */
} ilists;

typedef struct {
  int64_t npts, npot, ntgt;
  double *points;     /* npts x 3 */
  uint8_t *is_pot, *is_tgt;
  int32_t *pidx;      /* node id per point */
  int32_t count, alloc, cap, max_depth;
  double *corners;    /* 6 per node */
  int32_t *level, *parent, *children; /* children: 8 per node, 0 = none */
  ilists L[4];        /* U, V, W, X */
  /* nearest mode */
  int nearest;
  int64_t nfine;
  double *fine;               /* nfine x 3 */
  int64_t *nn_colptr, *nn_row; /* CSC over potential points (columns) */
  double *nn_val;
  /* derived: potential/target ordinal of each point (-1 if not) */
  int64_t *pot_ord, *tgt_ord;
} otree;

static int grow_nodes(otree *t, int need) {
  if (need <= t->alloc) return 0;
  int na = t->alloc ? t->alloc : 64;
  while (na < need) na *= 2;
  t->corners = realloc(t->corners, sizeof(double) * 6 * na);
  t->level = realloc(t->level, sizeof(int32_t) * na);
  t->parent = realloc(t->parent, sizeof(int32_t) * na);
  t->children = realloc(t->children, sizeof(int32_t) * 8 * na);
  if (!t->corners || !t->level || !t->parent || !t->children) return -1;
  t->alloc = na;
  return 0;
}

/* Per-node point membership for the recursive build. */
typedef struct { int64_t *pts; int64_t n; } plist;

static plist *g_plists; /* indexed by node id-1 during build */
static int g_plists_alloc;

static int dividenode(otree *t, int node) {
  /* OcTree.m:191-221 */
  int base = t->count;
  if (grow_nodes(t, base + 8)) return OR_ENOMEM;
  if (g_plists_alloc < base + 8) {
    int na = g_plists_alloc ? g_plists_alloc : 64;
    while (na < base + 8) na *= 2;
    g_plists = realloc(g_plists, sizeof(plist) * na);
    memset(g_plists + g_plists_alloc, 0, sizeof(plist) * (na - g_plists_alloc));
    g_plists_alloc = na;
  }
  const double *c = t->corners + 6 * (node - 1);
  double lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[3], c[4], c[5]}, ce[3];
  for (int d = 0; d < 3; d++) ce[d] = (lo[d] + hi[d]) / 2.0; /* mean([lo;hi]) */
  double co[9] = {lo[0], lo[1], lo[2], ce[0], ce[1], ce[2], hi[0], hi[1], hi[2]};
  static const int rows[8][6] = {
      {1, 2, 3, 4, 5, 6}, {1, 2, 6, 4, 5, 9}, {1, 5, 3, 4, 8, 6}, {1, 5, 6, 4, 8, 9},
      {4, 2, 3, 7, 5, 6}, {4, 2, 6, 7, 5, 9}, {4, 5, 3, 7, 8, 6}, {4, 5, 6, 7, 8, 9}};
  for (int k = 0; k < 8; k++) {
    int id = base + k + 1;
    for (int e = 0; e < 6; e++) t->corners[6 * (id - 1) + e] = co[rows[k][e] - 1];
    t->level[id - 1] = t->level[node - 1] + 1;
    t->parent[id - 1] = node;
    for (int e = 0; e < 8; e++) t->children[8 * (id - 1) + e] = 0;
    t->children[8 * (node - 1) + k] = id;
  }
  plist *pl = &g_plists[node - 1];
  int64_t cnt[8] = {0};
  uint8_t *loc = malloc(pl->n ? pl->n : 1);
  for (int64_t i = 0; i < pl->n; i++) {
    const double *p = t->points + 3 * pl->pts[i];
    int k = 4 * (p[0] > ce[0]) + 2 * (p[1] > ce[1]) + (p[2] > ce[2]);
    loc[i] = (uint8_t)k;
    cnt[k]++;
  }
  for (int k = 0; k < 8; k++) {
    plist *cp = &g_plists[base + k];
    cp->pts = malloc(sizeof(int64_t) * (cnt[k] ? cnt[k] : 1));
    cp->n = 0;
  }
  for (int64_t i = 0; i < pl->n; i++) {
    int k = loc[i];
    plist *cp = &g_plists[base + k];
    cp->pts[cp->n++] = pl->pts[i];
    t->pidx[pl->pts[i]] = base + k + 1;
  }
  free(loc);
  free(pl->pts);
  pl->pts = NULL;
  pl->n = 0;
  t->count = base + 8;
  return OR_OK;
}

static int divide(otree *t, const int *nodes, int nn) {
  /* OcTree.m:165-178: visit in order; recurse into the 8 new children. */
  for (int i = 0; i < nn; i++) {
    int node = nodes[i];
    if (t->level[node - 1] + 1 >= t->max_depth) {
      snprintf(g_err, sizeof g_err, "Max level reached: Increase Max Depth");
      return OR_EMAXDEPTH;
    }
    int old = t->count;
    if (g_plists[node - 1].n > t->cap) {
      int rc = dividenode(t, node);
      if (rc) return rc;
      int kids[8];
      for (int k = 0; k < 8; k++) kids[k] = old + k + 1;
      rc = divide(t, kids, 8);
      if (rc) return rc;
    }
  }
  return OR_OK;
}

/* closed-box intersection, Tools/Tree/nodeintersection.m:17-22 */
static inline int isect(const otree *t, int a, int b) {
  const double *c1 = t->corners + 6 * (a - 1), *c2 = t->corners + 6 * (b - 1);
  return !(c2[0] > c1[3] || c2[3] < c1[0] || c2[1] > c1[4] || c2[4] < c1[1] ||
           c2[2] > c1[5] || c2[5] < c1[2]);
}

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* sort + unique in place, returns new length (MATLAB setdiff/unique order) */
static int sort_unique(int32_t *v, int n) {
  if (n <= 1) return n;
  qsort(v, n, sizeof(int32_t), cmp_i32);
  int m = 1;
  for (int i = 1; i < n; i++)
    if (v[i] != v[m - 1]) v[m++] = v[i];
  return m;
}

/* dynamic int vector */
typedef struct { int32_t *v; int n, a; } ivec;
static void iv_push(ivec *x, int32_t val) {
  if (x->n == x->a) { x->a = x->a ? 2 * x->a : 16; x->v = realloc(x->v, sizeof(int32_t) * x->a); }
  x->v[x->n++] = val;
}

static int in_sorted(const int32_t *v, int n, int32_t key) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int m = (lo + hi) / 2;
    if (v[m] == key) return 1;
    if (v[m] < key) lo = m + 1; else hi = m - 1;
  }
  return 0;
}

static inline int is_leaf_all(const otree *t, int n) { return t->children[8 * (n - 1)] == 0; }

/* getChildren.m:10-17 as a set: if the list has no non-leaf -> empty;
 * otherwise its leaves plus every descendant of its non-leaves.        */
static void get_children(const otree *t, const int32_t *idx, int n, ivec *out) {
  int has_nonleaf = 0;
  for (int i = 0; i < n; i++) if (!is_leaf_all(t, idx[i])) { has_nonleaf = 1; break; }
  if (!has_nonleaf) return;
  for (int i = 0; i < n; i++) if (is_leaf_all(t, idx[i])) iv_push(out, idx[i]);
  ivec kids = {0};
  /* reshape(nodeChildren(nonLeaf,:),[],1) is column-major: child 1 of all,
   * then child 2 of all, ...  (order is irrelevant after setdiff).      */
  for (int k = 0; k < 8; k++)
    for (int i = 0; i < n; i++)
      if (!is_leaf_all(t, idx[i])) iv_push(&kids, t->children[8 * (idx[i] - 1) + k]);
  for (int i = 0; i < kids.n; i++) iv_push(out, kids.v[i]);
  get_children(t, kids.v, kids.n, out);
  free(kids.v);
}

static void build_lists(otree *t) {
  /* OcTree.m:265-366 computed from the set definitions, node by node. */
  int nc = t->count;
  uint8_t *tleaf = calloc(nc + 1, 1), *pleaf = calloc(nc + 1, 1);
  for (int64_t i = 0; i < t->npts; i++) {
    if (t->is_tgt[i]) tleaf[t->pidx[i]] = 1;
    if (t->is_pot[i]) pleaf[t->pidx[i]] = 1;
  }
  int maxlev = 0;
  for (int n = 1; n <= nc; n++) if (t->level[n - 1] > maxlev) maxlev = t->level[n - 1];
  ivec *U = calloc(nc + 1, sizeof(ivec)), *V = calloc(nc + 1, sizeof(ivec)),
       *W = calloc(nc + 1, sizeof(ivec));
  /* nodes per level in id order for the W coarsening loop */
  int *lev_ptr = calloc(maxlev + 2, sizeof(int)), *lev_nodes = malloc(sizeof(int) * nc);
  for (int n = 1; n <= nc; n++) lev_ptr[t->level[n - 1] + 1]++;
  for (int l = 0; l <= maxlev; l++) lev_ptr[l + 1] += lev_ptr[l];
  {
    int *fill = calloc(maxlev + 1, sizeof(int));
    for (int n = 1; n <= nc; n++) {
      int l = t->level[n - 1];
      lev_nodes[lev_ptr[l] + fill[l]++] = n;
    }
    free(fill);
  }
#pragma omp parallel
  {
    uint8_t *mark = calloc(nc + 1, 1);
    int32_t *tmp = malloc(sizeof(int32_t) * (nc + 1));
#pragma omp for schedule(dynamic, 16)
    for (int i = 2; i <= nc; i++) {
      int li = t->level[i - 1];
      /* U: OcTree.m:323-327 */
      if (tleaf[i])
        for (int n = 1; n <= nc; n++)
          if (isect(t, i, n) && pleaf[n]) iv_push(&U[i], n);
      /* V: OcTree.m:330-339 */
      {
        int P = t->parent[i - 1];
        int m = 0;
        for (int n = 1; n <= nc; n++)
          if (isect(t, P, n) && li - 1 >= t->level[n - 1]) tmp[m++] = n;
        /* setdiff(S, [parents(S) P]) */
        for (int k = 0; k < m; k++) { int pp = t->parent[tmp[k] - 1]; if (pp > 0) mark[pp] = 1; }
        mark[P] = 1;
        ivec keep = {0};
        for (int k = 0; k < m; k++) if (!mark[tmp[k]]) iv_push(&keep, tmp[k]);
        for (int k = 0; k < m; k++) { int pp = t->parent[tmp[k] - 1]; if (pp > 0) mark[pp] = 0; }
        mark[P] = 0;
        ivec ch = {0};
        for (int k = 0; k < 8; k++)
          for (int j = 0; j < keep.n; j++)
            if (!is_leaf_all(t, keep.v[j])) {
              int c = t->children[8 * (keep.v[j] - 1) + k];
              if (!isect(t, i, c)) iv_push(&ch, c);
            }
        int nv = sort_unique(ch.v, ch.n);
        for (int j = 0; j < nv; j++)
          if (!in_sorted(U[i].v, U[i].n, ch.v[j])) iv_push(&V[i], ch.v[j]);
        free(keep.v); free(ch.v);
      }
      /* W: OcTree.m:343-360 */
      if (tleaf[i]) {
        int m = 0;
        for (int n = 1; n <= nc; n++)
          if (isect(t, i, n) && li >= t->level[n - 1]) tmp[m++] = n;
        for (int k = 0; k < m; k++) { int pp = t->parent[tmp[k] - 1]; if (pp > 0) mark[pp] = 1; }
        uint8_t *drop = malloc(m ? m : 1);
        for (int k = 0; k < m; k++) drop[k] = mark[tmp[k]];
        for (int k = 0; k < m; k++) { int pp = t->parent[tmp[k] - 1]; if (pp > 0) mark[pp] = 0; }
        int mm = 0;
        for (int k = 0; k < m; k++) if (!drop[k]) tmp[mm++] = tmp[k];
        free(drop);
        /* tmp[0..mm) is sorted (ascending scan) */
        ivec gc = {0};
        get_children(t, tmp, mm, &gc);
        ivec w = {0};
        for (int k = 0; k < gc.n; k++) if (pleaf[gc.v[k]]) iv_push(&w, gc.v[k]);
        free(gc.v);
        int nw = sort_unique(w.v, w.n);
        int o = 0;
        for (int k = 0; k < nw; k++) if (!in_sorted(U[i].v, U[i].n, w.v[k])) w.v[o++] = w.v[k];
        w.n = o;
        /* coarsening, levels max..1, nodes in id order */
        for (int k = 0; k < w.n; k++) mark[w.v[k]] = 1;
        for (int l = maxlev; l >= 1; l--) {
          for (int q = lev_ptr[l]; q < lev_ptr[l + 1]; q++) {
            int p = lev_nodes[q];
            if (is_leaf_all(t, p)) continue; /* children are zeros */
            int all = 1;
            for (int k = 0; k < 8; k++) if (!mark[t->children[8 * (p - 1) + k]]) { all = 0; break; }
            if (!all) continue;
            /* W = [setdiff(W, children); p] */
            for (int k = 0; k < 8; k++) mark[t->children[8 * (p - 1) + k]] = 0;
            int o2 = 0;
            for (int k = 0; k < w.n; k++) if (mark[w.v[k]]) w.v[o2++] = w.v[k];
            w.n = o2;
            w.n = sort_unique(w.v, w.n);
            iv_push(&w, p);
            mark[p] = 1;
          }
        }
        for (int k = 0; k < w.n; k++) mark[w.v[k]] = 0;
        W[i] = w;
      }
    }
    free(mark); free(tmp);
  }
  /* X from find(Wpadded == i): column (position) outer, row (node) inner */
  int maxw = 0;
  for (int n = 1; n <= nc; n++) if (W[n].n > maxw) maxw = W[n].n;
  ivec *X = calloc(nc + 1, sizeof(ivec));
  for (int k = 0; k < maxw; k++)
    for (int j = 1; j <= nc; j++)
      if (W[j].n > k) iv_push(&X[W[j].v[k]], j);
  ivec *all[4] = {U, V, W, X};
  for (int q = 0; q < 4; q++) {
    ilists *L = &t->L[q];
    L->n = nc;
    L->ptr = calloc(nc + 2, sizeof(int32_t));
    for (int n = 1; n <= nc; n++) L->ptr[n + 1] = L->ptr[n] + (n >= 2 || q == 3 ? all[q][n].n : 0);
    L->idx = malloc(sizeof(int32_t) * (L->ptr[nc + 1] + 1));
    for (int n = 1; n <= nc; n++)
      if (n >= 2 || q == 3)
        memcpy(L->idx + L->ptr[n], all[q][n].v, sizeof(int32_t) * all[q][n].n);
    for (int n = 0; n <= nc; n++) free(all[q][n].v);
    free(all[q]);
  }
  free(lev_ptr); free(lev_nodes); free(tleaf); free(pleaf);
}

/* unique([pot;tgt],'rows','stable') + ismember masks, OcTree.m:80-82 */
static int row_eq(const double *a, const double *b) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

typedef struct { uint64_t h; int64_t i; } hent;
static uint64_t hrow(const double *p) {
  uint64_t h = 1469598103934665603ull;
  for (int d = 0; d < 3; d++) {
    double v = p[d] == 0.0 ? 0.0 : p[d]; /* -0 == 0 */
    uint64_t b;
    memcpy(&b, &v, 8);
    h ^= b + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  }
  return h;
}

typedef struct { int64_t *slot; int64_t mask; const double *base; } htab;
static void ht_init(htab *h, int64_t n, const double *base) {
  int64_t cap = 16;
  while (cap < 2 * n + 16) cap *= 2;
  h->slot = malloc(sizeof(int64_t) * cap);
  for (int64_t i = 0; i < cap; i++) h->slot[i] = -1;
  h->mask = cap - 1;
  h->base = base;
}
/* returns existing row index equal to p, or inserts idx and returns -1 */
static int64_t ht_find_or_insert(htab *h, const double *p, int64_t idx, int insert) {
  int64_t s = (int64_t)(hrow(p) & (uint64_t)h->mask);
  while (h->slot[s] >= 0) {
    if (row_eq(h->base + 3 * h->slot[s], p)) return h->slot[s];
    s = (s + 1) & h->mask;
  }
  if (insert) h->slot[s] = idx;
  return -1;
}

void or_tree_free(otree *t);

/* cube: NULL for OcTree.m's root [m m m M M M] (min/max over all
   coordinates, OcTree.m:97-98), or a fixed root [cube[0] x3, cube[1] x3]
   that must contain every point (an extension for time stepping, where a
   fixed root keeps the level widths, and so the operators, unchanged). */
int or_tree_build_cube(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
                       int cap, int max_depth, const double *cube, otree **out);

int or_tree_build(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
                  int cap, int max_depth, otree **out) {
  return or_tree_build_cube(pot, npot, tgt, ntgt, cap, max_depth, NULL, out);
}

int or_tree_build_cube(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
                       int cap, int max_depth, const double *cube, otree **out) {
  *out = NULL;
  if (npot + ntgt < 2) { snprintf(g_err, sizeof g_err, "need at least two points"); return OR_EARG; }
  otree *t = calloc(1, sizeof(otree));
  int64_t ntot = npot + ntgt;
  t->points = malloc(sizeof(double) * 3 * ntot);
  htab h;
  ht_init(&h, ntot, t->points);
  int64_t np = 0;
  for (int64_t i = 0; i < ntot; i++) {
    const double *p = i < npot ? pot + 3 * i : tgt + 3 * (i - npot);
    memcpy(t->points + 3 * np, p, 24);
    if (ht_find_or_insert(&h, t->points + 3 * np, np, 1) < 0) np++;
  }
  t->npts = np;
  t->is_pot = calloc(np, 1);
  t->is_tgt = calloc(np, 1);
  for (int64_t i = 0; i < npot; i++) {
    int64_t r = ht_find_or_insert(&h, pot + 3 * i, -1, 0);
    t->is_pot[r] = 1;
  }
  for (int64_t i = 0; i < ntgt; i++) {
    int64_t r = ht_find_or_insert(&h, tgt + 3 * i, -1, 0);
    t->is_tgt[r] = 1;
  }
  free(h.slot);
  t->npot = t->ntgt = 0;
  t->pot_ord = malloc(sizeof(int64_t) * np);
  t->tgt_ord = malloc(sizeof(int64_t) * np);
  for (int64_t i = 0; i < np; i++) {
    t->pot_ord[i] = t->is_pot[i] ? t->npot++ : -1;
    t->tgt_ord[i] = t->is_tgt[i] ? t->ntgt++ : -1;
  }
  double m = t->points[0], M = t->points[0];
  for (int64_t i = 0; i < 3 * np; i++) {
    if (t->points[i] < m) m = t->points[i];
    if (t->points[i] > M) M = t->points[i];
  }
  if (cube) {
    if (!(cube[0] <= m && M <= cube[1])) {
      snprintf(g_err, sizeof g_err, "root cube does not contain every point");
      or_tree_free(t);
      return OR_EARG;
    }
    m = cube[0];
    M = cube[1];
  }
  t->max_depth = max_depth;
  t->cap = cap < np - 1 ? cap : (int32_t)(np - 1);
  grow_nodes(t, 64);
  t->count = 1;
  double rc[6] = {m, m, m, M, M, M};
  memcpy(t->corners, rc, sizeof rc);
  t->level[0] = 0;
  t->parent[0] = 0;
  for (int k = 0; k < 8; k++) t->children[k] = 0;
  t->pidx = malloc(sizeof(int32_t) * np);
  for (int64_t i = 0; i < np; i++) t->pidx[i] = 1;
  g_plists_alloc = 64;
  g_plists = calloc(g_plists_alloc, sizeof(plist));
  g_plists[0].pts = malloc(sizeof(int64_t) * np);
  g_plists[0].n = np;
  for (int64_t i = 0; i < np; i++) g_plists[0].pts[i] = i;
  int root = 1;
  int rcode = divide(t, &root, 1);
  for (int i = 0; i < g_plists_alloc; i++) free(g_plists[i].pts);
  free(g_plists);
  g_plists = NULL;
  g_plists_alloc = 0;
  if (rcode) { or_tree_free(t); return rcode; }
  build_lists(t);
  *out = t;
  return OR_OK;
}

void or_tree_free(otree *t) {
  if (!t) return;
  free(t->points); free(t->is_pot); free(t->is_tgt); free(t->pidx);
  free(t->corners); free(t->level); free(t->parent); free(t->children);
  for (int q = 0; q < 4; q++) { free(t->L[q].ptr); free(t->L[q].idx); }
  free(t->fine); free(t->nn_colptr); free(t->nn_row); free(t->nn_val);
  free(t->pot_ord); free(t->tgt_ord);
  free(t);
}

/* accessors for the ctypes wrapper */
int64_t or_tree_npts(const otree *t) { return t->npts; }
int64_t or_tree_npot(const otree *t) { return t->npot; }
int64_t or_tree_ntgt(const otree *t) { return t->ntgt; }
int32_t or_tree_count(const otree *t) { return t->count; }
int32_t or_tree_cap(const otree *t) { return t->cap; }
void or_tree_export(const otree *t, double *points, uint8_t *is_pot, uint8_t *is_tgt,
                    int32_t *pidx, double *corners, int32_t *level, int32_t *parent,
                    int32_t *children) {
  memcpy(points, t->points, sizeof(double) * 3 * t->npts);
  memcpy(is_pot, t->is_pot, t->npts);
  memcpy(is_tgt, t->is_tgt, t->npts);
  memcpy(pidx, t->pidx, sizeof(int32_t) * t->npts);
  memcpy(corners, t->corners, sizeof(double) * 6 * t->count);
  memcpy(level, t->level, sizeof(int32_t) * t->count);
  memcpy(parent, t->parent, sizeof(int32_t) * t->count);
  memcpy(children, t->children, sizeof(int32_t) * 8 * t->count);
}
int32_t or_tree_list_total(const otree *t, int q) { return t->L[q].ptr[t->count + 1]; }
void or_tree_list_export(const otree *t, int q, int32_t *ptr, int32_t *idx) {
  /* ptr has count+1 entries: list of node n (1-based) is idx[ptr[n-1]:ptr[n]] */
  for (int n = 0; n <= t->count; n++) ptr[n] = t->L[q].ptr[n + 1];
  memcpy(idx, t->L[q].idx, sizeof(int32_t) * t->L[q].ptr[t->count + 1]);
}

/* attach NEAREST data: fine points and NN (Q x Npot) in CSC by columns */
int or_tree_set_nearest(otree *t, const double *fine, int64_t nfine,
                        const int64_t *colptr, const int64_t *row, const double *val) {
  t->nearest = 1;
  t->nfine = nfine;
  t->fine = malloc(sizeof(double) * 3 * (nfine ? nfine : 1));
  memcpy(t->fine, fine, sizeof(double) * 3 * nfine);
  int64_t nnz = colptr[t->npot];
  t->nn_colptr = malloc(sizeof(int64_t) * (t->npot + 1));
  memcpy(t->nn_colptr, colptr, sizeof(int64_t) * (t->npot + 1));
  t->nn_row = malloc(sizeof(int64_t) * (nnz ? nnz : 1));
  t->nn_val = malloc(sizeof(double) * (nnz ? nnz : 1));
  memcpy(t->nn_row, row, sizeof(int64_t) * nnz);
  memcpy(t->nn_val, val, sizeof(double) * nnz);
  return OR_OK;
}

/* ------------------------------------------------------------------ */
/* nearestMatrix.m:17-41: weight 1/#ties to every coarse point at the
 * minimum squared distance (summed x, then y, then z).  Output COO sorted
 * by (column, row) like MATLAB's sparse().                              */
int64_t or_nearest_matrix(const double *coarse, int64_t N, const double *fine,
                          int64_t Q, int64_t *rows, int64_t *cols, double *w,
                          int64_t maxnnz) {
  int64_t *cnt = calloc(Q, sizeof(int64_t));
  double *dmin = malloc(sizeof(double) * Q);
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < Q; q++) {
    const double *f = fine + 3 * q;
    double best = INFINITY;
    int64_t c = 0;
    for (int64_t j = 0; j < N; j++) {
      const double *p = coarse + 3 * j;
      double dx = f[0] - p[0], dy = f[1] - p[1], dz = f[2] - p[2];
      double d2 = (dx * dx + dy * dy) + dz * dz;
      if (d2 < best) { best = d2; c = 1; }
      else if (d2 == best) c++;
    }
    dmin[q] = best;
    cnt[q] = c;
  }
  int64_t total = 0;
  for (int64_t q = 0; q < Q; q++) total += cnt[q];
  if (total > maxnnz) { free(cnt); free(dmin); return -total; }
  /* emit column-major: for each coarse column j, rows ascending */
  int64_t k = 0;
  for (int64_t j = 0; j < N; j++) {
    const double *p = coarse + 3 * j;
    for (int64_t q = 0; q < Q; q++) {
      const double *f = fine + 3 * q;
      double dx = f[0] - p[0], dy = f[1] - p[1], dz = f[2] - p[2];
      double d2 = (dx * dx + dy * dy) + dz * dz;
      if (d2 == dmin[q]) { rows[k] = q; cols[k] = j; w[k] = 1.0 / (double)cnt[q]; k++; }
    }
  }
  free(cnt); free(dmin);
  return k;
}

/* ------------------------------------------------------------------ */
/* Per-call membership tables: potential / target point rows per node, and
 * for NEAREST mode the fine-point instances of every potential leaf.
 * Sources of one potential leaf are the potential points it holds, or, in
 * NEAREST mode, the fine points with a nonzero NN weight on one of its
 * coarse points, each carrying (NN_leaf (x) I3) f  (upwardPass.m:58-69,
 * downwardPass.m:84-116,213-246).                                        */
typedef struct { int64_t n; double *pos; double *q; } srcset;

typedef struct {
  int64_t *pot_ptr, *pot_rows; /* CSR over node ids (1-based, ptr[n]..ptr[n+1]) */
  int64_t *tgt_ptr, *tgt_rows;
  /* NEAREST: per node, instances [ins_ptr[n], ins_ptr[n+1]) with fine row
   * ins_fine[k] and contributions [ctr_ptr[k], ctr_ptr[k+1]) of
   * (potential ordinal ctr_col, weight ctr_w).                           */
  int64_t *ins_ptr, *ins_fine, *ctr_ptr, *ctr_col;
  double *ctr_w;
} prep_t;

typedef struct { int64_t row, col; double w; } trip;
static int cmp_trip(const void *a, const void *b) {
  const trip *x = a, *y = b;
  if (x->row != y->row) return (x->row > y->row) - (x->row < y->row);
  return (x->col > y->col) - (x->col < y->col);
}

static void prep_build(const otree *t, prep_t *p) {
  int nc = t->count;
  memset(p, 0, sizeof *p);
  p->pot_ptr = calloc(nc + 2, sizeof(int64_t));
  p->tgt_ptr = calloc(nc + 2, sizeof(int64_t));
  for (int64_t i = 0; i < t->npts; i++) {
    if (t->is_pot[i]) p->pot_ptr[t->pidx[i] + 1]++;
    if (t->is_tgt[i]) p->tgt_ptr[t->pidx[i] + 1]++;
  }
  for (int n = 1; n <= nc; n++) { p->pot_ptr[n + 1] += p->pot_ptr[n]; p->tgt_ptr[n + 1] += p->tgt_ptr[n]; }
  p->pot_rows = malloc(sizeof(int64_t) * (p->pot_ptr[nc + 1] + 1));
  p->tgt_rows = malloc(sizeof(int64_t) * (p->tgt_ptr[nc + 1] + 1));
  int64_t *fp = calloc(nc + 2, sizeof(int64_t)), *ft = calloc(nc + 2, sizeof(int64_t));
  for (int64_t i = 0; i < t->npts; i++) {
    int n = t->pidx[i];
    if (t->is_pot[i]) p->pot_rows[p->pot_ptr[n] + fp[n]++] = i;
    if (t->is_tgt[i]) p->tgt_rows[p->tgt_ptr[n] + ft[n]++] = i;
  }
  free(fp); free(ft);
  if (!t->nearest) return;
  p->ins_ptr = calloc(nc + 2, sizeof(int64_t));
  /* first pass: count instances per node */
  int64_t total_ins = 0, total_ctr = 0;
  trip **per = calloc(nc + 1, sizeof(trip *));
  int64_t *pern = calloc(nc + 1, sizeof(int64_t));
  for (int n = 1; n <= nc; n++) {
    int64_t m = 0;
    for (int64_t k = p->pot_ptr[n]; k < p->pot_ptr[n + 1]; k++) {
      int64_t c = t->pot_ord[p->pot_rows[k]];
      m += t->nn_colptr[c + 1] - t->nn_colptr[c];
    }
    if (!m) continue;
    trip *tr = malloc(sizeof(trip) * m);
    m = 0;
    for (int64_t k = p->pot_ptr[n]; k < p->pot_ptr[n + 1]; k++) {
      int64_t c = t->pot_ord[p->pot_rows[k]];
      for (int64_t e = t->nn_colptr[c]; e < t->nn_colptr[c + 1]; e++)
        if (t->nn_val[e] != 0.0) { tr[m].row = t->nn_row[e]; tr[m].col = c; tr[m].w = t->nn_val[e]; m++; }
    }
    qsort(tr, m, sizeof(trip), cmp_trip);
    int64_t ni = 0;
    for (int64_t e = 0; e < m; e++) if (e == 0 || tr[e].row != tr[e - 1].row) ni++;
    p->ins_ptr[n + 1] = ni;
    per[n] = tr; pern[n] = m;
    total_ins += ni; total_ctr += m;
  }
  for (int n = 1; n <= nc; n++) p->ins_ptr[n + 1] += p->ins_ptr[n];
  p->ins_fine = malloc(sizeof(int64_t) * (total_ins + 1));
  p->ctr_ptr = malloc(sizeof(int64_t) * (total_ins + 2));
  p->ctr_col = malloc(sizeof(int64_t) * (total_ctr + 1));
  p->ctr_w = malloc(sizeof(double) * (total_ctr + 1));
  int64_t ki = 0, kc = 0;
  p->ctr_ptr[0] = 0;
  for (int n = 1; n <= nc; n++) {
    trip *tr = per[n];
    for (int64_t e = 0; e < pern[n]; e++) {
      if (e == 0 || tr[e].row != tr[e - 1].row) {
        if (e) p->ctr_ptr[ki] = kc;
        p->ins_fine[ki++] = tr[e].row;
      }
      p->ctr_col[kc] = tr[e].col; p->ctr_w[kc] = tr[e].w; kc++;
    }
    if (pern[n]) p->ctr_ptr[ki] = kc;
    free(tr);
  }
  free(per); free(pern);
}

static void prep_free(prep_t *p) {
  free(p->pot_ptr); free(p->pot_rows); free(p->tgt_ptr); free(p->tgt_rows);
  free(p->ins_ptr); free(p->ins_fine); free(p->ctr_ptr); free(p->ctr_col); free(p->ctr_w);
}

static void leaf_sources(const otree *t, const prep_t *p, int node, const double *pot_q,
                         srcset *s) {
  if (!t->nearest) {
    int64_t b = p->pot_ptr[node], nc = p->pot_ptr[node + 1] - b;
    s->n = nc;
    s->pos = malloc(sizeof(double) * 3 * (nc ? nc : 1));
    s->q = malloc(sizeof(double) * 3 * (nc ? nc : 1));
    for (int64_t k = 0; k < nc; k++) {
      int64_t r = p->pot_rows[b + k];
      memcpy(s->pos + 3 * k, t->points + 3 * r, 24);
      memcpy(s->q + 3 * k, pot_q + 3 * t->pot_ord[r], 24);
    }
  } else {
    int64_t b = p->ins_ptr[node], m = p->ins_ptr[node + 1] - b;
    s->n = m;
    s->pos = malloc(sizeof(double) * 3 * (m ? m : 1));
    s->q = malloc(sizeof(double) * 3 * (m ? m : 1));
    for (int64_t k = 0; k < m; k++) {
      int64_t ins = b + k;
      memcpy(s->pos + 3 * k, t->fine + 3 * p->ins_fine[ins], 24);
      double g[3] = {0, 0, 0};
      for (int64_t e = p->ctr_ptr[ins]; e < p->ctr_ptr[ins + 1]; e++)
        for (int d = 0; d < 3; d++) g[d] += p->ctr_w[e] * pot_q[3 * p->ctr_col[e] + d];
      memcpy(s->q + 3 * k, g, 24);
    }
  }
}

static void src_free(srcset *s) { free(s->pos); free(s->q); s->pos = s->q = NULL; s->n = 0; }

/* target rows of a leaf, gathered */
static int64_t leaf_targets(const otree *t, const prep_t *p, int n, double **tp, int64_t **to) {
  int64_t b = p->tgt_ptr[n], nt = p->tgt_ptr[n + 1] - b;
  *tp = malloc(sizeof(double) * 3 * (nt ? nt : 1));
  *to = malloc(sizeof(int64_t) * (nt ? nt : 1));
  for (int64_t k = 0; k < nt; k++) {
    int64_t r = p->tgt_rows[b + k];
    memcpy(*tp + 3 * k, t->points + 3 * r, 24);
    (*to)[k] = t->tgt_ord[r];
  }
  return nt;
}

/* KIFMM matvec: upwardPass then downwardPass.  pot_q: npot x 3 (the
 * reference's (N,3) potentials), vel: ntgt x 3 in points(targets) order.  */
int or_fmm_matvec(const otree *t, double eps, double mu, int res, int shells,
                  const double *pot_q, double *vel) {
  int nc = t->count;
  int nq = or_surface_count(res, shells);
  int nq3 = 3 * nq;
  prep_t P;
  prep_build(t, &P);
  uint8_t *pleaf = calloc(nc + 1, 1), *tleaf = calloc(nc + 1, 1);
  for (int64_t i = 0; i < t->npts; i++) {
    if (t->is_pot[i]) pleaf[t->pidx[i]] = 1;
    if (t->is_tgt[i]) tleaf[t->pidx[i]] = 1;
  }
  int maxlev = 0;
  for (int n = 1; n <= nc; n++) if (t->level[n - 1] > maxlev) maxlev = t->level[n - 1];
  double *uppot = calloc((size_t)nc * nq3, sizeof(double));
  double *downpot = calloc((size_t)nc * nq3, sizeof(double));
  /* ---- upwardPass.m:32-98 ---- */
  for (int l = maxlev; l >= 0; l--) {
#pragma omp parallel
    {
      double *up = malloc(sizeof(double) * nq3), *dn = malloc(sizeof(double) * nq3);
      double *A = malloc(sizeof(double) * (size_t)nq3 * nq3);
      double *rhs = malloc(sizeof(double) * nq3);
      double *csurf = malloc(sizeof(double) * 8 * nq3);
      double *cpot = malloc(sizeof(double) * 8 * nq3);
#pragma omp for schedule(dynamic, 1)
      for (int n = 1; n <= nc; n++) {
        if (t->level[n - 1] != l) continue;
        int leaf = is_leaf_all(t, n);
        if (!(pleaf[n] || !leaf)) continue;
        or_surface(t->corners + 6 * (n - 1), res, shells, 0, up);
        or_surface(t->corners + 6 * (n - 1), res, shells, 1, dn);
        if (pleaf[n]) {
          srcset s;
          leaf_sources(t, &P, n, pot_q, &s);
          gemv_kernel(s.pos, s.n, dn, nq, s.q, eps, mu, rhs, 0);
          src_free(&s);
        } else {
          for (int k = 0; k < 8; k++) {
            int c = t->children[8 * (n - 1) + k];
            or_surface(t->corners + 6 * (c - 1), res, shells, 0, csurf + (size_t)k * nq3);
            memcpy(cpot + (size_t)k * nq3, uppot + (size_t)(c - 1) * nq3, sizeof(double) * nq3);
          }
          gemv_kernel(csurf, 8 * nq, dn, nq, cpot, eps, mu, rhs, 0);
        }
        or_kernel(up, nq, dn, nq, eps, mu, A); /* K(up -> down) */
        lu_solve(A, nq3, rhs);
        memcpy(uppot + (size_t)(n - 1) * nq3, rhs, sizeof(double) * nq3);
      }
      free(up); free(dn); free(A); free(rhs); free(csurf); free(cpot);
    }
  }
  /* ---- downwardPass.m:39-168 ---- */
  const ilists *LV = &t->L[1], *LX = &t->L[3], *LU = &t->L[0], *LW = &t->L[2];
  for (int l = 1; l <= maxlev; l++) {
#pragma omp parallel
    {
      double *up = malloc(sizeof(double) * nq3), *dn = malloc(sizeof(double) * nq3);
      double *pdn = malloc(sizeof(double) * nq3);
      double *A = malloc(sizeof(double) * (size_t)nq3 * nq3);
      double *rhs = malloc(sizeof(double) * nq3);
      double *vs = malloc(sizeof(double) * nq3);
#pragma omp for schedule(dynamic, 1)
      for (int n = 1; n <= nc; n++) {
        if (t->level[n - 1] != l) continue;
        or_surface(t->corners + 6 * (n - 1), res, shells, 0, up);
        or_surface(t->corners + 6 * (n - 1), res, shells, 1, dn);
        memset(rhs, 0, sizeof(double) * nq3);
        /* V list: K(vsurf -> upsurf) vpot */
        for (int e = LV->ptr[n]; e < LV->ptr[n + 1]; e++) {
          int v = LV->idx[e];
          or_surface(t->corners + 6 * (v - 1), res, shells, 0, vs);
          gemv_kernel(vs, nq, up, nq, uppot + (size_t)(v - 1) * nq3, eps, mu, rhs, 1);
        }
        /* X list: K(xsrc -> upsurf) xpot */
        for (int e = LX->ptr[n]; e < LX->ptr[n + 1]; e++) {
          int x = LX->idx[e];
          if (!pleaf[x]) continue; /* a target-only leaf has no sources */
          srcset s;
          leaf_sources(t, &P, x, pot_q, &s);
          gemv_kernel(s.pos, s.n, up, nq, s.q, eps, mu, rhs, 1);
          src_free(&s);
        }
        /* parent: K(parent downsurf -> upsurf) parentpot */
        int Pn = t->parent[n - 1];
        or_surface(t->corners + 6 * (Pn - 1), res, shells, 1, pdn);
        gemv_kernel(pdn, nq, up, nq, downpot + (size_t)(Pn - 1) * nq3, eps, mu, rhs, 1);
        or_kernel(dn, nq, up, nq, eps, mu, A); /* K(down -> up) */
        lu_solve(A, nq3, rhs);
        memcpy(downpot + (size_t)(n - 1) * nq3, rhs, sizeof(double) * nq3);
      }
      free(up); free(dn); free(pdn); free(A); free(rhs); free(vs);
    }
  }
  /* ---- downwardPass.m:170-305: target leaves ---- */
#pragma omp parallel
  {
    double *dn = malloc(sizeof(double) * nq3), *ws = malloc(sizeof(double) * nq3);
#pragma omp for schedule(dynamic, 1)
    for (int n = 1; n <= nc; n++) {
      if (!tleaf[n]) continue;
      double *tp; int64_t *to;
      int64_t nt = leaf_targets(t, &P, n, &tp, &to);
      double *u = calloc(3 * nt, sizeof(double));
      or_surface(t->corners + 6 * (n - 1), res, shells, 1, dn);
      gemv_kernel(dn, nq, tp, nt, downpot + (size_t)(n - 1) * nq3, eps, mu, u, 1);
      for (int e = LU->ptr[n]; e < LU->ptr[n + 1]; e++) {
        srcset s;
        leaf_sources(t, &P, LU->idx[e], pot_q, &s);
        gemv_kernel(s.pos, s.n, tp, nt, s.q, eps, mu, u, 1);
        src_free(&s);
      }
      for (int e = LW->ptr[n]; e < LW->ptr[n + 1]; e++) {
        int w = LW->idx[e];
        or_surface(t->corners + 6 * (w - 1), res, shells, 0, ws);
        gemv_kernel(ws, nq, tp, nt, uppot + (size_t)(w - 1) * nq3, eps, mu, u, 1);
      }
      for (int64_t k = 0; k < nt; k++) memcpy(vel + 3 * to[k], u + 3 * k, 24);
      free(tp); free(to); free(u);
    }
    free(dn); free(ws);
  }
  free(uppot); free(downpot); free(pleaf); free(tleaf);
  prep_free(&P);
  return OR_OK;
}

/* reducedPass.m:21-80: each target leaf sees only its own sources. */
int or_reduced_apply(const otree *t, double eps, double mu, const double *pot_q,
                     double *vel) {
  int nc = t->count;
  prep_t P;
  prep_build(t, &P);
#pragma omp parallel for schedule(dynamic, 4)
  for (int n = 1; n <= nc; n++) {
    if (P.tgt_ptr[n + 1] == P.tgt_ptr[n]) continue;
    double *tp; int64_t *to;
    int64_t nt = leaf_targets(t, &P, n, &tp, &to);
    double *u = calloc(3 * nt, sizeof(double));
    srcset s;
    leaf_sources(t, &P, n, pot_q, &s);
    gemv_kernel(s.pos, s.n, tp, nt, s.q, eps, mu, u, 1);
    src_free(&s);
    for (int64_t k = 0; k < nt; k++) memcpy(vel + 3 * to[k], u + 3 * k, 24);
    free(tp); free(to); free(u);
  }
  prep_free(&P);
  return OR_OK;
}

/* ReducedSolve.m:164-209: per target leaf, solve the leaf block
 * K(leaf sources -> leaf targets) [x kron(NN,I3)] for the leaf's
 * densities.  Requires potentials == targets, as the reference does.    */
int or_reduced_solve(const otree *t, double eps, double mu, const double *vel_in,
                     double *pot_out) {
  int nc = t->count;
  if (t->npot != t->ntgt) { snprintf(g_err, sizeof g_err, "reducedSolve needs potentials == targets"); return OR_EARG; }
  prep_t P;
  prep_build(t, &P);
#pragma omp parallel for schedule(dynamic, 4)
  for (int n = 1; n <= nc; n++) {
    if (P.tgt_ptr[n + 1] == P.tgt_ptr[n]) continue;
    double *tp; int64_t *to;
    int64_t nt = leaf_targets(t, &P, n, &tp, &to);
    int64_t n3 = 3 * nt;
    double *A = calloc((size_t)n3 * n3, sizeof(double));
    if (!t->nearest) {
      or_kernel(tp, nt, tp, nt, eps, mu, A);
    } else {
      /* column by column from unit coarse forces on this leaf's points */
      double *e = calloc(3 * t->npot, sizeof(double));
      double *col = malloc(sizeof(double) * n3);
      for (int64_t k = 0; k < nt; k++)
        for (int d = 0; d < 3; d++) {
          e[3 * to[k] + d] = 1.0;
          srcset s;
          leaf_sources(t, &P, n, e, &s);
          memset(col, 0, sizeof(double) * n3);
          gemv_kernel(s.pos, s.n, tp, nt, s.q, eps, mu, col, 1);
          src_free(&s);
          for (int64_t r = 0; r < n3; r++) A[(size_t)r * n3 + 3 * k + d] = col[r];
          e[3 * to[k] + d] = 0.0;
        }
      free(e); free(col);
    }
    double *b = malloc(sizeof(double) * n3);
    for (int64_t k = 0; k < nt; k++) memcpy(b + 3 * k, vel_in + 3 * to[k], 24);
    lu_solve(A, (int)n3, b);
    for (int64_t k = 0; k < nt; k++) memcpy(pot_out + 3 * to[k], b + 3 * k, 24);
    free(A); free(b); free(tp); free(to);
  }
  prep_free(&P);
  return OR_OK;
}

int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int n) { omp_set_num_threads(n); }
