// extern "C" boundary: status codes and a thread-local error message; no
// exception crosses it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <string>

#include "../../include/nnkifmm.h"
#include "nnk_batch.hpp"
#include "nnk_fmm.hpp"
#include "nnk_handles.hpp"
#include "nnk_mobility.hpp"
#include "nnk_tree.hpp"


struct nnk_mobility {
  std::unique_ptr<nnk::Mobility> m;
};
struct nnk_batch {
  std::unique_ptr<nnk::Batch> b;
};

namespace nnk {
int64_t nearest_matrix(const double *coarse, int64_t n, const double *fine, int64_t q,
                       int64_t *colptr, int64_t *row, double *val, int64_t cap);
void direct_device(const double *src, int64_t ns, const double *q, const double *tgt, int64_t nt,
                   double eps, double mu, double *out, hipStream_t s);
void direct_device_noalloc(const double *src, int64_t ns, const double *q, const double *tgt, int64_t nt,
                           double eps, double mu, double *out, hipStream_t s);
}

namespace {
thread_local std::string g_err;

int32_t fail(int32_t code, const std::string &m) {
  g_err = m;
  return code;
}

template <class F> int32_t guard(F &&fn) {
  try {
    g_err.clear();
    return fn();
  } catch (const nnk::Error &e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc &) {
    return fail(NNK_ENOMEM, "out of host memory");
  } catch (const std::exception &e) {
    return fail(NNK_ESTATE, e.what());
  } catch (...) {
    return fail(NNK_ESTATE, "unknown failure");
  }
}

// host-pointer variant: stage through device buffers on the null stream
int32_t apply_host(nnk_fmm *h, const double *f, double *u, bool reduced) {
  nnk::DBuf df, du;
  df.alloc(sizeof(double) * 3 * std::max<int64_t>(h->npot, 1));
  du.alloc(sizeof(double) * 3 * std::max<int64_t>(h->ntgt, 1));
  NNK_HIP(hipMemcpy(df.p, f, sizeof(double) * 3 * h->npot, hipMemcpyHostToDevice));
  du.zero();
  if (reduced) h->f->reduced_apply(df.as<double>(), du.as<double>(), 0);
  else h->apply(df.as<double>(), du.as<double>(), 0);
  NNK_HIP(hipMemcpy(u, du.p, sizeof(double) * 3 * h->ntgt, hipMemcpyDeviceToHost));
  return NNK_OK;
}
}  // namespace

extern "C" {

int32_t nnk_version(void) { return 100; }

const char *nnk_last_error(void) { return g_err.c_str(); }

int32_t nnk_tree_create(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
                        int32_t node_capacity, int32_t max_depth, nnk_tree **out) {
  return nnk_tree_create_in_cube(pot, npot, tgt, ntgt, node_capacity, max_depth, nullptr, out);
}

int32_t nnk_tree_create_in_cube(const double *pot, int64_t npot, const double *tgt, int64_t ntgt,
                                int32_t node_capacity, int32_t max_depth, const double *cube, nnk_tree **out) {
  return guard([&]() -> int32_t {
    if (!out) return fail(NNK_EARG, "nnk_tree_create: null out");
    *out = nullptr;
    if ((npot && !pot) || (ntgt && !tgt)) return fail(NNK_EARG, "nnk_tree_create: null points");
    if (node_capacity < 1) return fail(NNK_EARG, "nodeCapacity must be >= 1");
    if (cube && !(cube[0] < cube[1])) return fail(NNK_EARG, "nnk_tree_create_in_cube: empty root cube");
    auto t = std::make_shared<nnk::Tree>();
    std::string err;
    int rc = nnk::build_tree(pot, npot, tgt, ntgt, node_capacity, max_depth, *t, err, cube);
    if (rc) return fail(rc == nnk::NNK_EMAXDEPTH ? NNK_EMAXDEPTH : NNK_EARG, err);
    *out = new nnk_tree{t};
    return NNK_OK;
  });
}

int32_t nnk_tree_set_nearest(nnk_tree *t, const double *fine, int64_t nfine,
                             const int64_t *colptr, const int64_t *row, const double *val) {
  return guard([&]() -> int32_t {
    if (!t || !colptr || (nfine && !fine)) return fail(NNK_EARG, "nnk_tree_set_nearest: null argument");
    std::string err;
    int rc = nnk::set_nearest(*t->t, fine, nfine, colptr, row, val, err);
    if (rc) return fail(NNK_EARG, err);
    return NNK_OK;
  });
}

int32_t nnk_tree_sizes(const nnk_tree *t, int64_t *npts, int64_t *npot, int64_t *ntgt,
                       int32_t *node_count, int32_t *cap) {
  if (!t) return fail(NNK_EARG, "nnk_tree_sizes: null tree");
  if (npts) *npts = t->t->npts;
  if (npot) *npot = t->t->npot;
  if (ntgt) *ntgt = t->t->ntgt;
  if (node_count) *node_count = t->t->count;
  if (cap) *cap = t->t->capacity;
  return NNK_OK;
}

int32_t nnk_tree_build_times(const nnk_tree *t, double *ms4) {
  if (!t || !ms4) return fail(NNK_EARG, "nnk_tree_build_times: null argument");
  for (int k = 0; k < 4; k++) ms4[k] = t->t->build_ms[k];
  return NNK_OK;
}

int32_t nnk_tree_export(const nnk_tree *h, double *points, uint8_t *is_pot, uint8_t *is_tgt,
                        int32_t *point_index, double *corners, int32_t *level, int32_t *parent,
                        int32_t *children) {
  if (!h) return fail(NNK_EARG, "nnk_tree_export: null tree");
  const nnk::Tree &t = *h->t;
  if (points) std::memcpy(points, t.points.data(), sizeof(double) * 3 * t.npts);
  if (is_pot) std::memcpy(is_pot, t.is_pot.data(), t.npts);
  if (is_tgt) std::memcpy(is_tgt, t.is_tgt.data(), t.npts);
  if (point_index) std::memcpy(point_index, t.point_index.data(), sizeof(int32_t) * t.npts);
  if (corners) std::memcpy(corners, t.corners.data(), sizeof(double) * 6 * t.count);
  if (level) std::memcpy(level, t.level.data(), sizeof(int32_t) * t.count);
  if (parent) std::memcpy(parent, t.parent.data(), sizeof(int32_t) * t.count);
  if (children) std::memcpy(children, t.children.data(), sizeof(int32_t) * 8 * t.count);
  return NNK_OK;
}

int32_t nnk_tree_list_size(const nnk_tree *h, int32_t kind, int64_t *total) {
  if (!h || kind < 0 || kind > 3 || !total) return fail(NNK_EARG, "nnk_tree_list_size: bad argument");
  *total = (int64_t)h->t->lidx[kind].size();
  return NNK_OK;
}

int32_t nnk_tree_list_export(const nnk_tree *h, int32_t kind, int64_t *ptr, int32_t *idx) {
  if (!h || kind < 0 || kind > 3 || !ptr || !idx) return fail(NNK_EARG, "nnk_tree_list_export: bad argument");
  const nnk::Tree &t = *h->t;
  for (int n = 0; n <= t.count; n++) ptr[n] = t.lptr[kind][n + 1];
  std::memcpy(idx, t.lidx[kind].data(), sizeof(int32_t) * t.lidx[kind].size());
  return NNK_OK;
}

void nnk_tree_destroy(nnk_tree *t) { delete t; }

int32_t nnk_fmm_create(const nnk_tree *t, double eps, double mu, int32_t res, int32_t shells,
                       int32_t precision, double svd_tol, nnk_fmm **out) {
  return guard([&]() -> int32_t {
    if (!t || !out) return fail(NNK_EARG, "nnk_fmm_create: null argument");
    *out = nullptr;
    if (precision != NNK_FP64 && precision != NNK_MIXED && precision != NNK_MIXED3 && precision != NNK_MIXED_H &&
        precision != NNK_FP64_T)
      return fail(NNK_EARG, "nnk_fmm_create: unknown precision");
    auto f = std::make_unique<nnk::Fmm>(t->t, eps, mu, res, shells, precision,
                                        svd_tol > 0 ? svd_tol : 1e-8);
    *out = new nnk_fmm{std::move(f), t->t->npot, t->t->ntgt};
    return NNK_OK;
  });
}

int32_t nnk_fmm_apply(nnk_fmm *h, const double *f, double *u) {
  return guard([&]() -> int32_t {
    if (!h || (!f && h->npot) || (!u && h->ntgt)) return fail(NNK_EARG, "nnk_fmm_apply: null argument");
    return apply_host(h, f, u, false);
  });
}

int32_t nnk_fmm_apply_device(nnk_fmm *h, const double *d_f, double *d_u, void *stream) {
  return guard([&]() -> int32_t {
    if (!h || (!d_f && h->npot) || (!d_u && h->ntgt)) return fail(NNK_EARG, "nnk_fmm_apply_device: null argument");
    h->apply(d_f, d_u, (hipStream_t)stream);
    return NNK_OK;
  });
}

int32_t nnk_fmm_reduced_apply(nnk_fmm *h, const double *f, double *u) {
  return guard([&]() -> int32_t {
    if (!h || (!f && h->npot) || (!u && h->ntgt)) return fail(NNK_EARG, "nnk_fmm_reduced_apply: null argument");
    return apply_host(h, f, u, true);
  });
}

int32_t nnk_fmm_reduced_apply_device(nnk_fmm *h, const double *d_f, double *d_u, void *stream) {
  return guard([&]() -> int32_t {
    if (!h || (!d_f && h->npot) || (!d_u && h->ntgt)) return fail(NNK_EARG, "nnk_fmm_reduced_apply_device: null argument");
    h->f->reduced_apply(d_f, d_u, (hipStream_t)stream);
    return NNK_OK;
  });
}

void nnk_fmm_destroy(nnk_fmm *h) { delete h; }

int32_t nnk_fmm_reduced_solve(nnk_fmm *h, const double *v, double *f) {
  return guard([&]() -> int32_t {
    if (!h || !v || !f) return fail(NNK_EARG, "nnk_fmm_reduced_solve: null argument");
    if (h->npot != h->ntgt) return fail(NNK_EARG, "reducedSolve needs potentials == targets");
    nnk::DBuf dv, df;
    dv.alloc(sizeof(double) * 3 * std::max<int64_t>(h->ntgt, 1));
    df.alloc(sizeof(double) * 3 * std::max<int64_t>(h->npot, 1));
    NNK_HIP(hipMemcpy(dv.p, v, sizeof(double) * 3 * h->ntgt, hipMemcpyHostToDevice));
    h->f->reduced_solve(dv.as<double>(), df.as<double>(), 0);
    NNK_HIP(hipMemcpy(f, df.p, sizeof(double) * 3 * h->npot, hipMemcpyDeviceToHost));
    return NNK_OK;
  });
}

int32_t nnk_fmm_reduced_solve_device(nnk_fmm *h, const double *d_v, double *d_f, void *stream) {
  return guard([&]() -> int32_t {
    if (!h || !d_v || !d_f) return fail(NNK_EARG, "nnk_fmm_reduced_solve_device: null argument");
    h->f->reduced_solve(d_v, d_f, (hipStream_t)stream);
    return NNK_OK;
  });
}

int32_t nnk_nearest_matrix(const double *coarse, int64_t n, const double *fine, int64_t q,
                           int64_t *colptr, int64_t *rowidx, double *val, int64_t cap,
                           int64_t *nnz) {
  return guard([&]() -> int32_t {
    if (!coarse || n <= 0 || q < 0 || (q && !fine) || !colptr || !nnz)
      return fail(NNK_EARG, "nnk_nearest_matrix: bad argument");
    int64_t k = nnk::nearest_matrix(coarse, n, fine, q, colptr, rowidx, val, cap);
    *nnz = k;
    if (k > cap) return fail(NNK_EARG, "nnk_nearest_matrix: output capacity too small");
    return NNK_OK;
  });
}

int32_t nnk_mobility_create(nnk_fmm *fmm, nnk_fmm *fmm_sparse, int64_t nb6, const int64_t *Brp,
                            const int32_t *Bci, const double *Bv, const int64_t *BTrp,
                            const int32_t *BTci, const double *BTv, const double *A1,
                            const double *A2, nnk_mobility **out) {
  return guard([&]() -> int32_t {
    if (!fmm || !out || !Brp || !BTrp || !A1 || !A2) return fail(NNK_EARG, "nnk_mobility_create: null argument");
    *out = nullptr;
    auto m = std::make_unique<nnk::Mobility>(fmm->f.get(), fmm_sparse ? fmm_sparse->f.get() : nullptr, nb6,
                                             Brp, Bci, Bv, BTrp, BTci, BTv, A1, A2);
    *out = new nnk_mobility{std::move(m)};
    return NNK_OK;
  });
}

static int32_t mob_host(nnk_mobility *h, const double *x, double *y, bool pre) {
  return guard([&]() -> int32_t {
    if (!h || !x || !y) return fail(NNK_EARG, "nnk_mobility: null argument");
    const int64_t n = h->m->size();
    nnk::DBuf dx, dy;
    dx.alloc(sizeof(double) * n);
    dy.alloc(sizeof(double) * n);
    NNK_HIP(hipMemcpy(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
    if (pre) h->m->precon(dx.as<double>(), dy.as<double>(), 0);
    else h->m->apply(dx.as<double>(), dy.as<double>(), 0);
    NNK_HIP(hipMemcpy(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return NNK_OK;
  });
}

int32_t nnk_mobility_apply(nnk_mobility *h, const double *x, double *y) { return mob_host(h, x, y, false); }
int32_t nnk_mobility_precon(nnk_mobility *h, const double *x, double *y) { return mob_host(h, x, y, true); }

int32_t nnk_mobility_apply_device(nnk_mobility *h, const double *d_x, double *d_y, void *stream) {
  return guard([&]() -> int32_t {
    if (!h || !d_x || !d_y) return fail(NNK_EARG, "nnk_mobility_apply_device: null argument");
    h->m->apply(d_x, d_y, (hipStream_t)stream);
    return NNK_OK;
  });
}

int32_t nnk_mobility_precon_device(nnk_mobility *h, const double *d_x, double *d_y, void *stream) {
  return guard([&]() -> int32_t {
    if (!h || !d_x || !d_y) return fail(NNK_EARG, "nnk_mobility_precon_device: null argument");
    h->m->precon(d_x, d_y, (hipStream_t)stream);
    return NNK_OK;
  });
}

int32_t nnk_mobility_solve(nnk_mobility *h, const double *rhs, double tol, int32_t maxit,
                           int32_t precon, double *x, int32_t *flag, double *relres, int32_t *iters,
                           double *resvec) {
  return guard([&]() -> int32_t {
    if (!h || !rhs || !x || maxit < 1 || !(tol > 0)) return fail(NNK_EARG, "nnk_mobility_solve: bad argument");
    const int64_t n = h->m->size();
    nnk::DBuf db, dx;
    db.alloc(sizeof(double) * n);
    dx.alloc(sizeof(double) * n);
    NNK_HIP(hipMemcpy(db.p, rhs, sizeof(double) * n, hipMemcpyHostToDevice));
    std::vector<double> rv;
    nnk::GmresResult R = h->m->gmres(db.as<double>(), dx.as<double>(), tol, maxit, precon != 0, 0, &rv);
    NNK_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (flag) *flag = R.flag;
    if (relres) *relres = R.relres;
    if (iters) *iters = R.iter;
    if (resvec)
      for (size_t i = 0; i < rv.size() && i <= (size_t)maxit; i++) resvec[i] = rv[i];
    return NNK_OK;
  });
}

int32_t nnk_fmm_set_graph(nnk_fmm *h, int32_t on) {
  return guard([&]() -> int32_t {
    if (!h) return fail(NNK_EARG, "nnk_fmm_set_graph: null handle");
    h->graph = on != 0;
    return NNK_OK;
  });
}

int32_t nnk_fmm_set_autotune(nnk_fmm *h, int32_t on) {
  return guard([&]() -> int32_t {
    if (!h) return fail(NNK_EARG, "nnk_fmm_set_autotune: null handle");
    h->f->set_autotune(on != 0);
    return NNK_OK;
  });
}

int32_t nnk_op_cache_limit(int64_t bytes) {
  return guard([&]() -> int32_t {
    if (bytes < 0) return fail(NNK_EARG, "nnk_op_cache_limit: negative byte limit");
    nnk::OpCache::get().set_limit(bytes);
    return NNK_OK;
  });
}

int32_t nnk_op_cache_stats(int64_t *hits, int64_t *misses, int64_t *bytes, int64_t *entries) {
  return guard([&]() -> int32_t {
    if (!hits || !misses || !bytes || !entries) return fail(NNK_EARG, "nnk_op_cache_stats: null output");
    int64_t o[4];
    nnk::OpCache::get().stats(o);
    *hits = o[0];
    *misses = o[1];
    *bytes = o[2];
    *entries = o[3];
    return NNK_OK;
  });
}

int32_t nnk_fmm_launch_config(const nnk_fmm *h, int32_t *split, int32_t *depth, int32_t nlev) {
  return guard([&]() -> int32_t {
    if (!h || !split || !depth || nlev < 0) return fail(NNK_EARG, "nnk_fmm_launch_config: bad argument");
    for (int l = 0; l < nlev; l++) {
      int a = 0, b = 0;
      h->f->launch_config(l, &a, &b);
      split[l] = a;
      depth[l] = b;
    }
    return NNK_OK;
  });
}

int32_t nnk_fmm_set_launch_config(nnk_fmm *h, const int32_t *split, const int32_t *depth, int32_t nlev) {
  return guard([&]() -> int32_t {
    if (!h || !split || !depth || nlev < 0) return fail(NNK_EARG, "nnk_fmm_set_launch_config: bad argument");
    for (int l = 0; l < nlev; l++) {
      if (split[l] == 0) continue;
      h->f->set_launch_config(l, split[l], depth[l]);
    }
    return NNK_OK;
  });
}

int32_t nnk_fmm_leaf_threads(const nnk_fmm *h, int32_t *threads) {
  if (!h || !threads) return fail(NNK_EARG, "nnk_fmm_leaf_threads: bad argument");
  *threads = h->f->leaf_threads();
  return NNK_OK;
}

int32_t nnk_mobility_set_graph(nnk_mobility *h, int32_t on) {
  return guard([&]() -> int32_t {
    if (!h) return fail(NNK_EARG, "nnk_mobility_set_graph: null handle");
    h->m->set_graph(on != 0);
    return NNK_OK;
  });
}

int32_t nnk_mobility_solve_refined(nnk_mobility *h, nnk_fmm *inner, nnk_fmm *inner_sparse, const double *rhs,
                                   double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                   int32_t *flag, double *relres, int32_t *iters, int32_t *outer) {
  return guard([&]() -> int32_t {
    if (!h || !inner || !rhs || !x || maxit < 1 || !(tol > 0) || !(inner_tol > 0))
      return fail(NNK_EARG, "nnk_mobility_solve_refined: bad argument");
    if (inner->npot != h->m->points() || inner->ntgt != h->m->points() ||
        (inner_sparse && (inner_sparse->npot != h->m->points() || inner_sparse->ntgt != h->m->points())))
      return fail(NNK_EARG, "nnk_mobility_solve_refined: inner operator is not on the mobility tree");
    const int64_t n = h->m->size();
    nnk::DBuf db, dx;
    db.alloc(sizeof(double) * n);
    dx.alloc(sizeof(double) * n);
    NNK_HIP(hipMemcpy(db.p, rhs, sizeof(double) * n, hipMemcpyHostToDevice));
    h->m->set_inner(inner->f.get(), inner_sparse ? inner_sparse->f.get() : nullptr);
    int nout = 0;
    nnk::GmresResult R =
        h->m->refine(db.as<double>(), dx.as<double>(), tol, maxit, inner_tol, precon != 0, 0, nullptr, &nout);
    h->m->set_inner(nullptr);
    NNK_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (flag) *flag = R.flag;
    if (relres) *relres = R.relres;
    if (iters) *iters = R.iter;
    if (outer) *outer = nout;
    return NNK_OK;
  });
}

int32_t nnk_mobility_solve_fgmres(nnk_mobility *h, nnk_fmm *inner, nnk_fmm *inner_sparse, const double *rhs,
                                  double tol, int32_t maxit, double inner_tol, int32_t precon, double *x,
                                  int32_t *flag, double *relres, int32_t *iters, int32_t *outer) {
  return guard([&]() -> int32_t {
    if (!h || !inner || !rhs || !x || maxit < 1 || !(tol > 0) || !(inner_tol > 0))
      return fail(NNK_EARG, "nnk_mobility_solve_fgmres: bad argument");
    if (inner->npot != h->m->points() || inner->ntgt != h->m->points() ||
        (inner_sparse && (inner_sparse->npot != h->m->points() || inner_sparse->ntgt != h->m->points())))
      return fail(NNK_EARG, "nnk_mobility_solve_fgmres: inner operator is not on the mobility tree");
    const int64_t n = h->m->size();
    nnk::DBuf db, dx;
    db.alloc(sizeof(double) * n);
    dx.alloc(sizeof(double) * n);
    NNK_HIP(hipMemcpy(db.p, rhs, sizeof(double) * n, hipMemcpyHostToDevice));
    h->m->set_inner(inner->f.get(), inner_sparse ? inner_sparse->f.get() : nullptr);
    int nout = 0;
    nnk::GmresResult R;
    try {
      R = h->m->fgmres(db.as<double>(), dx.as<double>(), tol, maxit, inner_tol, precon != 0, 0, nullptr, &nout);
    } catch (...) {
      h->m->set_inner(nullptr);
      throw;
    }
    h->m->set_inner(nullptr);
    NNK_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (flag) *flag = R.flag;
    if (relres) *relres = R.relres;
    if (iters) *iters = R.iter;
    if (outer) *outer = nout;
    return NNK_OK;
  });
}

void nnk_mobility_destroy(nnk_mobility *h) { delete h; }

int32_t nnk_direct_apply(const double *src, int64_t ns, const double *q, const double *tgt,
                         int64_t nt, double eps, double mu, double *u) {
  return guard([&]() -> int32_t {
    if (ns < 0 || nt < 0 || (ns && (!src || !q)) || (nt && (!tgt || !u)))
      return fail(NNK_EARG, "nnk_direct_apply: bad argument");
    if (!(eps > 0) || !(mu > 0)) return fail(NNK_EARG, "nnk_direct_apply: eps and mu must be > 0");
    nnk::DBuf ds, dq, dt, du;
    ds.alloc(sizeof(double) * 3 * std::max<int64_t>(ns, 1));
    dq.alloc(sizeof(double) * 3 * std::max<int64_t>(ns, 1));
    dt.alloc(sizeof(double) * 3 * std::max<int64_t>(nt, 1));
    du.alloc(sizeof(double) * 3 * std::max<int64_t>(nt, 1));
    NNK_HIP(hipMemcpy(ds.p, src, sizeof(double) * 3 * ns, hipMemcpyHostToDevice));
    NNK_HIP(hipMemcpy(dq.p, q, sizeof(double) * 3 * ns, hipMemcpyHostToDevice));
    NNK_HIP(hipMemcpy(dt.p, tgt, sizeof(double) * 3 * nt, hipMemcpyHostToDevice));
    nnk::direct_device(ds.as<double>(), ns, dq.as<double>(), dt.as<double>(), nt, eps, mu,
                       du.as<double>(), 0);
    NNK_HIP(hipMemcpy(u, du.p, sizeof(double) * 3 * nt, hipMemcpyDeviceToHost));
    return NNK_OK;
  });
}

int32_t nnk_direct_apply_device(const double *d_src, int64_t ns, const double *d_q,
                                const double *d_tgt, int64_t nt, double eps, double mu,
                                double *d_u, void *stream) {
  return guard([&]() -> int32_t {
    if (ns < 0 || nt < 0 || (ns && (!d_src || !d_q)) || (nt && (!d_tgt || !d_u)))
      return fail(NNK_EARG, "nnk_direct_apply_device: bad argument");
    if (!(eps > 0) || !(mu > 0)) return fail(NNK_EARG, "nnk_direct_apply_device: eps and mu must be > 0");
    nnk::direct_device(d_src, ns, d_q, d_tgt, nt, eps, mu, d_u, (hipStream_t)stream);
    return NNK_OK;
  });
}

int32_t nnk_direct_apply_device_noalloc(const double *d_src, int64_t ns, const double *d_q,
                                        const double *d_tgt, int64_t nt, double eps, double mu,
                                        double *d_u, void *stream) {
  return guard([&]() -> int32_t {
    if (ns < 0 || nt < 0 || (ns && (!d_src || !d_q)) || (nt && (!d_tgt || !d_u)))
      return fail(NNK_EARG, "nnk_direct_apply_device_noalloc: bad argument");
    if (!(eps > 0) || !(mu > 0))
      return fail(NNK_EARG, "nnk_direct_apply_device_noalloc: eps and mu must be > 0");
    nnk::direct_device_noalloc(d_src, ns, d_q, d_tgt, nt, eps, mu, d_u, (hipStream_t)stream);
    return NNK_OK;
  });
}

int32_t nnk_fmm_set_profiling(nnk_fmm *h, int32_t on) {
  if (!h) return fail(NNK_EARG, "nnk_fmm_set_profiling: null handle");
  h->f->set_profiling(on != 0);
  return NNK_OK;
}

int32_t nnk_fmm_phase_times(nnk_fmm *h, double *ms, int64_t *count, int32_t nphases) {
  return guard([&]() -> int32_t {
    if (!h || !ms || !count || nphases < nnk::Fmm::kPhases)
      return fail(NNK_EARG, "nnk_fmm_phase_times: bad argument");
    h->f->phase_times(ms, count);
    return NNK_OK;
  });
}

int32_t nnk_fmm_stats(const nnk_fmm *h, int64_t *out, int32_t nstats) {
  if (!h || !out || nstats < nnk::Fmm::kStats) return fail(NNK_EARG, "nnk_fmm_stats: bad argument");
  h->f->stats(out);
  if (nstats > nnk::Fmm::kStats) out[nnk::Fmm::kStats] = h->f->K();
  if (nstats > nnk::Fmm::kStats + 1) out[nnk::Fmm::kStats + 1] = h->f->rank();
  return NNK_OK;
}

int32_t nnk_batch_create(int32_t nprob, const int64_t *pt_off, const double *pts, const int64_t *fine_off,
                         const double *fine, const int64_t *nn_ptr, const int32_t *nn_col, const double *nn_val,
                         const double *weight, const double *arm, const double *rot, const int32_t *pt_body,
                         const int64_t *body_off, const double *eps, double mu, int32_t precision,
                         nnk_batch **out) {
  return guard([&]() -> int32_t {
    if (!out) return fail(NNK_EARG, "nnk_batch_create: null output");
    *out = nullptr;
    if (nprob <= 0 || !pt_off || !pts || !fine_off || !nn_ptr || !weight || !arm || !rot || !pt_body ||
        !body_off || !eps)
      return fail(NNK_EARG, "nnk_batch_create: bad argument");
    if (fine_off[nprob] > 0 && (!fine || !nn_col || !nn_val))
      return fail(NNK_EARG, "nnk_batch_create: quadrature points without a nearest map");
    if (precision != NNK_FP64 && precision != NNK_MIXED)
      return fail(NNK_EARG, "nnk_batch_create: precision must be NNK_FP64 or NNK_MIXED");
    auto h = std::make_unique<nnk_batch>();
    h->b = std::make_unique<nnk::Batch>(nprob, pt_off, pts, fine_off, fine, nn_ptr, nn_col, nn_val, weight, arm,
                                        rot, pt_body, body_off, eps, mu, precision == NNK_MIXED ? 1 : 0);
    *out = h.release();
    return NNK_OK;
  });
}

int32_t nnk_batch_size(const nnk_batch *h, int64_t *n) {
  return guard([&]() -> int32_t {
    if (!h || !n) return fail(NNK_EARG, "nnk_batch_size: null argument");
    *n = h->b->size();
    return NNK_OK;
  });
}

static int32_t batch_apply_host(nnk_batch *h, const double *x, double *y, bool transpose) {
  const int64_t n = h->b->size();
  nnk::DBuf dx, dy;
  dx.alloc(sizeof(double) * n);
  dy.alloc(sizeof(double) * n);
  NNK_HIP(hipMemcpy(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
  h->b->apply(dx.as<double>(), dy.as<double>(), transpose, 0);
  NNK_HIP(hipMemcpy(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  return NNK_OK;
}

int32_t nnk_batch_apply(nnk_batch *h, const double *x, double *y) {
  return guard([&]() -> int32_t {
    if (!h || !x || !y) return fail(NNK_EARG, "nnk_batch_apply: null argument");
    return batch_apply_host(h, x, y, false);
  });
}

int32_t nnk_batch_apply_transpose(nnk_batch *h, const double *x, double *y) {
  return guard([&]() -> int32_t {
    if (!h || !x || !y) return fail(NNK_EARG, "nnk_batch_apply_transpose: null argument");
    return batch_apply_host(h, x, y, true);
  });
}

int32_t nnk_batch_apply_device(nnk_batch *h, const double *d_x, double *d_y, void *stream) {
  return guard([&]() -> int32_t {
    if (!h || !d_x || !d_y) return fail(NNK_EARG, "nnk_batch_apply_device: null argument");
    h->b->apply(d_x, d_y, false, (hipStream_t)stream);
    return NNK_OK;
  });
}

int32_t nnk_batch_set_graph(nnk_batch *h, int32_t on) {
  return guard([&]() -> int32_t {
    if (!h) return fail(NNK_EARG, "nnk_batch_set_graph: null handle");
    h->b->set_graph(on != 0);
    return NNK_OK;
  });
}

int32_t nnk_batch_set_autotune(nnk_batch *h, int32_t on) {
  return guard([&]() -> int32_t {
    if (!h) return fail(NNK_EARG, "nnk_batch_set_autotune: null handle");
    h->b->set_autotune(on != 0);
    return NNK_OK;
  });
}

int32_t nnk_batch_p2p_waves(const nnk_batch *h, int32_t *waves) {
  if (!h || !waves) return fail(NNK_EARG, "nnk_batch_p2p_waves: bad argument");
  *waves = h->b->p2p_waves();
  return NNK_OK;
}

int32_t nnk_batch_solve(nnk_batch *h, const double *rhs, double tol, int32_t maxit, int32_t transpose, double *x,
                        int32_t *flag, double *relres, int32_t *iters) {
  return guard([&]() -> int32_t {
    if (!h || !rhs || !x || !flag || !relres || !iters || maxit < 1 || !(tol > 0))
      return fail(NNK_EARG, "nnk_batch_solve: bad argument");
    const int64_t n = h->b->size();
    nnk::DBuf db, dx;
    db.alloc(sizeof(double) * n);
    dx.alloc(sizeof(double) * n);
    NNK_HIP(hipMemcpy(db.p, rhs, sizeof(double) * n, hipMemcpyHostToDevice));
    h->b->gmres(db.as<double>(), dx.as<double>(), tol, maxit, transpose != 0, 0, flag, relres, iters);
    NNK_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return NNK_OK;
  });
}

int32_t nnk_batch_gradient(nnk_batch *h, const double *z, const double *lambda, double *gx, double *gy,
                           double *geps) {
  return guard([&]() -> int32_t {
    if (!h || !z || !lambda || !gx || !geps || (h->b->nfine() && !gy))
      return fail(NNK_EARG, "nnk_batch_gradient: null argument");
    const int64_t n = h->b->size(), N = h->b->npoints(), Q = h->b->nfine();
    nnk::DBuf dz, dl, dgx, dgy;
    dz.alloc(sizeof(double) * n);
    dl.alloc(sizeof(double) * n);
    dgx.alloc(sizeof(double) * 3 * N);
    dgy.alloc(sizeof(double) * 3 * std::max<int64_t>(Q, 1));
    NNK_HIP(hipMemcpy(dz.p, z, sizeof(double) * n, hipMemcpyHostToDevice));
    NNK_HIP(hipMemcpy(dl.p, lambda, sizeof(double) * n, hipMemcpyHostToDevice));
    h->b->gradient(dz.as<double>(), dl.as<double>(), dgx.as<double>(), dgy.as<double>(), geps, 0);
    NNK_HIP(hipMemcpy(gx, dgx.p, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
    if (Q) NNK_HIP(hipMemcpy(gy, dgy.p, sizeof(double) * 3 * Q, hipMemcpyDeviceToHost));
    return NNK_OK;
  });
}

void nnk_batch_destroy(nnk_batch *h) { delete h; }

}  // extern "C"
