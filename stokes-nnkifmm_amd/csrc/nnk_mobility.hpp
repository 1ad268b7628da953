// Device mobility operator, block preconditioner and GMRES driver.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "nnk_common.hpp"
#include "nnk_fmm.hpp"

namespace nnk {

struct DevCsr {
  int64_t rows = 0, cols = 0;
  DBuf rowptr, col, val;
  void upload(int64_t r, int64_t c, const std::vector<int64_t> &rp, const std::vector<int32_t> &ci,
              const std::vector<double> &v);
  // y (+)= alpha A x
  void mul(const double *x, double *y, bool accumulate, double alpha, hipStream_t s) const;
};

struct GmresResult {
  int flag = 1;
  double relres = 0;
  int iter = 0;
};

class Mobility {
 public:
  Mobility(Fmm *fmm, Fmm *fmm_sparse, int64_t nb6, const int64_t *Brp, const int32_t *Bci,
           const double *Bv, const int64_t *BTrp, const int32_t *BTci, const double *BTv,
           const double *A1, const double *A2);
  int64_t size() const { return n_; }
  int64_t points() const { return N_; }
  // mobilityMatrixMulti: y = [S f + BT UO; B f]
  void apply(const double *x, double *y, hipStream_t s);
  // mobilityPrecon
  void precon(const double *x, double *y, hipStream_t s);
  // gmres(A (or A*P), b, [], tol, maxit); x = P(x) when preconditioned
  GmresResult gmres(const double *b, double *x, double tol, int maxit, bool use_precon, hipStream_t s,
                    std::vector<double> *resvec);
  // Iterative refinement: outer residuals r = b - A x with the FP64
  // operator, corrections from GMRES on the inner (mixed-form) operator to
  // inner_tol; stops when ||r|| <= tol ||b||.  Returns the total inner
  // iteration count in iter and the outer count in outer.
  GmresResult refine(const double *b, double *x, double tol, int maxit, double inner_tol, bool use_precon,
                     hipStream_t s, std::vector<double> *resvec, int *outer);
  // Flexible GMRES (Saad 1993): an FP64 Arnoldi process on the FP64
  // operator whose k-th basis vector is preconditioned by an inner GMRES
  // solve (to inner_tol, on the inner operator, itself right-preconditioned
  // when use_precon) that may differ from step to step; x = Z y from the
  // stored preconditioned vectors Z.  At most 50 outer steps (no restart)
  // and maxit inner iterations in total.  Returns the inner iteration count
  // in iter and the outer step count in outer.
  GmresResult fgmres(const double *b, double *x, double tol, int maxit, double inner_tol, bool use_precon,
                     hipStream_t s, std::vector<double> *resvec, int *outer);
  // inner operators of the refinement: the matvec and (optional) the
  // coronaRes-4 operator of the preconditioner used inside the inner solves
  void set_inner(Fmm *inner, Fmm *inner_sparse = nullptr) {
    if (inner != Fi_ || inner_sparse != F4i_) { gops_[2].reset(); gops_[3].reset(); }
    Fi_ = inner;
    F4i_ = inner_sparse;
  }
  // replay the GMRES operator (A or A P) from HIP graphs, one per
  // (preconditioned, inner operator) combination, captured on first use
  void set_graph(bool on) { graph_ = on; }

 private:
  void op(const double *x, double *y, bool use_precon, hipStream_t s);
  void op_any(const double *x, double *y, bool use_precon, hipStream_t s);
  void kifmm(Fmm *Fm, const double *f, double *u, bool accumulate, hipStream_t s);
  void reduced(const double *v, double *f, hipStream_t s);
  double dot(const double *a, const double *b, hipStream_t s);
  void cgs2(const double *V, int kk, double *w, double *y, double *hc, int m1, double *part, hipStream_t s);
  void read_column(const double *hc, int m1, int kk, std::vector<double> &h, std::vector<double> &hh,
                   hipStream_t s);

  Fmm *F_, *F4_, *Fi_ = nullptr, *F4i_ = nullptr;
  Fmm *Fcur_ = nullptr;  // operator used by apply(): F_ or Fi_
  int64_t nb6_, N_ = 0, n3_ = 0, n_ = 0;
  DevCsr B_, Bt_, BT_, BTt_;
  DevCsr A1inv_, A2inv_;  // inverses of A1, A2 (block diagonal: one 6 x 6 block per body)
  DBuf aos_a_, aos_b_, t3a_, t3b_, s6a_, s6b_, px_, part_;
  bool graph_ = false;
  GraphOp gops_[4];
};

}  // namespace nnk
