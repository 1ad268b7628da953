// Many independent mobility problems solved together, with their adjoints.
//
// Problem p is the system of mobilityProblem (Mobility/mobilityProblem.m:
// 35-57) for its own bodies (and optionally a fixed boundary):
//   [ S_p   BT_p ] [ f_p      ]   [ u_p ]
//   [ B_p   0    ] [ U_p; W_p ] = [ F_p; M_p ]
// with S_p the regularized-Stokeslet sum from the problem's quadrature
// points (nearest map) onto its force points, evaluated directly (what the
// KIFMM reduces to when the tree is the root and 8 leaves), and B_p, BT_p
// as genMatrixComponets builds them: per force point the nearest weight
// w_j, the weighted arm pm_j and the rotated arm r_j, and its body (-1 for
// boundary points, which carry no rigid-body unknowns).
//
// All problems share one device vector [x_0; x_1; ...], x_p =
// [f (3 N_p, all x, all y, all z); U (3 nsw_p: all x, all y, all z);
// Omega (3 nsw_p)].  The operator is three launches: the quadrature
// strengths g = (NN (x) I3) f, a persistent P2P kernel that walks
// (problem, 64-target tile) work items largest first and adds BT (U; W),
// and one workgroup per body for B f.  The transpose (adjoint) operator
// runs the P2P from force points onto quadrature points and gathers with
// NN^T.  GMRES runs all problems in lockstep (segmented dot products,
// per-problem Givens on the host) and retires each problem as it converges.
// gradient() forms -d(lambda^T A z)/d(theta) for theta = force-point
// positions, quadrature-point positions and eps (per problem): the
// gradient of J(z) when A z = b and A^T lambda = dJ/dz, with b and the
// nearest map held fixed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "nnk_common.hpp"
#include "nnk_graph.hpp"
#include "nnk_tree.hpp"  // status codes

namespace nnk {

class Batch {
 public:
  // pt_body: body of each force point, local to its problem (0..nsw_p-1)
  // or -1; the points of a body are contiguous.  body_off: nprob+1
  // offsets of the problems' bodies (nsw_p = body_off[p+1] - body_off[p]).
  Batch(int nprob, const int64_t *pt_off, const double *pts, const int64_t *fine_off, const double *fine,
        const int64_t *nn_ptr, const int32_t *nn_col, const double *nn_val, const double *w,
        const double *pm, const double *rot, const int32_t *pt_body, const int64_t *body_off,
        const double *eps, double mu, int f32);
  int nprob() const { return np_; }
  int64_t size() const { return n_; }
  // y = A x (transpose: y = A^T x) for every problem (device vectors)
  void apply(const double *x, double *y, bool transpose, hipStream_t s);
  void set_graph(bool on) { graph_ = on; }
  // time the P2P workgroup size (1, 2, 4 wavefronts) on the first apply
  void set_autotune(bool on) { tune_ = on; }
  int p2p_waves() const { return p2p_waves_; }
  // unrestarted GMRES from x = 0 per problem on A (or A^T); per-problem
  // flag (0 when the recomputed relative residual is <= tol), that residual
  // and the iteration count.  With FP32-math operators the Arnoldi
  // estimate is driven to tol / 2.
  void gmres(const double *b, double *x, double tol, int maxit, bool transpose, hipStream_t s, int32_t *flag,
             double *relres, int32_t *iters);
  // gx: 3 per force point, gy: 3 per quadrature point, geps: per problem
  void gradient(const double *z, const double *lambda, double *gx, double *gy, double *geps, hipStream_t s);
  int64_t npoints() const { return ntot_; }
  int64_t nfine() const { return qtot_; }

 private:
  void op(const double *x, double *y, hipStream_t s);
  void autotune_p2p(const double *x, double *y, hipStream_t s);
  void opT(const double *x, double *y, hipStream_t s);
  void segdots(int k, const double *V, const double *w, double *host_out, hipStream_t s);

  int np_ = 0, nbody_ = 0;
  int64_t ntot_ = 0, qtot_ = 0, n_ = 0;
  double mu_ = 1.0;
  int f32_ = 0;
  std::vector<int64_t> hpoff_, hvoff_;
  DBuf poff_, foff_, voff_, pts_, fine_, nnptr_, nncol_, nnval_, nntptr_, nntrow_, nntval_;
  DBuf w_, pm_, rot_, eps_, pbody_, fbody_, nsw_;
  DBuf bprob_, blocal_, bpt_;  // per body: problem, local index, point range
  DBuf xbase_, xstride_, pid_, pidf_, items_, itemsT_, counter_, g_, v_, active_, part_;
  int nitems_ = 0, nitemsT_ = 0;
  bool graph_ = false;
  bool tune_ = false, tuned_ = false;
  int p2p_waves_ = 4;
  GraphOp gop_[2];
};

}  // namespace nnk
