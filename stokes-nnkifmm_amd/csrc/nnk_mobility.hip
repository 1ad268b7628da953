// Saddle-point mobility operator, its block preconditioner and GMRES on the
// device (Mobility/MobilityMatrixMulti.m, MobilityPrecon.m, mobilityProblem.m).
//
// Vector layout (mobilityProblem.m:37-46): x = [f (3N, component-major:
// all x, then all y, then all z); U (3 nsw, same); Omega (3 nsw)].
//   apply:  y = [ S f + BT (U;Omega) ;  B f ]            (MobilityMatrixMulti.m:31-35)
//   precon: y_f  = R^-1 f,  y_UO = -A2 \ (BT^T S4(B^T (A1 \ UO))),
//           y_f -= R^-1 (BT y_UO)                          (MobilityPrecon.m:29-38)
// with S the KIFMM matvec, S4 the same at 4 points per corona edge and R^-1
// the leaf-block solve.  B, BT come from genMatrixComponets as CSR; A1 and
// A2 (6 nsw square) are LU-factored on the host once.
// GMRES: unrestarted, zero initial guess, classical Gram-Schmidt with one
// reorthogonalisation on the device, Givens rotations on the host
// (Saad & Schultz 1986), stopping on the estimated relative residual and
// reporting the true one, like MATLAB's gmres(A, b, [], tol, maxit).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nnk_fmm.hpp"
#include "nnk_mobility.hpp"

namespace nnk {

// ---------------------------------------------------------------------------
// kernels

// y[r] (+)= sum_e val[e] x[col[e]] over row r: one wavefront per row, lanes
// stride the row and combine in a fixed order.
__global__ __launch_bounds__(256) void k_spmv(int64_t rows, const int64_t *__restrict__ rp,
                                              const int32_t *__restrict__ ci,
                                              const double *__restrict__ v,
                                              const double *__restrict__ x, double *__restrict__ y,
                                              int accumulate, double alpha) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  double s = 0.0;
  for (int64_t e = rp[r] + lane; e < rp[r + 1]; e += 64) s += v[e] * x[ci[e]];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) y[r] = accumulate ? y[r] + alpha * s : alpha * s;
}

__global__ void k_soa2aos(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  b[3 * i] = a[i];
  b[3 * i + 1] = a[n + i];
  b[3 * i + 2] = a[2 * n + i];
}

__global__ void k_aos2soa(int64_t n, const double *__restrict__ b, double *__restrict__ a,
                          int accumulate) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < 3; d++) a[d * n + i] = accumulate ? a[d * n + i] + b[3 * i + d] : b[3 * i + d];
}

__global__ void k_axpby(int64_t n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

// partial[j * nblk + blk] = sum over this block's slice of V_j . w
__global__ __launch_bounds__(256) void k_multidot(int64_t n, int k, const double *__restrict__ V,
                                                  const double *__restrict__ w,
                                                  double *__restrict__ partial) {
  __shared__ double red[256];
  for (int j = 0; j < k; j++) {
    const double *vj = V + (size_t)j * n;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
      s += vj[i] * w[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)j * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

// out[j] = sum_{q < nb} partial[j * nb + q], q ascending (the order the
// host used when it summed the partial dot products itself)
__global__ void k_sum_partials(int k, int nb, const double *__restrict__ partial, double *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  double t = 0.0;
  for (int q = 0; q < nb; q++) t += partial[(size_t)j * nb + q];
  out[j] = t;
}

// *nrm = sqrt(sum of the nb partials of w . w); y = w / *nrm when it is > 0
__global__ void k_norm_from_partials(int nb, const double *__restrict__ partial, double *__restrict__ nrm) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double t = 0.0;
  for (int q = 0; q < nb; q++) t += partial[q];
  *nrm = sqrt(t);
}
__global__ void k_scale_by_norm(int64_t n, const double *__restrict__ nrm, const double *__restrict__ w,
                                double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double h = *nrm;
  if (i < n && h > 0) y[i] = (1.0 / h) * w[i];
}

// w -= sum_j h[j] V_j
__global__ void k_subtract_combo(int64_t n, int k, const double *__restrict__ V,
                                 const double *__restrict__ h, double *__restrict__ w) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; j++) s += h[j] * V[(size_t)j * n + i];
  w[i] -= s;
}

// ---------------------------------------------------------------------------

namespace {

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

void host_lu(std::vector<double> &A, std::vector<int32_t> &piv, int n) {
  piv.resize(n);
  for (int k = 0; k < n; k++) {
    int p = k;
    double amax = std::fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A[(size_t)i * n + k]) > amax) { amax = std::fabs(A[(size_t)i * n + k]); p = i; }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; j++) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
    double d = A[(size_t)k * n + k];
    if (d == 0.0) continue;
    for (int i = k + 1; i < n; i++) {
      double l = A[(size_t)i * n + k] / d;
      A[(size_t)i * n + k] = l;
      if (l != 0.0)
        for (int j = k + 1; j < n; j++) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
    }
  }
}

// CSR transpose on the host
void csr_transpose(int64_t rows, int64_t cols, const std::vector<int64_t> &rp,
                   const std::vector<int32_t> &ci, const std::vector<double> &v,
                   std::vector<int64_t> &trp, std::vector<int32_t> &tci, std::vector<double> &tv) {
  trp.assign(cols + 1, 0);
  for (size_t e = 0; e < ci.size(); e++) trp[ci[e] + 1]++;
  for (int64_t c = 0; c < cols; c++) trp[c + 1] += trp[c];
  tci.resize(ci.size());
  tv.resize(v.size());
  std::vector<int64_t> fill(trp.begin(), trp.end() - 1);
  for (int64_t r = 0; r < rows; r++)
    for (int64_t e = rp[r]; e < rp[r + 1]; e++) {
      int64_t k = fill[ci[e]]++;
      tci[k] = (int32_t)r;
      tv[k] = v[e];
    }
}

// Inverse of a dense n x n matrix (row-major) as a sparse matrix: the
// connected components of its nonzero pattern are found (union-find), each
// component's dense block is LU-factored with partial pivoting and inverted,
// and the inverse is emitted in CSR (rows ascending, columns ascending).
// A1 = B B^T and A2 = BT^T BT couple only the 6 unknowns of one body
// (genMatrixComponets.m:99-100), so their inverses hold 36 entries per body.
void block_inverse_csr(const double *A, int64_t n, const char *what, std::vector<int64_t> &rp,
                       std::vector<int32_t> &ci, std::vector<double> &v) {
  std::vector<int64_t> root(n);
  for (int64_t i = 0; i < n; i++) root[i] = i;
  auto find = [&](int64_t i) {
    while (root[i] != i) { root[i] = root[root[i]]; i = root[i]; }
    return i;
  };
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++)
      if (A[i * n + j] != 0.0) {
        int64_t a = find(i), b = find(j);
        if (a != b) root[std::max(a, b)] = std::min(a, b);
      }
  std::vector<std::vector<int64_t>> comps;
  std::vector<int64_t> slot(n, -1);
  int64_t total = 0;
  for (int64_t i = 0; i < n; i++) {
    int64_t r = find(i);
    if (slot[r] < 0) { slot[r] = (int64_t)comps.size(); comps.emplace_back(); }
    comps[slot[r]].push_back(i);
  }
  for (auto &c : comps) total += (int64_t)c.size() * (int64_t)c.size();
  if (total > (int64_t)1 << 26)
    throw Error(NNK_EARG, std::string("mobility: ") + what + " couples too many unknowns for its inverse (" +
                              std::to_string(total) + " entries > 2^26)");
  std::vector<int64_t> rowlen(n, 0);
  std::vector<std::vector<double>> inv(comps.size());
  for (size_t q = 0; q < comps.size(); q++) {
    const auto &c = comps[q];
    const int m = (int)c.size();
    std::vector<double> M((size_t)m * m);
    for (int a = 0; a < m; a++)
      for (int b = 0; b < m; b++) M[(size_t)a * m + b] = A[c[a] * n + c[b]];
    std::vector<int32_t> piv;
    host_lu(M, piv, m);
    for (int k = 0; k < m; k++)
      if (M[(size_t)k * m + k] == 0.0)
        throw Error(NNK_EARG, std::string("mobility: ") + what + " is singular");
    // columns of the inverse: solve L U x = P e_j
    std::vector<double> X((size_t)m * m, 0.0), y(m);
    for (int j = 0; j < m; j++) {
      std::fill(y.begin(), y.end(), 0.0);
      y[j] = 1.0;
      for (int k = 0; k < m; k++)
        if (piv[k] != k) std::swap(y[k], y[piv[k]]);
      for (int k = 0; k < m; k++)
        for (int i = k + 1; i < m; i++) y[i] -= M[(size_t)i * m + k] * y[k];
      for (int k = m - 1; k >= 0; k--) {
        y[k] /= M[(size_t)k * m + k];
        for (int i = 0; i < k; i++) y[i] -= M[(size_t)i * m + k] * y[k];
      }
      for (int i = 0; i < m; i++) X[(size_t)i * m + j] = y[i];
    }
    inv[q] = std::move(X);
    for (int64_t r : c) rowlen[r] = m;
  }
  rp.assign(n + 1, 0);
  for (int64_t i = 0; i < n; i++) rp[i + 1] = rp[i] + rowlen[i];
  ci.resize(rp[n]);
  v.resize(rp[n]);
  for (size_t q = 0; q < comps.size(); q++) {
    const auto &c = comps[q];
    const int m = (int)c.size();
    for (int a = 0; a < m; a++) {
      int64_t e = rp[c[a]];
      for (int b = 0; b < m; b++, e++) {
        ci[e] = (int32_t)c[b];
        v[e] = inv[q][(size_t)a * m + b];
      }
    }
  }
}

}  // namespace

void DevCsr::upload(int64_t r, int64_t c, const std::vector<int64_t> &rp, const std::vector<int32_t> &ci,
                    const std::vector<double> &v) {
  rows = r;
  cols = c;
  rowptr.upload(rp);
  col.upload(ci);
  val.upload(v);
}

void DevCsr::mul(const double *x, double *y, bool accumulate, double alpha, hipStream_t s) const {
  if (rows == 0) return;
  hipLaunchKernelGGL(k_spmv, dim3(nblk(rows, 4)), dim3(256), 0, s, rows, rowptr.as<int64_t>(),
                     col.as<int32_t>(), val.as<double>(), x, y, accumulate ? 1 : 0, alpha);
  NNK_HIP(hipGetLastError());
}

Mobility::Mobility(Fmm *fmm, Fmm *fmm_sparse, int64_t nb6, const int64_t *Brp, const int32_t *Bci,
                   const double *Bv, const int64_t *BTrp, const int32_t *BTci, const double *BTv,
                   const double *A1, const double *A2)
    : F_(fmm), F4_(fmm_sparse), nb6_(nb6) {
  if (!F_) throw Error(NNK_EARG, "mobility: null KIFMM operator");
  N_ = F_->npot();
  if (F_->ntgt() != N_ || (F4_ && (F4_->npot() != N_ || F4_->ntgt() != N_)))
    throw Error(NNK_EARG, "mobility: operators must map the N force points to themselves");
  if (nb6 < 0 || nb6 % 6) throw Error(NNK_EARG, "mobility: 6*nsw must be a non-negative multiple of 6");
  n3_ = 3 * N_;
  n_ = n3_ + nb6_;
  std::vector<int64_t> rp(Brp, Brp + nb6 + 1);
  std::vector<int32_t> ci(Bci, Bci + rp.back());
  std::vector<double> v(Bv, Bv + rp.back());
  for (int32_t c : ci)
    if (c < 0 || c >= n3_) throw Error(NNK_EARG, "mobility: B column out of range");
  std::vector<int64_t> trp;
  std::vector<int32_t> tci;
  std::vector<double> tv;
  B_.upload(nb6, n3_, rp, ci, v);
  csr_transpose(nb6, n3_, rp, ci, v, trp, tci, tv);
  Bt_.upload(n3_, nb6, trp, tci, tv);
  std::vector<int64_t> rp2(BTrp, BTrp + n3_ + 1);
  std::vector<int32_t> ci2(BTci, BTci + rp2.back());
  std::vector<double> v2(BTv, BTv + rp2.back());
  for (int32_t c : ci2)
    if (c < 0 || c >= nb6) throw Error(NNK_EARG, "mobility: BT column out of range");
  BT_.upload(n3_, nb6, rp2, ci2, v2);
  csr_transpose(n3_, nb6, rp2, ci2, v2, trp, tci, tv);
  BTt_.upload(nb6, n3_, trp, tci, tv);
  for (int q = 0; q < 2; q++) {
    block_inverse_csr(q == 0 ? A1 : A2, nb6, q == 0 ? "A1" : "A2", rp, ci, v);
    (q == 0 ? A1inv_ : A2inv_).upload(nb6, nb6, rp, ci, v);
  }
  aos_a_.alloc(sizeof(double) * n3_);
  aos_b_.alloc(sizeof(double) * n3_);
  t3a_.alloc(sizeof(double) * n3_);
  t3b_.alloc(sizeof(double) * n3_);
  s6a_.alloc(sizeof(double) * nb6_);
  s6b_.alloc(sizeof(double) * nb6_);
}

// u (3N, component-major) = S f (3N, component-major) with operator Fm
void Mobility::kifmm(Fmm *Fm, const double *f, double *u, bool accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_soa2aos, dim3(nblk(N_)), dim3(256), 0, s, N_, f, aos_a_.as<double>());
  Fm->apply(aos_a_.as<double>(), aos_b_.as<double>(), s);
  hipLaunchKernelGGL(k_aos2soa, dim3(nblk(N_)), dim3(256), 0, s, N_, aos_b_.as<double>(), u,
                     accumulate ? 1 : 0);
  NNK_HIP(hipGetLastError());
}

void Mobility::reduced(const double *v, double *f, hipStream_t s) {
  hipLaunchKernelGGL(k_soa2aos, dim3(nblk(N_)), dim3(256), 0, s, N_, v, aos_a_.as<double>());
  F_->reduced_solve(aos_a_.as<double>(), aos_b_.as<double>(), s);
  hipLaunchKernelGGL(k_aos2soa, dim3(nblk(N_)), dim3(256), 0, s, N_, aos_b_.as<double>(), f, 0);
  NNK_HIP(hipGetLastError());
}


void Mobility::apply(const double *x, double *y, hipStream_t s) {
  kifmm(Fcur_ ? Fcur_ : F_, x, y, false, s);  // S f
  BT_.mul(x + n3_, y, true, 1.0, s);          // + BT (U; Omega)
  B_.mul(x, y + n3_, false, 1.0, s);          // B f
}

void Mobility::precon(const double *x, double *y, hipStream_t s) {
  if (!F4_) throw Error(NNK_ESTATE, "mobility: no coronaRes-4 operator for the preconditioner");
  reduced(x, y, s);                                        // y_f = R^-1 f
  A1inv_.mul(x + n3_, s6a_.as<double>(), false, 1.0, s);          // A1 \ UO
  Bt_.mul(s6a_.as<double>(), t3a_.as<double>(), false, 1.0, s);     // B^T (.)
  kifmm((Fcur_ && F4i_) ? F4i_ : F4_, t3a_.as<double>(), t3b_.as<double>(), false, s);  // S4 (.)
  BTt_.mul(t3b_.as<double>(), s6b_.as<double>(), false, 1.0, s);    // BT^T (.)
  A2inv_.mul(s6b_.as<double>(), y + n3_, false, -1.0, s);         // -A2 \ (.)
  BT_.mul(y + n3_, t3a_.as<double>(), false, 1.0, s);               // BT out
  reduced(t3a_.as<double>(), t3b_.as<double>(), s);                 // R^-1 (.)
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n3_)), dim3(256), 0, s, n3_, -1.0, t3b_.as<double>(), 1.0, y);
  NNK_HIP(hipGetLastError());
}

void Mobility::op(const double *x, double *y, bool use_precon, hipStream_t s) {
  if (!use_precon) { apply(x, y, s); return; }
  if (!px_.p) px_.alloc(sizeof(double) * n_);
  precon(x, px_.as<double>(), s);
  apply(px_.as<double>(), y, s);
}

void Mobility::op_any(const double *x, double *y, bool use_precon, hipStream_t s) {
  if (!graph_) { op(x, y, use_precon, s); return; }
  GraphOp &G = gops_[(use_precon ? 1 : 0) + (Fcur_ ? 2 : 0)];
  if (!G.ready()) {
    op(x, y, use_precon, s);  // eager first call: settles lazy set-up (leaf LU, buffers)
    NNK_HIP(hipStreamSynchronize(s));
    G.capture(n_, n_, [&](const double *in, double *out, hipStream_t cs) { op(in, out, use_precon, cs); });
    return;
  }
  G.run(x, y, s);
}

double Mobility::dot(const double *a, const double *b, hipStream_t s) {
  const int nb = 256;
  if (!part_.p || part_.bytes < sizeof(double) * nb) part_.alloc(sizeof(double) * nb);
  hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(256), 0, s, n_, 1, a, b, part_.as<double>());
  NNK_HIP(hipGetLastError());
  std::vector<double> h(nb);
  NNK_HIP(hipMemcpyAsync(h.data(), part_.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  NNK_HIP(hipStreamSynchronize(s));
  double t = 0;
  for (double v : h) t += v;
  return t;
}

// w -= V[0..kk) c with c = V^T w, twice (classical Gram-Schmidt with one
// reorthogonalisation), then y = w / ||w||; nothing waits on the host.  The
// coefficients of the two passes land in hc[0..kk) and hc[m1..m1+kk), the
// norm in hc[2 m1].  The partial dot products are summed on the device in a
// fixed order (k_sum_partials).
void Mobility::cgs2(const double *V, int kk, double *w, double *y, double *hc, int m1, double *part,
                    hipStream_t s) {
  const int64_t n = n_;
  const int nb = 256;
  for (int pass = 0; pass < 2; pass++) {
    double *c = hc + (size_t)pass * m1;
    hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(256), 0, s, n, kk, V, w, part);
    hipLaunchKernelGGL(k_sum_partials, dim3((kk + 63) / 64), dim3(64), 0, s, kk, nb, part, c);
    hipLaunchKernelGGL(k_subtract_combo, dim3(nblk(n)), dim3(256), 0, s, n, kk, V, c, w);
  }
  hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(256), 0, s, n, 1, w, w, part);
  hipLaunchKernelGGL(k_norm_from_partials, dim3(1), dim3(64), 0, s, nb, part, hc + 2 * (size_t)m1);
  hipLaunchKernelGGL(k_scale_by_norm, dim3(nblk(n)), dim3(256), 0, s, n, hc + 2 * (size_t)m1, w, y);
  NNK_HIP(hipGetLastError());
}

// the Arnoldi column of step k from the two passes' coefficients and the
// norm (one host wait per GMRES iteration): h[j] = c1[j] + c2[j], h[k+1]
void Mobility::read_column(const double *hc, int m1, int kk, std::vector<double> &h, std::vector<double> &hh,
                           hipStream_t s) {
  NNK_HIP(hipMemcpyAsync(hh.data(), hc, sizeof(double) * (2 * (size_t)m1 + 1), hipMemcpyDeviceToHost, s));
  NNK_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < kk; j++) h[j] = hh[j] + hh[(size_t)m1 + j];
  h[kk] = hh[2 * (size_t)m1];
}

GmresResult Mobility::gmres(const double *b, double *x, double tol, int maxit, bool use_precon,
                            hipStream_t s, std::vector<double> *resvec) {
  GmresResult R{};
  const int64_t n = n_;
  const int m = std::max(1, (int)std::min<int64_t>(maxit, n));
  const int nb = 256;
  DBuf V, w, hdev, part, hc;
  V.alloc(sizeof(double) * (size_t)(m + 1) * n);
  w.alloc(sizeof(double) * n);
  hdev.alloc(sizeof(double) * (m + 1));
  part.alloc(sizeof(double) * (size_t)(m + 1) * nb);
  hc.alloc(sizeof(double) * (2 * (size_t)(m + 1) + 1));
  NNK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  const double bnorm = std::sqrt(dot(b, b, s));
  if (resvec) resvec->assign(1, bnorm);
  if (bnorm == 0.0) {  // MATLAB: x = 0, flag 0
    R.flag = 0; R.relres = 0; R.iter = 0;
    if (use_precon) { /* P(0) = 0 */ }
    return R;
  }
  double *Vp = V.as<double>();
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, 1.0 / bnorm, b, 0.0, Vp);
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0), h(m + 1),
      hh(2 * (size_t)(m + 1) + 1);
  g[0] = bnorm;
  int k = 0;
  bool done = false;
  for (k = 0; k < m && !done; k++) {
    op_any(Vp + (size_t)k * n, w.as<double>(), use_precon, s);
    cgs2(Vp, k + 1, w.as<double>(), Vp + (size_t)(k + 1) * n, hc.as<double>(), m + 1, part.as<double>(), s);
    read_column(hc.as<double>(), m + 1, k + 1, h, hh, s);
    const double hn = h[k + 1];
    for (int j = 0; j < k; j++) {  // previous rotations
      double t = cs[j] * h[j] + sn[j] * h[j + 1];
      h[j + 1] = -sn[j] * h[j] + cs[j] * h[j + 1];
      h[j] = t;
    }
    double r = std::hypot(h[k], h[k + 1]);
    cs[k] = r > 0 ? h[k] / r : 1.0;
    sn[k] = r > 0 ? h[k + 1] / r : 0.0;
    h[k] = r;
    h[k + 1] = 0.0;
    g[k + 1] = -sn[k] * g[k];
    g[k] = cs[k] * g[k];
    for (int j = 0; j <= k; j++) H[(size_t)j * m + k] = h[j];
    if (resvec) resvec->push_back(std::fabs(g[k + 1]));
    if (std::fabs(g[k + 1]) <= tol * bnorm || hn == 0.0) done = true;
  }
  const int kk = k;
  std::vector<double> y(kk, 0.0);
  for (int i = kk - 1; i >= 0; i--) {
    double t = g[i];
    for (int j = i + 1; j < kk; j++) t -= H[(size_t)i * m + j] * y[j];
    y[i] = H[(size_t)i * m + i] != 0 ? t / H[(size_t)i * m + i] : 0.0;
  }
  // x = V y (then the true residual)
  NNK_HIP(hipMemcpyAsync(hdev.p, y.data(), sizeof(double) * kk, hipMemcpyHostToDevice, s));
  NNK_HIP(hipMemsetAsync(w.p, 0, sizeof(double) * n, s));
  hipLaunchKernelGGL(k_subtract_combo, dim3(nblk(n)), dim3(256), 0, s, n, kk, Vp, hdev.as<double>(),
                     w.as<double>());
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, -1.0, w.as<double>(), 0.0, x);
  NNK_HIP(hipGetLastError());
  op_any(x, w.as<double>(), use_precon, s);
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, 1.0, b, -1.0, w.as<double>());
  R.relres = std::sqrt(dot(w.as<double>(), w.as<double>(), s)) / bnorm;
  R.iter = kk;
  R.flag = R.relres <= tol ? 0 : 1;
  if (use_precon) {  // x = P(x)  (mobilityProblem.m:54)
    precon(x, w.as<double>(), s);
    NNK_HIP(hipMemcpyAsync(x, w.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  }
  NNK_HIP(hipStreamSynchronize(s));
  return R;
}

GmresResult Mobility::refine(const double *b, double *x, double tol, int maxit, double inner_tol,
                            bool use_precon, hipStream_t s, std::vector<double> *resvec, int *outer) {
  if (!Fi_) throw Error(NNK_ESTATE, "mobility: no inner operator for refinement");
  const int64_t n = n_;
  DBuf r, d;
  r.alloc(sizeof(double) * n);
  d.alloc(sizeof(double) * n);
  NNK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  NNK_HIP(hipMemcpyAsync(r.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  const double bnorm = std::sqrt(dot(b, b, s));
  GmresResult R{};
  if (resvec) resvec->assign(1, bnorm);
  double rn = bnorm;
  int used = 0, it = 0;
  for (; it < 50 && rn > tol * bnorm && used < maxit; it++) {
    Fcur_ = Fi_;
    GmresResult g = gmres(r.as<double>(), d.as<double>(), std::max(inner_tol, tol * bnorm / rn), maxit - used,
                          use_precon, s, nullptr);
    Fcur_ = nullptr;
    used += std::max(g.iter, 1);
    hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, 1.0, d.as<double>(), 1.0, x);
    apply(x, r.as<double>(), s);  // FP64 operator
    hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, 1.0, b, -1.0, r.as<double>());
    NNK_HIP(hipGetLastError());
    rn = std::sqrt(dot(r.as<double>(), r.as<double>(), s));
    if (resvec) resvec->push_back(rn);
  }
  R.relres = rn / bnorm;
  R.iter = used;
  R.flag = R.relres <= tol ? 0 : 1;
  if (outer) *outer = it;
  NNK_HIP(hipStreamSynchronize(s));
  return R;
}

GmresResult Mobility::fgmres(const double *b, double *x, double tol, int maxit, double inner_tol,
                             bool use_precon, hipStream_t s, std::vector<double> *resvec, int *outer) {
  if (!Fi_) throw Error(NNK_ESTATE, "mobility: no inner operator for flexible GMRES");
  const int64_t n = n_;
  const int m = 50;  // outer Arnoldi steps (no restart)
  const int nb = 256;
  DBuf V, Z, w, hdev, part, hc;
  V.alloc(sizeof(double) * (size_t)(m + 1) * n);
  Z.alloc(sizeof(double) * (size_t)m * n);
  w.alloc(sizeof(double) * n);
  hdev.alloc(sizeof(double) * (m + 1));
  part.alloc(sizeof(double) * (size_t)(m + 1) * nb);
  hc.alloc(sizeof(double) * (2 * (size_t)(m + 1) + 1));
  GmresResult R{};
  NNK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
  const double bnorm = std::sqrt(dot(b, b, s));
  if (resvec) resvec->assign(1, bnorm);
  if (outer) *outer = 0;
  if (bnorm == 0.0) {
    R.flag = 0;
    NNK_HIP(hipStreamSynchronize(s));
    return R;
  }
  double *Vp = V.as<double>(), *Zp = Z.as<double>();
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, 1.0 / bnorm, b, 0.0, Vp);
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0), h(m + 1),
      hh(2 * (size_t)(m + 1) + 1);
  g[0] = bnorm;
  int k = 0, used = 0;
  bool done = false;
  for (k = 0; k < m && !done; k++) {
    // z_k: the inner (mixed-form) solve of A z = v_k, right-preconditioned
    Fcur_ = Fi_;
    GmresResult gi;
    try {
      gi = gmres(Vp + (size_t)k * n, Zp + (size_t)k * n, inner_tol, std::max(1, maxit - used), use_precon, s,
                 nullptr);
    } catch (...) {
      Fcur_ = nullptr;
      throw;
    }
    Fcur_ = nullptr;
    used += std::max(gi.iter, 1);
    apply(Zp + (size_t)k * n, w.as<double>(), s);  // FP64 operator
    cgs2(Vp, k + 1, w.as<double>(), Vp + (size_t)(k + 1) * n, hc.as<double>(), m + 1, part.as<double>(), s);
    read_column(hc.as<double>(), m + 1, k + 1, h, hh, s);
    const double hn = h[k + 1];
    for (int j = 0; j < k; j++) {
      const double t = cs[j] * h[j] + sn[j] * h[j + 1];
      h[j + 1] = -sn[j] * h[j] + cs[j] * h[j + 1];
      h[j] = t;
    }
    const double r = std::hypot(h[k], h[k + 1]);
    cs[k] = r > 0 ? h[k] / r : 1.0;
    sn[k] = r > 0 ? h[k + 1] / r : 0.0;
    h[k] = r;
    h[k + 1] = 0.0;
    g[k + 1] = -sn[k] * g[k];
    g[k] = cs[k] * g[k];
    for (int j = 0; j <= k; j++) H[(size_t)j * m + k] = h[j];
    if (resvec) resvec->push_back(std::fabs(g[k + 1]));
    if (std::fabs(g[k + 1]) <= tol * bnorm || hn == 0.0 || used >= maxit) done = true;
  }
  const int kk = k;
  std::vector<double> y(kk, 0.0);
  for (int i = kk - 1; i >= 0; i--) {
    double t = g[i];
    for (int j = i + 1; j < kk; j++) t -= H[(size_t)i * m + j] * y[j];
    y[i] = H[(size_t)i * m + i] != 0 ? t / H[(size_t)i * m + i] : 0.0;
  }
  // x = Z y, then the true FP64 residual
  NNK_HIP(hipMemcpyAsync(hdev.p, y.data(), sizeof(double) * kk, hipMemcpyHostToDevice, s));
  NNK_HIP(hipMemsetAsync(w.p, 0, sizeof(double) * n, s));
  hipLaunchKernelGGL(k_subtract_combo, dim3(nblk(n)), dim3(256), 0, s, n, kk, Zp, hdev.as<double>(),
                     w.as<double>());
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, -1.0, w.as<double>(), 0.0, x);
  apply(x, w.as<double>(), s);
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, s, n, 1.0, b, -1.0, w.as<double>());
  NNK_HIP(hipGetLastError());
  R.relres = std::sqrt(dot(w.as<double>(), w.as<double>(), s)) / bnorm;
  R.iter = used;
  R.flag = R.relres <= tol ? 0 : 1;
  if (outer) *outer = kk;
  NNK_HIP(hipStreamSynchronize(s));
  return R;
}

}  // namespace nnk
