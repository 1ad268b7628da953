// Leaf-block solve of the preconditioner (Mobility/ReducedSolve.m:164-209):
// for every target leaf, solve K(leaf sources -> leaf targets) x = v, or in
// NEAREST mode (K(fine -> targets) kron(NN_leaf, I3)) x = v (ReducedSolve.m
// :170-187).  The 3n x 3n blocks are formed and LU-factored with partial
// pivoting once per tree on the GPU (one workgroup per leaf), and each
// block's inverse is formed from its factors; each solve is then one
// block-diagonal matrix-vector product.  MATLAB's mldivide would use
// Cholesky on the symmetric non-NEAREST blocks; both give the solution up
// to rounding (relative error of order cond(A) times the FP64 epsilon).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "nnk_fmm.hpp"
#include "nnk_kernels.hpp"

namespace nnk {

struct RedArgs {
  const int32_t *leaves;
  const int32_t *tgt_ptr;
  const double *tgt_pos;
  const int32_t *tgt_out;
  const int32_t *src_ptr;
  const double *src_pos;
  const int64_t *ctr_ptr;   // NEAREST: contributions per source instance
  const int32_t *ctr_col;
  const double *ctr_w;
  const int32_t *ord2loc;   // target ordinal -> index within its leaf
  const int64_t *off;       // block offset (doubles) per leaf slot
  double *blk;
  int32_t *piv;             // 3n pivots per leaf at off3[leaf]
  int32_t *perm;            // the pivots as one permutation (same offsets)
  const int64_t *poff;
  double e2, c;
  int nearest;
};

// form the leaf block: rows 3t+a (targets), columns 3j+b (the leaf's coarse
// points in target order); block stored row-major
__global__ __launch_bounds__(256) void k_red_form(RedArgs a) {
  const int li = blockIdx.x;
  const int n = a.leaves[li];
  const int tb = a.tgt_ptr[n], nt = a.tgt_ptr[n + 1] - tb;
  const int n3 = 3 * nt;
  double *A = a.blk + a.off[li];
  for (int i = threadIdx.x; i < n3 * n3; i += 256) A[i] = 0.0;
  __syncthreads();
  const int sb = a.src_ptr[n], se = a.src_ptr[n + 1];
  for (int t = threadIdx.x; t < nt; t += 256) {
    const double *tp = a.tgt_pos + 3 * (size_t)(tb + t);
    double *row = A + (size_t)(3 * t) * n3;
    for (int s = sb; s < se; s++) {
      const double *sp = a.src_pos + 3 * (size_t)s;
      double rx = tp[0] - sp[0], ry = tp[1] - sp[1], rz = tp[2] - sp[2];
      double r2 = rx * rx + ry * ry + rz * rz;
      double inv = rsqrt(r2 + a.e2);
      double w3 = a.c * inv * inv * inv;
      double iso = r2 + 2.0 * a.e2;
      double r[3] = {rx, ry, rz};
      if (!a.nearest) {
        const int j = s - sb;  // sources and targets share the leaf order
        for (int p = 0; p < 3; p++)
          for (int q = 0; q < 3; q++)
            row[(size_t)p * n3 + 3 * j + q] = (r[p] * r[q] + (p == q ? iso : 0.0)) * w3;
      } else {
        for (int64_t e = a.ctr_ptr[s]; e < a.ctr_ptr[s + 1]; e++) {
          const int j = a.ord2loc[a.ctr_col[e]];
          const double w = a.ctr_w[e] * w3;
          for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++)
              row[(size_t)p * n3 + 3 * j + q] += (r[p] * r[q] + (p == q ? iso : 0.0)) * w;
        }
      }
    }
  }
}

// In-place LU with partial pivoting (first maximal |a_ik|), blocked
// right-looking (the order of LAPACK's getrf), one launch pair per 32-column
// panel for every leaf at once.  k_red_lu_panel (one workgroup per leaf)
// factors the panel column by column (row interchanges applied to whole
// rows) and solves the panel's rows of U right of it with its unit lower
// triangle; k_red_lu_update gives every 64 x 64 tile of every leaf's
// trailing matrix a workgroup of its own for the rank-32 update (staged
// through LDS), so that large leaves use the whole GPU instead of one CU.
constexpr int kLuThreads = 512, kLuNB = 32, kLuT = 64;
__global__ __launch_bounds__(kLuThreads) void k_red_lu_panel(RedArgs a, int k0) {
  const int li = blockIdx.x;
  const int n = a.leaves[li];
  const int n3 = 3 * (a.tgt_ptr[n + 1] - a.tgt_ptr[n]);
  if (k0 >= n3) return;  // the whole workgroup
  double *A = a.blk + a.off[li];
  int32_t *piv = a.piv + a.poff[li];
  __shared__ double sv[kLuThreads];
  __shared__ int si[kLuThreads];
  const int tid = threadIdx.x;
  const int k1 = min(k0 + kLuNB, n3);
  for (int k = k0; k < k1; k++) {
    double best = -1.0;
    int bi = n3;
    for (int i = k + tid; i < n3; i += kLuThreads) {
      const double v = fabs(A[(size_t)i * n3 + k]);
      if (v > best) { best = v; bi = i; }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int st = kLuThreads / 2; st > 0; st >>= 1) {
      if (tid < st) {
        const double v2 = sv[tid + st];
        const int i2 = si[tid + st];
        if (v2 > sv[tid] || (v2 == sv[tid] && i2 < si[tid])) {
          sv[tid] = v2;
          si[tid] = i2;
        }
      }
      __syncthreads();
    }
    const int p = si[0];
    if (tid == 0) piv[k] = p;
    if (p != k)
      for (int jj = tid; jj < n3; jj += kLuThreads) {
        const double t = A[(size_t)k * n3 + jj];
        A[(size_t)k * n3 + jj] = A[(size_t)p * n3 + jj];
        A[(size_t)p * n3 + jj] = t;
      }
    __syncthreads();
    const double d = A[(size_t)k * n3 + k];
    if (d != 0.0)
      for (int i = k + 1 + tid; i < n3; i += kLuThreads) A[(size_t)i * n3 + k] /= d;
    __syncthreads();
    const int w = k1 - k - 1;  // panel columns right of k
    if (w > 0) {
      const int m = n3 - k - 1;
      for (int idx = tid; idx < m * w; idx += kLuThreads) {
        const int i = k + 1 + idx / w, jj = k + 1 + idx % w;
        A[(size_t)i * n3 + jj] -= A[(size_t)i * n3 + k] * A[(size_t)k * n3 + jj];
      }
    }
    __syncthreads();
  }
  if (k1 >= n3) return;
  const int m = n3 - k1;  // trailing size
  // U12 = L11^-1 A12 (rows k0..k1-1, columns k1..)
  for (int k = k0; k < k1 - 1; k++) {
    const int h = k1 - k - 1;
    for (int idx = tid; idx < h * m; idx += kLuThreads) {
      const int i = k + 1 + idx / m, jj = k1 + idx % m;
      A[(size_t)i * n3 + jj] -= A[(size_t)i * n3 + k] * A[(size_t)k * n3 + jj];
    }
    __syncthreads();
  }
}

// A22 -= L21 U12 after the panel at k0: workgroup (leaf, t) updates tile t
// of the leaf's trailing matrix (row-major over its ceil(m / 64)^2 tiles)
__global__ __launch_bounds__(kLuThreads) void k_red_lu_update(RedArgs a, int k0) {
  const int li = blockIdx.x;
  const int n = a.leaves[li];
  const int n3 = 3 * (a.tgt_ptr[n + 1] - a.tgt_ptr[n]);
  const int k1 = k0 + kLuNB;
  if (k1 >= n3) return;  // the whole workgroup
  const int m = n3 - k1, nt = (m + kLuT - 1) / kLuT;
  const int t = blockIdx.y;
  if (t >= nt * nt) return;
  double *A = a.blk + a.off[li];
  __shared__ double Ls[kLuT][kLuNB + 1];
  __shared__ double Us[kLuNB][kLuT + 1];
  const int tid = threadIdx.x;
  const int r0 = k1 + (t / nt) * kLuT, c0 = k1 + (t % nt) * kLuT;
  for (int idx = tid; idx < kLuT * kLuNB; idx += kLuThreads) {
    const int r = idx / kLuNB, kk = idx % kLuNB;
    Ls[r][kk] = r0 + r < n3 ? A[(size_t)(r0 + r) * n3 + k0 + kk] : 0.0;
    const int kk2 = idx / kLuT, c = idx % kLuT;
    Us[kk2][c] = c0 + c < n3 ? A[(size_t)(k0 + kk2) * n3 + c0 + c] : 0.0;
  }
  __syncthreads();
  const int r = tid >> 3, cb = (tid & 7) * 8;
  if (r0 + r < n3) {
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (int kk = 0; kk < kLuNB; kk++) {
      const double l = Ls[r][kk];
#pragma unroll
      for (int c = 0; c < 8; c++) acc[c] += l * Us[kk][cb + c];
    }
#pragma unroll
    for (int c = 0; c < 8; c++)
      if (c0 + cb + c < n3) A[(size_t)(r0 + r) * n3 + c0 + cb + c] -= acc[c];
  }
}

// the factors of each leaf to column-major order (in place), so that the
// substitutions read whole columns contiguously
__global__ __launch_bounds__(256) void k_red_transpose(RedArgs a) {
  const int li = blockIdx.x;
  const int n = a.leaves[li];
  const int n3 = 3 * (a.tgt_ptr[n + 1] - a.tgt_ptr[n]);
  double *A = a.blk + a.off[li];
  const int64_t tot = (int64_t)n3 * n3;
  for (int64_t idx = threadIdx.x; idx < tot; idx += 256) {
    const int i = (int)(idx / n3), j = (int)(idx % n3);
    if (i < j) {
      const double t = A[(size_t)i * n3 + j];
      A[(size_t)i * n3 + j] = A[(size_t)j * n3 + i];
      A[(size_t)j * n3 + i] = t;
    }
  }
}

// the row interchanges of each leaf's LU, applied in order to the identity:
// perm[i] is the row of the right-hand side that lands in position i
__global__ void k_red_perm(RedArgs a, int nleaves) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= nleaves) return;
  const int n = a.leaves[li];
  const int n3 = 3 * (a.tgt_ptr[n + 1] - a.tgt_ptr[n]);
  const int32_t *piv = a.piv + a.poff[li];
  int32_t *perm = a.perm + a.poff[li];
  for (int i = 0; i < n3; i++) perm[i] = i;
  for (int k = 0; k < n3; k++) {
    const int p = piv[k];
    const int32_t t = perm[k];
    perm[k] = perm[p];
    perm[p] = t;
  }
}

// The inverse of each leaf block from its LU factors (set-up only), in the
// row-major working block W (W[i n3 + j]).  A workgroup owns 64 consecutive
// columns j of A^-1 = U^-1 L^-1 P (lane = column) and walks the rows in
// blocks of kRedInvB; each of its kRedInvWaves wavefronts keeps the
// column's entries of a quarter of the current row block in registers and
// subtracts the rows solved before (each kRedInvB x kRedInvB factor tile
// staged in LDS and read as broadcasts; the solved entries read from W,
// coalesced across the lanes).  The first wavefront then solves the
// diagonal tile in LDS, and every wavefront stores its rows.  Forward with
// the unit lower triangle (row blocks ascending), then back with the upper
// one (descending).  Each solved entry is read once per row block and
// wavefront instead of once per row.
constexpr int kRedInvCols = 64, kRedInvB = 32, kRedInvWaves = 4;
__global__ __launch_bounds__(kRedInvCols * kRedInvWaves) void k_red_inv(RedArgs a, double *__restrict__ W) {
  constexpr int B = kRedInvB, NT = kRedInvCols * kRedInvWaves, RW = B / kRedInvWaves;
  __shared__ double F[B][B + 1];           // F[kk][r]: factor entry (row i0 + r, column k0 + kk)
  __shared__ double Xs[B][kRedInvCols];    // the row block of the 64 columns (diagonal solve)
  const int li = blockIdx.x;
  const int n = a.leaves[li];
  const int n3 = 3 * (a.tgt_ptr[n + 1] - a.tgt_ptr[n]);
  if ((int)blockIdx.y * kRedInvCols >= n3) return;  // the whole workgroup
  const int tid = threadIdx.x, lane = tid % kRedInvCols, r0 = (tid / kRedInvCols) * RW;
  const int j = blockIdx.y * kRedInvCols + lane;
  const bool col = j < n3;
  const double *A = a.blk + a.off[li];  // column-major factors: A[k n3 + i] = LU[i][k]
  const int32_t *perm = a.perm + a.poff[li];
  double *X = W + a.off[li];
  const int nb = (n3 + B - 1) / B;
  // (the barriers also order the global stores of W before later reads by
  // the other wavefronts)
  auto stage = [&](int i0, int k0) {
    __syncthreads();  // the previous tile has been read
    for (int idx = tid; idx < B * B; idx += NT) {
      const int kk = idx / B, r = idx % B, k = k0 + kk, i = i0 + r;
      F[kk][r] = (k < n3 && i < n3) ? A[(size_t)k * n3 + i] : 0.0;
    }
    __syncthreads();
  };
  // this wavefront's rows r0 .. r0+RW-1 of the block, minus the already
  // solved rows k0 .. k0+B-1 of the column times the tile
  double acc[RW];
  auto update = [&](int k0) {
    for (int kk = 0; kk < B; kk++) {
      const double xk = (col && k0 + kk < n3) ? X[(size_t)(k0 + kk) * n3 + j] : 0.0;
#pragma unroll
      for (int r = 0; r < RW; r++) acc[r] -= F[kk][r0 + r] * xk;
    }
  };
  // hand the rows to the diagonal solve of the first wavefront (solve),
  // then store every wavefront's rows
  auto diag = [&](int i0, auto solve) {
#pragma unroll
    for (int r = 0; r < RW; r++) Xs[r0 + r][lane] = acc[r];
    __syncthreads();
    if (tid < kRedInvCols) solve();
    __syncthreads();
    if (col)
      for (int r = r0; r < r0 + RW && i0 + r < n3; r++) X[(size_t)(i0 + r) * n3 + j] = Xs[r][lane];
  };
  for (int I = 0; I < nb; I++) {  // L y = P e_j
    const int i0 = I * B;
#pragma unroll
    for (int r = 0; r < RW; r++) acc[r] = (col && i0 + r0 + r < n3 && perm[i0 + r0 + r] == j) ? 1.0 : 0.0;
    for (int K = 0; K < I; K++) {
      stage(i0, K * B);
      update(K * B);
    }
    stage(i0, i0);
    diag(i0, [&]() {
      for (int r = 1; r < B; r++) {  // unit lower diagonal tile
        double t = Xs[r][lane];
        for (int q = 0; q < r; q++) t -= F[q][r] * Xs[q][lane];
        Xs[r][lane] = t;
      }
    });
  }
  for (int I = nb - 1; I >= 0; I--) {  // U x = y
    const int i0 = I * B;
#pragma unroll
    for (int r = 0; r < RW; r++) acc[r] = (col && i0 + r0 + r < n3) ? X[(size_t)(i0 + r0 + r) * n3 + j] : 0.0;
    for (int K = I + 1; K < nb; K++) {
      stage(i0, K * B);
      update(K * B);
    }
    stage(i0, i0);
    diag(i0, [&]() {
      for (int r = B - 1; r >= 0; r--) {  // upper diagonal tile
        double t = Xs[r][lane];
        for (int q = r + 1; q < B; q++) t -= F[q][r] * Xs[q][lane];
        Xs[r][lane] = i0 + r < n3 ? t / F[r][r] : 0.0;
      }
    });
  }
}

// x = A^-1 v for each leaf with the column-major inverse
// (Inv[j n3 + i] = (A^-1)_ij): workgroup (leaf, row group) holds the leaf's
// right-hand side in LDS and gives each of its rows one thread; the columns
// are read in order (coalesced across the rows), the sum in column order.
// Large blocks spread over several workgroups instead of one CU.
constexpr int kRedApplyThreads = 256;
__global__ __launch_bounds__(kRedApplyThreads) void k_red_apply(RedArgs a, const double *__restrict__ v,
                                                                double *__restrict__ x) {
  extern __shared__ double b[];
  const int li = blockIdx.x;
  const int n = a.leaves[li];
  const int tb = a.tgt_ptr[n], nt = a.tgt_ptr[n + 1] - tb;
  const int n3 = 3 * nt;
  const int i0 = blockIdx.y * kRedApplyThreads;
  if (i0 >= n3) return;  // the whole workgroup
  const double *Inv = a.blk + a.off[li];
  for (int i = threadIdx.x; i < n3; i += kRedApplyThreads) b[i] = v[3 * (size_t)a.tgt_out[tb + i / 3] + i % 3];
  __syncthreads();
  const int i = i0 + threadIdx.x;
  if (i < n3) {
    double acc = 0.0;
    for (int jj = 0; jj < n3; jj++) acc += Inv[(size_t)jj * n3 + i] * b[jj];
    x[3 * (size_t)a.tgt_out[tb + i / 3] + i % 3] = acc;
  }
}

void Fmm::reduced_factor() {
  const Tree &T = tree_;
  if (T.npot != T.ntgt || T.npts != T.npot)
    throw Error(NNK_EARG, "reducedSolve needs potential points == target points");
  std::vector<int32_t> tl(n_tleaves_), tp(T.count + 2);
  NNK_HIP(hipMemcpy(tl.data(), tleaves_.p, sizeof(int32_t) * n_tleaves_, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(tp.data(), tgt_ptr_.p, sizeof(int32_t) * (T.count + 2), hipMemcpyDeviceToHost));
  std::vector<int64_t> off(n_tleaves_ + 1, 0), poff(n_tleaves_ + 1, 0);
  red_maxn_ = 0;
  for (int i = 0; i < n_tleaves_; i++) {
    int64_t n3 = 3 * (tp[tl[i] + 1] - tp[tl[i]]);
    off[i + 1] = off[i] + n3 * n3;
    poff[i + 1] = poff[i] + n3;
    red_maxn_ = std::max<int>(red_maxn_, (int)n3);
  }
  if ((size_t)red_maxn_ * sizeof(double) > 64 * 1024)
    throw Error(NNK_ESTATE, "reducedSolve: leaf block larger than 8192 unknowns");
  std::vector<int32_t> o2l(T.ntgt, 0);
  {
    std::vector<int32_t> to(T.ntgt);
    NNK_HIP(hipMemcpy(to.data(), tgt_out_.p, sizeof(int32_t) * T.ntgt, hipMemcpyDeviceToHost));
    for (int i = 0; i < n_tleaves_; i++)
      for (int k = tp[tl[i]]; k < tp[tl[i] + 1]; k++) o2l[to[k]] = k - tp[tl[i]];
  }
  red_off_.upload(off);
  DBuf dpoff;
  dpoff.upload(poff);
  red_ord2loc_.upload(o2l);
  red_blk_.alloc(sizeof(double) * std::max<int64_t>(off.back(), 1));
  red_piv_.alloc(sizeof(int32_t) * std::max<int64_t>(poff.back(), 1));
  red_perm_.alloc(sizeof(int32_t) * std::max<int64_t>(poff.back(), 1));
  red_poff_ = std::move(dpoff);
  RedArgs a = red_args();
  if (n_tleaves_) {
    hipLaunchKernelGGL(k_red_form, dim3(n_tleaves_), dim3(256), 0, 0, a);
    NNK_HIP(hipGetLastError());
    for (int k0 = 0; k0 < red_maxn_; k0 += kLuNB) {
      hipLaunchKernelGGL(k_red_lu_panel, dim3(n_tleaves_), dim3(kLuThreads), 0, 0, a, k0);
      const int m = red_maxn_ - k0 - kLuNB;
      if (m > 0) {
        const int nt = (m + kLuT - 1) / kLuT;
        hipLaunchKernelGGL(k_red_lu_update, dim3(n_tleaves_, nt * nt), dim3(kLuThreads), 0, 0, a, k0);
      }
    }
    NNK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_red_transpose, dim3(n_tleaves_), dim3(256), 0, 0, a);
    hipLaunchKernelGGL(k_red_perm, dim3((n_tleaves_ + 63) / 64), dim3(64), 0, 0, a, n_tleaves_);
    NNK_HIP(hipGetLastError());
    // the blocks' inverses replace their factors: each solve is then one
    // pass over them at HBM bandwidth instead of a chain of substitutions
    DBuf W;
    W.alloc(sizeof(double) * std::max<int64_t>(off.back(), 1));
    hipLaunchKernelGGL(k_red_inv, dim3(n_tleaves_, (red_maxn_ + kRedInvCols - 1) / kRedInvCols),
                       dim3(kRedInvCols * kRedInvWaves), 0, 0, a, W.as<double>());
    NNK_HIP(hipGetLastError());
    RedArgs w = a;
    w.blk = W.as<double>();
    hipLaunchKernelGGL(k_red_transpose, dim3(n_tleaves_), dim3(256), 0, 0, w);
    NNK_HIP(hipGetLastError());
    NNK_HIP(hipDeviceSynchronize());
    red_blk_ = std::move(W);
  }
  red_piv_ = DBuf();
  red_perm_ = DBuf();
  NNK_HIP(hipDeviceSynchronize());
  red_ready_ = true;
}

RedArgs Fmm::red_args() const {
  RedArgs a{};
  a.leaves = tleaves_.as<int32_t>();
  a.tgt_ptr = tgt_ptr_.as<int32_t>();
  a.tgt_pos = tgt_pos_.as<double>();
  a.tgt_out = tgt_out_.as<int32_t>();
  a.src_ptr = src_ptr_.as<int32_t>();
  a.src_pos = src_pos_.as<double>();
  a.ctr_ptr = ctr_ptr_.as<int64_t>();
  a.ctr_col = ctr_col_.as<int32_t>();
  a.ctr_w = ctr_w_.as<double>();
  a.ord2loc = red_ord2loc_.as<int32_t>();
  a.off = red_off_.as<int64_t>();
  a.blk = red_blk_.as<double>();
  a.piv = red_piv_.as<int32_t>();
  a.perm = red_perm_.as<int32_t>();
  a.poff = red_poff_.as<int64_t>();
  a.e2 = eps_ * eps_;
  a.c = 1.0 / (8.0 * kPi * mu_);
  a.nearest = tree_.nearest ? 1 : 0;
  return a;
}

void Fmm::reduced_solve(const double *d_v, double *d_f, hipStream_t s) {
  if (!red_ready_) reduced_factor();
  if (!n_tleaves_) return;
  RedArgs a = red_args();
  hipLaunchKernelGGL(k_red_apply, dim3(n_tleaves_, (red_maxn_ + kRedApplyThreads - 1) / kRedApplyThreads),
                     dim3(kRedApplyThreads), sizeof(double) * red_maxn_, s, a, d_v, d_f);
  NNK_HIP(hipGetLastError());
}

}  // namespace nnk
