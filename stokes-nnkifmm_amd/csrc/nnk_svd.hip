// SVD of the per-level check-to-equivalent matrices on the GPU: one-sided
// Jacobi (Hestenes 1958; Demmel & Veselic 1992) preconditioned by two QR
// factorizations (Drmac & Veselic 2008).  These matrices are square, dense
// and ill-conditioned (cond ~ 2e10 at 6 points per edge, ~1e15 at 9);
// applying the factors as x = V (S^-1 (U^T b)) is backward stable like the
// reference's pivoted LU (Up.m:31, DownNonleaf.m:46), whereas an explicit
// inverse is not.
//
//   A P = Q R        Householder QR with column pivoting (largest remaining
//                    column norm first, norms recomputed at every step)
//   R^T = Q1 R1      Householder QR, L = R1^T lower triangular
//   L V = U_L S      one-sided Jacobi on the columns of L, V accumulated
//   A = (Q U_L) S (P Q1 V)^T
// Jacobi on the raw matrix spends ~50 sweeps before its quadratic phase;
// on L it needs ~12-15 (measured on these matrices), and the singular values
// come out closer to LAPACK's.
//
// Layout: every matrix is column-major (column j contiguous), and the nb
// matrices of a batch (one per tree level) lie back to back; grid.y is the
// matrix.  QR: per column k one launch picks the pivot, swaps and forms the
// reflector (one workgroup per matrix), one launch applies it to the
// trailing columns (one workgroup per column, which also recomputes that
// column's remaining norm).  Jacobi: a round-robin tournament over m = n
// (+1 if odd) positions pairs every column with every other once per sweep
// in m - 1 steps of m/2 disjoint pairs (position 0 stays, 1..m-1 shift by
// one per step; pair k of a step is (ring[k], ring[m-1-k])); one launch per
// step, one workgroup per pair, both columns held in registers.  Sweeps stop
// when every pair of every matrix meets |g_p.g_q| <= sqrt(n) eps |g_p| |g_q|
// (the criterion of LAPACK's dgesvj), or after 60; the sweep's largest
// ratio is kept by an integer atomic max on its bit pattern (positive
// doubles order as their bits), read by the host once per sweep.  All sums
// are workgroup reductions in a fixed order.  Setup-only code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "nnk_common.hpp"

namespace nnk {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 12;  // n <= 3072 (12 points per edge: 3 Nq = 2184)
constexpr int kMaxSweeps = 60;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum of v over the 256 threads (same value in every thread)
__device__ __forceinline__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; k++) t += red[k];
  return t;
}

// Column k of matrix blockIdx.y: with pivoting, swap in the column of
// largest remaining norm (ties: the smallest index); then the Householder
// reflector H = I - tau v v^T (v_k = 1) with H x = beta e_k for the column's
// rows k..n-1: v below the diagonal, beta on it, tau in taus[k].
__global__ __launch_bounds__(kThreads) void k_qr_reflector(double *__restrict__ G, int n, int k,
                                                           double *__restrict__ norms, int *__restrict__ perm,
                                                           double *__restrict__ taus, int pivot) {
  __shared__ double red[kThreads / 64];
  __shared__ double bv[kThreads];
  __shared__ int bi[kThreads];
  const size_t nn = (size_t)n * n;
  G += nn * blockIdx.y;
  taus += (size_t)n * blockIdx.y;
  const int tid = threadIdx.x;
  if (pivot) {
    norms += (size_t)n * blockIdx.y;
    perm += (size_t)n * blockIdx.y;
    double best = -1.0;
    int bj = n;
    for (int j = k + tid; j < n; j += kThreads)
      if (norms[j] > best) { best = norms[j]; bj = j; }
    bv[tid] = best;
    bi[tid] = bj;
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
      if (tid < st) {
        const double o = bv[tid + st];
        const int oj = bi[tid + st];
        if (o > bv[tid] || (o == bv[tid] && oj < bi[tid])) { bv[tid] = o; bi[tid] = oj; }
      }
      __syncthreads();
    }
    const int p = bi[0];
    if (p != k && p < n) {
      double *ck = G + (size_t)k * n, *cp = G + (size_t)p * n;
      for (int i = tid; i < n; i += kThreads) {
        const double t = ck[i];
        ck[i] = cp[i];
        cp[i] = t;
      }
      if (tid == 0) {
        const double t = norms[k];
        norms[k] = norms[p];
        norms[p] = t;
        const int q = perm[k];
        perm[k] = perm[p];
        perm[p] = q;
      }
    }
    __syncthreads();
  }
  double *x = G + (size_t)k * n;
  double sig = 0;
  for (int i = k + 1 + tid; i < n; i += kThreads) sig += x[i] * x[i];
  sig = block_sum(sig, red);
  const double alpha = x[k];
  if (sig == 0.0) {  // already e_k times alpha: no reflection
    if (tid == 0) taus[k] = 0.0;
    return;
  }
  const double beta = -copysign(sqrt(alpha * alpha + sig), alpha);
  const double scale = 1.0 / (alpha - beta);
  for (int i = k + 1 + tid; i < n; i += kThreads) x[i] *= scale;
  if (tid == 0) {
    x[k] = beta;
    taus[k] = (beta - alpha) / beta;
  }
}

// M[:, c] = H_k M[:, c] for columns c = c0 + blockIdx.x of matrix
// blockIdx.y, H_k the reflector stored in column k of Vh (v_k = 1, v_i =
// Vh[i, k] below); with norms, the column's squared norm below row k is
// recomputed
__global__ __launch_bounds__(kThreads) void k_reflect(const double *__restrict__ Vh, int n, int k,
                                                      const double *__restrict__ taus, double *__restrict__ M,
                                                      int c0, double *__restrict__ norms) {
  __shared__ double red[kThreads / 64];
  const size_t nn = (size_t)n * n;
  const double tau = taus[(size_t)n * blockIdx.y + k];
  if (tau == 0.0 && !norms) return;
  const double *v = Vh + nn * blockIdx.y + (size_t)k * n;
  const int c = c0 + blockIdx.x;
  double *col = M + nn * blockIdx.y + (size_t)c * n;
  const int tid = threadIdx.x;
  double w = tid == 0 ? col[k] : 0.0;
  for (int i = k + 1 + tid; i < n; i += kThreads) w += v[i] * col[i];
  w = block_sum(w, red) * tau;
  if (tid == 0) col[k] -= w;
  double s2 = 0;
  for (int i = k + 1 + tid; i < n; i += kThreads) {
    const double y = col[i] - v[i] * w;
    col[i] = y;
    s2 += y * y;
  }
  if (norms) {
    s2 = block_sum(s2, red);
    if (tid == 0) norms[(size_t)n * blockIdx.y + c] = s2;
  }
}

// squared column norms of every matrix
__global__ __launch_bounds__(kThreads) void k_col_norms(const double *__restrict__ G, int n,
                                                        double *__restrict__ norms) {
  __shared__ double red[kThreads / 64];
  const double *col = G + (size_t)n * n * blockIdx.y + (size_t)blockIdx.x * n;
  double s2 = 0;
  for (int i = threadIdx.x; i < n; i += kThreads) s2 += col[i] * col[i];
  s2 = block_sum(s2, red);
  if (threadIdx.x == 0) norms[(size_t)n * blockIdx.y + blockIdx.x] = s2;
}

// X = (upper triangle of R)^T, lower triangular: X[i, j] = R[j, i] for
// i >= j, 0 above (column j of X per workgroup)
__global__ __launch_bounds__(kThreads) void k_upper_transpose(const double *__restrict__ R, int n,
                                                              double *__restrict__ X) {
  const size_t nn = (size_t)n * n;
  const double *r = R + nn * blockIdx.y;
  double *x = X + nn * blockIdx.y + (size_t)blockIdx.x * n;
  const int j = blockIdx.x;
  for (int i = threadIdx.x; i < n; i += kThreads) x[i] = i >= j ? r[(size_t)i * n + j] : 0.0;
}

// identity in every matrix (column per workgroup)
__global__ __launch_bounds__(kThreads) void k_identity(double *__restrict__ M, int n) {
  double *col = M + (size_t)n * n * blockIdx.y + (size_t)blockIdx.x * n;
  for (int i = threadIdx.x; i < n; i += kThreads) col[i] = i == (int)blockIdx.x ? 1.0 : 0.0;
}

__device__ __forceinline__ int ring_pos(int i, int step, int m) {
  if (i == 0) return 0;
  int r = (i - 1 - step) % (m - 1);
  if (r < 0) r += m - 1;
  return 1 + r;
}

__global__ __launch_bounds__(kThreads) void k_jacobi_step(double *__restrict__ G, double *__restrict__ Q, int n,
                                                          int m, int step, double tol,
                                                          unsigned long long *__restrict__ off) {
  __shared__ double red[3][kThreads / 64];
  int p = ring_pos(blockIdx.x, step, m), q = ring_pos(m - 1 - blockIdx.x, step, m);
  if (p >= n || q >= n) return;
  if (p > q) { const int t = p; p = q; q = t; }
  const size_t nn = (size_t)n * n;
  double *gp = G + nn * blockIdx.y + (size_t)p * n, *gq = G + nn * blockIdx.y + (size_t)q * n;
  const int tid = threadIdx.x;
  double x[kPerThread], y[kPerThread];
  double a = 0, b = 0, c = 0;
#pragma unroll
  for (int j = 0; j < kPerThread; j++) {
    const int i = tid + j * kThreads;
    x[j] = i < n ? gp[i] : 0.0;
    y[j] = i < n ? gq[i] : 0.0;
    a += x[j] * x[j];
    b += y[j] * y[j];
    c += x[j] * y[j];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int w = tid >> 6;
  if ((tid & 63) == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = c; }
  __syncthreads();
  a = b = c = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; k++) { a += red[0][k]; b += red[1][k]; c += red[2][k]; }
  if (c == 0.0) return;
  const double rel = fabs(c) / sqrt(a * b);
  if (tid == 0) atomicMax(off, (unsigned long long)__double_as_longlong(rel));
  if (rel <= tol) return;
  const double zeta = (b - a) / (2.0 * c);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll
  for (int j = 0; j < kPerThread; j++) {
    const int i = tid + j * kThreads;
    if (i < n) {
      gp[i] = cs * x[j] - sn * y[j];
      gq[i] = sn * x[j] + cs * y[j];
    }
  }
  double *vp = Q + nn * blockIdx.y + (size_t)p * n, *vq = Q + nn * blockIdx.y + (size_t)q * n;
#pragma unroll
  for (int j = 0; j < kPerThread; j++) {
    const int i = tid + j * kThreads;
    if (i < n) {
      const double u = vp[i], v = vq[i];
      vp[i] = cs * u - sn * v;
      vq[i] = sn * u + cs * v;
    }
  }
}

// column j of every matrix divided by its norm s[j] (left alone if zero)
__global__ __launch_bounds__(kThreads) void k_normalize(double *__restrict__ G, int n, double *__restrict__ s) {
  __shared__ double red[kThreads / 64];
  double *col = G + (size_t)n * n * blockIdx.y + (size_t)blockIdx.x * n;
  double s2 = 0;
  for (int i = threadIdx.x; i < n; i += kThreads) s2 += col[i] * col[i];
  const double nrm = sqrt(block_sum(s2, red));
  if (threadIdx.x == 0) s[(size_t)n * blockIdx.y + blockIdx.x] = nrm;
  if (nrm > 0) {
    const double inv = 1.0 / nrm;
    for (int i = threadIdx.x; i < n; i += kThreads) col[i] *= inv;
  }
}

// Householder QR of every matrix in G (reflectors below the diagonal, R on
// and above it); with pivoting, perm receives the column order
void house_qr(double *G, int n, int nb, double *norms, int *perm, double *taus, bool pivot) {
  if (pivot) {
    hipLaunchKernelGGL(k_col_norms, dim3(n, nb), dim3(kThreads), 0, 0, G, n, norms);
    NNK_HIP(hipGetLastError());
  }
  for (int k = 0; k < n; k++) {
    hipLaunchKernelGGL(k_qr_reflector, dim3(1, nb), dim3(kThreads), 0, 0, G, n, k, norms, perm, taus,
                       pivot ? 1 : 0);
    if (k + 1 < n)
      hipLaunchKernelGGL(k_reflect, dim3(n - k - 1, nb), dim3(kThreads), 0, 0, G, n, k, taus, G, k + 1,
                         pivot ? norms : nullptr);
  }
  NNK_HIP(hipGetLastError());
}

// M = H_0 H_1 ... H_{n-1} M (the Q of house_qr times M)
void apply_q(const double *Vh, const double *taus, int n, int nb, double *M) {
  for (int k = n - 1; k >= 0; k--)
    hipLaunchKernelGGL(k_reflect, dim3(n, nb), dim3(kThreads), 0, 0, Vh, n, k, taus, M, 0, nullptr);
  NNK_HIP(hipGetLastError());
}

}  // namespace

// SVD of nb matrices of order n at once.  A[b]: n x n row-major (host).  On
// return U[b] (n x n row-major, columns = left singular vectors), s[b] (n,
// descending), V[b] (n x n row-major, columns = right singular vectors),
// with A = U diag(s) V^T.  Returns the Jacobi sweeps used.
int svd_batch(const double *const *A, int n, int nb, double *const *U, double *const *s, double *const *V) {
  if (n <= 0 || nb <= 0) return 0;
  if (n > kThreads * kPerThread) throw Error(5, "svd: matrix order above 3072");
  const size_t nn = (size_t)n * n;
  std::vector<double> h(nn * nb);
  for (int b = 0; b < nb; b++)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) h[nn * b + (size_t)j * n + i] = A[b][(size_t)i * n + j];
  DBuf G, X, L, Vj, norms, taus1, taus2, perm, sv, dOff;
  G.upload(h);
  X.alloc(sizeof(double) * nn * nb);
  L.alloc(sizeof(double) * nn * nb);
  Vj.alloc(sizeof(double) * nn * nb);
  norms.alloc(sizeof(double) * n * nb);
  taus1.alloc(sizeof(double) * n * nb);
  taus2.alloc(sizeof(double) * n * nb);
  sv.alloc(sizeof(double) * n * nb);
  std::vector<int> p0((size_t)n * nb);
  for (int b = 0; b < nb; b++) std::iota(p0.begin() + (size_t)n * b, p0.begin() + (size_t)n * (b + 1), 0);
  perm.upload(p0);
  dOff.alloc(sizeof(unsigned long long) * kMaxSweeps);
  NNK_HIP(hipMemset(dOff.p, 0, sizeof(unsigned long long) * kMaxSweeps));
  // A P = Q R (G), then R^T = Q1 R1 (X), then L = R1^T
  house_qr(G.as<double>(), n, nb, norms.as<double>(), perm.as<int>(), taus1.as<double>(), true);
  hipLaunchKernelGGL(k_upper_transpose, dim3(n, nb), dim3(kThreads), 0, 0, G.as<double>(), n, X.as<double>());
  house_qr(X.as<double>(), n, nb, nullptr, nullptr, taus2.as<double>(), false);
  hipLaunchKernelGGL(k_upper_transpose, dim3(n, nb), dim3(kThreads), 0, 0, X.as<double>(), n, L.as<double>());
  hipLaunchKernelGGL(k_identity, dim3(n, nb), dim3(kThreads), 0, 0, Vj.as<double>(), n);
  NNK_HIP(hipGetLastError());
  // one-sided Jacobi on the columns of L
  const int m = n + (n & 1);
  const double tol = std::sqrt((double)n) * 0x1p-53;
  int sweep = 0;
  for (; sweep < kMaxSweeps; sweep++) {
    unsigned long long *off = dOff.as<unsigned long long>() + sweep;
    for (int step = 0; step < m - 1; step++)
      hipLaunchKernelGGL(k_jacobi_step, dim3(m / 2, nb), dim3(kThreads), 0, 0, L.as<double>(), Vj.as<double>(), n,
                         m, step, tol, off);
    NNK_HIP(hipGetLastError());
    unsigned long long bits = 0;
    NNK_HIP(hipMemcpy(&bits, off, sizeof(bits), hipMemcpyDeviceToHost));
    double o;
    std::memcpy(&o, &bits, sizeof(o));
    if (o <= tol) break;
  }
  // U_A = Q U_L (U_L = normalised columns of L), V_A = P Q1 V
  hipLaunchKernelGGL(k_normalize, dim3(n, nb), dim3(kThreads), 0, 0, L.as<double>(), n, sv.as<double>());
  NNK_HIP(hipGetLastError());
  apply_q(G.as<double>(), taus1.as<double>(), n, nb, L.as<double>());
  apply_q(X.as<double>(), taus2.as<double>(), n, nb, Vj.as<double>());
  std::vector<double> hu(nn * nb), hv(nn * nb), hs((size_t)n * nb);
  std::vector<int> hp((size_t)n * nb);
  NNK_HIP(hipMemcpy(hu.data(), L.p, sizeof(double) * nn * nb, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(hv.data(), Vj.p, sizeof(double) * nn * nb, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(hs.data(), sv.p, sizeof(double) * n * nb, hipMemcpyDeviceToHost));
  NNK_HIP(hipMemcpy(hp.data(), perm.p, sizeof(int) * n * nb, hipMemcpyDeviceToHost));
  for (int b = 0; b < nb; b++) {
    const double *gu = hu.data() + nn * b, *gv = hv.data() + nn * b, *gs = hs.data() + (size_t)n * b;
    const int *pr = hp.data() + (size_t)n * b;
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return gs[x] > gs[y]; });
    for (int k = 0; k < n; k++) {
      const int j = ord[k];
      s[b][k] = gs[j];
      for (int i = 0; i < n; i++) {
        U[b][(size_t)i * n + k] = gu[(size_t)j * n + i];
        V[b][(size_t)pr[i] * n + k] = gv[(size_t)j * n + i];
      }
    }
  }
  return sweep + 1;
}

int svd_jacobi(const double *A, int n, double *U, double *s, double *V) {
  return svd_batch(&A, n, 1, &U, &s, &V);
}

}  // namespace nnk
