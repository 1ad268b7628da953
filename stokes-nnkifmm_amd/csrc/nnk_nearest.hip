// Nearest-neighbour quadrature map on the GPU (Tools/Nearest/nearestMatrix.m):
// NN(q, j) = 1/#ties for every coarse point j at the minimum squared
// distance from fine point q.  Brute force like the reference, one thread
// per fine point with coarse points staged in LDS.  The squared distance is
// formed exactly as the reference sums it, ((dx^2 + dy^2) + dz^2) with each
// operation rounded (no fused multiply-add), so ties are decided on the same
// floating-point values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "nnk_common.hpp"

namespace nnk {

__device__ __forceinline__ double d2_exact(const double *f, double cx, double cy, double cz) {
  double dx = __dsub_rn(f[0], cx), dy = __dsub_rn(f[1], cy), dz = __dsub_rn(f[2], cz);
  return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// pass 1: minimum and tie count per fine point
__global__ __launch_bounds__(256) void k_nn_min(const double *__restrict__ coarse, int64_t n,
                                                const double *__restrict__ fine, int64_t q,
                                                double *__restrict__ dmin, int32_t *__restrict__ cnt) {
  __shared__ double cs[256 * 3];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double f[3] = {0, 0, 0};
  if (i < q) { f[0] = fine[3 * i]; f[1] = fine[3 * i + 1]; f[2] = fine[3 * i + 2]; }
  double best = __builtin_inf();
  int c = 0;
  for (int64_t j0 = 0; j0 < n; j0 += 256) {
    const int m = (int)min<int64_t>(256, n - j0);
    if (threadIdx.x < m)
      for (int d = 0; d < 3; d++) cs[3 * threadIdx.x + d] = coarse[3 * (j0 + threadIdx.x) + d];
    __syncthreads();
    if (i < q)
      for (int j = 0; j < m; j++) {
        double d2 = d2_exact(f, cs[3 * j], cs[3 * j + 1], cs[3 * j + 2]);
        if (d2 < best) { best = d2; c = 1; }
        else if (d2 == best) c++;
      }
    __syncthreads();
  }
  if (i < q) { dmin[i] = best; cnt[i] = c; }
}

// pass 2: emit the tied columns of each fine point, ascending
__global__ __launch_bounds__(256) void k_nn_emit(const double *__restrict__ coarse, int64_t n,
                                                 const double *__restrict__ fine, int64_t q,
                                                 const double *__restrict__ dmin,
                                                 const int64_t *__restrict__ off,
                                                 int64_t *__restrict__ col) {
  __shared__ double cs[256 * 3];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double f[3] = {0, 0, 0};
  double best = 0;
  int64_t o = 0;
  if (i < q) {
    f[0] = fine[3 * i]; f[1] = fine[3 * i + 1]; f[2] = fine[3 * i + 2];
    best = dmin[i];
    o = off[i];
  }
  for (int64_t j0 = 0; j0 < n; j0 += 256) {
    const int m = (int)min<int64_t>(256, n - j0);
    if (threadIdx.x < m)
      for (int d = 0; d < 3; d++) cs[3 * threadIdx.x + d] = coarse[3 * (j0 + threadIdx.x) + d];
    __syncthreads();
    if (i < q)
      for (int j = 0; j < m; j++)
        if (d2_exact(f, cs[3 * j], cs[3 * j + 1], cs[3 * j + 2]) == best) col[o++] = j0 + j;
    __syncthreads();
  }
}

// Returns nnz; fills CSC (colptr n+1, row, val) when nnz <= cap.
int64_t nearest_matrix(const double *coarse, int64_t n, const double *fine, int64_t q,
                       int64_t *colptr, int64_t *row, double *val, int64_t cap) {
  if (n <= 0) throw Error(1, "nearestMatrix: no coarse points");
  DBuf dc, df, dmin, dcnt;
  dc.alloc(sizeof(double) * 3 * n);
  df.alloc(sizeof(double) * 3 * std::max<int64_t>(q, 1));
  dmin.alloc(sizeof(double) * std::max<int64_t>(q, 1));
  dcnt.alloc(sizeof(int32_t) * std::max<int64_t>(q, 1));
  NNK_HIP(hipMemcpy(dc.p, coarse, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  if (q) NNK_HIP(hipMemcpy(df.p, fine, sizeof(double) * 3 * q, hipMemcpyHostToDevice));
  const unsigned nb = (unsigned)((q + 255) / 256);
  if (q) {
    hipLaunchKernelGGL(k_nn_min, dim3(nb), dim3(256), 0, 0, dc.as<double>(), n, df.as<double>(), q,
                       dmin.as<double>(), dcnt.as<int32_t>());
    NNK_HIP(hipGetLastError());
  }
  std::vector<int32_t> cnt(q);
  if (q) NNK_HIP(hipMemcpy(cnt.data(), dcnt.p, sizeof(int32_t) * q, hipMemcpyDeviceToHost));
  std::vector<int64_t> off(q + 1, 0);
  for (int64_t i = 0; i < q; i++) off[i + 1] = off[i] + cnt[i];
  const int64_t nnz = off[q];
  if (nnz > cap) return nnz;
  DBuf doff, dcol;
  doff.upload(off);
  dcol.alloc(sizeof(int64_t) * std::max<int64_t>(nnz, 1));
  if (q) {
    hipLaunchKernelGGL(k_nn_emit, dim3(nb), dim3(256), 0, 0, dc.as<double>(), n, df.as<double>(), q,
                       dmin.as<double>(), doff.as<int64_t>(), dcol.as<int64_t>());
    NNK_HIP(hipGetLastError());
  }
  std::vector<int64_t> cols(nnz);
  if (nnz) NNK_HIP(hipMemcpy(cols.data(), dcol.p, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost));
  // COO grouped by row (ascending) -> CSC, rows ascending within a column
  std::fill(colptr, colptr + n + 1, 0);
  for (int64_t e = 0; e < nnz; e++) colptr[cols[e] + 1]++;
  for (int64_t j = 0; j < n; j++) colptr[j + 1] += colptr[j];
  std::vector<int64_t> fill(colptr, colptr + n);
  for (int64_t i = 0; i < q; i++)
    for (int64_t e = off[i]; e < off[i + 1]; e++) {
      int64_t k = fill[cols[e]]++;
      row[k] = i;
      val[k] = 1.0 / (double)cnt[i];
    }
  return nnz;
}

}  // namespace nnk
