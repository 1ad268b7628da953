// Gathered-column block GEMM on the FP64 matrix cores (v_mfma_f64_16x16x4_f64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nnk_fmm.hpp"

namespace nnk {

typedef double d4 __attribute__((ext_vector_type(4)));

// C[:, cidx[n]] (+)= sum_b A_b X[:, gidx[b][n]] with A_b given k-major
// (At[k][m]) and node vectors of length K.  4 waves in a 2x2 layout, each
// computing a (BM/2) x (BN/2) tile with v_mfma_f64_16x16x4_f64.  Operand
// layout (measured): lane l holds A[l%16][l/16] and B[l/16][l%16];
// accumulator register r of lane l holds D[l/16 + 4r][l%16].
// TAG only separates launches by role in profiles (1 = M2L translations).
template <int BM, int BN, int BK, int TAG>
__global__ __launch_bounds__(256) void k_gemm_gather(int M, int K, int Mpad,
                                                     const double *__restrict__ X, int ldx,
                                                     double *__restrict__ C, int ldc, int beta,
                                                     const GemmJob *__restrict__ jobs) {
  const GemmJob J = jobs[blockIdx.z];
  const int col0 = blockIdx.x * BN;
  if (col0 >= J.N) return;
  const int m0 = blockIdx.y * BM;
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  __shared__ double As[BK][BM];
  __shared__ double Bs[BN][BK + 1];
  __shared__ int gcol[BN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  d4 acc[TM][TN];
  for (int i = 0; i < TM; i++)
    for (int j = 0; j < TN; j++) acc[i][j] = d4{0, 0, 0, 0};
  for (int b = 0; b < J.nb; b++) {
    int mine = 0;
    if (tid < BN) {
      const int c = col0 + tid;
      int g = -1;
      if (c < J.N) g = J.gidx ? J.gidx[(size_t)b * J.N + c] : c;
      gcol[tid] = g;
      mine = g >= 0;
    }
    if (!__syncthreads_or(mine)) continue;
    const double *__restrict__ Ab = J.A[b];
    for (int k0 = 0; k0 < K; k0 += BK) {
      for (int i = tid; i < BK * BM; i += 256) {
        const int kk = i / BM, mm = i - kk * BM;
        As[kk][mm] = Ab[(size_t)(k0 + kk) * Mpad + m0 + mm];
      }
      for (int i = tid; i < BK * BN; i += 256) {
        const int nn = i / BK, kk = i - nn * BK;
        const int g = gcol[nn];
        double v = 0.0;
        if (g >= 0 && k0 + kk < K) v = X[(size_t)g * ldx + k0 + kk];
        Bs[nn][kk] = v;
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        double av[TM], bv[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) av[i] = As[ks + lk][wm * WM + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < TN; j++) bv[j] = Bs[wn * WN + j * 16 + lr][ks + lk];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  for (int i = 0; i < TM; i++)
    for (int j = 0; j < TN; j++) {
      const int col = col0 + wn * WN + j * 16 + lr;
      if (col >= J.N) continue;
      const int cs = J.cidx ? J.cidx[col] : col;
      for (int r = 0; r < 4; r++) {
        const int row = m0 + wm * WM + i * 16 + lk + 4 * r;
        if (row >= M) continue;
        double *p = C + (size_t)cs * ldc + row;
        *p = beta ? *p + acc[i][j][r] : acc[i][j][r];
      }
    }
}

}  // namespace nnk
