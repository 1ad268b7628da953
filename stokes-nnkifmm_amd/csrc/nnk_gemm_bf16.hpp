// Split-BF16 gathered-column block GEMM on the BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16), for the mixed-precision M2L:
//
//   D[:, cidx[n]] = sum_b A_b X[:, gidx[b][n]]      (D in FP64)
//
// Every operand is stored as two BF16 parts, hi = bf16(v), lo = bf16(v - hi)
// (16 significant bits together).  Each tile product takes the three terms
// hi*hi + hi*lo + lo*hi with FP32 accumulation; after every operator block
// (one translation, K = R) the FP32 accumulators are added into FP64 ones,
// so rounding does not grow with the 189-block sum.
//
// A_b parts are row-major [Mrows][KP] (k contiguous), X parts node-major
// [node][KP]: both MFMA operands are then 16-byte contiguous per lane
// (lane l holds A[l%16][8(l/16) .. 8(l/16)+7] and B[8(l/16) .. +7][l%16]).
// Grid walk, active-block lists and the register-prefetch LDS ping-pong
// follow the FP64 kernel (nnk_gemm.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nnk_fmm.hpp"

#ifndef NNK_BF16_DROW
// accumulator register r of lane l holds D[NNK_BF16_DROW(l/16, r)][l%16]
#define NNK_BF16_DROW(g, r) (4 * (g) + (r))
#endif

namespace nnk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct GemmJobBf16 {
  const uint16_t *const *Ahi;  // nb operator parts, [Mrows][KP] bf16
  const uint16_t *const *Alo;
  const int32_t *gidx;         // nb x N (-1 = none)
  const int32_t *cidx;         // N output slots
  int nb, N;
};

template <int BM, int BN, int TAG>
__global__ __launch_bounds__(256, 2) void k_gemm_bf16x3(int M, int KP, const uint16_t *__restrict__ Xhi,
                                                        const uint16_t *__restrict__ Xlo, int ldx,
                                                        double *__restrict__ C, int ldc,
                                                        const GemmJobBf16 *__restrict__ jobs, int njobs,
                                                        int nrb) {
  constexpr int BK = 32;                       // one MFMA k-step
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int AROW = BK + 8;                 // padded LDS row (bf16)
  constexpr int AV = BM * BK / 8;              // 16-byte vectors per A part per chunk
  constexpr int BV = BN * BK / 8;
  constexpr int AR = AV / 256 > 0 ? AV / 256 : 1, BR = BV / 256 > 0 ? BV / 256 : 1;
  static_assert(AV % 256 == 0 || AV < 256, "A chunk split");
  static_assert(BV % 256 == 0 || BV < 256, "B chunk split");
  __shared__ __attribute__((aligned(16))) uint16_t As[2][2][BM][AROW];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][2][BN][AROW];
  __shared__ int act[512];
  __shared__ int nact_s;

  const int bid = blockIdx.x;
  const int jb = bid % njobs;
  const int rr = bid / njobs;
  const int rb = rr % nrb, ct = rr / nrb;
  const GemmJobBf16 J = jobs[jb];
  const int col0 = ct * BN;
  if (col0 >= J.N) return;
  const int m0 = rb * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lg = lane >> 4;

  if (tid == 0) nact_s = 0;
  __syncthreads();
  for (int b = tid; b < J.nb; b += 256) {
    bool any = false;
    const int32_t *g = J.gidx + (size_t)b * J.N;
    for (int n = 0; n < BN && !any; n++)
      if (col0 + n < J.N && g[col0 + n] >= 0) any = true;
    if (any) act[atomicAdd(&nact_s, 1)] = b;
  }
  __syncthreads();
  const int nact = nact_s;
  if (tid == 0)
    for (int i = 1; i < nact; i++) {
      int v = act[i], j = i - 1;
      while (j >= 0 && act[j] > v) { act[j + 1] = act[j]; j--; }
      act[j + 1] = v;
    }
  __syncthreads();

  f32x4 acc[TM][TN];
  double acc64[TM][TN][4];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      acc[i][j] = f32x4{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; r++) acc64[i][j][r] = 0.0;
    }

  const int KC = KP / BK;
  const int total = nact * KC;
  u32x4 ra[2][AR], rbv[2][BR];
  int gcol[BR];
  int cur_block = -1;

  auto fetch = [&](int c) {
    const int bi = c / KC, kc = c - bi * KC;
    const int b = act[bi];
    const int k0 = kc * BK;
    if (b != cur_block) {
      cur_block = b;
#pragma unroll
      for (int r = 0; r < BR; r++) {
        const int i = tid + 256 * r;
        const int n = i / (BK / 8);
        const int col = col0 + n;
        gcol[r] = (i < BV && col < J.N) ? J.gidx[(size_t)b * J.N + col] : -1;
      }
    }
#pragma unroll
    for (int p = 0; p < 2; p++) {
      const uint16_t *__restrict__ Ab = p ? J.Alo[b] : J.Ahi[b];
#pragma unroll
      for (int r = 0; r < AR; r++) {
        const int i = tid + 256 * r;
        if (i < AV) {
          const int m = i / (BK / 8), k8 = i % (BK / 8);
          ra[p][r] = *reinterpret_cast<const u32x4 *>(Ab + (size_t)(m0 + m) * KP + k0 + 8 * k8);
        }
      }
      const uint16_t *__restrict__ Xp = p ? Xlo : Xhi;
#pragma unroll
      for (int r = 0; r < BR; r++) {
        const int i = tid + 256 * r;
        const int k8 = i % (BK / 8);
        u32x4 v = {0, 0, 0, 0};
        if (gcol[r] >= 0) v = *reinterpret_cast<const u32x4 *>(Xp + (size_t)gcol[r] * ldx + k0 + 8 * k8);
        rbv[p][r] = v;
      }
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 2; p++) {
#pragma unroll
      for (int r = 0; r < AR; r++) {
        const int i = tid + 256 * r;
        if (i < AV) {
          const int m = i / (BK / 8), k8 = i % (BK / 8);
          *reinterpret_cast<u32x4 *>(&As[buf][p][m][8 * k8]) = ra[p][r];
        }
      }
#pragma unroll
      for (int r = 0; r < BR; r++) {
        const int i = tid + 256 * r;
        if (i < BV) {
          const int n = i / (BK / 8), k8 = i % (BK / 8);
          *reinterpret_cast<u32x4 *>(&Bs[buf][p][n][8 * k8]) = rbv[p][r];
        }
      }
    }
  };

  if (total > 0) {
    fetch(0);
    stash(0);
    __syncthreads();
    for (int c = 0; c < total; c++) {
      const int buf = c & 1;
      if (c + 1 < total) fetch(c + 1);
      bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) {
        ah[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][0][wm * WM + i * 16 + lr][8 * lg]);
        al[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][1][wm * WM + i * 16 + lr][8 * lg]);
      }
#pragma unroll
      for (int j = 0; j < TN; j++) {
        bh[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][0][wn * WN + j * 16 + lr][8 * lg]);
        bl[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][1][wn * WN + j * 16 + lr][8 * lg]);
      }
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      // end of an operator block: promote to FP64
      if ((c + 1) % KC == 0) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++) {
#pragma unroll
            for (int r = 0; r < 4; r++) acc64[i][j][r] += (double)acc[i][j][r];
            acc[i][j] = f32x4{0, 0, 0, 0};
          }
      }
      if (c + 1 < total) stash(buf ^ 1);
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int col = col0 + wn * WN + j * 16 + lr;
      if (col >= J.N) continue;
      const int cs = J.cidx ? J.cidx[col] : col;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = m0 + wm * WM + i * 16 + NNK_BF16_DROW(lg, r);
        if (row >= M) continue;
        C[(size_t)cs * ldc + row] = acc64[i][j][r];
      }
    }
}

}  // namespace nnk
