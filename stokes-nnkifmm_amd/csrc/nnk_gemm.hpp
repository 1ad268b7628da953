// Gathered-column block GEMM on the FP64 matrix cores (v_mfma_f64_16x16x4_f64):
//
//   C[:, cidx[n]] (+)= sum_b A_b X[:, gidx[b][n]]
//
// A_b are (3Nq x 3Nq) operators stored k-major (At[k][m], Kpad x Mpad, zero
// padded); X and C hold one density vector of length K per node.  Used for
// M2L (189 translation operators per child class), M2M (8 child
// operators), L2L and the two-factor check-to-equivalent solves.
//
// Work decomposition: the flattened grid is walked job-fastest, so with the
// hardware's round-robin dispatch of workgroups to the 8 XCDs the 8 M2L jobs
// (one per child class) each stay on one XCD and stream their operators
// through that XCD's L2.  A workgroup owns a BM x BN tile of C; its 4 waves
// form a 2x2 grid of (BM/2) x (BN/2) sub-tiles.  K is walked in BK-deep
// chunks over the operator blocks that have at least one source in the tile,
// through a two-stage LDS ping-pong fed by register prefetch (one barrier per
// chunk).  With nsplit > 1 the active blocks are divided over nsplit
// workgroups (split-K, for levels with few tiles); each writes its partial
// tile to P[(split, job, column)] and a fixed-order reduction forms C.
// MFMA operand layout (measured on gfx950): lane l holds A[l%16][l/16]
// and B[l/16][l%16]; accumulator register r of lane l holds D[l/16 + 4r][l%16].
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nnk_fmm.hpp"

namespace nnk {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

// TAG only separates launches by role in profiles (1 = M2L translations).
template <int BM, int BN, int BK, int TAG>
__global__ __launch_bounds__(256, 2) void k_gemm_gather(int M, int K, int Mpad, int Kpad,
                                                        const double *__restrict__ X, int ldx,
                                                        double *__restrict__ C, int ldc, int beta,
                                                        const GemmJob *__restrict__ jobs,
                                                        int njobs, int nrb, int nsplit,
                                                        double *__restrict__ P, int maxN) {
  static_assert(BK == 16 && BM % 32 == 0 && BN % 32 == 0, "tile shape");
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int A2 = BK * BM / 2;       // double2 per A chunk
  constexpr int B2 = BN * BK / 2;       // double2 per B chunk
  constexpr int AR = A2 / 256, BR = B2 / 256;
  static_assert(A2 % 256 == 0 && B2 % 256 == 0, "chunk must split evenly over 256 threads");
  constexpr int BKP = BK + 2;           // padded B row (keeps 16-byte alignment)
  __shared__ __attribute__((aligned(16))) double As[2][BK][BM];
  __shared__ __attribute__((aligned(16))) double Bs[2][BN][BKP];
  __shared__ int act[512];
  __shared__ int nact_s;

  const int bid = blockIdx.x;
  const int jb = bid % njobs;
  int rr = bid / njobs;
  const int sp = rr % nsplit;
  rr /= nsplit;
  const int rb = rr % nrb, ct = rr / nrb;
  const GemmJob J = jobs[jb];
  const int col0 = ct * BN;
  if (col0 >= J.N) return;
  const int m0 = rb * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  // ---- operator blocks with at least one source column in this tile ----
  int nact;
  if (J.tptr) {  // precomputed per tile
    const int t0 = J.tptr[ct];
    nact = J.tptr[ct + 1] - t0;
    for (int i = tid; i < nact; i += 256) act[i] = J.tlist[t0 + i];
    __syncthreads();
  } else {
    if (tid == 0) nact_s = 0;
    __syncthreads();
    for (int b = tid; b < J.nb; b += 256) {
      bool any = false;
      if (J.gidx) {
        const int32_t *g = J.gidx + (size_t)b * J.N;
        for (int n = 0; n < BN && !any; n++)
          if (col0 + n < J.N && g[col0 + n] >= 0) any = true;
      } else {
        any = true;
      }
      if (any) {
        int slot = atomicAdd(&nact_s, 1);
        act[slot] = b;
      }
    }
    __syncthreads();
    nact = nact_s;
    // keep the walk in block order (deterministic summation order); only
    // small jobs come here (at most 8 blocks)
    if (tid == 0) {
      for (int i = 1; i < nact; i++) {
        int v = act[i], j = i - 1;
        while (j >= 0 && act[j] > v) { act[j + 1] = act[j]; j--; }
        act[j + 1] = v;
      }
    }
    __syncthreads();
  }

  d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = d4{0, 0, 0, 0};

  const int KC = Kpad / BK;
  const int a0 = (int)((int64_t)nact * sp / nsplit), a1 = (int)((int64_t)nact * (sp + 1) / nsplit);
  const int total = (a1 - a0) * KC;
  d2v ra[AR], rbv[BR];
  int gcol[BR];
  int cur_block = -1;

  auto fetch = [&](int c) {
    const int q = c / KC, kc = c - q * KC;
    const int b = act[a0 + q];
    const int k0 = kc * BK;
    if (b != cur_block) {
      cur_block = b;
#pragma unroll
      for (int r = 0; r < BR; r++) {
        const int i = tid + 256 * r;
        const int n = i / (BK / 2);
        const int col = col0 + n;
        int g = -1;
        if (col < J.N) g = J.gidx ? J.gidx[(size_t)b * J.N + col] : col;
        gcol[r] = g;
      }
    }
    const double *__restrict__ Ab = J.A[b];
#pragma unroll
    for (int r = 0; r < AR; r++) {
      const int i = tid + 256 * r;
      const int kk = i / (BM / 2), m2 = i - kk * (BM / 2);
      ra[r] = *reinterpret_cast<const d2v *>(Ab + (size_t)(k0 + kk) * Mpad + m0 + 2 * m2);
    }
#pragma unroll
    for (int r = 0; r < BR; r++) {
      const int i = tid + 256 * r;
      const int kk2 = i % (BK / 2);
      const int k = k0 + 2 * kk2;
      d2v v = {0.0, 0.0};
      if (gcol[r] >= 0) {
        const double *xp = X + (size_t)gcol[r] * ldx + k;
        if (k + 1 < K) v = d2v{xp[0], xp[1]};
        else if (k < K) v = d2v{xp[0], 0.0};
      }
      rbv[r] = v;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int r = 0; r < AR; r++) {
      const int i = tid + 256 * r;
      const int kk = i / (BM / 2), m2 = i - kk * (BM / 2);
      *reinterpret_cast<d2v *>(&As[buf][kk][2 * m2]) = ra[r];
    }
#pragma unroll
    for (int r = 0; r < BR; r++) {
      const int i = tid + 256 * r;
      const int n = i / (BK / 2), kk2 = i % (BK / 2);
      *reinterpret_cast<d2v *>(&Bs[buf][n][2 * kk2]) = rbv[r];
    }
  };

  if (total > 0) {
    fetch(0);
    stash(0);
    __syncthreads();
    for (int c = 0; c < total; c++) {
      const int buf = c & 1;
      if (c + 1 < total) fetch(c + 1);
#pragma unroll
      for (int ks = 0; ks < BK; ks += 4) {
        double av[TM], bv[TN];
#pragma unroll
        for (int i = 0; i < TM; i++) av[i] = As[buf][ks + lk][wm * WM + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < TN; j++) bv[j] = Bs[buf][wn * WN + j * 16 + lr][ks + lk];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
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
        const int row = m0 + wm * WM + i * 16 + lk + 4 * r;
        if (row >= M) continue;
        if (nsplit > 1) {
          P[((size_t)(sp * njobs + jb) * maxN + col) * ldc + row] = acc[i][j][r];
        } else {
          double *p = C + (size_t)cs * ldc + row;
          *p = beta ? *p + acc[i][j][r] : acc[i][j][r];
        }
      }
    }
}

// C[slot[c]][row] = sum over splits of P[split][c][row], splits in order
// (c runs over job * maxN + column; slot -1 = no column)
__global__ void k_split_reduce(const int32_t *__restrict__ slot, int M, int ldc, int nsplit,
                               const double *__restrict__ P, size_t pstride, double *__restrict__ C);

}  // namespace nnk
