// Split-precision gathered-column block GEMM on the 16-bit matrix cores
// (v_mfma_f32_16x16x32_bf16 / _f16), for the mixed-precision M2L:
//
//   D[:, cidx[n]] = sum_b A_b X[:, gidx[b][n]]      (D in FP64)
//
// Every operand is held as NP 16-bit parts, v ~ p0 + p1 (+ p2), each part
// the rounding of what the previous ones leave.  Products keep the terms
// down to the order of the last part:
//   NP = 2:  p1*q0 + p0*q1 + p0*q0                      (3 MFMAs)
//   NP = 3:  p2*q0 + p1*q1 + p0*q2 + p1*q0 + p0*q1 + p0*q0   (6 MFMAs)
// accumulated in FP32 over one operator block (K = R) and then added into
// FP64 accumulators, so rounding does not grow with the 189-block sum.
// BF16 parts carry 8 bits each (16 or 24 in total).  FP16 parts (F16, NP =
// 2) carry 11 bits each (22 in total) but need their operands in range: each
// operator row is stored scaled by a power of two (its unscale factor per
// row in U_b, applied when the block's FP32 sums are promoted) and the
// coefficients by one power of two per matvec (*xun, applied at the end).
//
// Storage is in MFMA fragment order, so that every global load, LDS store
// and LDS load is a contiguous 1 KB per wavefront:
//   A_b:  [KC][NP][RT][64 lanes][8]   lane q of row tile t holds
//                                     A[16t + q%16][32kc + 8(q/16) + e]
//   X:    [node][KC][NP][32]          (the 64-byte k-slice of one node)
// A gather slot of -1 reads node `zrow`, an all-zero row.  With beta = 1
// (and no split) the tile is added to D instead of stored.
//
// Work decomposition: one 512-thread workgroup owns a 128 x 128 tile of D
// (8 waves as 2 x 4, 64 x 32 each).  The flattened grid is walked
// job-fastest, so with the round-robin dispatch of workgroups to the 8 XCDs
// the 8 child-class jobs each stay on one XCD.  The operator blocks with at
// least one source in the tile are walked in order in 32-deep chunks through
// a two-stage LDS ping-pong fed by a ring of D register stages (loads run
// D-1 chunks ahead of the MFMAs).  Optionally the active blocks are split
// over `nsplit` workgroups (split-K, for levels with few tiles); each split
// then writes its partial tile to P for a fixed-order reduction.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "nnk_fmm.hpp"

namespace nnk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// global-address-space views (pointers read from memory are otherwise
// generic, and generic loads also hold the LDS counter the MFMA loop waits on)
typedef const u32x4 __attribute__((address_space(1))) *gu32x4p;
typedef const int32_t __attribute__((address_space(1))) *gi32p;

struct GemmJobBf16 {
  const uint16_t *const *A;  // nb operators in fragment order
  const float *const *U;     // F16 form: per operator, the unscale factor of each row
  const int32_t *gidx;       // nb x N gather slots (-1 = none)
  const int32_t *cidx;       // N output slots
  int nb, N;
  const int32_t *tptr, *tlist;  // per 128-column tile: its active blocks, in order
};

template <int NP, int D, int TAG, bool F16>
__global__ __launch_bounds__(512, 1) void k_m2l_split(const int M, const int KC, const int RT,
                                                      const uint16_t *__restrict__ X, const int zrow,
                                                      double *__restrict__ C, const int ldc,
                                                      const GemmJobBf16 *__restrict__ jobs, const int njobs,
                                                      const int nrb, const int nsplit, double *__restrict__ P,
                                                      const int maxN, const double *__restrict__ xun,
                                                      const int beta) {
  static_assert(!F16 || NP == 2, "FP16 parts: two");
  constexpr int BN = 128, TM = 4, TN = 2, TILE = 512;
  typedef typename std::conditional<F16, f16x8, bf16x8>::type op8;
  __shared__ __attribute__((aligned(16))) uint16_t As[2][NP][8][TILE];
  __shared__ __attribute__((aligned(16))) uint16_t Xs[2][NP][8][TILE];
  __shared__ int act[512];
  __shared__ const uint16_t *aptr[512];
  __shared__ const float *uptr[F16 ? 512 : 1];

  const int bid = blockIdx.x;
  const int jb = bid % njobs;
  int rr = bid / njobs;
  const int sp = rr % nsplit;
  rr /= nsplit;
  const int rb = rr % nrb, ct = rr / nrb;
  const GemmJobBf16 J = jobs[jb];
  const int col0 = ct * BN;
  if (col0 >= J.N) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  const int lr = lane & 15, lg = lane >> 4;

  // ---- operator blocks with at least one source column in this tile ----
  const int t0 = J.tptr[ct];
  const int nact = J.tptr[ct + 1] - t0;
  for (int i = tid; i < nact; i += 512) {
    const int b = J.tlist[t0 + i];
    act[i] = b;
    aptr[i] = J.A[b];
    if constexpr (F16) uptr[i] = J.U[b];
  }
  __syncthreads();
  const int a0 = (int)((int64_t)nact * sp / nsplit), a1 = (int)((int64_t)nact * (sp + 1) / nsplit);
  const int total = (a1 - a0) * KC;

  // this thread's load slot: fragment tile ti (0..7), lane slot tq
  const int ti = tid >> 6, tq = tid & 63;
  const int my_col = col0 + ti * 16 + (tq & 15);
  const int kof = 8 * (tq >> 4);
  const size_t aoff = (size_t)(rb * 8 + ti) * TILE + (size_t)tq * 8;
  const size_t slab = (size_t)RT * TILE;  // one (kc, part) slab of an operator

  // gather slot of this thread's column in active block bi (clamped
  // indices, so the load is issued unconditionally); the raw slot is only
  // examined when the block is reached, a block after the load
  const gi32p gidx = (gi32p)J.gidx;
  auto gload = [&](int bi) -> int {
    const int bc = bi < a1 ? bi : a1 - 1;
    const int cc = my_col < J.N ? my_col : J.N - 1;
    return gidx[(size_t)act[bc] * J.N + cc];
  };
  auto gslot = [&](int raw, int bi) -> int {
    return (bi < a1 && my_col < J.N && raw >= 0) ? raw : zrow;
  };
  int fblk = -1, gcur = zrow, gnext = total > 0 ? gload(a0) : -1;

  // loads of chunk c (clamped to the last chunk, so that every stage issues
  // the same loads and the wait counts stay static)
  auto fetch = [&](int c, u32x4(&R)[2 * NP]) {
    const int cq = c < total ? c : total - 1;
    const int q = cq / KC, kc = cq - q * KC, bi = a0 + q;
    if (bi != fblk) {  // first chunk of a new block: its slot was loaded a block ahead
      fblk = bi;
      gcur = gslot(gnext, bi);
      gnext = gload(bi + 1);
    }
    const gu32x4p Ab = (gu32x4p)(aptr[bi] + aoff + (size_t)kc * NP * slab);
    const gu32x4p Xb = (gu32x4p)(X + ((size_t)gcur * KC + kc) * NP * 32 + kof);
#pragma unroll
    for (int p = 0; p < NP; p++) {
      R[p] = Ab[p * slab / 8];
      R[NP + p] = Xb[p * 4];
    }
  };
  auto stash = [&](const u32x4(&R)[2 * NP], int s) {
#pragma unroll
    for (int p = 0; p < NP; p++) {
      *reinterpret_cast<u32x4 *>(&As[s][p][ti][tq * 8]) = R[p];
      *reinterpret_cast<u32x4 *>(&Xs[s][p][ti][tq * 8]) = R[NP + p];
    }
  };

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

  auto mfma = [](const op8 &a, const op8 &b, const f32x4 &c) -> f32x4 {
    if constexpr (F16) return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  };
  // the products term by term over all tiles (an accumulator's MFMAs are
  // TM TN instructions apart; per accumulator the order is unchanged)
  auto compute = [&](int s) {
    op8 a[NP][TM], x[NP][TN];
#pragma unroll
    for (int p = NP - 1; p >= 0; p--) {
#pragma unroll
      for (int i = 0; i < TM; i++) a[p][i] = *reinterpret_cast<const op8 *>(&As[s][p][wm * 4 + i][lane * 8]);
#pragma unroll
      for (int j = 0; j < TN; j++)
        x[NP - 1 - p][j] = *reinterpret_cast<const op8 *>(&Xs[s][NP - 1 - p][wn * 2 + j][lane * 8]);
    }
    auto term = [&](int pa, int px) {
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = mfma(a[pa][i], x[px][j], acc[i][j]);
    };
    if constexpr (NP == 3) {
      term(2, 0);
      term(1, 1);
      term(0, 2);
    }
    term(1, 0);
    term(0, 1);
    term(0, 0);
  };
  const int mrow0 = rb * 128 + wm * 64 + 4 * lg;  // this lane's first output row
  // row unscale factors of block bi (FP16 form), loaded before the block's
  // last chunk is multiplied so that the MFMAs cover their latency
  auto unscale = [&](int bi, f32x4(&un)[TM]) {
#pragma unroll
    for (int i = 0; i < TM; i++) un[i] = *reinterpret_cast<const f32x4 *>(uptr[bi] + mrow0 + 16 * i);
  };
  auto promote = [&](const f32x4 *un) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          if constexpr (F16) acc64[i][j][r] += (double)acc[i][j][r] * (double)un[i][r];
          else acc64[i][j][r] += (double)acc[i][j][r];
        }
        acc[i][j] = f32x4{0, 0, 0, 0};
      }
  };

  if (total > 0) {
    u32x4 R[D][2 * NP];
#pragma unroll
    for (int d = 0; d < D; d++) fetch(d, R[d]);
    stash(R[0], 0);
    __syncthreads();
    // ring: at chunk cc, R[d] (chunk cc) is already in LDS and is refilled
    // with chunk cc + D; chunk cc + 1 moves from R[d+1] to the other stage
    for (int c = 0; c < total; c += D) {
#pragma unroll
      for (int d = 0; d < D; d++) {
        const int cc = c + d;
        fetch(cc + D, R[d]);
        if (cc < total) {
          if ((cc + 1) % KC == 0) {  // last chunk of an operator block
            if constexpr (F16) {
              f32x4 un[TM];
              unscale(a0 + cc / KC, un);
              compute(cc & 1);
              promote(un);
            } else {
              compute(cc & 1);
              promote(nullptr);
            }
          } else {
            compute(cc & 1);
          }
        }
        stash(R[(d + 1) % D], (cc + 1) & 1);
        __syncthreads();
      }
    }
  }

  const int m0 = rb * 128 + wm * 64;
  const double xs = F16 ? *xun : 1.0;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int col = col0 + wn * 32 + j * 16 + lr;
      if (col >= J.N) continue;
      double *out = nsplit > 1 ? P + ((size_t)(sp * njobs + jb) * maxN + col) * ldc
                               : C + (size_t)J.cidx[col] * ldc;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = m0 + i * 16 + 4 * lg + r;
        if (row < M) out[row] = (beta && nsplit == 1 ? out[row] : 0.0) + acc64[i][j][r] * xs;
      }
    }
}

// The same GEMM with its operand tiles staged by the LDS DMA path
// (global_load_lds_dwordx4: each lane's 16 bytes go straight from memory to
// LDS, lane q at byte 16q of a 1 KB fragment) instead of through a
// register ring and LDS stores.
//
// NS LDS stages (3, or 4 for 128-column tiles of two parts), each one chunk
// of A (8 row tiles) and X (4 TN column tiles) in fragment order, are
// separate arrays addressed with static names, so the compiler's wait
// insertion can tell the stage being read from the ones being filled and
// does not drain the DMA queue before the LDS reads.  At chunk c each wave
//   waits until its own DMAs of chunk c have landed (at most the LPW of
//   each of chunks c+1 .. c+NS-2 still outstanding), meets the others at a
//   barrier (so every wave's part of chunk c is in LDS and every wave is
//   done with chunk c-1), issues the DMAs of chunk c+NS-1 into the stage
//   chunk c-1 used, and multiplies chunk c.  Four stages keep three chunks
//   of loads in flight instead of two; their 128 KB leave room for an
//   exponent table of 192 blocks (class jobs).
// The barrier is s_barrier alone: __syncthreads would also wait for all
// outstanding loads and undo the look-ahead.
//
// Tile width: TN = 2 gives the 128 x 128 tile of k_m2l_split (8 waves as
// 2 x 4 of 64 x 32).  TN = 3 widens every wave to 64 x 48 (a 128 x 192
// tile): the A fragments each wave reads from LDS then feed 12 instead of 8
// MFMA tiles, which cuts the LDS traffic per MFMA by a sixth (the kernel is
// bound by LDS bandwidth, see below), at 48 more registers per wave.  The
// (8 + 4 TN) NP fragment loads of a chunk are dealt round-robin to the 8
// waves (load i to wave i mod 8: the A tiles first, then the X tiles), LPW
// per wave; with TN = 3 a wave's X loads gather three different column
// tiles, each with its own gather slot.  TN = 3 needs NP = 2 (LPW = 5).
//
// FP16 parts (F16, NP = 2): the row unscale factors of each operator block
// (powers of two, or zero for an all-zero row; 2^-126 at the smallest) are
// read once per workgroup into an LDS table of their exponent bytes, 128
// rows x (a1-a0) blocks, so that no global load sits in the main loop apart
// from the DMAs and the gather slots: a load whose result is needed at the
// end of a block would make the compiler wait for every DMA issued before
// it.  Requires a1 - a0 <= dma_max_blocks(TN) active blocks per split
// (checked on the host): 320 at TN = 2 (the 316 offsets of a level fit
// without a split), 192 at TN = 3 (a child class's 189; LDS holds no more).
// BF16 parts (NP = 2 or 3) have no unscale table; NP = 3 takes 144 KB of
// stages.  Results are bit-identical to k_m2l_split<NP, D, TAG, F16> and
// across TN: the same MFMAs in the same order per output element, and the
// unscale exact (a power of two).
constexpr int kDmaMaxBlocks = 320;
constexpr int dma_max_blocks(int tn, int ns = 3) { return tn == 3 || ns == 4 ? 192 : kDmaMaxBlocks; }

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  // s_waitcnt with vmcnt = N and the other counters at their maximum
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

template <int NP, int TAG, bool F16, int TN = 2, int NS = 3, bool PF = false>
__global__ __launch_bounds__(512, 1) void k_m2l_dma(const int M, const int KC, const int RT,
                                                    const uint16_t *__restrict__ X, const int zrow,
                                                    double *__restrict__ C, const int ldc,
                                                    const GemmJobBf16 *__restrict__ jobs, const int njobs,
                                                    const int nrb, const int nsplit, double *__restrict__ P,
                                                    const int maxN, const double *__restrict__ xun,
                                                    const int beta) {
  static_assert(!F16 || NP == 2, "FP16 parts: two");
  static_assert(TN == 2 || (TN == 3 && NP == 2), "64 x 48 wave tiles: two parts");
  static_assert(NS == 3 || (NS == 4 && TN == 2 && NP == 2), "four stages: 128-column tiles, two parts");
  static_assert(!PF || TN == 2, "fragment prefetch: 128-column tiles");
  constexpr int BN = 64 * TN, TM = 4, TILE = 512;
  constexpr int XT = 4 * TN;                  // 16-column X tiles of a workgroup tile
  constexpr int LPW = (8 + XT) * NP / 8;      // fragment loads per wave and chunk
  constexpr int NXS = TN == 2 ? 1 : 3;        // distinct X column tiles per wave
  constexpr int UB = dma_max_blocks(TN, NS);
  constexpr int MAXB = UB < 512 ? UB : 512;   // active blocks of a tile
  typedef typename std::conditional<F16, f16x8, bf16x8>::type op8;
  typedef const __attribute__((address_space(1))) void *gvp;
  typedef __attribute__((address_space(3))) void *lvp;
  typedef const __attribute__((address_space(1))) f32x4 *gf32x4p;
  struct Stage {
    uint16_t a[NP][8][TILE];   // [part][row tile][lane * 8 + e]
    uint16_t x[NP][XT][TILE];  // [part][column tile][lane * 8 + e]
  };
  typedef typename std::conditional<NS == 4, Stage, uint32_t[4]>::type Stage3;
  __shared__ __attribute__((aligned(1024))) Stage St0;
  __shared__ __attribute__((aligned(1024))) Stage St1;
  __shared__ __attribute__((aligned(1024))) Stage St2;
  __shared__ __attribute__((aligned(1024))) Stage3 St3;
  __shared__ __attribute__((aligned(16))) uint32_t ue[F16 ? UB * 32 : 1];
  __shared__ int act[MAXB];
  __shared__ const uint16_t *aptr[MAXB];
  __shared__ const float *uptr[F16 ? MAXB : 1];

  const int bid = blockIdx.x;
  const int jb = bid % njobs;
  int rr = bid / njobs;
  const int sp = rr % nsplit;
  rr /= nsplit;
  const int rb = rr % nrb, ct = rr / nrb;
  const GemmJobBf16 J = jobs[jb];
  const int col0 = ct * BN;
  if (col0 >= J.N) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  const int lr = lane & 15, lg = lane >> 4;

  const int t0 = J.tptr[ct];
  const int nact = J.tptr[ct + 1] - t0;
  for (int i = tid; i < nact; i += 512) {
    const int b = J.tlist[t0 + i];
    act[i] = b;
    aptr[i] = J.A[b];
    if constexpr (F16) uptr[i] = J.U[b];
  }
  __syncthreads();
  const int a0 = (int)((int64_t)nact * sp / nsplit), a1 = (int)((int64_t)nact * (sp + 1) / nsplit);
  const int total = (a1 - a0) * KC;

  // exponent table: word 32 q + w holds the exponent bytes of rows
  // 4w .. 4w+3 (of this 128-row tile) of active block a0 + q
  if constexpr (F16) {
    const int nq = (a1 - a0) * 32;
    for (int i0 = 0; i0 < nq; i0 += 4 * 512) {
      f32x4 v[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int i = i0 + k * 512 + tid;
        if (i < nq) v[k] = *(gf32x4p)(uptr[a0 + (i >> 5)] + rb * 128 + (i & 31) * 4);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int i = i0 + k * 512 + tid;
        if (i < nq) {
          uint32_t e4 = 0;
#pragma unroll
          for (int r = 0; r < 4; r++) e4 |= ((__float_as_uint(v[k][r]) >> 23) & 255u) << (8 * r);
          ue[i] = e4;
        }
      }
    }
    __syncthreads();
  }

  // this wave's fragment loads: load i = wave + 8 q of the chunk's list
  // [A part 0 tiles 0-7, A part 1, ..., X part 0 tiles 0..XT-1, ...].  The
  // first XQ0 are A loads (i < 8 NP); with TN = 2 the X loads all read the
  // column tile `wave` (one gather slot), with TN = 3 the three X loads read
  // tiles wave, (wave + 8) mod 12 and wave + 4 (gather slots 0, 1, 2).
  constexpr int XQ0 = TN == 2 ? NP : 2;
  const int tq = lane;
  const int kof = 8 * (tq >> 4);
  int lp[LPW], lt[LPW];  // part and tile of load q
#pragma unroll
  for (int q = 0; q < LPW; q++) {
    const int i = wave + 8 * q;
    const int j = q >= XQ0 ? i - 8 * NP : i;
    const int per = q >= XQ0 ? XT : 8;
    lp[q] = j / per;
    lt[q] = j - lp[q] * per;
  }
  int xtile[NXS];
#pragma unroll
  for (int s = 0; s < NXS; s++) xtile[s] = lt[XQ0 + (TN == 2 ? 0 : s)];
  int my_col[NXS];
#pragma unroll
  for (int s = 0; s < NXS; s++) my_col[s] = col0 + xtile[s] * 16 + (tq & 15);
  const size_t slab = (size_t)RT * TILE;
  const gi32p gidx = (gi32p)J.gidx;
  auto gload = [&](int bi, int s) -> int {
    const int bc = bi < a1 ? bi : a1 - 1;
    const int cc = my_col[s] < J.N ? my_col[s] : J.N - 1;
    return gidx[(size_t)act[bc] * J.N + cc];
  };
  auto gslot = [&](int raw, int bi, int s) -> int {
    return (bi < a1 && my_col[s] < J.N && raw >= 0) ? raw : zrow;
  };
  int fblk = -1, gcur[NXS], gnext[NXS];
#pragma unroll
  for (int s = 0; s < NXS; s++) {
    gcur[s] = zrow;
    gnext[s] = total > 0 ? gload(a0, s) : -1;
  }

  // DMAs of chunk c (clamped to the last chunk, so that every step issues
  // exactly LPW of them and the wait counts stay static)
  auto fetch = [&](int c, Stage &Sd) {
    const int cq = c < total ? c : total - 1;
    const int q = cq / KC, kc = cq - q * KC, bi = a0 + q;
    if (bi != fblk) {
      fblk = bi;
#pragma unroll
      for (int s = 0; s < NXS; s++) {
        gcur[s] = gslot(gnext[s], bi, s);
        gnext[s] = gload(bi + 1, s);
      }
    }
#pragma unroll
    for (int l = 0; l < LPW; l++) {
      if (l < XQ0) {
        const uint16_t *Ab = aptr[bi] + (size_t)(rb * 8 + lt[l]) * TILE + (size_t)tq * 8 +
                             ((size_t)kc * NP + lp[l]) * slab;
        __builtin_amdgcn_global_load_lds((gvp)Ab, (lvp)&Sd.a[lp[l]][lt[l]][0], 16, 0, 0);
      } else {
        const int g = gcur[TN == 2 ? 0 : l - XQ0];
        const uint16_t *Xb = X + (((size_t)g * KC + kc) * NP + lp[l]) * 32 + kof;
        __builtin_amdgcn_global_load_lds((gvp)Xb, (lvp)&Sd.x[lp[l]][lt[l]][0], 16, 0, 0);
      }
    }
  };

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
  auto mfma = [](const op8 &a, const op8 &b, const f32x4 &c) -> f32x4 {
    if constexpr (F16) return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  };
  // All fragments of the chunk are read first; the products then run term
  // by term over all TM x TN tiles, so an accumulator's next MFMA is
  // TM TN instructions after its previous one instead of depending on the
  // one just issued.  Each accumulator still takes the terms in the same
  // order (bit-identical results).
  struct Frag {
    op8 a[NP][TM], x[NP][TN];
  };
  auto readfrag = [&](const Stage &Sc, Frag &F) {
#pragma unroll
    for (int p = NP - 1; p >= 0; p--) {
#pragma unroll
      for (int i = 0; i < TM; i++) F.a[p][i] = *reinterpret_cast<const op8 *>(&Sc.a[p][wm * 4 + i][lane * 8]);
#pragma unroll
      for (int j = 0; j < TN; j++)
        F.x[NP - 1 - p][j] = *reinterpret_cast<const op8 *>(&Sc.x[NP - 1 - p][wn * TN + j][lane * 8]);
    }
  };
  auto mul = [&](const Frag &F) {
    auto term = [&](int pa, int px) {
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = mfma(F.a[pa][i], F.x[px][j], acc[i][j]);
    };
    if constexpr (NP == 3) {
      term(2, 0);
      term(1, 1);
      term(0, 2);
    }
    term(1, 0);
    term(0, 1);
    term(0, 0);
  };
  auto compute = [&](const Stage &Sc) {
    Frag F;
    readfrag(Sc, F);
    mul(F);
  };
  // promotion of block q's FP32 sums (FP16 parts: each row times its
  // unscale factor)
  auto promote = [&](int q) {
#pragma unroll
    for (int i = 0; i < TM; i++) {
      const uint32_t e4 = F16 ? ue[q * 32 + wm * 16 + lg + 4 * i] : 0u;
#pragma unroll
      for (int j = 0; j < TN; j++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          if constexpr (F16)
            acc64[i][j][r] += (double)acc[i][j][r] * (double)__uint_as_float(((e4 >> (8 * r)) & 255u) << 23);
          else
            acc64[i][j][r] += (double)acc[i][j][r];
        }
        acc[i][j] = f32x4{0, 0, 0, 0};
      }
    }
  };
  // step cc: chunks cc+1 .. cc+NS-2 may still be in flight; chunk cc+NS-1
  // goes into the stage chunk cc-1 used
  auto step = [&](int cc, const Stage &Sc, Stage &Sd) {
    wait_vmcnt<(NS - 2) * LPW>();
    __builtin_amdgcn_s_barrier();
    fetch(cc + NS - 1, Sd);
    if (cc < total) {
      compute(Sc);
      if ((cc + 1) % KC == 0) promote(cc / KC);
    }
  };
  // fragment prefetch (PF): step cc reads the fragments of chunk cc+1 from
  // LDS into one register set while the MFMAs of chunk cc run on the other,
  // so that the LDS reads of one chunk overlap the products of the previous
  // one instead of following the barrier.  The DMAs of chunk cc+NS go into
  // the stage chunk cc was read from (every wave has waited for its reads of
  // chunk cc before the barrier), so NS stages still hold NS-1 chunks of
  // loads in flight.
  constexpr int PW = (NS - 2) * LPW;
  auto pstep = [&](int cc, const Stage &Sn, Stage &Sd, const Frag &Fc, Frag &Fn) {
    // vmcnt: this wave's loads of chunk cc+1 have landed (those of cc+2 ..
    // cc+NS-1 may be outstanding); lgkmcnt(0): its LDS reads of chunk cc
    // are done
    __builtin_amdgcn_s_waitcnt((PW & 15) | ((PW >> 4) << 14) | (7 << 4));
    __builtin_amdgcn_s_barrier();
    fetch(cc + NS, Sd);
    readfrag(Sn, Fn);
    if (cc < total) {
      mul(Fc);
      if ((cc + 1) % KC == 0) promote(cc / KC);
    }
  };
  if (total > 0 && PF) {
    Frag F0, F1;
    fetch(0, St0);
    fetch(1, St1);
    if constexpr (NS == 4) fetch(2, St2);
    wait_vmcnt<PW>();
    __builtin_amdgcn_s_barrier();
    readfrag(St0, F0);
    if constexpr (NS == 4) {
      fetch(3, St3);
      for (int c = 0; c < total; c += 4) {
        pstep(c, St1, St0, F0, F1);
        pstep(c + 1, St2, St1, F1, F0);
        pstep(c + 2, St3, St2, F0, F1);
        pstep(c + 3, St0, St3, F1, F0);
      }
    } else {
      fetch(2, St2);
      for (int c = 0; c < total; c += 6) {
        pstep(c, St1, St0, F0, F1);
        pstep(c + 1, St2, St1, F1, F0);
        pstep(c + 2, St0, St2, F0, F1);
        pstep(c + 3, St1, St0, F1, F0);
        pstep(c + 4, St2, St1, F0, F1);
        pstep(c + 5, St0, St2, F1, F0);
      }
    }
    wait_vmcnt<0>();  // no DMA may land after the workgroup ends
  } else if (total > 0) {
    fetch(0, St0);
    fetch(1, St1);
    if constexpr (NS == 4) {
      fetch(2, St2);
      for (int c = 0; c < total; c += 4) {
        step(c, St0, St3);
        step(c + 1, St1, St0);
        step(c + 2, St2, St1);
        step(c + 3, St3, St2);
      }
    } else {
      for (int c = 0; c < total; c += 3) {
        step(c, St0, St2);
        step(c + 1, St1, St0);
        step(c + 2, St2, St1);
      }
    }
    wait_vmcnt<0>();  // no DMA may land after the workgroup ends
  }

  const int m0 = rb * 128 + wm * 64;
  const double xs = F16 ? *xun : 1.0;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int col = col0 + wn * 16 * TN + j * 16 + lr;
      if (col >= J.N) continue;
      double *out = nsplit > 1 ? P + ((size_t)(sp * njobs + jb) * maxN + col) * ldc
                               : C + (size_t)J.cidx[col] * ldc;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = m0 + i * 16 + 4 * lg + r;
        if (row < M) out[row] = (beta && nsplit == 1 ? out[row] : 0.0) + acc64[i][j][r] * xs;
      }
    }
}

}  // namespace nnk
