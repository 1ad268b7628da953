// Mixed-precision form of the KIFMM matvec.
//
// Check-to-equivalent solves use a truncated SVD pseudo-inverse (singular
// values below tau * s_max dropped; a common rank R over the levels), in
// FP64.  Densities are carried as coefficients in the singular bases:
//   up:   c = S_R^-1 U_R^T chk_up   (equivalent density V_R c)
//   down: d = S_R^-1 V_R^T chk_dn   (equivalent density U_R d)
// so every translation becomes an R x R operator:
//   M2M_k <- Pup(l-1) K(child up -> parent down) V_R(l)
//   L2L_k <- Pdn(l)   K(parent down -> child up) U_R(l-1)
//   M2L_o <- Pdn(l)   K(up at offset o -> up)    V_R(l)
// formed once on the device with the FP64 GEMM.  M2L (the dominant work)
// runs as split-BF16 MFMA GEMMs (two or three BF16 parts per operand) with
// FP32 accumulation promoted to FP64 per translation; M2M, L2L and the up
// expansion stay FP64 (their S^-1 factor amplifies input rounding), the
// down expansion U_R d (orthonormal U_R) runs split-FP16 in the FP16 form;
// the
// direct sums (S2M, P2L, L2T, P2P, M2T) evaluate the kernel in FP32 from
// FP64 separations and accumulate in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "nnk_fmm.hpp"
#include "nnk_gemm_bf16.hpp"
#include "nnk_kernels.hpp"

namespace nnk {

namespace {

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// v ~ parts[0] + ... + parts[np-1], each part the BF16 rounding of the
// remainder (the remainders are exact in FP64)
__device__ __forceinline__ uint16_t split_part(double v, int np, int p) {
  uint16_t h = 0;
  for (int i = 0; i <= p && i < np; i++) {
    h = f2bf((float)v);
    v -= (double)bf2f(h);
  }
  return h;
}

__device__ __forceinline__ uint16_t f2h(double v) {
  const _Float16 h = (_Float16)(float)v;
  return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ double h2d(uint16_t b) { return (double)(float)__builtin_bit_cast(_Float16, b); }

// FP16 parts: v ~ h0 + h1
__device__ __forceinline__ uint16_t split_part_h(double v, int p) {
  const uint16_t h0 = f2h(v);
  return p == 0 ? h0 : f2h(v - h2d(h0));
}

// 2^k with k = floor(log2(a)), a > 0 (exact power of two)
__device__ __forceinline__ double pow2_floor(double a) {
  int e;
  (void)frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
  return ldexp(1.0, e - 1);
}

}  // namespace

// max |X[rows[i]][k]| over k < R into *amax (bit pattern of a non-negative
// double, so unsigned max orders it)
__global__ void k_absmax(const int32_t *__restrict__ rows, int nrows, const double *__restrict__ X, int ldx, int R,
                         unsigned long long *amax) {
  __shared__ double red[256];
  double m = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (int64_t)nrows * R;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / R), k = (int)(idx % R);
    m = fmax(m, fabs(X[(int64_t)rows[i] * ldx + k]));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(amax, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_amax_reset(unsigned long long *amax) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *amax = 0ull;
}

// FP16 coefficient parts with one power-of-two scale per matvec: x 2^14 /
// 2^floor(log2 max|X|); *xun receives the inverse
__global__ void k_split_rows_h(const int32_t *__restrict__ rows, int nrows, const double *__restrict__ X, int ldx,
                               int R, int KC, const unsigned long long *__restrict__ amax, uint16_t *__restrict__ out,
                               double *__restrict__ xun) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double a = __longlong_as_double((long long)*amax);
  const double sc = a > 0 ? 16384.0 / pow2_floor(a) : 1.0;
  if (idx == 0) *xun = 1.0 / sc;
  const int kk = KC * 32;
  if (idx >= (int64_t)nrows * kk) return;
  const int i = (int)(idx / kk), k = (int)(idx % kk);
  const int64_t r = rows[i];
  const double v = k < R ? X[r * ldx + k] * sc : 0.0;
  const int kc = k >> 5, e = k & 31;
  for (int p = 0; p < 2; p++) out[((r * KC + kc) * 2 + p) * 32 + e] = split_part_h(v, p);
}

// per operator row: unscale factor 2^floor(log2 max_n |C_j[m][n]|) / 2^14
// (0 for rows beyond R or all-zero rows)
__global__ void k_op_rowscale(const double *__restrict__ CT, int ldct, int R, int nops, int RPm,
                              float *__restrict__ un) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nops * RPm) return;
  const int j = (int)(idx / RPm), m = (int)(idx % RPm);
  double a = 0.0;
  if (m < R)
    for (int n = 0; n < R; n++) a = fmax(a, fabs(CT[((size_t)j * R + n) * ldct + m]));
  // a power of two, normal in FP32 (the DMA-staged kernel keeps only its
  // exponent byte)
  un[idx] = a > 0 ? (float)fmax(pow2_floor(a) / 16384.0, 0x1p-126) : 0.0f;
}

// FP16 operator parts in fragment order (as k_split_op), rows scaled by
// the inverse of their unscale factor
__global__ void k_split_op_h(const double *__restrict__ CT, int ldct, int R, int nops, int KC, int RT,
                             const float *__restrict__ un, uint16_t *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)KC * 2 * RT * 512;
  if (idx >= per * nops) return;
  const int j = (int)(idx / per);
  int64_t rem = idx % per;
  const int e = (int)(rem & 7), q = (int)((rem >> 3) & 63);
  rem >>= 9;
  const int t = (int)(rem % RT);
  rem /= RT;
  const int p = (int)(rem % 2), kc = (int)(rem / 2);
  const int m = 16 * t + (q & 15), k = 32 * kc + 8 * (q >> 4) + e;
  const float u = un[(size_t)j * RT * 16 + m];
  const double v = (m < R && k < R && u > 0) ? CT[((size_t)j * R + k) * ldct + m] / (double)u : 0.0;
  out[idx] = split_part_h(v, p);
}

// the same for rectangular operators: A_j(m, n) = A[j * ostride + n * ld + m]
// (m < Mr rows, n < Kc columns)
__global__ void k_op_rowscale_g(const double *__restrict__ A, int64_t ostride, int ld, int Mr, int Kc, int nops,
                                int RPm, float *__restrict__ un) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nops * RPm) return;
  const int j = (int)(idx / RPm), m = (int)(idx % RPm);
  double a = 0.0;
  if (m < Mr)
    for (int n = 0; n < Kc; n++) a = fmax(a, fabs(A[j * ostride + (int64_t)n * ld + m]));
  // a power of two, normal in FP32 (the DMA-staged kernel keeps only its
  // exponent byte)
  un[idx] = a > 0 ? (float)fmax(pow2_floor(a) / 16384.0, 0x1p-126) : 0.0f;
}

__global__ void k_split_op_hg(const double *__restrict__ A, int64_t ostride, int ld, int Mr, int Kc, int nops,
                              int KC, int RT, const float *__restrict__ un, uint16_t *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)KC * 2 * RT * 512;
  if (idx >= per * nops) return;
  const int j = (int)(idx / per);
  int64_t rem = idx % per;
  const int e = (int)(rem & 7), q = (int)((rem >> 3) & 63);
  rem >>= 9;
  const int t = (int)(rem % RT);
  rem /= RT;
  const int p = (int)(rem % 2), kc = (int)(rem / 2);
  const int m = 16 * t + (q & 15), k = 32 * kc + 8 * (q >> 4) + e;
  const float u = un[(size_t)j * RT * 16 + m];
  const double v = (m < Mr && k < Kc && u > 0) ? A[j * ostride + (int64_t)k * ld + m] / (double)u : 0.0;
  out[idx] = split_part_h(v, p);
}

// rows[i] of X (length R, stride ldx) -> parts in [node][KC][np][32] order
// (zero beyond R)
__global__ void k_split_rows(const int32_t *__restrict__ rows, int nrows, const double *__restrict__ X,
                             int ldx, int R, int KC, int np, uint16_t *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = KC * 32;
  if (idx >= (int64_t)nrows * kk) return;
  const int i = (int)(idx / kk), k = (int)(idx % kk);
  const int64_t r = rows[i];
  const double v = k < R ? X[r * ldx + k] : 0.0;
  const int kc = k >> 5, e = k & 31;
  for (int p = 0; p < np; p++) out[((r * KC + kc) * np + p) * 32 + e] = split_part(v, np, p);
}

// operator j: CT[(j*R + n) * ldct + m] = C_j[m][n] -> parts in MFMA fragment
// order [KC][np][RT][64][8]: lane q of row tile t holds
// C_j[16t + q%16][32kc + 8(q/16) + e]
__global__ void k_split_op(const double *__restrict__ CT, int ldct, int R, int nops, int KC, int np, int RT,
                           uint16_t *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)KC * np * RT * 512;
  if (idx >= per * nops) return;
  const int j = (int)(idx / per);
  int64_t rem = idx % per;
  const int e = (int)(rem & 7), q = (int)((rem >> 3) & 63);
  rem >>= 9;
  const int t = (int)(rem % RT);
  rem /= RT;
  const int p = (int)(rem % np), kc = (int)(rem / np);
  const int m = 16 * t + (q & 15), k = 32 * kc + 8 * (q >> 4) + e;
  const double v = (m < R && k < R) ? CT[((size_t)j * R + k) * ldct + m] : 0.0;
  out[idx] = split_part(v, np, p);
}

constexpr int MBM = 128, MBN = 128;   // split-BF16 M2L tile

void Fmm::build_mixed() {
  const int K = K_;
  R_ = 0;
  for (int l = 2; l <= maxlev_; l++) {
    const auto &S = ops_[l]->svdS;
    int r = 0;
    while (r < K && S[r] >= tau_ * S[0]) r++;
    R_ = std::max(R_, r);
  }
  if (R_ == 0) R_ = 32;
  np_ = prec_ == 2 ? 3 : 2;
  f16_ = prec_ == 3;
  RP_ = (R_ + 31) / 32 * 32;
  RPm_ = (R_ + MBM - 1) / MBM * MBM;
  KCm_ = RP_ / 32;
  RT_ = RPm_ / 16;
  MpadR_ = (R_ + 159) / 160 * 160;
  KpadR_ = (R_ + 15) / 16 * 16;
  RTk_ = (K_ + MBM - 1) / MBM * MBM / 16;
  const int R = R_;
  const size_t matKR = (size_t)Kpad_ * MpadR_;   // R x K operators (Pup, Pdn)
  const size_t matRK = (size_t)KpadR_ * Mpad_;   // K x R operators (Vr, Ur)
  const size_t matRR = (size_t)KpadR_ * MpadR_;  // R x R operators
  OpCache &cache = OpCache::get();
  mops_.assign(maxlev_ + 1, nullptr);
  std::vector<std::shared_ptr<MixedOps>> fresh(maxlev_ + 1);
  for (int l = 2; l <= maxlev_; l++) {
    if (auto hit = cache.find(key(2 + prec_, l))) {
      mops_[l] = std::static_pointer_cast<const MixedOps>(hit);
      continue;
    }
    fresh[l] = std::make_shared<MixedOps>();
    mops_[l] = fresh[l];
    MixedOps &X = *fresh[l];
    const auto &U = ops_[l]->svdU, &S = ops_[l]->svdS, &V = ops_[l]->svdV;
    std::vector<double> Pup(matKR, 0.0), Pdn(matKR, 0.0), Vr(matRK, 0.0), Ur(matRK, 0.0),
        VrT((size_t)R * K), UrT((size_t)R * K);
    for (int k = 0; k < K; k++)
      for (int m = 0; m < R; m++) {
        Pup[(size_t)k * MpadR_ + m] = U[(size_t)k * K + m] / S[m];
        Pdn[(size_t)k * MpadR_ + m] = V[(size_t)k * K + m] / S[m];
        Vr[(size_t)m * Mpad_ + k] = V[(size_t)k * K + m];
        Ur[(size_t)m * Mpad_ + k] = U[(size_t)k * K + m];
        VrT[(size_t)m * K + k] = V[(size_t)k * K + m];
        UrT[(size_t)m * K + k] = U[(size_t)k * K + m];
      }
    X.Pup.upload(Pup);
    X.Pdn.upload(Pdn);
    X.Vr.upload(Vr);
    X.Ur.upload(Ur);
    X.VrT.upload(VrT);
    X.UrT.upload(UrT);
  }
  // compressed translations; T_o, M2M_k, L2L_k are the FP64 kernel matrices
  auto batch = [&](int njobs, const std::vector<const double *> &A, int N,
                   const std::vector<std::vector<int32_t>> *g, const std::vector<std::vector<int32_t>> *c) {
    GemmBatch B;
    std::vector<GemmJob> jobs(njobs);
    for (int j = 0; j < njobs; j++) {
      GemmJob &J = jobs[j];
      B.owned.emplace_back();
      B.owned.back().upload(std::vector<const double *>{A[j]});
      J.A = B.owned.back().as<const double *>();
      if (g) { B.owned.emplace_back(); B.owned.back().upload((*g)[j]); J.gidx = B.owned.back().as<int32_t>(); }
      if (c) { B.owned.emplace_back(); B.owned.back().upload((*c)[j]); J.cidx = B.owned.back().as<int32_t>(); }
      J.nb = 1;
      J.N = N;
    }
    B.njobs = njobs;
    B.maxN = N;
    B.jobs.upload(jobs);
    return B;
  };
  auto offs_ids = [&](int nj, int stride) {
    std::vector<std::vector<int32_t>> v(nj, std::vector<int32_t>(R));
    for (int j = 0; j < nj; j++)
      for (int n = 0; n < R; n++) v[j][n] = j * stride + n;
    return v;
  };
  for (int l = 2; l <= maxlev_; l++) {
    if (!fresh[l]) continue;
    const LevelOps &O = *ops_[l];
    MixedOps &X = *fresh[l];
    // ---- M2L: C_o = Pdn T_o V_R, with T_o formed for this level only ----
    DBuf Ttmp;
    std::vector<double *> tslot = O.tslot;
    if (!O.T.p) form_T(l, Ttmp, tslot);
    std::vector<const double *> Ts;
    std::vector<int> slots;
    for (int sl = 0; sl < 343; sl++)
      if (tslot[sl]) { Ts.push_back(tslot[sl]); slots.push_back(sl); }
    const int no = (int)Ts.size();
    {
      DBuf Y, CT;
      Y.alloc(sizeof(double) * (size_t)no * R * K);
      CT.alloc(sizeof(double) * (size_t)no * R * MpadR_);
      auto ids = offs_ids(no, R);
      GemmBatch b1 = batch(no, Ts, R, nullptr, &ids);
      gemm_dims(b1, X.VrT.as<double>(), K, Y.as<double>(), K, 0, 0, K, K, Mpad_, Kpad_);
      std::vector<const double *> P(no, X.Pdn.as<double>());
      GemmBatch b2 = batch(no, P, R, &ids, &ids);
      gemm_dims(b2, Y.as<double>(), K, CT.as<double>(), MpadR_, 0, 0, R, K, MpadR_, Kpad_);
      if (prec_ == 4) {  // FP64 operators, At layout padded to KpadR rows
        const size_t opd = (size_t)KpadR_ * MpadR_;
        X.Cd.alloc(sizeof(double) * opd * no);
        X.Cd.zero();
        for (int j = 0; j < no; j++)
          NNK_HIP(hipMemcpyAsync(X.Cd.as<double>() + opd * j, CT.as<double>() + (size_t)j * R * MpadR_,
                                 sizeof(double) * (size_t)R * MpadR_, hipMemcpyDeviceToDevice, 0));
        NNK_HIP(hipDeviceSynchronize());
        X.dop.assign(343, nullptr);
        for (int j = 0; j < no; j++) X.dop[slots[j]] = X.Cd.as<double>() + opd * j;
      } else {
        const size_t per = (size_t)KCm_ * np_ * RT_ * 512;
        X.Cf.alloc(sizeof(uint16_t) * per * no);
        const int64_t tot = (int64_t)per * no;
        if (f16_) {
          X.Uf.alloc(sizeof(float) * (size_t)RPm_ * no);
          const int64_t nr = (int64_t)RPm_ * no;
          hipLaunchKernelGGL(k_op_rowscale, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, 0, CT.as<double>(),
                             MpadR_, R, no, RPm_, X.Uf.as<float>());
          hipLaunchKernelGGL(k_split_op_h, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, CT.as<double>(),
                             MpadR_, R, no, KCm_, RT_, X.Uf.as<float>(), X.Cf.as<uint16_t>());
        } else {
          hipLaunchKernelGGL(k_split_op, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, CT.as<double>(),
                             MpadR_, R, no, KCm_, np_, RT_, X.Cf.as<uint16_t>());
        }
        NNK_HIP(hipGetLastError());
        NNK_HIP(hipDeviceSynchronize());
        X.cop.assign(343, nullptr);
        X.uop.assign(343, nullptr);
        for (int j = 0; j < no; j++) {
          X.cop[slots[j]] = X.Cf.as<uint16_t>() + per * j;
          if (f16_) X.uop[slots[j]] = X.Uf.as<float>() + (size_t)RPm_ * j;
        }
      }
    }
    // ---- M2M (this level is the child level; parent level l-1 >= 2) ----
    if (l >= 3) {
      const size_t mat = (size_t)Kpad_ * Mpad_;
      std::vector<const double *> Ms(8);
      for (int k = 0; k < 8; k++) Ms[k] = O.M2M.as<double>() + k * mat;
      DBuf Y;
      Y.alloc(sizeof(double) * (size_t)8 * R * K);
      auto ids = offs_ids(8, R);
      GemmBatch b1 = batch(8, Ms, R, nullptr, &ids);
      gemm_dims(b1, X.VrT.as<double>(), K, Y.as<double>(), K, 0, 0, K, K, Mpad_, Kpad_);
      X.Mc.alloc(sizeof(double) * matRR * 8);
      X.Mc.zero();
      auto outs = offs_ids(8, KpadR_);
      std::vector<const double *> P(8, mops_[l - 1]->Pup.as<double>());
      GemmBatch b2 = batch(8, P, R, &ids, &outs);
      gemm_dims(b2, Y.as<double>(), K, X.Mc.as<double>(), MpadR_, 0, 0, R, K, MpadR_, Kpad_);
      // ---- L2L: Lc_a = Pdn(l) L2L_a U_R(l-1) ----
      std::vector<const double *> Ls(8);
      for (int k = 0; k < 8; k++) Ls[k] = O.L2L.as<double>() + k * mat;
      GemmBatch b3 = batch(8, Ls, R, nullptr, &ids);
      gemm_dims(b3, mops_[l - 1]->UrT.as<double>(), K, Y.as<double>(), K, 0, 0, K, K, Mpad_, Kpad_);
      X.Lc.alloc(sizeof(double) * matRR * 8);
      X.Lc.zero();
      std::vector<const double *> P2(8, X.Pdn.as<double>());
      GemmBatch b4 = batch(8, P2, R, &ids, &outs);
      gemm_dims(b4, Y.as<double>(), K, X.Lc.as<double>(), MpadR_, 0, 0, R, K, MpadR_, Kpad_);
      NNK_HIP(hipDeviceSynchronize());
    }
    if (f16_) {  // the down expansion U_R as FP16 parts
      auto frag = [&](const DBuf &A, int64_t ostride, int ld, int Mr, int Kc, int nops, int KC, int RT, DBuf &F,
                      DBuf &U) {
        const int64_t nu = (int64_t)nops * RT * 16, per = (int64_t)KC * 2 * RT * 512;
        U.alloc(sizeof(float) * nu);
        F.alloc(sizeof(uint16_t) * per * nops);
        hipLaunchKernelGGL(k_op_rowscale_g, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, 0, A.as<double>(),
                           ostride, ld, Mr, Kc, nops, RT * 16, U.as<float>());
        hipLaunchKernelGGL(k_split_op_hg, dim3((unsigned)((per * nops + 255) / 256)), dim3(256), 0, 0,
                           A.as<double>(), ostride, ld, Mr, Kc, nops, KC, RT, U.as<float>(), F.as<uint16_t>());
        NNK_HIP(hipGetLastError());
      };
      frag(X.Ur, 0, Mpad_, K, R, 1, KCm_, RTk_, X.UrF, X.UrU);
      NNK_HIP(hipDeviceSynchronize());
    }
    cache.put(key(2 + prec_, l), fresh[l], X.bytes());
  }
  const int nc = tree_.count;
  chat_.alloc(sizeof(double) * (size_t)nc * RP_);
  dhat_.alloc(sizeof(double) * (size_t)nc * RP_);
  // split coefficients, one all-zero row (slot nc) for absent sources
  if (m16()) {
    chat_x_.alloc(sizeof(uint16_t) * (size_t)(nc + 1) * np_ * RP_);
    chat_x_.zero();
  }
  chat_.zero();
  dhat_.zero();
  amax_.alloc(sizeof(unsigned long long));
  xun_.alloc(sizeof(double));
  if (f16_) {
    dhat_x_.alloc(sizeof(uint16_t) * (size_t)(nc + 1) * KCm_ * 2 * 32);
    dhat_x_.zero();
    dun_lv_.alloc(sizeof(double) * (maxlev_ + 2));
    std::vector<double> ones(maxlev_ + 2, 1.0);
    NNK_HIP(hipMemcpy(dun_lv_.p, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice));
  }
  {
    const double one = 1.0;
    NNK_HIP(hipMemcpy(xun_.p, &one, sizeof(double), hipMemcpyHostToDevice));
  }
  NNK_HIP(hipDeviceSynchronize());
}

void Fmm::plan_mixed() {
  const size_t matRR = (size_t)KpadR_ * MpadR_;
  std::vector<int32_t> upall;
  // one job per level of a batch over all levels
  std::vector<GemmJob> jobs_upP, jobs_wV, jobs_downU;
  mx_upP_all_ = GemmBatch();
  mx_wV_all_ = GemmBatch();
  mx_downU_all_ = GemmBatch();
  auto jobl = [&](GemmBatch &B, std::vector<GemmJob> &jobs, const double *A, const std::vector<int32_t> &g) {
    if (g.empty()) return;
    GemmJob J{};
    B.owned.emplace_back();
    B.owned.back().upload(std::vector<const double *>{A});
    J.A = B.owned.back().as<const double *>();
    B.owned.emplace_back();
    B.owned.back().upload(g);
    J.gidx = B.owned.back().as<int32_t>();
    J.cidx = J.gidx;  // the same node list in and out
    J.nb = 1;
    J.N = (int)g.size();
    jobs.push_back(J);
    B.maxN = std::max(B.maxN, J.N);
  };
  auto job1 = [&](GemmBatch &B, const double *A, const std::vector<int32_t> &g,
                  const std::vector<int32_t> &c) {
    if (g.empty()) return;
    std::vector<GemmJob> jobs(1);
    B.owned.emplace_back();
    B.owned.back().upload(std::vector<const double *>{A});
    jobs[0].A = B.owned.back().as<const double *>();
    B.owned.emplace_back();
    B.owned.back().upload(g);
    jobs[0].gidx = B.owned.back().as<int32_t>();
    B.owned.emplace_back();
    B.owned.back().upload(c);
    jobs[0].cidx = B.owned.back().as<int32_t>();
    jobs[0].nb = 1;
    jobs[0].N = (int)c.size();
    B.njobs = 1;
    B.maxN = jobs[0].N;
    B.jobs.upload(jobs);
  };
  for (int l = 2; l <= maxlev_; l++) {
    LevelPlan &P = plans_[l];
    HostLevel &H = hl_[l];
    const MixedOps &X = *mops_[l];
    upall.insert(upall.end(), H.up_leaves.begin(), H.up_leaves.end());
    upall.insert(upall.end(), H.up_nonleaves.begin(), H.up_nonleaves.end());
    if (!H.up_nonleaves.empty()) {
      const int N = (int)H.up_nonleaves.size();
      std::vector<GemmJob> jobs(1);
      std::vector<const double *> A(8);
      for (int k = 0; k < 8; k++) A[k] = mops_[l + 1]->Mc.as<double>() + k * matRR;
      P.mx_m2m.owned.emplace_back();
      P.mx_m2m.owned.back().upload(A);
      jobs[0].A = P.mx_m2m.owned.back().as<const double *>();
      P.mx_m2m.owned.emplace_back();
      P.mx_m2m.owned.back().upload(H.m2m_g);
      jobs[0].gidx = P.mx_m2m.owned.back().as<int32_t>();
      P.mx_m2m.owned.emplace_back();
      P.mx_m2m.owned.back().upload(H.up_nonleaves);
      jobs[0].cidx = P.mx_m2m.owned.back().as<int32_t>();
      jobs[0].nb = 8;
      jobs[0].N = N;
      P.mx_m2m.njobs = 1;
      P.mx_m2m.maxN = N;
      P.mx_m2m.jobs.upload(jobs);
      const int64_t tiles = (int64_t)(MpadR_ / GBM) * (((int64_t)N + GBN0 - 1) / GBN0);
      P.mx_m2m_split = split_m2m(tiles, (size_t)N * RP_);
      if (P.mx_m2m_split > 1) {
        P.mx_m2m_slot.upload(H.up_nonleaves);
        reserve_part((size_t)P.mx_m2m_split * N * RP_);
      }
    }
    jobl(mx_upP_all_, jobs_upP, X.Pup.as<double>(), H.up_leaves);
    jobl(mx_wV_all_, jobs_wV, X.Vr.as<double>(), H.wnodes);
    if (!f16_) jobl(mx_downU_all_, jobs_downU, X.Ur.as<double>(), H.tleaves);
    {
      std::vector<int32_t> xs;
      for (int32_t n : H.xnodes) xs.push_back(n - 1);
      job1(P.mx_xP, X.Pdn.as<double>(), xs, xs);
    }
    // split-BF16 M2L jobs (one per child class) and FP64 L2L jobs
    std::vector<GemmJobBf16> mj;
    std::vector<GemmJob> lj;
    // the V-list offsets of child class a, in its gather table's block order
    auto class_offsets = [&](int a) {
      std::vector<int> sls = class_offset_slots(a);
      for (int sl : sls)
        if (!X.cop[sl]) throw Error(NNK_ESTATE, "mixed M2L: missing translation operator");
      return sls;
    };
    // Levels with few nodes leave most of a 128-column tile empty in every
    // class job; one job over all 316 offsets then walks fewer (tile,
    // operator block) pairs (plan_m2l_level decides and checks the tables).
    M2lLevelPlan mp;
    if (m16()) plan_m2l_level(H.cls, H.gv, tree_.count, MBN, mp);
    const bool merged = mp.merged;
    const std::vector<int> &msl = mp.msl;
    const std::vector<int32_t> &mcols = mp.mcols;
    const int Nall = (int)mcols.size();
    if (merged) {
      std::vector<const uint16_t *> ops;
      std::vector<const float *> uops;
      for (int sl : msl) {
        if (!X.cop[sl]) throw Error(NNK_ESTATE, "mixed M2L: missing translation operator");
        ops.push_back(X.cop[sl]);
        uops.push_back(f16_ ? X.uop[sl] : nullptr);
      }
      GemmJobBf16 J{};
      P.mx_owned.emplace_back();
      P.mx_owned.back().upload(ops);
      J.A = P.mx_owned.back().as<const uint16_t *>();
      P.mx_owned.emplace_back();
      P.mx_owned.back().upload(uops);
      J.U = P.mx_owned.back().as<const float *>();
      P.mx_owned.emplace_back();
      P.mx_owned.back().upload(mp.mg);
      J.gidx = P.mx_owned.back().as<int32_t>();
      P.mx_owned.emplace_back();
      P.mx_owned.back().upload(mcols);
      J.cidx = P.mx_owned.back().as<int32_t>();
      J.nb = (int)ops.size();
      J.N = Nall;
      P.mx_owned.emplace_back();
      P.mx_owned.back().upload(mp.mtp);
      J.tptr = P.mx_owned.back().as<int32_t>();
      P.mx_owned.emplace_back();
      P.mx_owned.back().upload(mp.mtl);
      J.tlist = P.mx_owned.back().as<int32_t>();
      mj.push_back(J);
      P.mx_m2l_maxN = Nall;
    }
    for (int a = 0; a < 8; a++) {
      if (H.cls[a].empty()) continue;
      if (m16() && !merged) {
        std::vector<const uint16_t *> ops;
        std::vector<const float *> uops;
        for (int sl : class_offsets(a)) {
          ops.push_back(X.cop[sl]);
          uops.push_back(f16_ ? X.uop[sl] : nullptr);
        }
        GemmJobBf16 J{};
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(ops);
        J.A = P.mx_owned.back().as<const uint16_t *>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(uops);
        J.U = P.mx_owned.back().as<const float *>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(H.gv[a]);
        J.gidx = P.mx_owned.back().as<int32_t>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(H.cls[a]);
        J.cidx = P.mx_owned.back().as<int32_t>();
        J.nb = (int)ops.size();
        J.N = (int)H.cls[a].size();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(mp.ctp[a]);
        J.tptr = P.mx_owned.back().as<int32_t>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(mp.ctl[a]);
        J.tlist = P.mx_owned.back().as<int32_t>();
        mj.push_back(J);
        P.mx_m2l_maxN = std::max(P.mx_m2l_maxN, J.N);
      }
      if (l >= 3) {
        GemmJob L{};
        P.mx_l2l.owned.emplace_back();
        P.mx_l2l.owned.back().upload(std::vector<const double *>{X.Lc.as<double>() + a * matRR});
        L.A = P.mx_l2l.owned.back().as<const double *>();
        P.mx_l2l.owned.emplace_back();
        P.mx_l2l.owned.back().upload(H.gp[a]);
        L.gidx = P.mx_l2l.owned.back().as<int32_t>();
        P.mx_l2l.owned.emplace_back();
        P.mx_l2l.owned.back().upload(H.cls[a]);
        L.cidx = P.mx_l2l.owned.back().as<int32_t>();
        L.nb = 1;
        L.N = (int)H.cls[a].size();
        lj.push_back(L);
        P.mx_l2l.maxN = std::max(P.mx_l2l.maxN, L.N);
      }
    }
    P.mx_m2l_jobs = (int)mj.size();
    if (!mj.empty()) {
      P.mx_m2l.upload(mj);
      const int nj = P.mx_m2l_jobs, mN = P.mx_m2l_maxN;
      std::vector<int32_t> slot((size_t)nj * mN, -1);
      int minnb = 512;
      for (int j = 0; j < nj; j++) {
        minnb = std::min(minnb, mj[j].nb);
      }
      if (merged) {
        for (int c = 0; c < Nall; c++) slot[c] = mcols[c];
      } else {
        for (int j = 0, a = 0; a < 8; a++) {
          if (H.cls[a].empty()) continue;
          for (size_t c = 0; c < H.cls[a].size(); c++) slot[(size_t)j * mN + c] = H.cls[a][c];
          j++;
        }
      }
      P.mx_maxnb = 0;
      for (int j = 0; j < nj; j++) P.mx_maxnb = std::max(P.mx_maxnb, mj[j].nb);
      P.mx_split = split_for((int64_t)nj * (RPm_ / MBM) * ((mN + MBN - 1) / MBN), minnb);
      P.mx_depth = dma_ok(P, P.mx_split) ? 4 : (np_ == 3 && !f16_) ? 2 : 3;
      // class jobs of two-part forms: also as 192-column tiles for the
      // 64 x 48 wave-tile DMA kernel (its tile lists checked like the others)
      P.mx_tn3 = !merged && np_ == 2 && P.mx_maxnb <= dma_max_blocks(3);
      if (P.mx_tn3) {
        std::vector<GemmJobBf16> mj3 = mj;
        for (int j = 0, a = 0; a < 8; a++) {
          if (H.cls[a].empty()) continue;
          std::vector<int32_t> tp, tl;
          tile_lists(H.gv[a], mj3[j].nb, mj3[j].N, 3 * 64, tp, tl);
          check_gather_job(H.gv[a], mj3[j].nb, mj3[j].N, 3 * 64, tp, tl, tree_.count, "M2L job (192-column tiles)");
          P.mx_owned.emplace_back();
          P.mx_owned.back().upload(tp);
          mj3[j].tptr = P.mx_owned.back().as<int32_t>();
          P.mx_owned.emplace_back();
          P.mx_owned.back().upload(tl);
          mj3[j].tlist = P.mx_owned.back().as<int32_t>();
          j++;
        }
        P.mx_m2l3.upload(mj3);
      }
      P.mx_slot.upload(slot);  // also for the autotuner's split candidates
      reserve_part((size_t)std::min(64, 2 * P.mx_split) * nj * mN * RP_);
      // sparse levels: the per-offset pair jobs where they do less padded
      // work than the class or merged jobs
      plan_m2l_pairs16(P, H, X, (merged ? mp.merged_pairs : mp.class_pairs) * MBN);
    }
    P.mx_l2l.njobs = (int)lj.size();
    if (!lj.empty()) P.mx_l2l.jobs.upload(lj);
    if (prec_ == 4 && P.n_down) {  // FP64 R x R M2L jobs, one per child class, as in the FP64 form
      P.m2l = GemmBatch();
      std::vector<GemmJob> jobs;
      std::vector<const double *> cops[8];  // per class: its operators in block order
      int minnb = 512;
      for (int a = 0; a < 8; a++) {
        if (H.cls[a].empty()) continue;
        std::vector<const double *> &ops = cops[a];
        for (int sl : class_offset_slots(a)) {
          if (!X.dop[sl]) throw Error(NNK_ESTATE, "FP64 truncated M2L: missing translation operator");
          ops.push_back(X.dop[sl]);
        }
        const int nb = (int)ops.size(), N = (int)H.cls[a].size();
        GemmJob J{};
        P.m2l.owned.emplace_back();
        P.m2l.owned.back().upload(ops);
        J.A = P.m2l.owned.back().as<const double *>();
        P.m2l.owned.emplace_back();
        P.m2l.owned.back().upload(H.gv[a]);
        J.gidx = P.m2l.owned.back().as<int32_t>();
        P.m2l.owned.emplace_back();
        P.m2l.owned.back().upload(H.cls[a]);
        J.cidx = P.m2l.owned.back().as<int32_t>();
        J.nb = nb;
        J.N = N;
        std::vector<int32_t> tp, tl;
        tile_lists(H.gv[a], nb, N, GBN1, tp, tl);
        check_gather_job(H.gv[a], nb, N, GBN1, tp, tl, tree_.count, "FP64 truncated M2L job");
        P.m2l.owned.emplace_back();
        P.m2l.owned.back().upload(tp);
        J.tptr = P.m2l.owned.back().as<int32_t>();
        P.m2l.owned.emplace_back();
        P.m2l.owned.back().upload(tl);
        J.tlist = P.m2l.owned.back().as<int32_t>();
        jobs.push_back(J);
        P.m2l.maxN = std::max(P.m2l.maxN, N);
        minnb = std::min(minnb, nb);
      }
      P.m2l.njobs = (int)jobs.size();
      P.m2l.jobs.upload(jobs);
      const int nj = P.m2l.njobs, mN = P.m2l.maxN;
      P.m2l_split = split_for((int64_t)nj * (MpadR_ / GBM) * ((mN + GBN1 - 1) / GBN1), minnb);
      reserve_part((size_t)std::min(64, 2 * P.m2l_split) * nj * mN * RP_);
      plan_m2l_pairs(P, H.cls, H.gv, cops, R_);
    }
    if (f16_) {  // split-FP16 down expansion
      auto hjob = [&](std::vector<GemmJobBf16> &jobs, int &maxN, const std::vector<const uint16_t *> &ops,
                      const std::vector<const float *> &uops, const std::vector<int32_t> &g,
                      const std::vector<int32_t> &c) {
        GemmJobBf16 J{};
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(ops);
        J.A = P.mx_owned.back().as<const uint16_t *>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(uops);
        J.U = P.mx_owned.back().as<const float *>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(g);
        J.gidx = P.mx_owned.back().as<int32_t>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(c);
        J.cidx = P.mx_owned.back().as<int32_t>();
        J.nb = (int)ops.size();
        J.N = (int)c.size();
        std::vector<int32_t> tp, tl;
        tile_lists(g, J.nb, J.N, MBN, tp, tl);
        check_gather_job(g, J.nb, J.N, MBN, tp, tl, tree_.count, "FP16 down expansion job");
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(tp);
        J.tptr = P.mx_owned.back().as<int32_t>();
        P.mx_owned.emplace_back();
        P.mx_owned.back().upload(tl);
        J.tlist = P.mx_owned.back().as<int32_t>();
        jobs.push_back(J);
        maxN = std::max(maxN, J.N);
      };
      std::vector<GemmJobBf16> jd;
      if (!H.tleaves.empty())
        hjob(jd, P.hx_dn_maxN, {X.UrF.as<uint16_t>()}, {X.UrU.as<float>()}, H.tleaves, H.tleaves);
      P.hx_dn_n = (int)jd.size();
      if (!jd.empty()) P.hx_dn.upload(jd);
      P.n_dn_rows = (int)H.dn.size();
      P.dn_rows.upload(H.dn);
    }
  }
  n_upall_ = (int)upall.size();
  upall_.upload(upall);
  auto fin = [](GemmBatch &B, const std::vector<GemmJob> &jobs) {
    B.njobs = (int)jobs.size();
    if (B.njobs) B.jobs.upload(jobs);
  };
  fin(mx_upP_all_, jobs_upP);
  fin(mx_wV_all_, jobs_wV);
  fin(mx_downU_all_, jobs_downU);
}

// whether the DMA-staged M2L kernel can run level plan P with split-K
// factor `split`: depth 4 (128-column tiles) holds kDmaMaxBlocks blocks per
// split in its FP16 unscale table; depths 5 (192-column tiles) and 6 (four
// LDS stages) run the class jobs of two-part forms and hold 192 blocks of
// a tile in LDS
void Fmm::plan_m2l_pairs16(LevelPlan &P, const HostLevel &H, const MixedOps &X, int64_t base_work) {
  P.m2l_npairs = 0;
  P.m2l_pair_on = false;
  P.mx_m2lp_jobs = 0;
  M2lPairPlan HP;
  plan_m2l_pairs_host(H.cls, H.gv, tree_.count, MBN, R_, kPairBufMaxElems, kPairRt16, base_work, HP);
  if (!HP.use) return;
  const int nj = (int)HP.job_slot.size();
  std::vector<const uint16_t *> jop(nj, nullptr);
  std::vector<const float *> juop(nj, nullptr);
  for (int j = 0; j < nj; j++) {
    const int sl = HP.job_slot[j];
    if (sl < 0) continue;
    if (!(jop[j] = X.cop[sl]) || (f16_ && !(juop[j] = X.uop[sl])))
      throw Error(NNK_ESTATE, "per-offset mixed M2L: missing translation operator");
  }
  std::vector<int32_t> col(HP.npairs);
  for (int64_t p = 0; p < HP.npairs; p++) col[p] = (int32_t)p;
  const std::vector<int32_t> tp = {0, 1}, tl = {0};  // one tile, its one block
  auto up = [&](const auto &v) {
    P.mx_owned.emplace_back();
    P.mx_owned.back().upload(v);
    return P.mx_owned.back().p;
  };
  const uint16_t *const *dA = (const uint16_t *const *)up(jop);
  const float *const *dU = (const float *const *)up(juop);
  const int32_t *dsrc = (const int32_t *)up(HP.src);
  const int32_t *dcol = (const int32_t *)up(col);
  const int32_t *dtp = (const int32_t *)up(tp);
  const int32_t *dtl = (const int32_t *)up(tl);
  std::vector<GemmJobBf16> jobs(nj);
  for (int j = 0; j < nj; j++) {
    GemmJobBf16 &J = jobs[j];
    J.A = dA + j;
    J.U = dU + j;
    J.gidx = dsrc + HP.job_p0[j];  // empty jobs (N = 0) return before any read
    J.cidx = dcol + HP.job_p0[j];
    J.nb = 1;
    J.N = HP.job_n[j];
    J.tptr = dtp;
    J.tlist = dtl;
  }
  P.mx_m2lp.upload(jobs);
  P.mx_m2lp_jobs = nj;
  P.m2lp_nodes = (int)HP.node.size();
  P.m2lp_node.upload(HP.node);
  P.m2lp_ptr.upload(HP.ptr);
  P.m2lp_idx.upload(HP.idx);
  if ((size_t)HP.npairs * R_ > pairbuf_elems_) {
    pairbuf_.alloc(sizeof(double) * (size_t)HP.npairs * R_);
    pairbuf_elems_ = (size_t)HP.npairs * R_;
  }
  P.m2l_npairs = (int)HP.npairs;
  P.m2l_pair_on = true;
  P.mx_pair_depth = 4;
}

// the per-offset jobs of level P through the 16-bit M2L kernel of depth 2,
// 3 or 4 (no split-K: one block per job), then the per-target sums
void Fmm::launch_m2l_pairs16(const LevelPlan &P, int depth, hipStream_t s) {
  if (!P.m2l_npairs || !P.mx_m2lp_jobs) throw Error(NNK_ESTATE, "per-offset mixed M2L: no plan for this level");
  if (depth < 2 || depth > 7 || (depth >= 5 && depth <= 6) || ((depth == 3 || depth == 7) && np_ == 3 && !f16_))
    throw Error(NNK_ESTATE, "per-offset mixed M2L: depth not available for this form");
  const int nrb = RPm_ / MBM, nj = P.mx_m2lp_jobs;
  const dim3 grid((unsigned)(nj * nrb));  // one 128-column tile per job
  const GemmJobBf16 *jobs = P.mx_m2lp.as<GemmJobBf16>();
  double *Y = pairbuf_.as<double>();
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(512), 0, s, R_, KCm_, RT_, chat_x_.as<uint16_t>(), (int)tree_.count, Y,
                       R_, jobs, nj, nrb, 1, nullptr, MBN, xun_.as<double>(), 0);
  };
  if (f16_) {
    if (depth == 7) launch(k_m2l_dma<2, 1, true, 2, 3, true>);
    else if (depth == 4) launch(k_m2l_dma<2, 1, true>);
    else if (depth == 2) launch(k_m2l_split<2, 2, 1, true>);
    else launch(k_m2l_split<2, 3, 1, true>);
  } else if (np_ == 3) {
    if (depth == 4) launch(k_m2l_dma<3, 1, false>);
    else launch(k_m2l_split<3, 2, 1, false>);
  } else {
    if (depth == 7) launch(k_m2l_dma<2, 1, false, 2, 3, true>);
    else if (depth == 4) launch(k_m2l_dma<2, 1, false>);
    else if (depth == 2) launch(k_m2l_split<2, 2, 1, false>);
    else launch(k_m2l_split<2, 3, 1, false>);
  }
  NNK_HIP(hipGetLastError());
  pair_sum(P, R_, dhat_.as<double>(), RP_, s);
}

bool Fmm::dma_ok(const LevelPlan &P, int split, int depth) const {
  if (split < 1) return false;
  if (depth == 5) return P.mx_tn3 && P.mx_maxnb <= dma_max_blocks(3);
  if (depth == 6 || depth == 8) return P.mx_tn3 && P.mx_maxnb <= dma_max_blocks(2, 4);
  return !f16_ || (P.mx_maxnb + split - 1) / split <= kDmaMaxBlocks;
}

// M2L of level l in the mixed form with register-ring depth `depth` (2 or
// 3; 4 selects the DMA-staged kernel, 5 the DMA-staged kernel with 64 x 48
// wave tiles, 6 the DMA-staged kernel with four LDS stages, 7 and 8 the
// DMA-staged kernel with fragment prefetch, three and four LDS stages) and
// split-K
// factor `split` (the part buffer must hold it)
void Fmm::launch_m2l_mixed(int l, int depth, int split, hipStream_t s) {
  LevelPlan &P = plans_[l];
  if (!P.mx_m2l_jobs) return;
  if (depth > 10) {  // 12-14: the per-offset pair jobs, kernel depth 2-4
    if (split != 1) throw Error(NNK_ESTATE, "per-offset mixed M2L: no split-K");
    launch_m2l_pairs16(P, depth - 10, s);
    return;
  }
  const int bn = depth == 5 ? 3 * 64 : MBN;
  const int nrb = RPm_ / MBM, nct = (P.mx_m2l_maxN + bn - 1) / bn;
  const int nj = P.mx_m2l_jobs, mN = P.mx_m2l_maxN;
  if (split > 1 && (size_t)split * nj * mN * RP_ > part_elems_) throw Error(NNK_ESTATE, "mixed M2L: split buffer too small");
  const dim3 grid((unsigned)(nj * split * nrb * nct));
  const int zrow = tree_.count;
  double *d = dhat_.as<double>();
  double *pp = split > 1 ? part_.as<double>() : nullptr;
  const GemmJobBf16 *jobs = (depth == 5 ? P.mx_m2l3 : P.mx_m2l).as<GemmJobBf16>();
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(512), 0, s, R_, KCm_, RT_, chat_x_.as<uint16_t>(), zrow, d, RP_, jobs, nj,
                       nrb, split, pp, mN, xun_.as<double>(), 0);
  };
  if (depth >= 4 && !dma_ok(P, split, depth))
    throw Error(NNK_ESTATE, "mixed M2L: too many operator blocks for the DMA-staged kernel");
  if (depth >= 5 && np_ != 2) throw Error(NNK_ESTATE, "mixed M2L: this DMA-staged kernel needs two parts");
  if (f16_) {
    if (depth == 8) launch(k_m2l_dma<2, 1, true, 2, 4, true>);
    else if (depth == 7) launch(k_m2l_dma<2, 1, true, 2, 3, true>);
    else if (depth == 6) launch(k_m2l_dma<2, 1, true, 2, 4>);
    else if (depth == 5) launch(k_m2l_dma<2, 1, true, 3>);
    else if (depth == 4) launch(k_m2l_dma<2, 1, true>);
    else if (depth == 2) launch(k_m2l_split<2, 2, 1, true>);
    else launch(k_m2l_split<2, 3, 1, true>);
  } else if (np_ == 3) {
    if (depth == 4) launch(k_m2l_dma<3, 1, false>);
    else launch(k_m2l_split<3, 2, 1, false>);  // a third register stage would spill
  } else {
    if (depth == 8) launch(k_m2l_dma<2, 1, false, 2, 4, true>);
    else if (depth == 7) launch(k_m2l_dma<2, 1, false, 2, 3, true>);
    else if (depth == 6) launch(k_m2l_dma<2, 1, false, 2, 4>);
    else if (depth == 5) launch(k_m2l_dma<2, 1, false, 3>);
    else if (depth == 4) launch(k_m2l_dma<2, 1, false>);
    else if (depth == 2) launch(k_m2l_split<2, 2, 1, false>);
    else launch(k_m2l_split<2, 3, 1, false>);
  }
  NNK_HIP(hipGetLastError());
  if (split > 1) split_reduce(P.mx_slot, nj * mN, R_, RP_, split, d, s);
}

void Fmm::launch_m2l_t(int l, int split, hipStream_t s) {
  LevelPlan &P = plans_[l];
  gemm_dims(P.m2l, chat_.as<double>(), RP_, dhat_.as<double>(), RP_, 0, s, R_, R_, MpadR_, KpadR_, split,
            &P.m2l_slot, 1);
}

// one launch of the split-FP16 GEMM (no split-K, register-ring depth 2):
// C[:, cidx] (+)= sum_b A_b X[:, gidx_b] * (*xun)
void Fmm::launch_h16(const DBuf &jobs, int njobs, int maxN, int M, int KC, const uint16_t *X, double *C, int ldc,
                     const double *xun, int beta, hipStream_t s) {
  if (!njobs) return;
  const int nrb = (M + MBM - 1) / MBM, nct = (maxN + MBN - 1) / MBN;
  const int RT = nrb * MBM / 16;
  // TAG 2: a kernel symbol of its own, apart from the M2L launches (TAG 1)
  hipLaunchKernelGGL((k_m2l_split<2, 2, 2, true>), dim3((unsigned)(njobs * nrb * nct)), dim3(512), 0, s, M, KC, RT,
                     X, tree_.count, C, ldc, jobs.as<GemmJobBf16>(), njobs, nrb, 1, nullptr, maxN, xun, beta);
  NNK_HIP(hipGetLastError());
}

// rows of X (length len, stride ldx) -> FP16 parts scaled by one power of
// two from their max |value|; *xun receives the unscale factor
void Fmm::split_h(const DBuf &rows, int nrows, const double *X, int ldx, int len, int KC, uint16_t *out,
                  double *xun, hipStream_t s) {
  if (!nrows) return;
  const int64_t tot = (int64_t)nrows * KC * 32;
  hipLaunchKernelGGL(k_amax_reset, dim3(1), dim3(64), 0, s, amax_.as<unsigned long long>());
  hipLaunchKernelGGL(k_absmax, dim3(std::min<int64_t>(1024, (tot + 255) / 256)), dim3(256), 0, s, rows.as<int32_t>(),
                     nrows, X, ldx, len, amax_.as<unsigned long long>());
  hipLaunchKernelGGL(k_split_rows_h, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, rows.as<int32_t>(), nrows,
                     X, ldx, len, KC, amax_.as<unsigned long long>(), out, xun);
  NNK_HIP(hipGetLastError());
}

void Fmm::apply_mixed(const double *d_f, double *d_u, hipStream_t s) {
  const int K = K_, R = R_;
  double *chk = chk_.as<double>(), *c = chat_.as<double>(), *d = dhat_.as<double>();
  mark(0, true, s);
  strengths(d_f, s);
  mark(0, false, s);
  if (maxlev_ >= 2 && n_pleaves_) {
    mark(1, true, s);
    launch_surf(pleaves_.as<int32_t>(), n_pleaves_, nullptr, nullptr, 1, 0, chk, s);
    mark(1, false, s);
  }
  mark(2, true, s);
  // the leaves' up expansions (every level at once), then M2M level by
  // level: a parent's children at the level below are leaves or parents
  // done in the previous step
  gemm_dims(mx_upP_all_, chk, K, c, RP_, 0, s, R, K, MpadR_, Kpad_);
  for (int l = maxlev_; l >= 2; l--) {
    LevelPlan &P = plans_[l];
    gemm_dims(P.mx_m2m, c, RP_, c, RP_, 0, s, R, R, MpadR_, KpadR_, P.mx_m2m_split, &P.mx_m2m_slot);
  }
  if (n_upall_ && m16()) {
    const int64_t tot = (int64_t)n_upall_ * KCm_ * 32;
    if (f16_) {
      hipLaunchKernelGGL(k_amax_reset, dim3(1), dim3(64), 0, s, amax_.as<unsigned long long>());
      hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, upall_.as<int32_t>(), n_upall_, c, RP_, R,
                         amax_.as<unsigned long long>());
      hipLaunchKernelGGL(k_split_rows_h, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                         upall_.as<int32_t>(), n_upall_, c, RP_, R, KCm_, amax_.as<unsigned long long>(),
                         chat_x_.as<uint16_t>(), xun_.as<double>());
    } else {
      hipLaunchKernelGGL(k_split_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                         upall_.as<int32_t>(), n_upall_, c, RP_, R, KCm_, np_, chat_x_.as<uint16_t>());
    }
    NNK_HIP(hipGetLastError());
  }
  gemm_dims(mx_wV_all_, c, RP_, uppot_.as<double>(), K, 0, s, K, R, Mpad_, KpadR_);
  mark(2, false, s);
  if (n_x_all_) {  // P2L of every level (the up pass is done with the check potentials)
    mark(5, true, s);
    launch_surf(x_all_.as<int32_t>(), n_x_all_, xptr_.as<int64_t>(), xidx_.as<int32_t>(), 0, 0, chk, s);
    mark(5, false, s);
  }
  for (int l = 2; l <= maxlev_; l++) {
    LevelPlan &P = plans_[l];
    if (P.mx_m2l_jobs) {
      mark(3, true, s);
      if (P.m2l_pair_on) launch_m2l_pairs16(P, P.mx_pair_depth, s);
      else launch_m2l_mixed(l, P.mx_depth, P.mx_split, s);
      mark(3, false, s);
    } else if (prec_ == 4 && P.m2l.njobs) {
      mark(3, true, s);
      if (P.m2l_pair_on) launch_m2l_pairs(P, chat_.as<double>(), RP_, dhat_.as<double>(), RP_, R_, MpadR_, KpadR_, s);
      else launch_m2l_t(l, P.m2l_split, s);
      mark(3, false, s);
    }
    mark(4, true, s);
    gemm_dims(P.mx_l2l, d, RP_, d, RP_, 1, s, R, R, MpadR_, KpadR_);
    mark(4, false, s);
    if (P.n_x) {
      mark(5, true, s);
      gemm_dims(P.mx_xP, chk, K, d, RP_, 1, s, R, K, MpadR_, Kpad_);
      mark(5, false, s);
    }
    if (f16_) {  // U_R is orthonormal: the down expansion keeps its accuracy in split FP16
      mark(6, true, s);
      split_h(P.dn_rows, P.n_dn_rows, d, RP_, R, KCm_, dhat_x_.as<uint16_t>(), dun_lv_.as<double>() + l, s);
      launch_h16(P.hx_dn, P.hx_dn_n, P.hx_dn_maxN, K, KCm_, dhat_x_.as<uint16_t>(), downpot_.as<double>(), K,
                 dun_lv_.as<double>() + l, 0, s);
      mark(6, false, s);
    }
  }
  if (!f16_) {  // FP64 down expansion of the target leaves of every level at once
    mark(6, true, s);
    gemm_dims(mx_downU_all_, d, RP_, downpot_.as<double>(), K, 0, s, K, R, Mpad_, KpadR_);
    mark(6, false, s);
  }
  mark(7, true, s);
  launch_leaf(1, d_u, s);
  mark(7, false, s);
}

}  // namespace nnk
