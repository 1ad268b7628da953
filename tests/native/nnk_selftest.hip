// Test-only entry points (not part of the C ABI): exercise the GEMM kernel
// against a host product, read back the FMM's per-node arrays, probe MFMA
// lane layouts, and run the host-only planners and input checks of the
// library (the CPU tests call these without a GPU).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "nnk_gemm.hpp"
#include "nnk_handles.hpp"

using namespace nnk;

// bn: the tile width the library launches (GBN1 = 64 for M2L, GBN0 = 32
// for the other GEMMs)
extern "C" int nnk_selftest_gemm(int M, int K, int Mpad, int Kpad, const double *At, int nb,
                                 const double *X, int nx, int ldx, const int *gidx,
                                 const int *cidx, int N, double *C, int nc, int ldc, int beta, int bn) {
  try {
    DBuf dA, dX, dC, dG, dCi, dP, dJ;
    size_t mat = (size_t)Kpad * Mpad;
    dA.alloc(sizeof(double) * mat * nb);
    NNK_HIP(hipMemcpy(dA.p, At, sizeof(double) * mat * nb, hipMemcpyHostToDevice));
    dX.alloc(sizeof(double) * (size_t)nx * ldx);
    NNK_HIP(hipMemcpy(dX.p, X, sizeof(double) * (size_t)nx * ldx, hipMemcpyHostToDevice));
    dC.alloc(sizeof(double) * (size_t)nc * ldc);
    NNK_HIP(hipMemcpy(dC.p, C, sizeof(double) * (size_t)nc * ldc, hipMemcpyHostToDevice));
    std::vector<const double *> ptrs(nb);
    for (int b = 0; b < nb; b++) ptrs[b] = dA.as<double>() + b * mat;
    dP.upload(ptrs);
    GemmJob j{};
    j.A = dP.as<const double *>();
    if (gidx) { dG.upload(std::vector<int>(gidx, gidx + (size_t)nb * N)); j.gidx = dG.as<int32_t>(); }
    if (cidx) { dCi.upload(std::vector<int>(cidx, cidx + N)); j.cidx = dCi.as<int32_t>(); }
    j.nb = nb;
    j.N = N;
    dJ.upload(std::vector<GemmJob>{j});
    const int nrb = Mpad / GBM;
    if (Kpad % GBK || (bn != GBN0 && bn != GBN1)) return 2;
    dim3 grid(nrb * ((N + bn - 1) / bn));
    if (bn == GBN1)
      hipLaunchKernelGGL((k_gemm_gather<GBM, GBN1, GBK, 0>), grid, dim3(256), 0, 0, M, K, Mpad, Kpad,
                         dX.as<double>(), ldx, dC.as<double>(), ldc, beta, dJ.as<GemmJob>(), 1, nrb, 1,
                         nullptr, N);
    else
      hipLaunchKernelGGL((k_gemm_gather<GBM, GBN0, GBK, 0>), grid, dim3(256), 0, 0, M, K, Mpad, Kpad,
                         dX.as<double>(), ldx, dC.as<double>(), ldc, beta, dJ.as<GemmJob>(), 1, nrb, 1,
                         nullptr, N);
    NNK_HIP(hipGetLastError());
    NNK_HIP(hipDeviceSynchronize());
    NNK_HIP(hipMemcpy(C, dC.p, sizeof(double) * (size_t)nc * ldc, hipMemcpyDeviceToHost));
    return 0;
  } catch (...) {
    return 1;
  }
}

extern "C" int nnk_selftest_fmm_dump(nnk_fmm *h, int which, double *out) {
  try {
    h->f->debug_copy(which, out);
    return 0;
  } catch (...) {
    return 1;
  }
}

extern "C" int nnk_selftest_fmm_K(nnk_fmm *h) { return h->f->K(); }

// One v_mfma_f64_16x16x4_f64 per trial on a single wave: lane l supplies
// a[t*64+l], b[t*64+l]; d[(t*64+l)*4+r] receives accumulator r.
__global__ void k_mfma_probe(const double *a, const double *b, double *d, int trials) {
  const int l = threadIdx.x;
  for (int t = 0; t < trials; t++) {
    d4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t * 64 + l], b[t * 64 + l], c, 0, 0, 0);
    for (int r = 0; r < 4; r++) d[(t * 64 + l) * 4 + r] = c[r];
  }
}

extern "C" int nnk_selftest_mfma_probe(const double *a, const double *b, double *d, int trials) {
  try {
    DBuf da, db, dd;
    da.alloc(sizeof(double) * 64 * trials);
    db.alloc(sizeof(double) * 64 * trials);
    dd.alloc(sizeof(double) * 256 * trials);
    NNK_HIP(hipMemcpy(da.p, a, sizeof(double) * 64 * trials, hipMemcpyHostToDevice));
    NNK_HIP(hipMemcpy(db.p, b, sizeof(double) * 64 * trials, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, da.as<double>(), db.as<double>(),
                       dd.as<double>(), trials);
    NNK_HIP(hipGetLastError());
    NNK_HIP(hipDeviceSynchronize());
    NNK_HIP(hipMemcpy(d, dd.p, sizeof(double) * 256 * trials, hipMemcpyDeviceToHost));
    return 0;
  } catch (...) {
    return 1;
  }
}

// One v_mfma_f32_16x16x32_bf16 per trial on a single wave: lane l supplies
// 8 bf16 values a[(t*64+l)*8 + j], b[...]; d[(t*64+l)*4+r] receives
// accumulator r.  Inputs are bf16 bit patterns (uint16).
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ void k_mfma_bf16_probe(const uint16_t *a, const uint16_t *b, float *d, int trials) {
  const int l = threadIdx.x;
  for (int t = 0; t < trials; t++) {
    bf8v av, bv;
    for (int j = 0; j < 8; j++) {
      uint16_t x = a[(t * 64 + l) * 8 + j], y = b[(t * 64 + l) * 8 + j];
      av[j] = __builtin_bit_cast(__bf16, x);
      bv[j] = __builtin_bit_cast(__bf16, y);
    }
    f4v c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
    for (int r = 0; r < 4; r++) d[(t * 64 + l) * 4 + r] = c[r];
  }
}

extern "C" int nnk_selftest_mfma_bf16_probe(const uint16_t *a, const uint16_t *b, float *d, int trials) {
  try {
    DBuf da, db, dd;
    da.alloc(sizeof(uint16_t) * 512 * trials);
    db.alloc(sizeof(uint16_t) * 512 * trials);
    dd.alloc(sizeof(float) * 256 * trials);
    NNK_HIP(hipMemcpy(da.p, a, sizeof(uint16_t) * 512 * trials, hipMemcpyHostToDevice));
    NNK_HIP(hipMemcpy(db.p, b, sizeof(uint16_t) * 512 * trials, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_bf16_probe, dim3(1), dim3(64), 0, 0, da.as<uint16_t>(), db.as<uint16_t>(),
                       dd.as<float>(), trials);
    NNK_HIP(hipGetLastError());
    NNK_HIP(hipDeviceSynchronize());
    NNK_HIP(hipMemcpy(d, dd.p, sizeof(float) * 256 * trials, hipMemcpyDeviceToHost));
    return 0;
  } catch (...) {
    return 1;
  }
}

static void copy_err(const std::string &m, char *err, int errlen) {
  if (err && errlen > 0) {
    std::strncpy(err, m.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = 0;
  }
}

// Host only: the nearest-map check of nnk_tree_set_nearest.
extern "C" int nnk_selftest_validate_nn(int64_t ncol, int64_t nfine, const double *fine, const int64_t *colptr,
                                        const int64_t *row, const double *val, char *err, int errlen) {
  std::string m;
  int rc = nnk::validate_nn_csc(ncol, nfine, fine, colptr, row, val, m);
  copy_err(m, err, errlen);
  return rc;
}

// Host only: the mixed-form M2L plan of one level.  sizes[8] = nodes per
// child class, gv = the classes' gather tables concatenated (189 x sizes[a]
// each), count = node count.  Returns 0 with merged[0] and pairs[0..1] =
// (class pairs, merged pairs or -1), or nonzero with the refusal in err.
extern "C" int nnk_selftest_m2l_plan(const int32_t *sizes, const int32_t *gv, int64_t count, int BN,
                                     int32_t *merged, int64_t *pairs, char *err, int errlen) {
  try {
    std::vector<int32_t> cls[8], g[8];
    size_t off = 0;
    for (int a = 0; a < 8; a++) {
      for (int c = 0; c < sizes[a]; c++) cls[a].push_back(c);
      const size_t n = (size_t)nnk::class_offset_slots(a).size() * sizes[a];
      g[a].assign(gv + off, gv + off + n);
      off += n;
    }
    nnk::M2lLevelPlan P;
    nnk::plan_m2l_level(cls, g, count, BN, P);
    merged[0] = P.merged ? 1 : 0;
    pairs[0] = P.class_pairs;
    pairs[1] = P.merged_pairs;
    copy_err("", err, errlen);
    return 0;
  } catch (const std::exception &e) {
    copy_err(e.what(), err, errlen);
    return 1;
  }
}

// Host only: the per-offset M2L plan of one level (FP64 forms).  sizes[8]
// and gv as for nnk_selftest_m2l_plan, but class a's column c is the node
// slot base_a + c (base_a = sizes[0] + ... + sizes[a-1], count >= their
// sum); M and cap: product rows and pair-buffer limit in doubles.
// info[0..4] = use, class work, pair work, pairs, jobs.  With use and
// room (maxjobs, maxpairs, maxnodes) the tables are copied out: job_slot,
// job_p0, job_n (jobs); src, idx (pairs); node, ptr (nodes, nodes + 1).
// Returns 0, 2 if the tables did not fit, or 1 with the refusal in err.
extern "C" int nnk_selftest_m2l_pairs(const int32_t *sizes, const int32_t *gv, int64_t count, int BN, int64_t M,
                                      int64_t cap, int64_t *info, int32_t *job_slot, int32_t *job_p0,
                                      int32_t *job_n, int64_t maxjobs, int32_t *src, int32_t *idx,
                                      int64_t maxpairs, int32_t *node, int32_t *ptr, int64_t maxnodes, char *err,
                                      int errlen) {
  try {
    std::vector<int32_t> cls[8], g[8];
    size_t off = 0;
    int32_t base = 0;
    for (int a = 0; a < 8; a++) {
      for (int c = 0; c < sizes[a]; c++) cls[a].push_back(base + c);
      base += sizes[a];
      const size_t n = (size_t)nnk::class_offset_slots(a).size() * sizes[a];
      g[a].assign(gv + off, gv + off + n);
      off += n;
    }
    nnk::M2lPairPlan P;
    nnk::plan_m2l_pairs_host(cls, g, count, BN, M, cap, 90, -1, P);
    info[0] = P.use ? 1 : 0;
    info[1] = P.class_work;
    info[2] = P.pair_work;
    info[3] = P.npairs;
    info[4] = (int64_t)P.job_slot.size();
    copy_err("", err, errlen);
    if (!P.use) return 0;
    if ((int64_t)P.job_slot.size() > maxjobs || P.npairs > maxpairs || (int64_t)P.node.size() > maxnodes)
      return 2;
    std::copy(P.job_slot.begin(), P.job_slot.end(), job_slot);
    std::copy(P.job_p0.begin(), P.job_p0.end(), job_p0);
    std::copy(P.job_n.begin(), P.job_n.end(), job_n);
    std::copy(P.src.begin(), P.src.end(), src);
    std::copy(P.idx.begin(), P.idx.end(), idx);
    std::copy(P.node.begin(), P.node.end(), node);
    std::copy(P.ptr.begin(), P.ptr.end(), ptr);
    return 0;
  } catch (const std::exception &e) {
    copy_err(e.what(), err, errlen);
    return 1;
  }
}

namespace nnk {
int svd_jacobi(const double *A, int n, double *U, double *s, double *V);
}

// The check-to-equivalent SVD (GPU one-sided Jacobi) of an n x n row-major
// matrix; returns the sweeps used, or -1 with the message in err.
extern "C" int nnk_selftest_svd(const double *A, int n, double *U, double *s, double *V, char *err, int errlen) {
  try {
    return nnk::svd_jacobi(A, n, U, s, V);
  } catch (const std::exception &e) {
    copy_err(e.what(), err, errlen);
    return -1;
  }
}
