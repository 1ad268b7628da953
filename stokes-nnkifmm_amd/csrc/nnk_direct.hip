// Direct regularized-Stokeslet sum u = K(src -> tgt) q on the GPU: the
// reference's blockcomputation (Tools/Other/blockcomputation.m:17-28)
// without materialising any block of K.  One thread per target; source
// tiles staged in LDS and shared by the 256 targets of a block.
#include <hip/hip_runtime.h>

#include "nnk_common.hpp"
#include "nnk_kernels.hpp"

namespace nnk {

__global__ __launch_bounds__(256) void k_direct(const double *__restrict__ src, int64_t ns,
                                                const double *__restrict__ q,
                                                const double *__restrict__ tgt, int64_t nt,
                                                double e2, double c, double *__restrict__ out) {
  __shared__ double sp[256 * 6];
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double tx = 0, ty = 0, tz = 0;
  if (t < nt) { tx = tgt[3 * t]; ty = tgt[3 * t + 1]; tz = tgt[3 * t + 2]; }
  double ux = 0, uy = 0, uz = 0;
  for (int64_t s0 = 0; s0 < ns; s0 += 256) {
    const int cnt = (int)min<int64_t>(256, ns - s0);
    if (threadIdx.x < cnt) {
      const int64_t s = s0 + threadIdx.x;
      for (int d = 0; d < 3; d++) {
        sp[6 * threadIdx.x + d] = src[3 * s + d];
        sp[6 * threadIdx.x + 3 + d] = q[3 * s + d];
      }
    }
    __syncthreads();
    if (t < nt)
      for (int j = 0; j < cnt; j++)
        stokeslet_acc(tx, ty, tz, sp[6 * j], sp[6 * j + 1], sp[6 * j + 2], sp[6 * j + 3],
                      sp[6 * j + 4], sp[6 * j + 5], e2, ux, uy, uz);
    __syncthreads();
  }
  if (t < nt) {
    out[3 * t] = c * ux;
    out[3 * t + 1] = c * uy;
    out[3 * t + 2] = c * uz;
  }
}

void direct_device(const double *src, int64_t ns, const double *q, const double *tgt, int64_t nt,
                   double eps, double mu, double *out, hipStream_t s) {
  if (nt == 0) return;
  const double c = 1.0 / (8.0 * kPi * mu);
  hipLaunchKernelGGL(k_direct, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, src, ns, q, tgt,
                     nt, eps * eps, c, out);
  NNK_HIP(hipGetLastError());
}

}  // namespace nnk
