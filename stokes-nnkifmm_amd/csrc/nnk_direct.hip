// Direct regularized-Stokeslet sum u = K(src -> tgt) q on the GPU: the
// reference's blockcomputation (Tools/Other/blockcomputation.m:17-28)
// without materialising any block of K.  One thread per target; source
// tiles staged in LDS and shared by the 256 targets of a block.
//
// Few targets against many sources (an error sample of 4,096 targets over
// 2^20 sources) would fill 16 workgroups, so the sources are split into S
// contiguous parts (grid.y), each part's sums go to a partial buffer
// P[S][3 nt] and a second kernel adds the parts in order s = 0..S-1.  The
// split depends on (ns, nt) only, so a given input is always summed in the
// same order.  The partial buffer comes from the stream-ordered allocator
// (nnk_direct_apply_device) or from a static device array of fixed size
// (nnk_direct_apply_device_noalloc); both run the same kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "nnk_common.hpp"
#include "nnk_kernels.hpp"

namespace nnk {

// partial sums of the no-allocation entry point: 16 MiB on every GPU the
// library runs on, held until the process ends
constexpr int64_t kDirectScratchDoubles = int64_t(1) << 21;
static __device__ __align__(256) double scratch[kDirectScratchDoubles];

template <bool PART>
__global__ __launch_bounds__(256) void k_direct(const double *__restrict__ src, int64_t ns,
                                                const double *__restrict__ q,
                                                const double *__restrict__ tgt, int64_t nt,
                                                double e2, double c, int nparts, double *__restrict__ out) {
  __shared__ double sp[256 * 6];
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int part = blockIdx.y;
  const int64_t b0 = ns * part / nparts, b1 = ns * (part + 1) / nparts;
  double tx = 0, ty = 0, tz = 0;
  if (t < nt) { tx = tgt[3 * t]; ty = tgt[3 * t + 1]; tz = tgt[3 * t + 2]; }
  double ux = 0, uy = 0, uz = 0;
  for (int64_t s0 = b0; s0 < b1; s0 += 256) {
    const int cnt = (int)min<int64_t>(256, b1 - s0);
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
    double *o = PART ? out + (size_t)part * 3 * nt : out;
    const double f = PART ? 1.0 : c;
    o[3 * t] = f * ux;
    o[3 * t + 1] = f * uy;
    o[3 * t + 2] = f * uz;
  }
}

// out[i] = c * sum_{s = 0..nparts-1} P[s][i], in order
__global__ __launch_bounds__(256) void k_direct_reduce(const double *__restrict__ P, int nparts, int64_t n3,
                                                       double c, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n3) return;
  double a = 0.0;
  for (int s = 0; s < nparts; s++) a += P[(size_t)s * n3 + i];
  out[i] = c * a;
}

// Source parts for (ns, nt): enough workgroups to fill the chip (about
// 2,048), at least 1,024 sources per part, none when the targets alone give
// 512 workgroups.  Then parts * nt < 655,360 (tb < 512 target blocks:
// parts <= 2048/tb + 1), so the partials always fit the 2^21 doubles of the
// static array.
int direct_parts(int64_t ns, int64_t nt) {
  const int64_t tb = (nt + 255) / 256;
  if (tb == 0 || tb >= 512) return 1;
  const int64_t want = (2048 + tb - 1) / tb;
  const int64_t most = std::max<int64_t>(1, ns / 1024);
  return (int)std::max<int64_t>(1, std::min(want, most));
}

int64_t direct_scratch_doubles(int64_t ns, int64_t nt) {
  const int parts = direct_parts(ns, nt);
  return parts > 1 ? (int64_t)parts * 3 * nt : 0;
}

static void direct_launch(const double *src, int64_t ns, const double *q, const double *tgt, int64_t nt,
                          double eps, double mu, double *out, double *P, int parts, hipStream_t s) {
  const double c = 1.0 / (8.0 * kPi * mu);
  const unsigned tb = (unsigned)((nt + 255) / 256);
  if (parts == 1) {
    hipLaunchKernelGGL(k_direct<false>, dim3(tb), dim3(256), 0, s, src, ns, q, tgt, nt, eps * eps, c, 1, out);
    NNK_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_direct<true>, dim3(tb, (unsigned)parts), dim3(256), 0, s, src, ns, q, tgt, nt, eps * eps, c,
                     parts, P);
  NNK_HIP(hipGetLastError());
  const int64_t n3 = 3 * nt;
  hipLaunchKernelGGL(k_direct_reduce, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, P, parts, n3, c, out);
  NNK_HIP(hipGetLastError());
}

void direct_device(const double *src, int64_t ns, const double *q, const double *tgt, int64_t nt,
                   double eps, double mu, double *out, hipStream_t s) {
  if (nt == 0) return;
  const int parts = direct_parts(ns, nt);
  if (parts == 1) {
    direct_launch(src, ns, q, tgt, nt, eps, mu, out, nullptr, 1, s);
    return;
  }
  void *P = nullptr;
  NNK_HIP(hipMallocAsync(&P, sizeof(double) * (size_t)direct_scratch_doubles(ns, nt), s));
  try {
    direct_launch(src, ns, q, tgt, nt, eps, mu, out, (double *)P, parts, s);
  } catch (...) {
    (void)hipFreeAsync(P, s);
    throw;
  }
  NNK_HIP(hipFreeAsync(P, s));
}

// Device address of `scratch` on the current device, looked up on the first
// call per device (the host-side name is not a device address).
static double *scratch_ptr() {
  constexpr int kMaxDev = 64;
  static std::mutex mu;
  static double *ptr[kMaxDev] = {};
  int dev = 0;
  NNK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) throw Error(3, "device id out of range for the static scratch table");
  std::lock_guard<std::mutex> g(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    NNK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = static_cast<double *>(p);
  }
  return ptr[dev];
}

void direct_device_noalloc(const double *src, int64_t ns, const double *q, const double *tgt, int64_t nt,
                           double eps, double mu, double *out, hipStream_t s) {
  if (nt == 0) return;
  const int parts = direct_parts(ns, nt);
  const int64_t need = direct_scratch_doubles(ns, nt);
  if (need > kDirectScratchDoubles)
    throw Error(1, "nnk_direct_apply_device_noalloc: " + std::to_string(need * 8) +
                       " bytes of partial sums exceed the 16 MiB static scratch");
  double *P = parts > 1 ? scratch_ptr() : nullptr;
  direct_launch(src, ns, q, tgt, nt, eps, mu, out, P, parts, s);
}

}  // namespace nnk
