// Nearest-neighbour quadrature map on the GPU (Tools/Nearest/nearestMatrix.m):
// NN(q, j) = 1/#ties for every coarse point j at the minimum squared
// distance from fine point q.
//
// The reference compares every fine point with every coarse point
// (O(Q N), :23-33).  Here the coarse points are binned once into a uniform
// grid (counting sort on the host, cells of about two points) and each fine
// point (one thread, fine points visited in cell order so that a wavefront
// walks neighbouring cells) scans the grid in Chebyshev rings around its
// cell until the ring boundary is provably farther than the best distance
// found.  The result is exact, ties included: the squared distance is formed
// as the reference sums it, ((dx^2 + dy^2) + dz^2) with each operation
// rounded (fused multiply-add disabled), the stop test keeps a margin for the
// rounding of the binning and of the distances, and a second pass emits
// every coarse point whose distance equals the minimum bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "nnk_common.hpp"

namespace nnk {

// HIP's __dmul_rn / __dadd_rn are plain operators, which the compiler may
// fuse into FMAs (contraction is on by default); forbid that here so that
// every product and sum is rounded on its own, as in the reference
__device__ __forceinline__ double d2_exact(const double *f, double cx, double cy, double cz) {
#pragma clang fp contract(off)
  const double dx = f[0] - cx, dy = f[1] - cy, dz = f[2] - cz;
  return (dx * dx + dy * dy) + dz * dz;
}

struct NnGrid {
  const double *cpos;     // coarse points in cell order, xyz
  const int32_t *cidx;    // their indices
  const int32_t *cstart;  // cells + 1 offsets (cell = (a * g1 + b) * g2 + c)
  int g[3];
  double lo[3], h, delta;
};

// pass 1 (EMIT = false): minimum squared distance and tie count per fine
// point; pass 2 (EMIT = true): the columns at that minimum, at off[fine].
template <bool EMIT>
__global__ __launch_bounds__(256) void k_nn_grid(NnGrid G, const double *__restrict__ fine,
                                                 const int32_t *__restrict__ forder, int64_t q,
                                                 double *__restrict__ dmin, int32_t *__restrict__ cnt,
                                                 const int64_t *__restrict__ off, int64_t *__restrict__ col) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= q) return;
  const int64_t fi = forder[i];
  const double f[3] = {fine[3 * fi], fine[3 * fi + 1], fine[3 * fi + 2]};
  int c[3];
  for (int d = 0; d < 3; d++) {
    double t = floor((f[d] - G.lo[d]) / G.h);
    t = fmin(fmax(t, 0.0), (double)(G.g[d] - 1));
    c[d] = (int)t;
  }
  double best = EMIT ? dmin[fi] : __builtin_inf();
  int count = 0;
  int64_t o = EMIT ? off[fi] : 0;
  const int64_t oend = EMIT ? off[fi] + cnt[fi] : 0;   // never write past the counted ties
  auto scan = [&](int a, int b, int cc) {
    const int cell = (a * G.g[1] + b) * G.g[2] + cc;
    for (int p = G.cstart[cell]; p < G.cstart[cell + 1]; p++) {
      const double d2 = d2_exact(f, G.cpos[3 * p], G.cpos[3 * p + 1], G.cpos[3 * p + 2]);
      if (EMIT) {
        if (d2 == best && o < oend) col[o++] = G.cidx[p];
      } else if (d2 < best) {
        best = d2;
        count = 1;
      } else if (d2 == best) {
        count++;
      }
    }
  };
  for (int r = 0;; r++) {
    const int a0 = max(c[0] - r, 0), a1 = min(c[0] + r, G.g[0] - 1);
    const int b0 = max(c[1] - r, 0), b1 = min(c[1] + r, G.g[1] - 1);
    for (int a = a0; a <= a1; a++)
      for (int b = b0; b <= b1; b++) {
        if (abs(a - c[0]) == r || abs(b - c[1]) == r) {
          for (int cc = max(c[2] - r, 0); cc <= min(c[2] + r, G.g[2] - 1); cc++) scan(a, b, cc);
        } else {
          if (c[2] - r >= 0) scan(a, b, c[2] - r);
          if (r > 0 && c[2] + r < G.g[2]) scan(a, b, c[2] + r);
        }
      }
    // distance from f to the cells outside rings 0..r (infinite on a side
    // where the grid ends)
    double dr = __builtin_inf();
    bool covered = true;
    for (int d = 0; d < 3; d++) {
      if (c[d] - r > 0) { dr = fmin(dr, f[d] - (G.lo[d] + (c[d] - r) * G.h)); covered = false; }
      if (c[d] + r < G.g[d] - 1) { dr = fmin(dr, G.lo[d] + (c[d] + r + 1) * G.h - f[d]); covered = false; }
    }
    if (covered) break;
    const double m = dr - G.delta;
    if (m > 0 && best < m * m * (1.0 - 1e-12)) break;
  }
  if (!EMIT) {
    dmin[fi] = best;
    cnt[fi] = count;
  }
}

namespace {

// cells per axis and cell width for about `target` cells over extents ext
void grid_shape(const double ext[3], int64_t target, int g[3], double &h) {
  const double emax = std::max({ext[0], ext[1], ext[2]});
  if (!(emax > 0)) {
    g[0] = g[1] = g[2] = 1;
    h = 1.0;
    return;
  }
  auto cells = [&](double hh) {
    double p = 1;
    for (int d = 0; d < 3; d++) p *= std::max(1.0, std::ceil(ext[d] / hh));
    return p;
  };
  double lo = emax / 4.0e6, hi = emax;   // cells(hi) = 1
  for (int it = 0; it < 100; it++) {
    double mid = std::sqrt(lo * hi);
    if (cells(mid) > (double)target) lo = mid;
    else hi = mid;
  }
  h = hi;
  for (int d = 0; d < 3; d++) g[d] = (int)std::max(1.0, std::ceil(ext[d] / h));
}

inline int cell_1d(double x, double lo, double h, int g) {
  double t = std::floor((x - lo) / h);
  t = std::min(std::max(t, 0.0), (double)(g - 1));
  return (int)t;
}

}  // namespace

// Returns nnz; fills CSC (colptr n+1, row, val) when nnz <= cap.
int64_t nearest_matrix(const double *coarse, int64_t n, const double *fine, int64_t q,
                       int64_t *colptr, int64_t *row, double *val, int64_t cap) {
  if (n <= 0) throw Error(1, "nearestMatrix: no coarse points");
  if (n > INT32_MAX) throw Error(1, "nearestMatrix: too many coarse points");
  for (int64_t k = 0; k < 3 * n; k++)
    if (!std::isfinite(coarse[k])) throw Error(1, "nearestMatrix: non-finite coarse point");
  for (int64_t k = 0; k < 3 * q; k++)
    if (!std::isfinite(fine[k])) throw Error(1, "nearestMatrix: non-finite fine point");
  // ---- grid over the coarse points ----
  double lo[3], hi[3], ext[3], amax = 0;
  for (int d = 0; d < 3; d++) lo[d] = hi[d] = coarse[d];
  for (int64_t j = 0; j < n; j++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], coarse[3 * j + d]);
      hi[d] = std::max(hi[d], coarse[3 * j + d]);
    }
  for (int d = 0; d < 3; d++) {
    ext[d] = hi[d] - lo[d];
    amax = std::max({amax, std::fabs(lo[d]), std::fabs(hi[d])});
  }
  int g[3];
  double h;
  grid_shape(ext, std::max<int64_t>(1, std::min<int64_t>(n / 2, 1 << 22)), g, h);
  const int64_t ncell = (int64_t)g[0] * g[1] * g[2];
  std::vector<int32_t> cstart(ncell + 1, 0), ccell(n);
  for (int64_t j = 0; j < n; j++) {
    const double *p = coarse + 3 * j;
    int32_t cell = (cell_1d(p[0], lo[0], h, g[0]) * g[1] + cell_1d(p[1], lo[1], h, g[1])) * g[2] +
                   cell_1d(p[2], lo[2], h, g[2]);
    ccell[j] = cell;
    cstart[cell + 1]++;
  }
  for (int64_t k = 0; k < ncell; k++) cstart[k + 1] += cstart[k];
  std::vector<double> cpos(3 * n);
  std::vector<int32_t> cidx(n);
  {
    std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
    for (int64_t j = 0; j < n; j++) {
      int32_t s = fill[ccell[j]]++;
      cidx[s] = (int32_t)j;
      for (int d = 0; d < 3; d++) cpos[3 * s + d] = coarse[3 * j + d];
    }
  }
  // fine points in cell order (neighbouring threads, neighbouring cells)
  std::vector<int32_t> forder(q);
  {
    std::vector<int64_t> fst(ncell + 1, 0);
    std::vector<int32_t> fcell(q);
    for (int64_t i = 0; i < q; i++) {
      const double *p = fine + 3 * i;
      fcell[i] = (cell_1d(p[0], lo[0], h, g[0]) * g[1] + cell_1d(p[1], lo[1], h, g[1])) * g[2] +
                 cell_1d(p[2], lo[2], h, g[2]);
      fst[fcell[i] + 1]++;
    }
    for (int64_t k = 0; k < ncell; k++) fst[k + 1] += fst[k];
    for (int64_t i = 0; i < q; i++) forder[fst[fcell[i]]++] = (int32_t)i;
  }
  DBuf dcpos, dcidx, dcstart, df, dord, dmin, dcnt;
  dcpos.upload(cpos);
  dcidx.upload(cidx);
  dcstart.upload(cstart);
  df.alloc(sizeof(double) * 3 * std::max<int64_t>(q, 1));
  if (q) NNK_HIP(hipMemcpy(df.p, fine, sizeof(double) * 3 * q, hipMemcpyHostToDevice));
  dord.upload(forder);
  dmin.alloc(sizeof(double) * std::max<int64_t>(q, 1));
  dcnt.alloc(sizeof(int32_t) * std::max<int64_t>(q, 1));
  NnGrid G{};
  G.cpos = dcpos.as<double>();
  G.cidx = dcidx.as<int32_t>();
  G.cstart = dcstart.as<int32_t>();
  for (int d = 0; d < 3; d++) {
    G.g[d] = g[d];
    G.lo[d] = lo[d];
  }
  G.h = h;
  G.delta = 1e-6 * h + 1e-12 * amax;
  const unsigned nb = (unsigned)((q + 255) / 256);
  if (q) {
    hipLaunchKernelGGL(k_nn_grid<false>, dim3(nb), dim3(256), 0, 0, G, df.as<double>(), dord.as<int32_t>(), q,
                       dmin.as<double>(), dcnt.as<int32_t>(), nullptr, nullptr);
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
    hipLaunchKernelGGL(k_nn_grid<true>, dim3(nb), dim3(256), 0, 0, G, df.as<double>(), dord.as<int32_t>(), q,
                       dmin.as<double>(), dcnt.as<int32_t>(), doff.as<int64_t>(), dcol.as<int64_t>());
    NNK_HIP(hipGetLastError());
  }
  std::vector<int64_t> cols(nnz);
  if (nnz) NNK_HIP(hipMemcpy(cols.data(), dcol.p, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost));
  // COO grouped by fine point (ascending) -> CSC, rows ascending within a
  // column whatever the order of the ties within a fine point
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
