// KIFMM matvec on gfx950, FP64 form.
//
// Pass structure follows Tools/Passes/upwardPass.m and downwardPass.m:
//   S2M   (Up.m:31-33)          check potential of a leaf's sources on its
//                               down surface                  k_surf_eval
//   M2M   (upwardPass.m:81-93)  children's up densities onto the parent's
//                               down surface                  gemm (8 ops)
//   up solve  K(up->down) \ chk per level                     gemm x2 (SVD)
//   M2L   (DownNonleaf.m:42)    V list onto the up surface    gemm (189 ops)
//   L2L   (DownNonleaf.m:44)    parent down density           gemm (1 op, +=)
//   P2L   (DownNonleaf.m:43)    X-list sources                k_surf_eval
//   down solve  K(down->up) \ chk per level                   gemm x2 (SVD)
//   L2T+P2P+M2T (Downleaf.m:37-40) per target leaf            k_leaf
// Per-level operators use the level's canonical box geometry (all boxes of a
// level share their width; their absolute corners differ from the canonical
// ones only by rounding).  The check-to-equivalent solves apply one-sided
// Jacobi SVD factors, x = V (S^-1 (U^T b)), which is backward stable like
// the reference's pivoted LU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "nnk_fmm.hpp"
#include "nnk_gemm.hpp"
#include "nnk_kernels.hpp"

namespace nnk {

int svd_batch(const double *const *A, int n, int nb, double *const *U, double *const *s, double *const *V);


// ---------------------------------------------------------------------------
// kernels

__global__ void k_strength(int64_t nsrc, const int32_t *__restrict__ from,
                           const int64_t *__restrict__ cptr, const int32_t *__restrict__ ccol,
                           const double *__restrict__ cw, const double *__restrict__ f,
                           double *__restrict__ q) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsrc) return;
  if (from) {
    int64_t j = from[k];
    q[3 * k] = f[3 * j];
    q[3 * k + 1] = f[3 * j + 1];
    q[3 * k + 2] = f[3 * j + 2];
  } else {
    double gx = 0, gy = 0, gz = 0;
    for (int64_t e = cptr[k]; e < cptr[k + 1]; e++) {
      int64_t j = ccol[e];
      double w = cw[e];
      gx += w * f[3 * j];
      gy += w * f[3 * j + 1];
      gz += w * f[3 * j + 2];
    }
    q[3 * k] = gx;
    q[3 * k + 1] = gy;
    q[3 * k + 2] = gz;
  }
}

// At[(3q+b)*Mpad + 3p+a] = c * S_ab(t_p - s_q)    (dense kernel.m block)
__global__ void k_form_At(const double *__restrict__ sp, int ns, const double *__restrict__ tp,
                          int nt, double e2, double c, double *__restrict__ At, int Mpad) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ns * nt) return;
  int p = idx % nt, q = idx / nt;
  double rx = tp[3 * p] - sp[3 * q], ry = tp[3 * p + 1] - sp[3 * q + 1],
         rz = tp[3 * p + 2] - sp[3 * q + 2];
  double r2 = rx * rx + ry * ry + rz * rz;
  double re = r2 + e2;
  double inv = rsqrt(re);
  double s = c * inv * inv * inv;
  double iso = r2 + 2.0 * e2;
  double r[3] = {rx, ry, rz};
  for (int b = 0; b < 3; b++)
    for (int a = 0; a < 3; a++)
      At[(size_t)(3 * q + b) * Mpad + 3 * p + a] = (r[a] * r[b] + (a == b ? iso : 0.0)) * s;
}

struct SurfArgs {
  const int32_t *nodes;       // work items (node ids)
  const int64_t *lptr;        // per-node source-leaf lists (X); null = the node itself
  const int32_t *lidx;
  int down, accumulate;
  const double *corners;
  const int32_t *surf_idx;
  int nq, n1, K;
  const int32_t *src_ptr;     // per node id, [src_ptr[n], src_ptr[n+1])
  const double *src_pos, *src_q;
  double e2, c;
  double *chk;
};

// Check potential on a node surface from the sources of a set of leaves
// (S2M: the leaf itself onto its down surface; P2L: the X list onto the up
// surface).  One 512-thread block per node; sources staged in LDS 512 at a
// time.  The nq surface points are taken in chunks of c, each point shared
// by nsub = 512 / c lanes that split the sources; their partial sums are
// added in a fixed order.  Every chunk restages the sources (a global-load
// latency and two barriers), so c is chosen by a cost model in units of
// one pair per lane: chunks x (sources per lane + 24).  A leaf's ~32
// sources onto 386 points (res 9) then take one chunk of 386 lanes rather
// than seven chunks of 56 points x 9 lanes.
constexpr int kSurfThreads = 512;

template <bool F32>
__global__ __launch_bounds__(kSurfThreads) void k_surf_eval(SurfArgs a) {
  constexpr int T = kSurfThreads;
  const int n = a.nodes[blockIdx.x];
  // FP32 math: positions as FP32 offsets from this node's centre (the
  // smaller box of an X-list pair; S2M pairs share the box)
  typedef typename std::conditional<F32, SrcStageRel<T>, SrcStage<false, T>>::type Stage;
  __shared__ Stage st;
  __shared__ double red[T * 3];
  const double *box = a.corners + 6 * (size_t)(n - 1);
  double cx, cy, cz;
  box_centre(box, cx, cy, cz);
  const float e2f = (float)a.e2;
  const int tid = threadIdx.x;
  int64_t lb = 0, le = 1;
  if (a.lptr) { lb = a.lptr[n]; le = a.lptr[n + 1]; }
  // chunk size: the c = ceil(nq / k), k = 1..8, of least modelled cost
  int nsrc = 0;
  for (int64_t e = lb; e < le; e++) {
    const int leaf = a.lptr ? a.lidx[e] : n;
    nsrc += a.src_ptr[leaf + 1] - a.src_ptr[leaf];
  }
  int csz = min(a.nq, T);
  int64_t best = -1;
  for (int k = 1; k <= 8; k++) {
    const int c = (a.nq + k - 1) / k;
    if (c > T) continue;
    const int nsub = T / c;
    const int64_t cost = (int64_t)((a.nq + c - 1) / c) * ((nsrc + nsub - 1) / nsub + 24);
    if (best < 0 || cost < best) { best = cost; csz = c; }
  }
  for (int p0 = 0; p0 < a.nq; p0 += csz) {
    const int ntc = min(csz, a.nq - p0);
    const int nsub = T / ntc;
    const bool act = tid < ntc * nsub;
    const int ti = tid % ntc, sub = tid / ntc;
    const int p = p0 + ti;
    double tx = 0, ty = 0, tz = 0;
    if (act) surf_point(box, a.down, a.surf_idx[p], a.n1, tx, ty, tz);
    double ux = 0, uy = 0, uz = 0;
    for (int64_t e = lb; e < le; e++) {
      const int leaf = a.lptr ? a.lidx[e] : n;
      const int sb = a.src_ptr[leaf], se = a.src_ptr[leaf + 1];
      for (int s0 = sb; s0 < se; s0 += T) {
        const int cnt = min(T, se - s0);
        if (tid < cnt) {
          const size_t s = (size_t)(s0 + tid);
          if constexpr (F32)
            st.put(tid, a.src_pos[3 * s] - cx, a.src_pos[3 * s + 1] - cy, a.src_pos[3 * s + 2] - cz,
                   a.src_q[3 * s], a.src_q[3 * s + 1], a.src_q[3 * s + 2]);
          else
            st.put(tid, a.src_pos[3 * s], a.src_pos[3 * s + 1], a.src_pos[3 * s + 2], a.src_q[3 * s],
                   a.src_q[3 * s + 1], a.src_q[3 * s + 2]);
        }
        __syncthreads();
        if (act) {
          if constexpr (F32)
            st.run(sub, nsub, cnt, (float)(tx - cx), (float)(ty - cy), (float)(tz - cz), e2f, ux, uy, uz);
          else
            st.run(sub, nsub, cnt, tx, ty, tz, a.e2, ux, uy, uz);
        }
        __syncthreads();
      }
    }
    if (act) {
      red[3 * tid] = ux;
      red[3 * tid + 1] = uy;
      red[3 * tid + 2] = uz;
    }
    __syncthreads();
    if (tid < ntc) {
      double v[3] = {0, 0, 0};
      for (int q = 0; q < nsub; q++)
        for (int d = 0; d < 3; d++) v[d] += red[3 * (tid + q * ntc) + d];
      double *o = a.chk + (size_t)(n - 1) * a.K + 3 * (p0 + tid);
      if (a.accumulate) {
        o[0] += a.c * v[0]; o[1] += a.c * v[1]; o[2] += a.c * v[2];
      } else {
        o[0] = a.c * v[0]; o[1] = a.c * v[1]; o[2] = a.c * v[2];
      }
    }
    __syncthreads();
  }
}

struct LeafArgs {
  const int32_t *leaves;
  const double *corners;
  const int32_t *surf_idx;
  int nq, n1, K;
  const int32_t *tgt_ptr;
  const double *tgt_pos;
  const int32_t *tgt_out;
  const int32_t *src_ptr;
  const double *src_pos, *src_q;
  const int64_t *uptr; const int32_t *uidx;
  const int64_t *wptr; const int32_t *widx;
  const double *uppot, *downpot;
  double e2, c;
  int far;          // 0: sources of the leaf itself only (reducedPass.m)
  int tc;           // targets per workgroup and pass (<= T); blockIdx.y picks the chunk
  double *out;      // ntgt x 3
};

// Velocity at the targets of one leaf: L2T from its down density, P2P from
// its U-list sources, M2T from its W-list up densities (Downleaf.m:37-40).
// Each target is shared by nsub lanes that split the sources; partial sums
// are combined in a fixed order.  T threads per workgroup (the autotuner
// picks 128, 256 or 512: more lanes per target, or more targets in flight).
// The sources of the U-list leaves are staged as one concatenated list, T
// at a time (a leaf holds ~32: staging leaf by leaf would pay a load
// latency and two barriers per leaf); lists longer than kLeafMaxU leaves
// are staged leaf by leaf.
constexpr int kLeafMaxU = 512;

template <bool F32, int T>
__global__ __launch_bounds__(T) void k_leaf(LeafArgs a) {
  const int n = a.leaves[blockIdx.x];
  // FP32 math: the surface sources (L2T, M2T) go through the offset stage
  // with the origin at the centre of the surface's node (this leaf for
  // L2T, the smaller W node for M2T); the U-list sources keep the FP64
  // separation
  __shared__ union {
    SrcStage<F32, T> pts;
    SrcStageRel<T> srf;
  } st;
  __shared__ double red[T * 3];
  __shared__ int upre[kLeafMaxU + 1];  // source-count prefix over the U list
  const int tid = threadIdx.x;
  const int tb = a.tgt_ptr[n], nt = a.tgt_ptr[n + 1] - tb;
  const int64_t u0 = a.far ? a.uptr[n] : 0, u1 = a.far ? a.uptr[n + 1] : 0;
  const int nu = (int)(u1 - u0);
  const bool ucat = a.far && nu <= kLeafMaxU;
  if (ucat) {
    for (int e = tid; e < nu; e += T) {
      const int lf = a.uidx[u0 + e];
      upre[e + 1] = a.src_ptr[lf + 1] - a.src_ptr[lf];
    }
    __syncthreads();
    if (tid == 0) {
      upre[0] = 0;
      for (int e = 1; e <= nu; e++) upre[e] += upre[e - 1];
    }
    __syncthreads();
  }
  for (int t0 = blockIdx.y * a.tc; t0 < nt; t0 += gridDim.y * a.tc) {
    const int ntc = min(a.tc, nt - t0);
    const int nsub = T / ntc;
    const bool act = tid < ntc * nsub;
    const int ti = tid % ntc, sub = tid / ntc;
    double tx = 0, ty = 0, tz = 0;
    if (act) {
      const double *pp = a.tgt_pos + 3 * (size_t)(tb + t0 + ti);
      tx = pp[0]; ty = pp[1]; tz = pp[2];
    }
    double ux = 0, uy = 0, uz = 0;
    auto consume = [&](int cnt) {
      if (act)
        st.pts.run(sub, nsub, cnt, tx, ty, tz, a.e2, ux, uy, uz);
    };
    auto surface = [&](int node, int down, const double *dens) {
      const double *box = a.corners + 6 * (size_t)(node - 1);
      double cx = 0, cy = 0, cz = 0;
      if constexpr (F32) box_centre(box, cx, cy, cz);
      const float txf = (float)(tx - cx), tyf = (float)(ty - cy), tzf = (float)(tz - cz);
      for (int p0 = 0; p0 < a.nq; p0 += T) {
        const int cnt = min(T, a.nq - p0);
        if (tid < cnt) {
          double x, y, z;
          surf_point(box, down, a.surf_idx[p0 + tid], a.n1, x, y, z);
          const double *q = dens + (size_t)(node - 1) * a.K + 3 * (p0 + tid);
          if constexpr (F32) st.srf.put(tid, x - cx, y - cy, z - cz, q[0], q[1], q[2]);
          else st.pts.put(tid, x, y, z, q[0], q[1], q[2]);
        }
        __syncthreads();
        if constexpr (F32) {
          if (act) st.srf.run(sub, nsub, cnt, txf, tyf, tzf, (float)a.e2, ux, uy, uz);
        } else {
          consume(cnt);
        }
        __syncthreads();
      }
    };
    auto sources = [&](int leaf) {
      const int sb = a.src_ptr[leaf], se = a.src_ptr[leaf + 1];
      for (int s0 = sb; s0 < se; s0 += T) {
        const int cnt = min(T, se - s0);
        if (tid < cnt) {
          const size_t s = (size_t)(s0 + tid);
          st.pts.put(tid, a.src_pos[3 * s], a.src_pos[3 * s + 1], a.src_pos[3 * s + 2], a.src_q[3 * s],
                     a.src_q[3 * s + 1], a.src_q[3 * s + 2]);
        }
        __syncthreads();
        consume(cnt);
        __syncthreads();
      }
    };
    // the U-list sources as one list: entry g lies in leaf e with
    // upre[e] <= g < upre[e + 1] (binary search; empty leaves are skipped)
    auto usources = [&]() {
      const int tot = upre[nu];
      for (int g0 = 0; g0 < tot; g0 += T) {
        const int cnt = min(T, tot - g0);
        if (tid < cnt) {
          const int g = g0 + tid;
          int lo = 0, hi = nu;
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (upre[mid] <= g) lo = mid;
            else hi = mid;
          }
          const size_t s = (size_t)(a.src_ptr[a.uidx[u0 + lo]] + (g - upre[lo]));
          st.pts.put(tid, a.src_pos[3 * s], a.src_pos[3 * s + 1], a.src_pos[3 * s + 2], a.src_q[3 * s],
                     a.src_q[3 * s + 1], a.src_q[3 * s + 2]);
        }
        __syncthreads();
        consume(cnt);
        __syncthreads();
      }
    };
    if (a.far) {
      surface(n, 1, a.downpot);
      for (int64_t e = a.wptr[n]; e < a.wptr[n + 1]; e++) surface(a.widx[e], 0, a.uppot);
      if (ucat) usources();
      else
        for (int64_t e = u0; e < u1; e++) sources(a.uidx[e]);
    } else {
      sources(n);
    }
    red[3 * tid] = ux; red[3 * tid + 1] = uy; red[3 * tid + 2] = uz;
    __syncthreads();
    if (tid < ntc) {
      double sx = 0, sy = 0, sz = 0;
      for (int s = 0; s < nsub; s++) {
        sx += red[3 * (s * ntc + tid)];
        sy += red[3 * (s * ntc + tid) + 1];
        sz += red[3 * (s * ntc + tid) + 2];
      }
      const int o = a.tgt_out[tb + t0 + tid];
      a.out[3 * (size_t)o] = a.c * sx;
      a.out[3 * (size_t)o + 1] = a.c * sy;
      a.out[3 * (size_t)o + 2] = a.c * sz;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host side

namespace {

std::vector<int32_t> surface_codes(int res, int shells) {
  std::vector<int32_t> c;
  for (int iz = 0; iz < res; iz++)
    for (int ix = 0; ix < res; ix++)
      for (int iy = 0; iy < res; iy++) {
        bool inside = iy >= shells && iy < res - shells && ix >= shells && ix < res - shells &&
                      iz >= shells && iz < res - shells;
        if (!inside) c.push_back(ix | (iy << 8) | (iz << 16));
      }
  return c;
}

double lin_h(double d1, double d2, int k, int n1) {
  if (k == 0) return d1;
  if (k == n1) return d2;
  return d1 + ((double)k * (d2 - d1)) / (double)n1;
}

// surface points of a box (lo, width w per axis) as in genupsurf/gendownsurf
std::vector<double> surface_pts(const std::vector<int32_t> &codes, int res, const double lo[3],
                                double w, bool down) {
  double l[3], h[3];
  for (int d = 0; d < 3; d++) {
    l[d] = lo[d];
    h[d] = lo[d] + w;
    if (down) {
      double ww = h[d] - l[d];
      l[d] -= ww;
      h[d] += ww;
    }
  }
  std::vector<double> p(3 * codes.size());
  for (size_t i = 0; i < codes.size(); i++) {
    p[3 * i] = lin_h(l[0], h[0], codes[i] & 255, res - 1);
    p[3 * i + 1] = lin_h(l[1], h[1], (codes[i] >> 8) & 255, res - 1);
    p[3 * i + 2] = lin_h(l[2], h[2], (codes[i] >> 16) & 255, res - 1);
  }
  return p;
}

inline int offset_slot(int ox, int oy, int oz) { return (ox + 3) * 49 + (oy + 3) * 7 + (oz + 3); }

// the 189 V-list offsets available to a child of class a (a = 4gx+2gy+gz)
// the M2L visiting order of the 316 offsets (m2l_slot_order)
const std::vector<int> &visit_order() {
  static const std::vector<int> order = [] {
    std::vector<int> o;
    for (int p = 0; p < 8; p++) {
      auto vals = [](int par) {
        std::vector<int> v;
        for (int x = -3; x <= 3; x++)
          if ((x & 1) == par) v.push_back(x);
        return v;
      };
      const std::vector<int> X = vals((p >> 2) & 1), Y = vals((p >> 1) & 1), Z = vals(p & 1);
      int row = 0;
      for (size_t i = 0; i < X.size(); i++)
        for (size_t jj = 0; jj < Y.size(); jj++, row++) {
          const size_t j = (i & 1) ? Y.size() - 1 - jj : jj;
          for (size_t kk = 0; kk < Z.size(); kk++) {
            const size_t k = (row & 1) ? Z.size() - 1 - kk : kk;
            if (std::max({std::abs(X[i]), std::abs(Y[j]), std::abs(Z[k])}) < 2) continue;
            o.push_back(offset_slot(X[i], Y[j], Z[k]));
          }
        }
    }
    if (o.size() != 316) throw Error(NNK_ESTATE, "M2L offset order: not 316 offsets");
    return o;
  }();
  return order;
}

std::vector<std::array<int, 3>> class_offsets(int a) {
  const int g[3] = {(a >> 2) & 1, (a >> 1) & 1, a & 1};
  std::vector<std::array<int, 3>> out;
  for (int sl : visit_order()) {
    const int o[3] = {sl / 49 - 3, (sl / 7) % 7 - 3, sl % 7 - 3};
    bool in = true;
    for (int d = 0; d < 3; d++) in = in && o[d] >= -2 - g[d] && o[d] <= 3 - g[d];
    if (in) out.push_back({o[0], o[1], o[2]});
  }
  return out;
}

// v in decreasing order of w[v[i]] (ties in increasing order of v[i])
void order_by_work(std::vector<int32_t> &v, const std::vector<int64_t> &w) {
  std::sort(v.begin(), v.end(), [&](int32_t a, int32_t b) { return w[a] != w[b] ? w[a] > w[b] : a < b; });
}

template <class T> DBuf make_dbuf(const std::vector<T> &v) {
  DBuf d;
  d.upload(v);
  return d;
}

}  // namespace

// tile configurations of the gather GEMM (BM x BN); BK = 16
template <int BN, int TAG>
static void launch_gemm(int nrb, int M, int K, int Mpad, int Kpad, const double *X, int ldx,
                        double *C, int ldc, int beta, const GemmBatch &b, hipStream_t s, int nsplit = 1,
                        double *P = nullptr) {
  const int nct = (b.maxN + BN - 1) / BN;
  dim3 grid((unsigned)(b.njobs * nsplit * nrb * nct));
  hipLaunchKernelGGL((k_gemm_gather<GBM, BN, GBK, TAG>), grid, dim3(256), 0, s, M, K, Mpad, Kpad,
                     X, ldx, C, ldc, beta, b.jobs.as<GemmJob>(), b.njobs, nrb, nsplit, P, b.maxN);
  NNK_HIP(hipGetLastError());
}

__global__ void k_split_reduce(const int32_t *__restrict__ slot, int M, int ldc, int nsplit,
                               const double *__restrict__ P, size_t pstride, double *__restrict__ C) {
  const int c = blockIdx.x;
  const int cs = slot[c];
  if (cs < 0) return;
  for (int row = threadIdx.x; row < M; row += blockDim.x) {
    double v = 0.0;
    for (int sp = 0; sp < nsplit; sp++) v += P[sp * pstride + (size_t)c * ldc + row];
    C[(size_t)cs * ldc + row] = v;
  }
}

// per-offset M2L: C[:, node[j]] = the sum of the product columns of target
// j's pairs (Y column p = rows 0..M-1 at p * M), in their listed order
__global__ void k_pair_sum(const int32_t *__restrict__ node, const int32_t *__restrict__ ptr,
                           const int32_t *__restrict__ idx, const double *__restrict__ Y, int M,
                           double *__restrict__ C, int ldc) {
  const int j = blockIdx.x;
  const int t = node[j], e0 = ptr[j], e1 = ptr[j + 1];
  for (int r = threadIdx.x; r < M; r += blockDim.x) {
    double v = 0.0;
    for (int e = e0; e < e1; e++) v += Y[(size_t)idx[e] * M + r];
    C[(size_t)t * ldc + r] = v;
  }
}

void tile_lists(const std::vector<int32_t> &gidx, int nb, int N, int BN, std::vector<int32_t> &ptr,
                std::vector<int32_t> &list) {
  const int nct = (N + BN - 1) / BN;
  ptr.assign(nct + 1, 0);
  list.clear();
  for (int ct = 0; ct < nct; ct++) {
    const int c0 = ct * BN, c1 = std::min(N, c0 + BN);
    for (int b = 0; b < nb; b++) {
      const int32_t *g = gidx.data() + (size_t)b * N;
      for (int c = c0; c < c1; c++)
        if (g[c] >= 0) { list.push_back(b); break; }
    }
    ptr[ct + 1] = (int32_t)list.size();
  }
}

void check_gather_job(const std::vector<int32_t> &gidx, int nb, int N, int BN, const std::vector<int32_t> &tptr,
                      const std::vector<int32_t> &tlist, int64_t count, const char *what) {
  auto fail = [&](const char *why) { throw Error(NNK_ESTATE, std::string(what) + ": " + why); };
  if (nb < 0 || N < 0 || BN <= 0 || gidx.size() != (size_t)nb * (size_t)N) fail("gather table size");
  for (int32_t g : gidx)
    if (g < -1 || g >= count) fail("gather index out of range");
  const size_t nct = ((size_t)N + BN - 1) / BN;
  if (tptr.size() != nct + 1 || tptr[0] != 0 || (size_t)tptr[nct] != tlist.size()) fail("tile pointers");
  for (size_t ct = 0; ct < nct; ct++)
    if (tptr[ct + 1] < tptr[ct]) fail("tile pointers not monotone");
  for (int32_t b : tlist)
    if (b < 0 || b >= nb) fail("tile-list block out of range");
}

const std::vector<int> &m2l_slot_order() { return visit_order(); }

std::vector<int> class_offset_slots(int a) {
  std::vector<int> sls;
  for (const auto &o : class_offsets(a)) sls.push_back(offset_slot(o[0], o[1], o[2]));
  return sls;
}

void plan_m2l_level(const std::vector<int32_t> *cls, const std::vector<int32_t> *gv, int64_t count, int BN,
                    M2lLevelPlan &P) {
  P = M2lLevelPlan{};
  for (int a = 0; a < 8; a++) {
    if (cls[a].empty()) continue;
    const int nb = (int)class_offset_slots(a).size(), Na = (int)cls[a].size();
    tile_lists(gv[a], nb, Na, BN, P.ctp[a], P.ctl[a]);
    check_gather_job(gv[a], nb, Na, BN, P.ctp[a], P.ctl[a], count, "M2L class job");
    P.class_pairs += (int64_t)P.ctl[a].size();
  }
  std::vector<int> moff(343, -1);
  for (int sl : m2l_slot_order()) {
    moff[sl] = (int)P.msl.size();
    P.msl.push_back(sl);
  }
  for (int a = 0; a < 8; a++) P.mcols.insert(P.mcols.end(), cls[a].begin(), cls[a].end());
  const int Nall = (int)P.mcols.size();
  if (Nall == 0 || Nall > 8 * BN) return;
  P.mg.assign(P.msl.size() * (size_t)Nall, -1);
  for (int a = 0, off = 0; a < 8; a++) {
    const int Na = (int)cls[a].size();
    if (!Na) continue;
    const std::vector<int> sls = class_offset_slots(a);
    for (size_t b = 0; b < sls.size(); b++)
      for (int c = 0; c < Na; c++) P.mg[(size_t)moff[sls[b]] * Nall + off + c] = gv[a][b * Na + c];
    off += Na;
  }
  tile_lists(P.mg, (int)P.msl.size(), Nall, BN, P.mtp, P.mtl);
  P.merged_pairs = (int64_t)P.mtl.size();
  P.merged = P.merged_pairs < P.class_pairs;
  if (P.merged) check_gather_job(P.mg, (int)P.msl.size(), Nall, BN, P.mtp, P.mtl, count, "merged M2L job");
}

void plan_m2l_pairs_host(const std::vector<int32_t> *cls, const std::vector<int32_t> *gv, int64_t count, int BN,
                         int64_t M, int64_t cap_elems, int64_t rt, int64_t base_work, M2lPairPlan &P) {
  P = M2lPairPlan{};
  if (BN <= 0) throw Error(NNK_ESTATE, "per-offset M2L plan: bad tile width");
  // class tiles: one (column tile, operator block) product per active block
  // of each BN-column tile
  std::vector<int> nslot(343, 0);
  for (int a = 0; a < 8; a++) {
    if (cls[a].empty()) continue;
    const std::vector<int> sls = class_offset_slots(a);
    const int nb = (int)sls.size(), Na = (int)cls[a].size();
    std::vector<int32_t> tp, tl;
    tile_lists(gv[a], nb, Na, BN, tp, tl);
    check_gather_job(gv[a], nb, Na, BN, tp, tl, count, "M2L class job");
    P.class_work += (int64_t)tl.size() * BN;
    for (int b = 0; b < nb; b++)
      for (int c = 0; c < Na; c++)
        if (gv[a][(size_t)b * Na + c] >= 0) nslot[sls[b]]++;
  }
  int64_t tiles = 0;
  for (int sl = 0; sl < 343; sl++) {
    P.npairs += nslot[sl];
    tiles += (nslot[sl] + BN - 1) / BN;
  }
  // every pair's product column (M doubles) is written once and read once
  // by the sums: 16 M bytes, counted as rt / M of a column's product
  P.pair_work = tiles * BN + (P.npairs * rt + M - 1) / std::max<int64_t>(M, 1);
  const int64_t other = base_work >= 0 ? base_work : P.class_work;
  P.use = P.npairs > 0 && P.pair_work < other && P.npairs * M <= cap_elems;
  if (!P.use) return;
  // pairs grouped by offset slot (in the M2L visiting order,
  // m2l_slot_order), within a slot by class and column; each target's list
  // then comes out in that order, as its class and merged jobs add them
  std::vector<int64_t> base(343, 0), fill(343, 0);
  {
    int64_t at = 0;
    for (int sl : m2l_slot_order()) {
      base[sl] = at;
      at += nslot[sl];
    }
  }
  fill = base;
  P.src.assign(P.npairs, -1);
  std::vector<int32_t> pair_tgt(P.npairs, -1);
  for (int sl : m2l_slot_order()) {
    if (!nslot[sl]) continue;
    for (int a = 0; a < 8; a++) {
      if (cls[a].empty()) continue;
      const std::vector<int> sls = class_offset_slots(a);
      const int Na = (int)cls[a].size();
      for (size_t b = 0; b < sls.size(); b++) {
        if (sls[b] != sl) continue;
        for (int c = 0; c < Na; c++) {
          const int32_t g = gv[a][b * Na + c];
          if (g < 0) continue;
          const int64_t p = fill[sl]++;
          P.src[p] = g;
          pair_tgt[p] = cls[a][c];
        }
      }
    }
  }
  // jobs: the BN-pair tiles of each slot in slot order, dealt to the 8
  // XCDs in contiguous runs (tile k of run x becomes job 8k + x), so that
  // the tiles of one operator run side by side on one XCD and share its
  // rows through that L2 (workgroup b of the launch goes to XCD b mod 8,
  // and the job count is a multiple of 8)
  std::vector<std::array<int64_t, 3>> tl;  // slot, first pair, pairs
  for (int sl : m2l_slot_order())
    for (int64_t p0 = base[sl]; p0 < base[sl] + nslot[sl]; p0 += BN)
      tl.push_back({sl, p0, std::min<int64_t>(BN, base[sl] + nslot[sl] - p0)});
  const int64_t T = (int64_t)tl.size();
  int64_t runmax = 0;
  for (int x = 0; x < 8; x++) runmax = std::max(runmax, (x + 1) * T / 8 - x * T / 8);
  const int64_t nj = 8 * runmax;
  if (nj > INT32_MAX || P.npairs > INT32_MAX) throw Error(NNK_ESTATE, "per-offset M2L plan: too many pairs");
  P.job_slot.assign(nj, -1);
  P.job_p0.assign(nj, 0);
  P.job_n.assign(nj, 0);
  for (int x = 0; x < 8; x++)
    for (int64_t k = 0, t = x * T / 8; t < (x + 1) * T / 8; k++, t++) {
      P.job_slot[8 * k + x] = (int32_t)tl[t][0];
      P.job_p0[8 * k + x] = (int32_t)tl[t][1];
      P.job_n[8 * k + x] = (int32_t)tl[t][2];
    }
  // per target (the level's class columns, each once): its pairs
  std::vector<int32_t> pos(count, -1);
  for (int a = 0; a < 8; a++)
    for (int32_t t : cls[a]) {
      if (t < 0 || t >= count || pos[t] >= 0) throw Error(NNK_ESTATE, "per-offset M2L plan: bad target list");
      pos[t] = (int32_t)P.node.size();
      P.node.push_back(t);
    }
  P.ptr.assign(P.node.size() + 1, 0);
  for (int64_t p = 0; p < P.npairs; p++) P.ptr[pos[pair_tgt[p]] + 1]++;
  for (size_t j = 0; j < P.node.size(); j++) P.ptr[j + 1] += P.ptr[j];
  std::vector<int32_t> at(P.ptr.begin(), P.ptr.end() - 1);
  P.idx.assign(P.npairs, -1);
  for (int64_t p = 0; p < P.npairs; p++) P.idx[at[pos[pair_tgt[p]]]++] = (int32_t)p;
}

int Fmm::split_for(int64_t tiles, int nb) {
  // aim at ~4 workgroups per CU, with at least 4 operator blocks per split
  const int64_t want = 1024;
  if (tiles >= want / 2) return 1;
  int s = (int)((want + tiles - 1) / tiles);
  s = std::min(s, std::max(1, nb / 4));
  return std::max(1, std::min(s, 64));
}

int Fmm::split_m2m(int64_t tiles, size_t part_elems) {
  // the levels near the root have few parents: their M2M is a handful of
  // workgroups each walking all 8 children; split the children over up to
  // 8 parts (aiming at ~1,024 workgroups), keeping the partial tiles under
  // 256 MB
  if (tiles >= 512) return 1;
  int s = (int)std::min<int64_t>(8, (1024 + tiles - 1) / tiles);
  while (s > 1 && (size_t)s * part_elems * sizeof(double) > ((size_t)256 << 20)) s--;
  return std::max(1, s);
}

void Fmm::reserve_part(size_t elems) {
  if (elems <= part_elems_) return;
  part_.alloc(sizeof(double) * elems);
  part_elems_ = elems;
}


void Fmm::plan_m2l_pairs(LevelPlan &P, const std::vector<int32_t> *cls, const std::vector<int32_t> *gv,
                         const std::vector<const double *> *ops, int M) {
  P.m2l_npairs = 0;
  P.m2l_pair_on = false;
  P.m2lp = GemmBatch();
  M2lPairPlan H;
  plan_m2l_pairs_host(cls, gv, tree_.count, GBN1, M, kPairBufMaxElems, kPairRtFP64, -1, H);
  if (!H.use) return;
  std::vector<const double *> slot_op(343, nullptr);
  for (int a = 0; a < 8; a++) {
    if (cls[a].empty()) continue;
    const std::vector<int> sls = class_offset_slots(a);
    if (ops[a].size() != sls.size()) throw Error(NNK_ESTATE, "per-offset M2L plan: operator count");
    for (size_t b = 0; b < sls.size(); b++) slot_op[sls[b]] = ops[a][b];
  }
  const int nj = (int)H.job_slot.size();
  std::vector<const double *> jop(nj, nullptr);
  for (int j = 0; j < nj; j++)
    if (H.job_slot[j] >= 0 && !(jop[j] = slot_op[H.job_slot[j]]))
      throw Error(NNK_ESTATE, "per-offset M2L plan: missing translation operator");
  std::vector<int32_t> col(H.npairs);
  for (int64_t p = 0; p < H.npairs; p++) col[p] = (int32_t)p;
  const std::vector<int32_t> tp = {0, 1}, tl = {0};  // one tile, its one block
  GemmBatch &B = P.m2lp;
  B.owned.emplace_back(make_dbuf(jop));
  const double *const *dop = B.owned.back().as<const double *>();
  B.owned.emplace_back(make_dbuf(H.src));
  const int32_t *dsrc = B.owned.back().as<int32_t>();
  B.owned.emplace_back(make_dbuf(col));
  const int32_t *dcol = B.owned.back().as<int32_t>();
  B.owned.emplace_back(make_dbuf(tp));
  const int32_t *dtp = B.owned.back().as<int32_t>();
  B.owned.emplace_back(make_dbuf(tl));
  const int32_t *dtl = B.owned.back().as<int32_t>();
  std::vector<GemmJob> jobs(nj);
  for (int j = 0; j < nj; j++) {
    GemmJob &J = jobs[j];
    J.A = dop + j;
    J.gidx = dsrc + H.job_p0[j];  // empty jobs (N = 0) return before any read
    J.cidx = dcol + H.job_p0[j];
    J.nb = 1;
    J.N = H.job_n[j];
    J.tptr = dtp;
    J.tlist = dtl;
    B.maxN = std::max(B.maxN, J.N);
  }
  B.njobs = nj;
  B.jobs.upload(jobs);
  P.m2lp_nodes = (int)H.node.size();
  P.m2lp_node.upload(H.node);
  P.m2lp_ptr.upload(H.ptr);
  P.m2lp_idx.upload(H.idx);
  if ((size_t)H.npairs * M > pairbuf_elems_) {
    pairbuf_.alloc(sizeof(double) * (size_t)H.npairs * M);
    pairbuf_elems_ = (size_t)H.npairs * M;
  }
  P.m2l_npairs = (int)H.npairs;
  P.m2l_pair_on = true;
}

void Fmm::launch_m2l_pairs(const LevelPlan &P, const double *X, int ldx, double *C, int ldc, int M, int Mpad,
                           int Kpad, hipStream_t s) {
  if (!P.m2l_npairs) throw Error(NNK_ESTATE, "per-offset M2L: no plan for this level");
  gemm_dims(P.m2lp, X, ldx, pairbuf_.as<double>(), M, 0, s, M, M, Mpad, Kpad, 1, nullptr, 1);
  pair_sum(P, M, C, ldc, s);
}

void Fmm::pair_sum(const LevelPlan &P, int M, double *C, int ldc, hipStream_t s) {
  if ((size_t)P.m2l_npairs * M > pairbuf_elems_) throw Error(NNK_ESTATE, "per-offset M2L: pair buffer too small");
  hipLaunchKernelGGL(k_pair_sum, dim3((unsigned)P.m2lp_nodes), dim3(256), 0, s, P.m2lp_node.as<int32_t>(),
                     P.m2lp_ptr.as<int32_t>(), P.m2lp_idx.as<int32_t>(), pairbuf_.as<double>(), M, C, ldc);
  NNK_HIP(hipGetLastError());
}

void Fmm::split_reduce(const DBuf &slot, int ncols, int M, int ldc, int nsplit, double *C, hipStream_t s) {
  hipLaunchKernelGGL(k_split_reduce, dim3((unsigned)ncols), dim3(256), 0, s, slot.as<int32_t>(), M, ldc,
                     nsplit, part_.as<double>(), (size_t)ncols * ldc, C);
  NNK_HIP(hipGetLastError());
}

void Fmm::gemm_dims(const GemmBatch &b, const double *X, int ldx, double *C, int ldc, int beta,
                    hipStream_t s, int M, int K, int Mpad, int Kpad, int nsplit, const DBuf *slot, int tag) {
  if (b.njobs == 0 || b.maxN == 0) return;
  if (Mpad % GBM || Kpad % GBK) throw Error(NNK_ESTATE, "gemm_dims: bad padding");
  // tag 1: the M2L's own kernel symbol (profiles tell it from the other GEMMs)
  auto go = [&](int bt, double *P, int ns) {
    if (tag == 1) launch_gemm<GBN1, 1>(Mpad / GBM, M, K, Mpad, Kpad, X, ldx, C, ldc, bt, b, s, ns, P);
    else launch_gemm<GBN0, 0>(Mpad / GBM, M, K, Mpad, Kpad, X, ldx, C, ldc, bt, b, s, ns, P);
  };
  if (nsplit > 1) {
    if (beta != 0 || !slot) throw Error(NNK_ESTATE, "split GEMM needs beta = 0 and a slot list");
    if ((size_t)nsplit * b.njobs * b.maxN * ldc > part_elems_) throw Error(NNK_ESTATE, "split buffer too small");
    go(0, part_.as<double>(), nsplit);
    split_reduce(*slot, b.njobs * b.maxN, M, ldc, nsplit, C, s);
    return;
  }
  go(beta, nullptr, 1);
}

void Fmm::gemm(const GemmBatch &b, const double *X, int ldx, double *C, int ldc, int beta,
               hipStream_t s, int tag, int nsplit, const DBuf *slot) {
  if (b.njobs == 0 || b.maxN == 0) return;
  const int nrb = Mpad_ / GBM;
  if (nsplit > 1) {
    // partial tiles, then the fixed-order reduction (stores: beta must be 0)
    if (beta != 0 || !slot) throw Error(NNK_ESTATE, "split GEMM needs beta = 0 and a slot list");
    if ((size_t)nsplit * b.njobs * b.maxN * ldc > part_elems_) throw Error(NNK_ESTATE, "split buffer too small");
    if (tag == 1) launch_gemm<GBN1, 1>(nrb, K_, K_, Mpad_, Kpad_, X, ldx, C, ldc, 0, b, s, nsplit, part_.as<double>());
    else launch_gemm<GBN0, 0>(nrb, K_, K_, Mpad_, Kpad_, X, ldx, C, ldc, 0, b, s, nsplit, part_.as<double>());
    split_reduce(*slot, b.njobs * b.maxN, K_, ldc, nsplit, C, s);
    return;
  }
  if (tag == 1) {
    launch_gemm<GBN1, 1>(nrb, K_, K_, Mpad_, Kpad_, X, ldx, C, ldc, beta, b, s);
  } else {
    launch_gemm<GBN0, 0>(nrb, K_, K_, Mpad_, Kpad_, X, ldx, C, ldc, beta, b, s);
  }
}

Fmm::~Fmm() {
  for (auto &m : marks_) {
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
}

Fmm::Fmm(std::shared_ptr<const Tree> t, double eps, double mu, int res, int shells, int precision,
         double tau)
    : tp_(std::move(t)), tree_(*tp_), eps_(eps), mu_(mu), res_(res), shells_(shells),
      prec_(precision), tau_(tau) {
  const Tree &T = tree_;
  if (precision < 0 || precision > 4) throw Error(NNK_EARG, "unknown precision");
  if (precision >= 1 && !(tau > 0 && tau < 1)) throw Error(NNK_EARG, "truncation tolerance must be in (0,1)");
  if (res < 2 || res > 16 || shells < 1 || 2 * shells > res)
    throw Error(NNK_EARG, "coronaRes must be in [2,16] and 1 <= coronaShells <= res/2");
  if (!(eps > 0) || !(mu > 0)) throw Error(NNK_EARG, "kernel parameters eps and mu must be > 0");
  auto codes = surface_codes(res, shells);
  nq_ = (int)codes.size();
  K_ = 3 * nq_;
  Kpad_ = (K_ + GBK - 1) / GBK * GBK;
  Mpad_ = (K_ + GBM - 1) / GBM * GBM;
  npot_ = T.npot;
  ntgt_ = T.ntgt;
  maxlev_ = T.max_level;
  w0_ = T.corners[3] - T.corners[0];
  surf_idx_.upload(codes);
  corners_.upload(T.corners);
  const int nc = T.count;
  // ---- sources in leaf order ----
  std::vector<int32_t> src_ptr(nc + 2, 0);
  std::vector<double> src_pos;
  std::vector<int32_t> src_from;
  std::vector<int64_t> ctr_ptr{0};
  std::vector<int32_t> ctr_col;
  std::vector<double> ctr_w;
  if (!T.nearest) {
    for (int n = 1; n <= nc; n++) {
      src_ptr[n + 1] = src_ptr[n];
      for (int64_t k = T.pt_ptr[n]; k < T.pt_ptr[n + 1]; k++) {
        int64_t r = T.pt_rows[k];
        if (!T.is_pot[r]) continue;
        src_pos.insert(src_pos.end(), &T.points[3 * r], &T.points[3 * r + 3]);
        src_from.push_back((int32_t)T.pot_ord[r]);
        src_ptr[n + 1]++;
      }
    }
  } else {
    // fine points with a nonzero NN weight on one of the leaf's coarse
    // points, each once per leaf, ascending fine index (upwardPass.m:58-69)
    struct Tr { int64_t row; int32_t col; double w; };
    std::vector<Tr> tr;
    for (int n = 1; n <= nc; n++) {
      src_ptr[n + 1] = src_ptr[n];
      tr.clear();
      for (int64_t k = T.pt_ptr[n]; k < T.pt_ptr[n + 1]; k++) {
        int64_t r = T.pt_rows[k];
        if (!T.is_pot[r]) continue;
        int64_t c = T.pot_ord[r];
        for (int64_t e = T.nn_colptr[c]; e < T.nn_colptr[c + 1]; e++)
          if (T.nn_val[e] != 0.0) tr.push_back({T.nn_row[e], (int32_t)c, T.nn_val[e]});
      }
      std::sort(tr.begin(), tr.end(), [](const Tr &x, const Tr &y) {
        return x.row != y.row ? x.row < y.row : x.col < y.col;
      });
      for (size_t e = 0; e < tr.size(); e++) {
        if (e == 0 || tr[e].row != tr[e - 1].row) {
          if (e) ctr_ptr.push_back((int64_t)ctr_col.size());
          src_pos.insert(src_pos.end(), &T.fine[3 * tr[e].row], &T.fine[3 * tr[e].row + 3]);
          src_ptr[n + 1]++;
        }
        ctr_col.push_back(tr[e].col);
        ctr_w.push_back(tr[e].w);
      }
      if (!tr.empty()) ctr_ptr.push_back((int64_t)ctr_col.size());
    }
  }
  nsrc_ = src_pos.size() / 3;
  src_ptr_.upload(src_ptr);
  src_ptr_h_ = src_ptr;
  src_pos_.upload(src_pos);
  src_q_.alloc(sizeof(double) * 3 * std::max<int64_t>(nsrc_, 1));
  if (!T.nearest) src_from_.upload(src_from);
  else {
    ctr_ptr_.upload(ctr_ptr);
    ctr_col_.upload(ctr_col);
    ctr_w_.upload(ctr_w);
  }
  // ---- targets in leaf order ----
  std::vector<int32_t> tgt_ptr(nc + 2, 0), tgt_out;
  std::vector<double> tgt_pos;
  for (int n = 1; n <= nc; n++) {
    tgt_ptr[n + 1] = tgt_ptr[n];
    for (int64_t k = T.pt_ptr[n]; k < T.pt_ptr[n + 1]; k++) {
      int64_t r = T.pt_rows[k];
      if (!T.is_tgt[r]) continue;
      tgt_pos.insert(tgt_pos.end(), &T.points[3 * r], &T.points[3 * r + 3]);
      tgt_out.push_back((int32_t)T.tgt_ord[r]);
      tgt_ptr[n + 1]++;
    }
  }
  tgt_ptr_.upload(tgt_ptr);
  tgt_pos_.upload(tgt_pos);
  tgt_out_.upload(tgt_out);
  uptr_.upload(T.lptr[0]);
  uidx_.upload(T.lidx[0]);
  wptr_.upload(T.lptr[2]);
  widx_.upload(T.lidx[2]);
  xptr_.upload(T.lptr[3]);
  xidx_.upload(T.lidx[3]);
  std::vector<int32_t> pl, tl;
  for (int n = 1; n <= nc; n++) {
    if (T.pot_leaf[n]) pl.push_back(n);
    if (T.tgt_leaf[n]) tl.push_back(n);
  }
  // One workgroup runs each leaf of these lists.  They are launched in
  // decreasing order of their pair work (S2M: sources; leaf kernel:
  // targets x (U-list sources + surface points of the leaf and its W list)),
  // so that the heaviest start first instead of finishing last on a few
  // CUs.  The results do not depend on the order.
  {
    auto nsrc = [&](int n) -> int64_t { return src_ptr[n + 1] - src_ptr[n]; };
    std::vector<int64_t> wp(nc + 1, 0), wt(nc + 1, 0);
    for (int n : pl) wp[n] = nsrc(n);
    for (int n : tl) {
      int64_t u = 0;
      for (int64_t e = T.lptr[0][n]; e < T.lptr[0][n + 1]; e++) u += nsrc(T.lidx[0][e]);
      wt[n] = (int64_t)(tgt_ptr[n + 1] - tgt_ptr[n]) * (u + (int64_t)nq_ * (1 + T.lptr[2][n + 1] - T.lptr[2][n]));
    }
    order_by_work(pl, wp);
    order_by_work(tl, wt);
  }
  n_pleaves_ = (int)pl.size();
  n_tleaves_ = (int)tl.size();
  max_leaf_tgt_ = 0;
  for (int n : tl) max_leaf_tgt_ = std::max(max_leaf_tgt_, tgt_ptr[n + 1] - tgt_ptr[n]);
  {
    int dev = 0;
    NNK_HIP(hipGetDevice(&dev));
    NNK_HIP(hipDeviceGetAttribute(&ncu_, hipDeviceAttributeMultiprocessorCount, dev));
  }
  pleaves_.upload(pl);
  tleaves_.upload(tl);
  chk_.alloc(sizeof(double) * (size_t)nc * K_);
  uppot_.alloc(sizeof(double) * (size_t)nc * K_);
  downpot_.alloc(sizeof(double) * (size_t)nc * K_);
  chk_.zero();
  uppot_.zero();
  downpot_.zero();
  ops_.assign(maxlev_ + 1, nullptr);
  {
    OpCache &C = OpCache::get();
    std::vector<int> miss;
    for (int l = 2; l <= maxlev_; l++) {
      // FP64 form: operators with T.  Mixed forms: any entry (T unused).
      auto hit = C.find(key(0, l));
      if (!hit && prec_ >= 1) hit = C.find(key(1, l));
      if (hit) ops_[l] = std::static_pointer_cast<const LevelOps>(hit);
      else miss.push_back(l);
    }
    // the check-to-equivalent SVDs of all missing levels in one batch
    const int n = K_, nb = (int)miss.size();
    std::vector<std::vector<double>> A(nb), U(nb, std::vector<double>((size_t)n * n)),
        S(nb, std::vector<double>(n)), V(nb, std::vector<double>((size_t)n * n));
    std::vector<const double *> pa(nb);
    std::vector<double *> pu(nb), ps(nb), pv(nb);
    for (int b = 0; b < nb; b++) {
      A[b] = up_to_down(miss[b]);
      pa[b] = A[b].data();
      pu[b] = U[b].data();
      ps[b] = S[b].data();
      pv[b] = V[b].data();
    }
    if (nb) svd_batch(pa.data(), n, nb, pu.data(), ps.data(), pv.data());
    for (int b = 0; b < nb; b++) {
      const int l = miss[b];
      ops_[l] = build_ops(l, prec_ == 0, std::move(U[b]), std::move(S[b]), std::move(V[b]));
      C.put(key(prec_ == 0 ? 0 : 1, l), ops_[l], ops_[l]->bytes());
    }
  }
  plan();
  if (prec_ >= 1) {
    build_mixed();
    plan_mixed();
  }
  NNK_HIP(hipDeviceSynchronize());
}

size_t LevelOps::bytes() const {
  return Wup.bytes + Vup.bytes + Wdn.bytes + Udn.bytes + M2M.bytes + L2L.bytes + T.bytes +
         sizeof(double) * (svdU.size() + svdS.size() + svdV.size());
}

size_t MixedOps::bytes() const {
  return Pup.bytes + Vr.bytes + Pdn.bytes + Ur.bytes + VrT.bytes + UrT.bytes + Mc.bytes + Lc.bytes + Cf.bytes +
         Uf.bytes + UrF.bytes + UrU.bytes + Cd.bytes;
}

bool OpKey::operator<(const OpKey &o) const {
  return std::tie(dev, kind, res, shells, level, rank, eps, mu, w, tau) <
         std::tie(o.dev, o.kind, o.res, o.shells, o.level, o.rank, o.eps, o.mu, o.w, o.tau);
}

static std::mutex g_cache_mu;

// never destroyed: kept operators must not be freed by a static destructor
// after the HIP runtime has shut down (the process exit reclaims them)
OpCache &OpCache::get() {
  static OpCache *c = new OpCache;
  return *c;
}

std::shared_ptr<const void> OpCache::find(const OpKey &k) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  auto it = map_.find(k);
  std::shared_ptr<const void> v;
  if (it != map_.end()) v = it->second.live.lock();
  if (v) {
    hits_++;
    it->second.tick = ++tick_;
  } else {
    misses_++;
    if (it != map_.end()) map_.erase(it);
  }
  return v;
}

void OpCache::put(const OpKey &k, std::shared_ptr<const void> v, size_t bytes) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  Ent &e = map_[k];
  e.live = v;
  e.kept = limit_ > 0 ? v : nullptr;
  e.bytes = bytes;
  e.tick = ++tick_;
  trim();
}

// drop kept references, least recently used first, until the entries
// without a live owner outside the cache fit in the limit; forget entries
// that nothing owns any more
void OpCache::trim() {
  for (;;) {
    int64_t held = 0;
    auto lru = map_.end();
    for (auto it = map_.begin(); it != map_.end();) {
      if (it->second.live.expired()) { it = map_.erase(it); continue; }
      if (it->second.kept && it->second.kept.use_count() == 1) {
        held += (int64_t)it->second.bytes;
        if (lru == map_.end() || it->second.tick < lru->second.tick) lru = it;
      }
      ++it;
    }
    if (held <= limit_ || lru == map_.end()) return;
    lru->second.kept.reset();
  }
}

void OpCache::set_limit(int64_t bytes) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  limit_ = std::max<int64_t>(0, bytes);
  if (limit_ == 0)
    for (auto &kv : map_) kv.second.kept.reset();
  else
    for (auto &kv : map_) if (!kv.second.kept) kv.second.kept = kv.second.live.lock();
  trim();
}

void OpCache::stats(int64_t *out) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  trim();
  int64_t held = 0, n = 0;
  for (auto &kv : map_)
    if (kv.second.kept) { held += (int64_t)kv.second.bytes; n++; }
  out[0] = hits_;
  out[1] = misses_;
  out[2] = held;
  out[3] = n;
}

void OpCache::clear() {
  std::lock_guard<std::mutex> g(g_cache_mu);
  map_.clear();
  hits_ = misses_ = 0;
}

OpKey Fmm::key(int kind, int l) const {
  int dev = 0;
  NNK_HIP(hipGetDevice(&dev));
  OpKey k{};
  k.dev = dev;
  k.kind = kind;
  k.res = res_;
  k.shells = shells_;
  k.level = l;
  k.rank = kind >= 2 ? R_ : 0;
  k.eps = eps_;
  k.mu = mu_;
  k.w = std::ldexp(w0_, -l);
  k.tau = kind >= 2 ? tau_ : 0.0;
  return k;
}

// k_form_At of one source/target surface pair into dst (At layout)
static void form_op(const std::vector<double> &src, const std::vector<double> &tgt, int nq, double e2, double c,
                    double *dst, int Mpad) {
  DBuf ds, dt;
  ds.upload(src);
  dt.upload(tgt);
  int tot = nq * nq;
  hipLaunchKernelGGL(k_form_At, dim3((tot + 255) / 256), dim3(256), 0, 0, ds.as<double>(), nq, dt.as<double>(), nq,
                     e2, c, dst, Mpad);
  NNK_HIP(hipGetLastError());
  NNK_HIP(hipDeviceSynchronize());
}

// K(up surface -> down surface) of the canonical box of level l (3Nq x
// 3Nq row-major; rows: down-surface points, the system Up.m:31 solves)
std::vector<double> Fmm::up_to_down(int l) const {
  const auto codes = surface_codes(res_, shells_);
  const double w = std::ldexp(w0_, -l);
  const double e2 = eps_ * eps_, c = 1.0 / (8.0 * kPi * mu_);
  const double zero[3] = {0, 0, 0};
  auto up = surface_pts(codes, res_, zero, w, false);
  auto dn = surface_pts(codes, res_, zero, w, true);
  const int n = K_;
  std::vector<double> A((size_t)n * n);
  for (int p = 0; p < nq_; p++)      // row point (down surface, target)
    for (int q = 0; q < nq_; q++) {  // column point (up surface, source)
      double r[3] = {dn[3 * p] - up[3 * q], dn[3 * p + 1] - up[3 * q + 1], dn[3 * p + 2] - up[3 * q + 2]};
      double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double invs = (1.0 / std::pow(r2 + e2, 1.5)) * c;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          A[(size_t)(3 * p + a) * n + 3 * q + b] = (r[a] * r[b] + (a == b ? r2 + 2 * e2 : 0.0)) * invs;
    }
  return A;
}

std::shared_ptr<const LevelOps> Fmm::build_ops(int l, bool with_T, std::vector<double> U, std::vector<double> S,
                                               std::vector<double> V) {
  auto op = std::make_shared<LevelOps>();
  LevelOps &O = *op;
  const auto codes = surface_codes(res_, shells_);
  const double w = std::ldexp(w0_, -l);
  const double e2 = eps_ * eps_, c = 1.0 / (8.0 * kPi * mu_);
  const size_t mat = (size_t)Kpad_ * Mpad_;
  // ---- check-to-equivalent factors from the SVD of K(up -> down) ----
  {
    const int n = K_;
    // up:   x = V (S^-1 U^T b);  down (K(down->up) = A^T): x = U (S^-1 V^T b)
    std::vector<double> Wup(mat, 0.0), Vup(mat, 0.0), Wdn(mat, 0.0), Udn(mat, 0.0);
    for (int k = 0; k < n; k++)
      for (int m = 0; m < n; m++) {
        Wup[(size_t)k * Mpad_ + m] = U[(size_t)k * n + m] / S[m];  // W[m][k] = U[k][m]/s_m
        Vup[(size_t)k * Mpad_ + m] = V[(size_t)m * n + k];         // V[m][k]
        Wdn[(size_t)k * Mpad_ + m] = V[(size_t)k * n + m] / S[m];
        Udn[(size_t)k * Mpad_ + m] = U[(size_t)m * n + k];
      }
    O.Wup.upload(Wup);
    O.Vup.upload(Vup);
    O.Wdn.upload(Wdn);
    O.Udn.upload(Udn);
    O.svdU = std::move(U);
    O.svdS = std::move(S);
    O.svdV = std::move(V);
  }
  // ---- kernel-matrix operators formed on the device ----
  O.M2M.alloc(sizeof(double) * mat * 8);
  O.L2L.alloc(sizeof(double) * mat * 8);
  O.M2M.zero();
  O.L2L.zero();
  NNK_HIP(hipDeviceSynchronize());
  {
    const double plo[3] = {0, 0, 0};
    auto pdn = surface_pts(codes, res_, plo, 2 * w, true);  // parent (width 2w) down surface
    for (int k = 0; k < 8; k++) {
      double lo[3] = {((k >> 2) & 1) * w, ((k >> 1) & 1) * w, (k & 1) * w};
      auto cup = surface_pts(codes, res_, lo, w, false);
      form_op(cup, pdn, nq_, e2, c, O.M2M.as<double>() + k * mat, Mpad_);  // child up -> parent down
      form_op(pdn, cup, nq_, e2, c, O.L2L.as<double>() + k * mat, Mpad_);  // parent down -> child up
    }
  }
  if (with_T) form_T(l, O.T, O.tslot);
  else O.tslot.assign(343, nullptr);
  return op;
}

void Fmm::form_T(int l, DBuf &T, std::vector<double *> &tslot) {
  const auto codes = surface_codes(res_, shells_);
  const double w = std::ldexp(w0_, -l);
  const double e2 = eps_ * eps_, c = 1.0 / (8.0 * kPi * mu_);
  const size_t mat = (size_t)Kpad_ * Mpad_;
  const double zero[3] = {0, 0, 0};
  T.alloc(sizeof(double) * mat * 316);
  T.zero();
  NNK_HIP(hipDeviceSynchronize());
  tslot.assign(343, nullptr);
  auto tup = surface_pts(codes, res_, zero, w, false);
  int k = 0;
  for (int ox = -3; ox <= 3; ox++)
    for (int oy = -3; oy <= 3; oy++)
      for (int oz = -3; oz <= 3; oz++) {
        if (std::max({std::abs(ox), std::abs(oy), std::abs(oz)}) < 2) continue;
        double lo[3] = {ox * w, oy * w, oz * w};
        auto sup = surface_pts(codes, res_, lo, w, false);
        double *dst = T.as<double>() + (size_t)k * mat;
        form_op(sup, tup, nq_, e2, c, dst, Mpad_);
        tslot[offset_slot(ox, oy, oz)] = dst;
        k++;
      }
}

void Fmm::plan() {
  const Tree &T = tree_;
  const int nc = T.count;
  // which nodes carry a nonzero up density / need a down density
  std::vector<uint8_t> pot_below(nc + 1, 0), tgt_below(nc + 1, 0);
  {
    std::vector<int> order(nc);
    for (int i = 0; i < nc; i++) order[i] = i + 1;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return T.level[a - 1] > T.level[b - 1]; });
    for (int n : order) {
      if (T.is_leaf(n)) {
        pot_below[n] = T.pot_leaf[n];
        tgt_below[n] = T.tgt_leaf[n];
      } else {
        for (int k = 0; k < 8; k++) {
          int ch = T.children[8 * (n - 1) + k];
          pot_below[n] |= pot_below[ch];
          tgt_below[n] |= tgt_below[ch];
        }
      }
    }
  }
  plans_.clear();
  plans_.resize(maxlev_ + 1);
  hl_.assign(maxlev_ + 1, HostLevel());
  {
    std::vector<uint8_t> inw(nc + 1, 0);
    for (int32_t w : T.lidx[2]) inw[w] = 1;
    for (int n = 1; n <= nc; n++) {
      const int l = T.level[n - 1];
      if (l < 2) continue;
      if (inw[n]) hl_[l].wnodes.push_back(n - 1);
      if (T.tgt_leaf[n]) hl_[l].tleaves.push_back(n - 1);
    }
  }
  size_t maxn = 1;
  std::vector<std::vector<int>> bylev(maxlev_ + 1);
  for (int n = 1; n <= nc; n++) bylev[T.level[n - 1]].push_back(n);
  auto add_job = [&](GemmBatch &B, std::vector<GemmJob> &jobs, const std::vector<const double *> &A,
                     const std::vector<int32_t> *gidx, const std::vector<int32_t> *cidx, int N) {
    GemmJob j{};
    B.owned.emplace_back(make_dbuf(A));
    j.A = B.owned.back().as<const double *>();
    if (gidx) { B.owned.emplace_back(make_dbuf(*gidx)); j.gidx = B.owned.back().as<int32_t>(); }
    if (cidx) { B.owned.emplace_back(make_dbuf(*cidx)); j.cidx = B.owned.back().as<int32_t>(); }
    j.nb = (int)A.size();
    if (j.nb > 512) throw Error(NNK_ESTATE, "too many operator blocks in one GEMM job");
    j.N = N;
    jobs.push_back(j);
    B.maxN = std::max(B.maxN, N);
  };
  auto finish = [&](GemmBatch &B, std::vector<GemmJob> &jobs) {
    B.njobs = (int)jobs.size();
    B.jobs.upload(jobs);
  };
  for (int l = 2; l <= maxlev_; l++) {
    LevelPlan &P = plans_[l];
    const LevelOps &O = *ops_[l];
    P.level = l;
    // ---- upward ----
    std::vector<int32_t> up, nl;
    for (int n : bylev[l]) {
      if (T.is_leaf(n)) {
        if (T.pot_leaf[n]) up.push_back(n - 1);
      } else if (pot_below[n]) {
        nl.push_back(n - 1);
      }
    }
    if (!nl.empty()) {
      // M2M: the children (level l+1) carry class-k operators of level l+1
      std::vector<const double *> A(8);
      const size_t mat = (size_t)Kpad_ * Mpad_;
      for (int k = 0; k < 8; k++) A[k] = ops_[l + 1]->M2M.as<double>() + k * mat;
      std::vector<int32_t> g(8 * nl.size());
      for (size_t j = 0; j < nl.size(); j++)
        for (int k = 0; k < 8; k++) {
          int ch = T.children[8 * nl[j] + k];
          g[k * nl.size() + j] = pot_below[ch] ? ch - 1 : -1;
        }
      std::vector<GemmJob> jobs;
      add_job(P.m2m, jobs, A, &g, &nl, (int)nl.size());
      finish(P.m2m, jobs);
      hl_[l].m2m_g = g;
      const int64_t tiles = (int64_t)(Mpad_ / GBM) * (((int64_t)nl.size() + GBN0 - 1) / GBN0);
      P.m2m_split = split_m2m(tiles, nl.size() * (size_t)K_);
      if (P.m2m_split > 1) {
        P.m2m_slot.upload(nl);
        reserve_part((size_t)P.m2m_split * nl.size() * K_);
      }
    }
    hl_[l].up_leaves = up;
    hl_[l].up_nonleaves = nl;
    std::vector<int32_t> upall = up;
    upall.insert(upall.end(), nl.begin(), nl.end());
    P.n_up = (int)upall.size();
    maxn = std::max(maxn, upall.size());
    if (P.n_up) {
      std::vector<GemmJob> j1, j2;
      add_job(P.up_solve1, j1, {O.Wup.as<double>()}, &upall, nullptr, P.n_up);
      finish(P.up_solve1, j1);
      add_job(P.up_solve2, j2, {O.Vup.as<double>()}, nullptr, &upall, P.n_up);
      finish(P.up_solve2, j2);
    }
    P.up_nodes.upload(upall);
    // ---- downward ----
    std::vector<int32_t> dn;
    std::vector<int32_t> cls[8];
    std::vector<int> clsnode[8];
    for (int n : bylev[l]) {
      if (!tgt_below[n]) continue;
      dn.push_back(n - 1);
      int par = T.parent[n - 1];
      int a = n - T.children[8 * (par - 1)];
      cls[a].push_back(n - 1);
      clsnode[a].push_back(n);
    }
    P.n_down = (int)dn.size();
    maxn = std::max(maxn, dn.size());
    if (P.n_down) {
      std::vector<GemmJob> jobs, ljobs;
      std::vector<const double *> cops[8];  // per class: its operators in block order
      const size_t mat = (size_t)Kpad_ * Mpad_;
      for (int a = 0; a < 8; a++) {
        if (cls[a].empty()) continue;
        auto offs = class_offsets(a);
        std::map<int, int> pos;
        std::vector<const double *> &A = cops[a];
        for (size_t b = 0; b < offs.size(); b++) {
          pos[offset_slot(offs[b][0], offs[b][1], offs[b][2])] = (int)b;
          A.push_back(O.tslot[offset_slot(offs[b][0], offs[b][1], offs[b][2])]);
        }
        const int nb = (int)A.size(), N = (int)cls[a].size();
        std::vector<int32_t> g((size_t)nb * N, -1);
        for (int j = 0; j < N; j++) {
          int n = clsnode[a][j];
          const double *bn = T.box(n);
          for (int64_t e = T.lptr[1][n]; e < T.lptr[1][n + 1]; e++) {
            int v = T.lidx[1][e];
            if (!pot_below[v]) continue;
            const double *bv = T.box(v);
            int o[3];
            for (int d = 0; d < 3; d++) o[d] = (int)std::lround((bv[d] - bn[d]) / std::ldexp(w0_, -l));
            auto it = pos.find(offset_slot(o[0], o[1], o[2]));
            if (it == pos.end() || std::max({std::abs(o[0]), std::abs(o[1]), std::abs(o[2])}) > 3)
              throw Error(NNK_ESTATE, "V-list node outside the class offset set");
            g[(size_t)it->second * N + j] = v - 1;
            stats_[0]++;
          }
        }
        add_job(P.m2l, jobs, A, &g, &cls[a], N);
        {  // active blocks per 64-column tile of the M2L launch
          std::vector<int32_t> tp, tl;
          tile_lists(g, nb, N, GBN1, tp, tl);
          check_gather_job(g, nb, N, GBN1, tp, tl, T.count, "M2L job");
          P.m2l.owned.emplace_back(make_dbuf(tp));
          jobs.back().tptr = P.m2l.owned.back().as<int32_t>();
          P.m2l.owned.emplace_back(make_dbuf(tl));
          jobs.back().tlist = P.m2l.owned.back().as<int32_t>();
        }
        std::vector<int32_t> gp(N);
        for (int j = 0; j < N; j++) gp[j] = T.parent[clsnode[a][j] - 1] - 1;
        if (l >= 3)  // parent's down density onto the child's up surface
          add_job(P.l2l, ljobs, {O.L2L.as<double>() + a * mat}, &gp, &cls[a], N);
        hl_[l].cls[a] = cls[a];
        hl_[l].gv[a] = std::move(g);
        hl_[l].gp[a] = std::move(gp);
      }
      finish(P.m2l, jobs);
      finish(P.l2l, ljobs);
      {  // split-K for levels with few output tiles
        const int nj = P.m2l.njobs, mN = P.m2l.maxN;
        std::vector<int32_t> slot((size_t)nj * mN, -1);
        int j = 0, minnb = 512;
        for (int a = 0; a < 8; a++) {
          if (cls[a].empty()) continue;
          for (size_t c = 0; c < cls[a].size(); c++) slot[(size_t)j * mN + c] = cls[a][c];
          minnb = std::min(minnb, (int)class_offsets(a).size());
          j++;
        }
        P.m2l_split = split_for((int64_t)nj * (Mpad_ / GBM) * ((mN + GBN1 - 1) / GBN1), minnb);
        P.m2l_slot.upload(slot);  // also for the autotuner's split candidates
        reserve_part((size_t)std::min(64, 2 * P.m2l_split) * nj * mN * K_);
      }
      if (prec_ == 0) plan_m2l_pairs(P, hl_[l].cls, hl_[l].gv, cops, K_);
      std::vector<GemmJob> j1, j2;
      add_job(P.dn_solve1, j1, {O.Wdn.as<double>()}, &dn, nullptr, P.n_down);
      finish(P.dn_solve1, j1);
      add_job(P.dn_solve2, j2, {O.Udn.as<double>()}, nullptr, &dn, P.n_down);
      finish(P.dn_solve2, j2);
    }
    P.down_nodes.upload(dn);
    std::vector<int32_t> xs;
    for (int n : bylev[l])
      if (tgt_below[n] && T.lptr[3][n + 1] > T.lptr[3][n]) xs.push_back(n);
    P.n_x = (int)xs.size();
    P.x_nodes.upload(xs);
    hl_[l].dn = dn;
    hl_[l].xnodes = xs;
  }
  {
    // node ids (from 1) of every level's X-list owners, heaviest P2L first
    // (one workgroup each; see the leaf lists)
    std::vector<int32_t> xall;
    for (int l = 2; l <= maxlev_; l++) xall.insert(xall.end(), hl_[l].xnodes.begin(), hl_[l].xnodes.end());
    std::vector<int64_t> wx(T.count + 1, 0);
    for (int n : xall)
      for (int64_t e = T.lptr[3][n]; e < T.lptr[3][n + 1]; e++)
        wx[n] += src_ptr_h_[T.lidx[3][e] + 1] - src_ptr_h_[T.lidx[3][e]];
    order_by_work(xall, wx);
    n_x_all_ = (int)xall.size();
    x_all_.upload(xall);
    if (n_x_all_ && prec_ == 0) xchk_.alloc(sizeof(double) * (size_t)nc * K_);
    else xchk_ = DBuf();
  }
  tmp_.alloc(sizeof(double) * maxn * K_);
  // remaining work counts
  std::vector<int64_t> srcn(nc + 2, 0), tgtn(nc + 2, 0);
  {
    std::vector<int32_t> sp(nc + 2), tp(nc + 2);
    NNK_HIP(hipMemcpy(sp.data(), src_ptr_.p, sizeof(int32_t) * (nc + 2), hipMemcpyDeviceToHost));
    NNK_HIP(hipMemcpy(tp.data(), tgt_ptr_.p, sizeof(int32_t) * (nc + 2), hipMemcpyDeviceToHost));
    for (int n = 1; n <= nc; n++) { srcn[n] = sp[n + 1] - sp[n]; tgtn[n] = tp[n + 1] - tp[n]; }
  }
  for (int n = 1; n <= nc; n++) {
    stats_[10] += srcn[n];
    stats_[11] += tgtn[n];
    if (T.pot_leaf[n] && T.level[n - 1] >= 2) stats_[2] += srcn[n] * nq_;
    if (T.tgt_leaf[n]) {
      stats_[3] += tgtn[n] * nq_;
      for (int64_t e = T.lptr[0][n]; e < T.lptr[0][n + 1]; e++) stats_[1] += tgtn[n] * srcn[T.lidx[0][e]];
      stats_[4] += tgtn[n] * nq_ * (T.lptr[2][n + 1] - T.lptr[2][n]);
    }
    if (tgt_below[n] && T.level[n - 1] >= 2)
      for (int64_t e = T.lptr[3][n]; e < T.lptr[3][n + 1]; e++) stats_[5] += nq_ * srcn[T.lidx[3][e]];
  }
  for (int l = 2; l <= maxlev_; l++) {
    stats_[6] += plans_[l].n_up;
    stats_[7] += plans_[l].n_down;
    for (int n : bylev[l]) {
      if (!T.is_leaf(n) && pot_below[n]) stats_[8]++;
      if (l >= 3 && tgt_below[n]) stats_[9]++;
    }
  }
}

void Fmm::stats(int64_t *out) const {
  for (int i = 0; i < kStats; i++) out[i] = stats_[i];
}

void Fmm::set_profiling(bool on) { prof_ = on; }

void Fmm::mark(int phase, bool begin, hipStream_t s) {
  if (!prof_) return;
  hipEvent_t e;
  NNK_HIP(hipEventCreate(&e));
  NNK_HIP(hipEventRecord(e, s));
  if (begin) open_ = e;
  else marks_.push_back({phase, open_, e});
}

void Fmm::phase_times(double *ms, int64_t *count) {
  for (int i = 0; i < kPhases; i++) { ms[i] = 0; count[i] = 0; }
  for (auto &m : marks_) {
    NNK_HIP(hipEventSynchronize(m.b));
    float t = 0;
    NNK_HIP(hipEventElapsedTime(&t, m.a, m.b));
    ms[m.phase] += t;
    count[m.phase]++;
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  marks_.clear();
}

void Fmm::strengths(const double *d_f, hipStream_t s) {
  if (nsrc_ == 0) return;
  hipLaunchKernelGGL(k_strength, dim3((nsrc_ + 255) / 256), dim3(256), 0, s, nsrc_,
                     tree_.nearest ? nullptr : src_from_.as<int32_t>(), ctr_ptr_.as<int64_t>(),
                     ctr_col_.as<int32_t>(), ctr_w_.as<double>(), d_f, src_q_.as<double>());
  NNK_HIP(hipGetLastError());
}

// C[node] += X[node] for the listed node ids (from 1), rows of K
__global__ __launch_bounds__(256) void k_add_rows(const int32_t *__restrict__ nodes, const double *__restrict__ X,
                                                  double *__restrict__ C, int K) {
  const size_t o = (size_t)(nodes[blockIdx.x] - 1) * K;
  for (int i = threadIdx.x; i < K; i += 256) C[o + i] += X[o + i];
}

void Fmm::launch_surf(const int32_t *nodes, int n, const int64_t *lptr, const int32_t *lidx, int down,
                      int accumulate, double *chk, hipStream_t s) {
  if (n <= 0) return;
  SurfArgs a{};
  a.nodes = nodes;
  a.lptr = lptr;
  a.lidx = lidx;
  a.down = down;
  a.accumulate = accumulate;
  a.corners = corners_.as<double>();
  a.surf_idx = surf_idx_.as<int32_t>();
  a.nq = nq_; a.n1 = res_ - 1; a.K = K_;
  a.src_ptr = src_ptr_.as<int32_t>();
  a.src_pos = src_pos_.as<double>();
  a.src_q = src_q_.as<double>();
  a.e2 = eps_ * eps_;
  a.c = 1.0 / (8.0 * kPi * mu_);
  a.chk = chk;
  if (fp32_math()) hipLaunchKernelGGL(k_surf_eval<true>, dim3(n), dim3(kSurfThreads), 0, s, a);
  else hipLaunchKernelGGL(k_surf_eval<false>, dim3(n), dim3(kSurfThreads), 0, s, a);
  NNK_HIP(hipGetLastError());
}

void Fmm::launch_leaf(int far, double *d_u, hipStream_t s) {
  if (!n_tleaves_) return;
  LeafArgs a{};
  a.leaves = tleaves_.as<int32_t>();
  a.corners = corners_.as<double>();
  a.surf_idx = surf_idx_.as<int32_t>();
  a.nq = nq_; a.n1 = res_ - 1; a.K = K_;
  a.tgt_ptr = tgt_ptr_.as<int32_t>();
  a.tgt_pos = tgt_pos_.as<double>();
  a.tgt_out = tgt_out_.as<int32_t>();
  a.src_ptr = src_ptr_.as<int32_t>();
  a.src_pos = src_pos_.as<double>();
  a.src_q = src_q_.as<double>();
  a.uptr = uptr_.as<int64_t>(); a.uidx = uidx_.as<int32_t>();
  a.wptr = wptr_.as<int64_t>(); a.widx = widx_.as<int32_t>();
  a.uppot = uppot_.as<double>();
  a.downpot = downpot_.as<double>();
  a.e2 = eps_ * eps_;
  a.c = 1.0 / (8.0 * kPi * mu_);
  a.far = far;
  a.out = d_u;
  const int T = far ? leaf_threads_ : 256;
  // Few target leaves (small trees, large leaves) would leave most CUs idle
  // with one workgroup per leaf: their targets are then split into chunks
  // of tc over a second grid dimension (the lanes of a target split its
  // sources T / tc ways), until there are about two workgroups per CU
  a.tc = T;
  while (a.tc > 32 && (int64_t)n_tleaves_ * ((max_leaf_tgt_ + a.tc - 1) / a.tc) < 2 * (int64_t)ncu_) a.tc /= 2;
  const dim3 grid((unsigned)n_tleaves_, (unsigned)std::max(1, (max_leaf_tgt_ + a.tc - 1) / a.tc));
  auto go = [&](auto kern, int threads) { hipLaunchKernelGGL(kern, grid, dim3(threads), 0, s, a); };
  if (fp32_math()) {
    if (T == 128) go(k_leaf<true, 128>, 128);
    else if (T == 512) go(k_leaf<true, 512>, 512);
    else go(k_leaf<true, 256>, 256);
  } else {
    if (T == 128) go(k_leaf<false, 128>, 128);
    else if (T == 512) go(k_leaf<false, 512>, 512);
    else go(k_leaf<false, 256>, 256);
  }
  NNK_HIP(hipGetLastError());
}

void Fmm::apply(const double *d_f, double *d_u, hipStream_t s) {
  if (tune_ && !tuned_) {
    apply_impl(d_f, d_u, s);  // fills the densities the candidates run on
    autotune(s);
    autotune_leaf(s);
    tuned_ = true;
  }
  apply_impl(d_f, d_u, s);
}

void Fmm::launch_config(int level, int *split, int *depth) const {
  *split = 0;
  *depth = 0;
  if (level < 2 || level > maxlev_) return;
  const LevelPlan &P = plans_[level];
  if (m16()) {
    if (P.mx_m2l_jobs) {
      *split = P.m2l_pair_on ? 1 : P.mx_split;
      *depth = P.m2l_pair_on ? 10 + P.mx_pair_depth : P.mx_depth;
    }
  } else if (P.n_down) {
    *split = P.m2l_pair_on ? 1 : P.m2l_split;
    *depth = P.m2l_pair_on ? 1 : 0;
  }
}

void Fmm::set_launch_config(int level, int split, int depth) {
  if (level < 2 || level > maxlev_) throw Error(NNK_EARG, "launch config: level without M2L");
  LevelPlan &P = plans_[level];
  const bool mixed = m16();
  if (mixed ? !P.mx_m2l_jobs : !P.n_down) throw Error(NNK_EARG, "launch config: level without M2L");
  const size_t per_split = mixed ? (size_t)P.mx_m2l_jobs * P.mx_m2l_maxN * RP_
                                 : (size_t)P.m2l.njobs * P.m2l.maxN * m2l_ld();
  if (split < 1 || split > 64 || (split > 1 && ((size_t)split * per_split > part_elems_ ||
                                                !(mixed ? P.mx_slot.p : P.m2l_slot.p))))
    throw Error(NNK_EARG, "launch config: split-K factor out of range");
  if (!mixed) {
    if (depth == 1) {
      if (!P.m2l_npairs) throw Error(NNK_EARG, "launch config: no per-offset M2L plan at this level");
      if (split != 1) throw Error(NNK_EARG, "launch config: the per-offset M2L takes split 1");
      P.m2l_pair_on = true;
    } else if (depth == 0) {
      P.m2l_split = split;
      P.m2l_pair_on = false;
    } else {
      throw Error(NNK_EARG, "launch config: the FP64 forms take depth 0 or 1");
    }
  } else if ((depth >= 12 && depth <= 14) || depth == 17) {  // per-offset pair jobs, kernel depth 2-4, 7
    if (!P.m2l_npairs) throw Error(NNK_EARG, "launch config: no per-offset M2L plan at this level");
    if (split != 1) throw Error(NNK_EARG, "launch config: the per-offset M2L takes split 1");
    if ((depth == 13 || depth == 17) && np_ == 3 && !f16_)
      throw Error(NNK_EARG, "launch config: depth not available for this form");
    P.mx_pair_depth = depth - 10;
    P.m2l_pair_on = true;
  } else {
    const bool ok = depth == 2 || (depth == 3 && !(np_ == 3 && !f16_)) ||
                    (depth >= 4 && depth <= 8 && dma_ok(P, split, depth) && !(depth >= 7 && np_ == 3 && !f16_));
    if (!ok) throw Error(NNK_EARG, "launch config: depth not available for this form");
    P.mx_split = split;
    P.mx_depth = depth;
    P.m2l_pair_on = false;
  }
  gop_.reset();
}

void Fmm::autotune(hipStream_t s) {
  hipEvent_t a, b;
  NNK_HIP(hipEventCreate(&a));
  NNK_HIP(hipEventCreate(&b));
  const bool prof = prof_;
  prof_ = false;
  try {
    for (int l = 2; l <= maxlev_; l++) {
      LevelPlan &P = plans_[l];
      const bool mixed = m16();
      if (mixed ? !P.mx_m2l_jobs : !P.n_down) continue;
      const int s0 = mixed ? P.mx_split : P.m2l_split;
      const size_t per_split = mixed ? (size_t)P.mx_m2l_jobs * P.mx_m2l_maxN * RP_
                                     : (size_t)P.m2l.njobs * P.m2l.maxN * m2l_ld();
      std::vector<int> splits = {s0};
      if (s0 > 1) splits.push_back(std::max(1, s0 / 2));
      if ((size_t)std::min(64, 2 * s0) * per_split <= part_elems_) splits.push_back(std::min(64, 2 * s0));
      // FP64 forms: depth 0 = class tiles, 1 = per-offset pair jobs (split
      // 1); mixed forms: 12-14, 17 = per-offset pair jobs through the depth 2-4, 7
      // kernel (split 1)
      std::vector<int> depths = !mixed ? (P.m2l_npairs ? std::vector<int>{0, 1} : std::vector<int>{0})
                                : np_ == 3 && !f16_ ? std::vector<int>{2, 4}
                                                    : std::vector<int>{2, 3, 4, 5, 6, 7, 8};
      if (mixed && P.m2l_npairs) {
        if (np_ == 3 && !f16_) depths.insert(depths.end(), {12, 14});
        else depths.insert(depths.end(), {12, 13, 14, 17});
      }
      auto time_cand = [&](int sp, int dp) {
        auto run = [&]() {
          if (mixed) launch_m2l_mixed(l, dp, dp > 10 ? 1 : sp, s);
          else if (dp == 1 && prec_ == 4)
            launch_m2l_pairs(P, chat_.as<double>(), RP_, dhat_.as<double>(), RP_, R_, MpadR_, KpadR_, s);
          else if (dp == 1) launch_m2l_pairs(P, uppot_.as<double>(), K_, chk_.as<double>(), K_, K_, Mpad_, Kpad_, s);
          else if (prec_ == 4) launch_m2l_t(l, sp, s);
          else gemm(P.m2l, uppot_.as<double>(), K_, chk_.as<double>(), K_, 0, s, 1, sp, &P.m2l_slot);
        };
        run();  // warm
        NNK_HIP(hipEventRecord(a, s));
        run();
        run();
        NNK_HIP(hipEventRecord(b, s));
        NNK_HIP(hipEventSynchronize(b));
        float ms = 0;
        NNK_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
      };
      struct Cand { int sp, dp; float ms; };
      std::vector<Cand> cands;
      for (int sp : splits) {
        if (sp > 1 && !(mixed ? P.mx_slot.p : P.m2l_slot.p)) continue;
        for (int dp : depths) {
          if (dp >= 4 && dp <= 8 && !dma_ok(P, sp, dp)) continue;
          if ((dp == 1 || dp > 10) && sp != s0) continue;  // one candidate per pair-job launch
          cands.push_back({sp, dp, time_cand(sp, dp)});
        }
      }
      if (cands.empty()) continue;
      // One timing can be off by a few percent (clocks, the L2's contents),
      // and candidates that close would be chosen at random.  The ones within
      // 5 % of the fastest are timed again, interleaved, three more times;
      // each keeps its best time, and of two within 0.5 % the earlier one in
      // the candidate order wins.
      float best = 1e30f;
      for (const Cand &c : cands) best = std::min(best, c.ms);
      std::vector<Cand> top;
      for (const Cand &c : cands)
        if (c.ms <= 1.05f * best) top.push_back(c);
      for (int rep = 0; rep < 3 && top.size() > 1; rep++)
        for (Cand &c : top) c.ms = std::min(c.ms, time_cand(c.sp, c.dp));
      float tb = 1e30f;
      for (const Cand &c : top) tb = std::min(tb, c.ms);
      int bs = top[0].sp, bd = top[0].dp;
      for (const Cand &c : top)
        if (c.ms <= 1.005f * tb) { bs = c.sp; bd = c.dp; break; }
      if (mixed && bd > 10) { P.m2l_pair_on = true; P.mx_pair_depth = bd - 10; }
      else if (mixed) { P.mx_split = bs; P.mx_depth = bd; P.m2l_pair_on = false; }
      else if (bd == 1) P.m2l_pair_on = true;
      else { P.m2l_split = bs; P.m2l_pair_on = false; }
    }
  } catch (...) {
    prof_ = prof;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  prof_ = prof;
  NNK_HIP(hipEventDestroy(a));
  NNK_HIP(hipEventDestroy(b));
}

// time the leaf kernel (L2T + P2P + M2T) at 128, 256 and 512 threads per
// target leaf on the current densities and keep the fastest
void Fmm::autotune_leaf(hipStream_t s) {
  if (!n_tleaves_) return;
  DBuf out;
  out.alloc(sizeof(double) * 3 * std::max<int64_t>(ntgt_, 1));
  hipEvent_t a, b;
  NNK_HIP(hipEventCreate(&a));
  NNK_HIP(hipEventCreate(&b));
  float best = 1e30f;
  int bt = leaf_threads_;
  try {
    for (int T : {128, 256, 512}) {
      leaf_threads_ = T;
      launch_leaf(1, out.as<double>(), s);  // warm
      NNK_HIP(hipEventRecord(a, s));
      launch_leaf(1, out.as<double>(), s);
      launch_leaf(1, out.as<double>(), s);
      NNK_HIP(hipEventRecord(b, s));
      NNK_HIP(hipEventSynchronize(b));
      float ms = 0;
      NNK_HIP(hipEventElapsedTime(&ms, a, b));
      if (ms < best) { best = ms; bt = T; }
    }
  } catch (...) {
    leaf_threads_ = bt;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  leaf_threads_ = bt;
  NNK_HIP(hipEventDestroy(a));
  NNK_HIP(hipEventDestroy(b));
}

void Fmm::apply_impl(const double *d_f, double *d_u, hipStream_t s) {
  if (prec_ >= 1) { apply_mixed(d_f, d_u, s); return; }
  mark(0, true, s);
  strengths(d_f, s);
  mark(0, false, s);
  double *chk = chk_.as<double>(), *up = uppot_.as<double>(), *dn = downpot_.as<double>(),
         *tmp = tmp_.as<double>();
  // S2M for every potential leaf (up densities of levels 0-1 are never used,
  // but computing them is harmless and keeps one launch)
  if (maxlev_ >= 2 && n_pleaves_) {
    mark(1, true, s);
    launch_surf(pleaves_.as<int32_t>(), n_pleaves_, nullptr, nullptr, 1, 0, chk, s);
    mark(1, false, s);
  }
  mark(2, true, s);
  for (int l = maxlev_; l >= 2; l--) {
    LevelPlan &P = plans_[l];
    if (!P.n_up) continue;
    gemm(P.m2m, up, K_, chk, K_, 0, s, 0, P.m2m_split, &P.m2m_slot);
    gemm(P.up_solve1, chk, K_, tmp, K_, 0, s);
    gemm(P.up_solve2, tmp, K_, up, K_, 0, s);
  }
  mark(2, false, s);
  if (n_x_all_) {  // P2L of every level into its own buffer
    mark(5, true, s);
    launch_surf(x_all_.as<int32_t>(), n_x_all_, xptr_.as<int64_t>(), xidx_.as<int32_t>(), 0, 0,
                xchk_.as<double>(), s);
    mark(5, false, s);
  }
  for (int l = 2; l <= maxlev_; l++) {
    LevelPlan &P = plans_[l];
    if (!P.n_down) continue;
    mark(3, true, s);
    if (P.m2l_pair_on) launch_m2l_pairs(P, up, K_, chk, K_, K_, Mpad_, Kpad_, s);
    else gemm(P.m2l, up, K_, chk, K_, 0, s, 1, P.m2l_split, &P.m2l_slot);
    mark(3, false, s);
    mark(4, true, s);
    gemm(P.l2l, dn, K_, chk, K_, 1, s);
    mark(4, false, s);
    if (P.n_x) {  // chk += the P2L check potentials (as one addition each)
      mark(5, true, s);
      hipLaunchKernelGGL(k_add_rows, dim3(P.n_x), dim3(256), 0, s, P.x_nodes.as<int32_t>(), xchk_.as<double>(), chk,
                         K_);
      NNK_HIP(hipGetLastError());
      mark(5, false, s);
    }
    mark(6, true, s);
    gemm(P.dn_solve1, chk, K_, tmp, K_, 0, s);
    gemm(P.dn_solve2, tmp, K_, dn, K_, 0, s);
    mark(6, false, s);
  }
  mark(7, true, s);
  launch_leaf(1, d_u, s);
  mark(7, false, s);
}

void Fmm::apply_graph(const double *d_f, double *d_u, hipStream_t s) {
  if (tune_ && !tuned_) apply(d_f, d_u, s);  // tune outside the capture (it times and allocates)
  if (!gop_.ready()) {
    const bool prof = prof_;
    prof_ = false;  // no event records inside the graph
    try {
      gop_.capture(3 * npot_, 3 * ntgt_, [&](const double *in, double *out, hipStream_t cs) { apply(in, out, cs); });
    } catch (...) {
      prof_ = prof;
      throw;
    }
    prof_ = prof;
  }
  gop_.run(d_f, d_u, s);
}

void Fmm::debug_copy(int which, double *host) const {
  const DBuf &b = which == 0 ? chk_ : which == 1 ? uppot_ : downpot_;
  NNK_HIP(hipDeviceSynchronize());
  NNK_HIP(hipMemcpy(host, b.p, sizeof(double) * (size_t)tree_.count * K_, hipMemcpyDeviceToHost));
}

void Fmm::reduced_apply(const double *d_f, double *d_u, hipStream_t s) {
  strengths(d_f, s);
  launch_leaf(0, d_u, s);
}

}  // namespace nnk
