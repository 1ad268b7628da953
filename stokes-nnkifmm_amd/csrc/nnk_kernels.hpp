// Device kernels for the KIFMM matvec (FP64 form).
//
// Notation: a node's "up surface" is the res^3 grid on its box minus the
// interior, its "down surface" the same on the box widened by its width on
// every side (Tools/KIFMM/genupsurf.m, gendownsurf.m).  Surface point p has
// grid indices packed in surf_idx[p] = ix | iy<<8 | iz<<16, ordered y
// fastest, then x, then z.  Densities are stored per node as 3*Nq doubles,
// interleaved [p0x p0y p0z p1x ...] as in the reference's uppot(node,:).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nnk {

constexpr double kPi = 3.14159265358979323846;

// linspace(d1,d2,n)(k): d1 + (k*(d2-d1))/(n-1) with exact endpoints, as
// MATLAB's linspace computes it.
__device__ __forceinline__ double lin(double d1, double d2, int k, int n1) {
  if (k == 0) return d1;
  if (k == n1) return d2;
  return d1 + ((double)k * (d2 - d1)) / (double)n1;
}

__device__ __forceinline__ void surf_point(const double *__restrict__ box, int down, int code,
                                           int n1, double &x, double &y, double &z) {
  double lo0 = box[0], lo1 = box[1], lo2 = box[2], hi0 = box[3], hi1 = box[4], hi2 = box[5];
  if (down) {
    double w0 = hi0 - lo0, w1 = hi1 - lo1, w2 = hi2 - lo2;
    lo0 = lo0 - w0; hi0 = hi0 + w0;
    lo1 = lo1 - w1; hi1 = hi1 + w1;
    lo2 = lo2 - w2; hi2 = hi2 + w2;
  }
  x = lin(lo0, hi0, code & 255, n1);
  y = lin(lo1, hi1, (code >> 8) & 255, n1);
  z = lin(lo2, hi2, (code >> 16) & 255, n1);
}

// u += S(t <- s) q without the 1/(8 pi mu) factor:
// S q = [(r^2 + 2 eps^2) q + r (r.q)] / (r^2 + eps^2)^{3/2},  r = t - s.
__device__ __forceinline__ void stokeslet_acc(double tx, double ty, double tz, double sx, double sy,
                                              double sz, double qx, double qy, double qz, double e2,
                                              double &ux, double &uy, double &uz) {
  double rx = tx - sx, ry = ty - sy, rz = tz - sz;
  double r2 = rx * rx + ry * ry + rz * rz;
  double re = r2 + e2;
  double inv = rsqrt(re);
  double inv3 = inv * inv * inv;
  double iso = (re + e2) * inv3;
  double rq = (rx * qx + ry * qy + rz * qz) * inv3;
  ux += iso * qx + rq * rx;
  uy += iso * qy + rq * ry;
  uz += iso * qz + rq * rz;
}

// Mixed form: the separation r = t - s is formed in FP64 (so near pairs keep
// their digits), the kernel is evaluated in FP32 with the hardware reciprocal
// square root, and the result is accumulated in FP64.  The strength comes in
// FP32 (rounded once, where the source is staged).
__device__ __forceinline__ void stokeslet_acc_f32q(double tx, double ty, double tz, double sx,
                                                   double sy, double sz, float fx, float fy,
                                                   float fz, double e2, double &ux, double &uy,
                                                   double &uz) {
  const float rx = (float)(tx - sx), ry = (float)(ty - sy), rz = (float)(tz - sz);
  const float e2f = (float)e2;
  const float r2 = rx * rx + ry * ry + rz * rz;
  const float re = r2 + e2f;
  const float inv = __builtin_amdgcn_rsqf(re);
  const float inv3 = inv * inv * inv;
  const float iso = __builtin_fmaf(e2f, inv3, inv);  // (re + e2) / re^1.5
  const float rq = (rx * fx + ry * fy + rz * fz) * inv3;
  ux += (double)(iso * fx + rq * rx);
  uy += (double)(iso * fy + rq * ry);
  uz += (double)(iso * fz + rq * rz);
}

__device__ __forceinline__ void stokeslet_acc_f32(double tx, double ty, double tz, double sx,
                                                  double sy, double sz, double qx, double qy,
                                                  double qz, double e2, double &ux, double &uy,
                                                  double &uz) {
  stokeslet_acc_f32q(tx, ty, tz, sx, sy, sz, (float)qx, (float)qy, (float)qz, e2, ux, uy, uz);
}

// LDS stage of up to T sources: positions in FP64; strengths in FP64, or,
// for the FP32-math form, rounded to FP32 once when staged (the values the
// pair routine would otherwise round per pair).  acc(j, ...) adds source j.
template <bool F32, int T> struct SrcStage;
template <int T> struct SrcStage<false, T> {
  double p[3 * T], q[3 * T];
  __device__ __forceinline__ void put(int i, double x, double y, double z, double qx, double qy, double qz) {
    p[3 * i] = x; p[3 * i + 1] = y; p[3 * i + 2] = z;
    q[3 * i] = qx; q[3 * i + 1] = qy; q[3 * i + 2] = qz;
  }
  __device__ __forceinline__ void acc(int j, double tx, double ty, double tz, double e2, double &ux,
                                      double &uy, double &uz) const {
    stokeslet_acc(tx, ty, tz, p[3 * j], p[3 * j + 1], p[3 * j + 2], q[3 * j], q[3 * j + 1], q[3 * j + 2], e2,
                  ux, uy, uz);
  }
  // sources j = j0, j0 + step, ... < cnt, four per iteration (the same
  // sequence of FP64 accumulations as one at a time)
  __device__ __forceinline__ void run(int j0, int step, int cnt, double tx, double ty, double tz, double e2,
                                      double &ux, double &uy, double &uz) const {
    int j = j0;
    for (; j + 3 * step < cnt; j += 4 * step) {
#pragma unroll
      for (int k = 0; k < 4; k++) acc(j + k * step, tx, ty, tz, e2, ux, uy, uz);
    }
    for (; j < cnt; j += step) acc(j, tx, ty, tz, e2, ux, uy, uz);
  }
};
template <int T> struct SrcStage<true, T> {
  double p[3 * T];
  __attribute__((aligned(16))) float q[4 * T];
  __device__ __forceinline__ void put(int i, double x, double y, double z, double qx, double qy, double qz) {
    p[3 * i] = x; p[3 * i + 1] = y; p[3 * i + 2] = z;
    q[4 * i] = (float)qx; q[4 * i + 1] = (float)qy; q[4 * i + 2] = (float)qz;
  }
  __device__ __forceinline__ void acc(int j, double tx, double ty, double tz, double e2, double &ux,
                                      double &uy, double &uz) const {
    stokeslet_acc_f32q(tx, ty, tz, p[3 * j], p[3 * j + 1], p[3 * j + 2], q[4 * j], q[4 * j + 1], q[4 * j + 2],
                       e2, ux, uy, uz);
  }
  // sources j = j0, j0 + step, ... < cnt, four per iteration: the four
  // FP32 contributions are added in FP32 (in order) and then once into the
  // FP64 sums, which saves three conversions and FP64 adds per pair
  __device__ __forceinline__ void run(int j0, int step, int cnt, double tx, double ty, double tz, double e2,
                                      double &ux, double &uy, double &uz) const {
    const float e2f = (float)e2;
    int j = j0;
    for (; j + 3 * step < cnt; j += 4 * step) {
      float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int i = j + k * step;
        const float4 f = *reinterpret_cast<const float4 *>(&q[4 * i]);
        const float rx = (float)(tx - p[3 * i]), ry = (float)(ty - p[3 * i + 1]), rz = (float)(tz - p[3 * i + 2]);
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float re = r2 + e2f;
        const float inv = __builtin_amdgcn_rsqf(re);
        const float inv3 = inv * inv * inv;
        const float iso = __builtin_fmaf(e2f, inv3, inv);  // (re + e2) / re^1.5
        const float rq = (rx * f.x + ry * f.y + rz * f.z) * inv3;
        gx += iso * f.x + rq * rx;
        gy += iso * f.y + rq * ry;
        gz += iso * f.z + rq * rz;
      }
      ux += (double)gx;
      uy += (double)gy;
      uz += (double)gz;
    }
    for (; j < cnt; j += step) acc(j, tx, ty, tz, e2, ux, uy, uz);
  }
};

// FP32-math stage for the evaluations that involve a surface (S2M, P2L,
// L2T, M2T): positions are FP32 offsets from an origin o, formed in FP64
// and rounded once (at staging for the sources, per target for the
// targets).  The separation t - s is then one FP32 subtraction.  Its
// rounding error is at most 2^-24 (|t - o| + |s - o|) per component.
// The callers put o at the centre of the smaller of the two boxes (width
// w).  The surface geometry keeps every such pair at least w apart and the
// point nearer o within 1.5 sqrt(3) w of it, so the error stays within a
// few FP32 ulps of |r|, the size of the kernel's own FP32 rounding.  (P2P
// pairs can be arbitrarily close and keep the FP64 separation of
// SrcStage.)
template <int T> struct SrcStageRel {
  float4 p[T];  // x, y, z offset from the origin; w unused
  float4 q[T];  // strength; w unused
  __device__ __forceinline__ void put(int i, double dx, double dy, double dz, double qx, double qy,
                                      double qz) {
    p[i] = float4{(float)dx, (float)dy, (float)dz, 0.f};
    q[i] = float4{(float)qx, (float)qy, (float)qz, 0.f};
  }
  __device__ __forceinline__ static void pair(const float4 &s, const float4 &f, float tx, float ty, float tz,
                                              float e2f, float &gx, float &gy, float &gz) {
    const float rx = tx - s.x, ry = ty - s.y, rz = tz - s.z;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float re = r2 + e2f;
    const float inv = __builtin_amdgcn_rsqf(re);
    const float inv3 = inv * inv * inv;
    const float iso = __builtin_fmaf(e2f, inv3, inv);  // (re + e2) / re^1.5
    const float rq = (rx * f.x + ry * f.y + rz * f.z) * inv3;
    gx = __builtin_fmaf(rq, rx, __builtin_fmaf(iso, f.x, gx));
    gy = __builtin_fmaf(rq, ry, __builtin_fmaf(iso, f.y, gy));
    gz = __builtin_fmaf(rq, rz, __builtin_fmaf(iso, f.z, gz));
  }
  // sources j = j0, j0 + step, ... < cnt onto the target offset (tx, ty,
  // tz): each group of four is summed in FP32, then added to the FP64 sums
  __device__ __forceinline__ void run(int j0, int step, int cnt, float tx, float ty, float tz, float e2f,
                                      double &ux, double &uy, double &uz) const {
    int j = j0;
    for (; j + 3 * step < cnt; j += 4 * step) {
      float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
      for (int k = 0; k < 4; k++) pair(p[j + k * step], q[j + k * step], tx, ty, tz, e2f, gx, gy, gz);
      ux += (double)gx;
      uy += (double)gy;
      uz += (double)gz;
    }
    for (; j < cnt; j += step) {
      float gx = 0.f, gy = 0.f, gz = 0.f;
      pair(p[j], q[j], tx, ty, tz, e2f, gx, gy, gz);
      ux += (double)gx;
      uy += (double)gy;
      uz += (double)gz;
    }
  }
};

// centre of box [lo0 lo1 lo2 hi0 hi1 hi2]
__device__ __forceinline__ void box_centre(const double *__restrict__ box, double &cx, double &cy, double &cz) {
  cx = 0.5 * (box[0] + box[3]);
  cy = 0.5 * (box[1] + box[4]);
  cz = 0.5 * (box[2] + box[5]);
}

template <bool F32>
__device__ __forceinline__ void stk(double tx, double ty, double tz, double sx, double sy, double sz,
                                    double qx, double qy, double qz, double e2, double &ux, double &uy,
                                    double &uz) {
  if constexpr (F32) stokeslet_acc_f32(tx, ty, tz, sx, sy, sz, qx, qy, qz, e2, ux, uy, uz);
  else stokeslet_acc(tx, ty, tz, sx, sy, sz, qx, qy, qz, e2, ux, uy, uz);
}

}  // namespace nnk
