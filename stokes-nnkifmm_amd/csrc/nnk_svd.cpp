// One-sided Jacobi SVD (Hestenes 1958; Demmel & Veselic 1992) for the
// per-level check-to-equivalent matrices.  These are square, dense and
// ill-conditioned (cond ~ 2e10 at 6 points per edge); applying the factors
// as x = V (S^-1 (U^T b)) is backward stable like the reference's pivoted
// LU (Up.m:31, DownNonleaf.m:46), whereas an explicit inverse is not.
//
// Columns are rotated pairwise until they are mutually orthogonal to
// working precision; a round-robin ordering lets n/2 disjoint pairs rotate
// in parallel.  Setup-only code: it runs once per tree level.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace nnk {

// A: n x n row-major.  On return U (n x n row-major, columns = left
// singular vectors), s (n, descending), V (n x n row-major, columns = right
// singular vectors), with A = U diag(s) V^T.  Returns sweeps used.
int svd_jacobi(const double *A, int n, double *U, double *s, double *V) {
  // work on columns stored contiguously: G[j] = column j of A
  std::vector<double> G((size_t)n * n), Q((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) G[(size_t)j * n + i] = A[(size_t)i * n + j];
  for (int j = 0; j < n; j++) Q[(size_t)j * n + j] = 1.0;
  const int m = n + (n & 1);  // round-robin over an even count (dummy if odd)
  std::vector<int> ring(m);
  std::iota(ring.begin(), ring.end(), 0);
  const double tol = 1e-15;
  int sweep = 0;
  for (; sweep < 60; sweep++) {
    double off = 0.0;
    for (int step = 0; step < m - 1; step++) {
#pragma omp parallel for schedule(static) reduction(max : off)
      for (int k = 0; k < m / 2; k++) {
        int p = ring[k], q = ring[m - 1 - k];
        if (p >= n || q >= n) continue;
        if (p > q) std::swap(p, q);
        double *gp = &G[(size_t)p * n], *gq = &G[(size_t)q * n];
        double a = 0, b = 0, c = 0;
        for (int i = 0; i < n; i++) {
          a += gp[i] * gp[i];
          b += gq[i] * gq[i];
          c += gp[i] * gq[i];
        }
        if (c == 0.0) continue;
        double rel = std::fabs(c) / std::sqrt(a * b);
        off = std::max(off, rel);
        if (rel <= tol) continue;
        double zeta = (b - a) / (2.0 * c);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < n; i++) {
          double x = gp[i], y = gq[i];
          gp[i] = cs * x - sn * y;
          gq[i] = sn * x + cs * y;
        }
        double *vp = &Q[(size_t)p * n], *vq = &Q[(size_t)q * n];
        for (int i = 0; i < n; i++) {
          double x = vp[i], y = vq[i];
          vp[i] = cs * x - sn * y;
          vq[i] = sn * x + cs * y;
        }
      }
      // rotate the ring, keeping ring[0] fixed
      int last = ring[m - 1];
      for (int k = m - 1; k > 1; k--) ring[k] = ring[k - 1];
      ring[1] = last;
    }
    if (off <= tol) break;
  }
  std::vector<double> norms(n);
  for (int j = 0; j < n; j++) {
    double a = 0;
    for (int i = 0; i < n; i++) a += G[(size_t)j * n + i] * G[(size_t)j * n + i];
    norms[j] = std::sqrt(a);
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return norms[x] > norms[y]; });
  for (int k = 0; k < n; k++) {
    int j = ord[k];
    s[k] = norms[j];
    double inv = norms[j] > 0 ? 1.0 / norms[j] : 0.0;
    for (int i = 0; i < n; i++) {
      U[(size_t)i * n + k] = G[(size_t)j * n + i] * inv;
      V[(size_t)i * n + k] = Q[(size_t)j * n + i];
    }
  }
  return sweep + 1;
}

}  // namespace nnk
