// Eigenvalues and eigenvectors of a small dense non-symmetric matrix
// (host code; the Davidson subspace matrices of the EOM solver, n <= ~64).
//
// Eigenvalues: reduction to upper Hessenberg form by Householder reflections,
// then single-shift QR iterations in complex arithmetic with the Wilkinson
// shift of the trailing 2x2 block, deflating from the bottom.  Eigenvectors of
// an eigenvalue: inverse iteration with LU (partial pivoting) on the shifted
// original matrix.  Plain C++ (no GPU code), so it can be built by any host
// compiler.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <stdexcept>
#include <vector>

namespace mpqc_hip {

// all eigenvalues of the row-major n x n matrix A
inline std::vector<std::complex<double>> small_eigenvalues(int n, const std::vector<double>& A) {
  typedef std::complex<double> cd;
  if ((int)A.size() < n * n) throw std::invalid_argument("small_eigenvalues: matrix too small");
  std::vector<cd> H(A.begin(), A.begin() + (size_t)n * n);
  auto h = [&](int i, int j) -> cd& { return H[(size_t)i * n + j]; };
  double scale = 0.0;
  for (const cd& x : H) scale = std::max(scale, std::abs(x));
  if (scale == 0.0) return std::vector<cd>(n, cd(0.0, 0.0));
  // Householder reduction to Hessenberg form: zero column k below row k + 1
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += std::norm(h(i, k));
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    std::vector<cd> u(n, cd(0.0, 0.0));
    const cd x0 = h(k + 1, k);
    const cd ph = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : cd(1.0, 0.0);
    u[k + 1] = x0 + ph * alpha;
    for (int i = k + 2; i < n; ++i) u[i] = h(i, k);
    double un = 0.0;
    for (int i = k + 1; i < n; ++i) un += std::norm(u[i]);
    if (un == 0.0) continue;
    // H <- (I - 2 u u^H / |u|^2) H (I - 2 u u^H / |u|^2)
    for (int j = 0; j < n; ++j) {
      cd s(0.0, 0.0);
      for (int i = k + 1; i < n; ++i) s += std::conj(u[i]) * h(i, j);
      s *= 2.0 / un;
      for (int i = k + 1; i < n; ++i) h(i, j) -= u[i] * s;
    }
    for (int i = 0; i < n; ++i) {
      cd s(0.0, 0.0);
      for (int j = k + 1; j < n; ++j) s += h(i, j) * u[j];
      s *= 2.0 / un;
      for (int j = k + 1; j < n; ++j) h(i, j) -= s * std::conj(u[j]);
    }
    for (int i = k + 2; i < n; ++i) h(i, k) = cd(0.0, 0.0);
  }
  // shifted QR on the active block [0, hi]
  std::vector<cd> w(n);
  const double eps = 2.2e-16;
  int hi = n - 1, iter = 0;
  while (hi >= 0) {
    if (hi == 0) {
      w[0] = h(0, 0);
      break;
    }
    // deflation point: the last negligible subdiagonal element
    int lo = hi;
    while (lo > 0) {
      const double s = std::abs(h(lo, lo)) + std::abs(h(lo - 1, lo - 1));
      if (std::abs(h(lo, lo - 1)) <= eps * (s > 0.0 ? s : scale)) {
        h(lo, lo - 1) = cd(0.0, 0.0);
        break;
      }
      --lo;
    }
    if (lo == hi) {
      w[hi] = h(hi, hi);
      --hi;
      iter = 0;
      continue;
    }
    if (++iter > 60 * n) throw std::runtime_error("small_eigenvalues: QR iterations did not converge");
    // Wilkinson shift: eigenvalue of the trailing 2x2 block closest to h(hi,hi)
    const cd a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
    const cd tr = 0.5 * (a + d);
    const cd disc = std::sqrt(0.25 * (a - d) * (a - d) + b * c);
    cd mu = (std::abs(tr + disc - d) < std::abs(tr - disc - d)) ? tr + disc : tr - disc;
    if (iter % 11 == 0) mu += cd(std::abs(h(hi, hi - 1)), 0.0);  // exceptional shift
    // QR step on rows/cols [lo, hi] by Givens rotations: H - mu I = Q R, H <- R Q + mu I
    for (int i = lo; i <= hi; ++i) h(i, i) -= mu;
    std::vector<cd> cs(hi - lo), sn(hi - lo);
    for (int k = lo; k < hi; ++k) {
      const cd x = h(k, k), y = h(k + 1, k);
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      cd cc(1.0, 0.0), ss(0.0, 0.0);
      if (r > 0.0) {
        cc = x / r;
        ss = y / r;
      }
      cs[k - lo] = cc;
      sn[k - lo] = ss;
      // rows k, k+1 <- G^H [rows]  with G = [[c, -conj(s)], [s, conj(c)]]
      for (int j = k; j < n; ++j) {
        const cd p = h(k, j), q = h(k + 1, j);
        h(k, j) = std::conj(cc) * p + std::conj(ss) * q;
        h(k + 1, j) = -ss * p + cc * q;
      }
    }
    for (int k = lo; k < hi; ++k) {
      const cd cc = cs[k - lo], ss = sn[k - lo];
      // columns k, k+1 <- [cols] G
      for (int i = 0; i <= std::min(hi, k + 1); ++i) {
        const cd p = h(i, k), q = h(i, k + 1);
        h(i, k) = p * cc + q * ss;
        h(i, k + 1) = -p * std::conj(ss) + q * std::conj(cc);
      }
    }
    for (int i = lo; i <= hi; ++i) h(i, i) += mu;
  }
  return w;
}

// right eigenvector of the (possibly complex) eigenvalue lambda of the row-major
// n x n matrix A, by inverse iteration with LU (partial pivoting) of A - sigma I,
// sigma = lambda shifted by 1e-12 of the largest element; unit 2-norm, phase
// fixed so that the largest component is real and positive
inline std::vector<std::complex<double>> small_eigenvector(int n, const std::vector<double>& A,
                                                           std::complex<double> lambda) {
  typedef std::complex<double> cd;
  double scale = 0.0;
  for (int i = 0; i < n * n; ++i) scale = std::max(scale, std::fabs(A[i]));
  if (scale == 0.0) {  // the zero matrix: every vector is an eigenvector
    std::vector<cd> e(n, cd(0.0, 0.0));
    e[0] = cd(1.0, 0.0);
    return e;
  }
  const cd sigma = lambda + cd(1e-12 * scale, 0.0);
  std::vector<cd> M(A.begin(), A.begin() + (size_t)n * n);
  for (int i = 0; i < n; ++i) M[(size_t)i * n + i] -= sigma;
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(M[(size_t)i * n + k]) > std::abs(M[(size_t)p * n + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(M[(size_t)k * n + j], M[(size_t)p * n + j]);
    cd pv = M[(size_t)k * n + k];
    if (std::abs(pv) < 1e-14 * scale) {
      pv = cd(1e-14 * scale, 0.0);
      M[(size_t)k * n + k] = pv;
    }
    for (int i = k + 1; i < n; ++i) {
      const cd f = M[(size_t)i * n + k] / pv;
      M[(size_t)i * n + k] = f;
      for (int j = k + 1; j < n; ++j) M[(size_t)i * n + j] -= f * M[(size_t)k * n + j];
    }
  }
  std::vector<cd> x(n, cd(1.0, 0.0));
  for (int it = 0; it < 3; ++it) {
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) std::swap(x[k], x[piv[k]]);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) x[i] -= M[(size_t)i * n + j] * x[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) x[i] -= M[(size_t)i * n + j] * x[j];
      x[i] /= M[(size_t)i * n + i];
    }
    double nr = 0.0;
    for (const cd& e : x) nr += std::norm(e);
    nr = std::sqrt(nr);
    if (!(nr > 0.0) || !std::isfinite(nr)) throw std::runtime_error("small_eigenvector: breakdown");
    for (cd& e : x) e /= nr;
  }
  int im = 0;
  for (int i = 1; i < n; ++i)
    if (std::abs(x[i]) > std::abs(x[im])) im = i;
  const cd ph = std::conj(x[im]) / std::abs(x[im]);
  for (cd& e : x) e *= ph;
  return x;
}

}  // namespace mpqc_hip
