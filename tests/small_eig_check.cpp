// Host driver for tests/test_small_eig.py: reads n and a row-major n x n
// matrix from stdin, prints every eigenvalue (real, imaginary) and then, for
// each eigenvalue in the same order, its eigenvector (real and imaginary parts).
#include <cstdio>
#include <vector>

#include "../mpqc_amd/csrc/small_eig.hpp"

int main() {
  int n = 0;
  if (std::scanf("%d", &n) != 1 || n < 1) return 2;
  std::vector<double> A((size_t)n * n);
  for (double& a : A)
    if (std::scanf("%lf", &a) != 1) return 2;
  const auto w = mpqc_hip::small_eigenvalues(n, A);
  for (const auto& z : w) std::printf("%.17g %.17g\n", z.real(), z.imag());
  for (const auto& z : w) {
    const auto x = mpqc_hip::small_eigenvector(n, A, z);
    for (const auto& e : x) std::printf("%.17g %.17g ", e.real(), e.imag());
    std::printf("\n");
  }
  return 0;
}
