// Hierarchical-reconstruction helpers in the Taylor basis x^n/n! on [-1,1]
// (limiting_kernels.cu:1470-1486 minmod, 1534-1583 limit_monomial,
// 1748-1790 integrals and average_monomial).  Every body is compiled without
// FMA contraction so that the limiter rounds like the CPU path whichever
// translation unit includes this header.
#pragma once

namespace dgx {

__device__ inline int signum_d(double v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

__device__ inline double minmod_d(double a, double b) {
#pragma clang fp contract(off)
  int sa = signum_d(a);
  if (sa != signum_d(b)) return 0;
  double fa = fabs(a), fb = fabs(b);
  if (fa < fb) return sa * fa;
  return sa * fb;
}

// skip_zero_terms: the loops below leave out the products with a zero
// integral (known once the loops are unrolled).  Each left-out product of a
// finite value is +-0, and adding +-0 to a sum that started at +0.0 does not
// change it: a non-zero sum stays as it is, and in round-to-nearest the sum
// is never -0 (x + (-x) = +0 and +0 + (-0) = +0).  The results are therefore
// the bits of the full sums the reference forms.
__host__ __device__ constexpr int ifact(int n) { return n <= 1 ? 1 : n * ifact(n - 1); }
__host__ __device__ constexpr long ipow(long b, int n) { return n == 0 ? 1 : b * ipow(b, n - 1); }

// integral over [-1,1] of the k-th derivative of x^n/n! (limiting_kernels.cu:1748-1760)
__device__ inline double integ_deriv(int k, int n) {
#pragma clang fp contract(off)
  int num = n - k + 1;
  if (num % 2) return 2.0 / (double)ifact(num);
  return 0.0;
}
// over [a,b] with integer a,b (limiting_kernels.cu:1762-1775); powers are exact
__device__ inline double integ_deriv_bounds(int k, int n, int a, int b) {
#pragma clang fp contract(off)
  int num = n - k + 1;
  return ((double)ipow(b, num) - (double)ipow(a, num)) / (double)ifact(num);
}

template <int N>
__device__ inline void limit_monomial_d(const double* AL, const double* AC, const double* AR, double* Alim) {
#pragma clang fp contract(off)
  double avgLC = 0;
#pragma unroll
  for (int m = N; m > 0; m--) {
    double avgdUL = 0, avgdUC = 0, avgdUR = 0, avgRL = 0, avgRC = 0, avgRR = 0;
#pragma unroll
    for (int n = m - 1; n <= N; n++) {
      // the zero integrals (n - m even) are left out: see skip_zero_terms
      double integral = integ_deriv(m - 1, n);
      if (integral != 0.0) {
        avgdUL += AL[n] * integral;
        avgdUC += AC[n] * integral;
        avgdUR += AR[n] * integral;
      }
      if (n >= m + 1) {
        double alim = Alim[n];
        avgRL += alim * integ_deriv_bounds(m - 1, n, -3, -1);
        if (integral != 0.0) avgRC += alim * integral;
        avgRR += alim * integ_deriv_bounds(m - 1, n, 1, 3);
      }
    }
    double avgLL = 0.5 * (avgdUL - avgRL);
    avgLC = 0.5 * (avgdUC - avgRC);
    double avgLR = 0.5 * (avgdUR - avgRR);
    double c1 = 0.5 * (avgLC - avgLL);
    double c2 = 0.5 * (avgLR - avgLC);
    Alim[m] = minmod_d(c1, c2);
  }
  Alim[0] = avgLC;
}

template <int N>
__device__ inline void average_monomial_d(const double* A, double* Alim) {
#pragma clang fp contract(off)
  double avg = 0;
#pragma unroll
  for (int n = 0; n <= N; n++) {
    if (integ_deriv(0, n) != 0.0) avg += A[n] * integ_deriv(0, n);
    Alim[n] = 0;
  }
  Alim[0] = 0.5 * avg;
}

}  // namespace dgx
