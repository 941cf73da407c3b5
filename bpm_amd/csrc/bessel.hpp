// Bessel functions of the first kind J_0, J_1, J_2 in float64, host and device.
//
// Methods (Abramowitz & Stegun, Handbook of Mathematical Functions, 1964):
//   |x| <  14: ascending power series, A&S 9.1.10
//       J_n(x) = (x/2)^n sum_k (-x^2/4)^k / (k! (n+k)!)
//     (largest term at x=14 is ~2.7e4, so ~4 digits of the 16 are lost to
//      cancellation -- well beyond float32 needs);
//   |x| >= 14: Hankel asymptotic expansion, A&S 9.2.5 with 9.2.9/9.2.10
//       J_n(x) = sqrt(2/(pi x)) (P cos chi - Q sin chi),  chi = x - (n/2+1/4) pi
//       P = sum_k (-1)^k a_{2k}/x^{2k},  Q = sum_k (-1)^k a_{2k+1}/x^{2k+1},
//       a_k = prod_{j=1..k} (4n^2 - (2j-1)^2) / (k! 8^k),
//     summed until the terms stop decreasing (error ~ e^{-2x} < 1e-12).
// J_n(-x) = (-1)^n J_n(x).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace bpmbessel {

__host__ __device__ inline double jn_series(int n, double x) {
  const double h = 0.5 * x;
  double lead = 1.0;  // (x/2)^n / n!
  for (int i = 1; i <= n; ++i) lead *= h / i;
  const double q = -h * h;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 80; ++k) {
    term *= q / ((double)k * (double)(k + n));
    sum += term;
    if (fabs(term) < 1e-17 * fabs(sum) && k > 2) break;
  }
  return lead * sum;
}

__host__ __device__ inline double jn_asymptotic(int n, double x) {
  const double mu = 4.0 * n * n;
  double P = 0.0, Q = 0.0;
  double a = 1.0;       // a_k / x^k
  double prev = 1e300;
  for (int k = 0; k < 60; ++k) {
    if (k > 0) {
      const double odd = 2.0 * k - 1.0;
      a *= (mu - odd * odd) / (8.0 * k * x);
    }
    const double mag = fabs(a);
    if (k > 1 && mag > prev) break;  // asymptotic series started to diverge
    // sign (-1)^(k/2) for both P (k even) and Q (k odd)
    const double sgn = ((k >> 1) & 1) ? -1.0 : 1.0;
    if (k & 1) Q += sgn * a;
    else P += sgn * a;
    if (mag < 1e-17) break;
    prev = mag;
  }
  const double chi = x - (0.5 * n + 0.25) * M_PI;
  return sqrt(2.0 / (M_PI * x)) * (P * cos(chi) - Q * sin(chi));
}

// J_n(x) for n >= 0.
__host__ __device__ inline double jn(int n, double x) {
  const double ax = fabs(x);
  const double v = (ax < 14.0) ? jn_series(n, ax) : jn_asymptotic(n, ax);
  return (x < 0.0 && (n & 1)) ? -v : v;
}

}  // namespace bpmbessel
