// kg_special.cpp -- special functions for the spectral tables, evaluated once
// on the host when a spectral projector is created.
//
// The reference takes them from GSL (tools/FFT_sph.f90:22-34,290-297:
// gsl_sf_legendre_Plm, gsl_sf_bessel_jl, gsl_sf_bessel_zero_Jnu).  Here:
//  - ylm_norm_plm: N_lm P_l^m(x) with N_lm = sqrt((2l+1)/4pi (l-m)!/(l+m)!)
//    and the Condon-Shortley phase, from the recurrence of the normalised
//    functions (no factorials, no overflow for large l);
//  - sph_jl: j_l(x) by upward recurrence when x > l, Miller's downward
//    recurrence (normalised by the larger of j_0, j_1) otherwise, series for
//    small x;
//  - jl_zeros: the first nk positive zeros of j_l (= of J_{l+1/2}) for every
//    l <= lmax, each bracketed by two zeros of j_{l-1} (interlacing) and found
//    by safeguarded Newton iteration.
#include <cmath>
#include <utility>
#include <vector>

#include "kg_internal.h"

namespace kg {

double ylm_norm_plm(int l, int m, double x) {
  const double pi = std::acos(-1.0);
  // normalised P_m^m: sqrt((2m+1)/4pi (2m)!) / (2^m m!) * (-1)^m (1-x^2)^(m/2)
  double p = std::sqrt(1.0 / (4.0 * pi));
  const double s = std::sqrt((1.0 - x) * (1.0 + x));
  for (int k = 1; k <= m; ++k) p *= -std::sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
  if (l == m) return p;
  double p1 = x * std::sqrt(2.0 * m + 3.0) * p;
  if (l == m + 1) return p1;
  double p0 = p;
  for (int ll = m + 2; ll <= l; ++ll) {
    const double a = std::sqrt((4.0 * ll * ll - 1.0) / ((double)ll * ll - (double)m * m));
    const double b = std::sqrt(((ll - 1.0) * (ll - 1.0) - (double)m * m) / (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
    const double p2 = a * (x * p1 - b * p0);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

double sph_jl(int l, double x) {
  if (x == 0.0) return l == 0 ? 1.0 : 0.0;
  const double ax = std::fabs(x);
  if (ax < 1e-4) {
    double df = 1.0;
    for (int i = 3; i <= 2 * l + 1; i += 2) df *= i;
    const double x2 = x * x;
    return std::pow(x, l) / df * (1.0 - x2 / (2.0 * (2 * l + 3)));
  }
  const double s = std::sin(x), c = std::cos(x);
  const double j0 = s / x;
  if (l == 0) return j0;
  if (ax > l) {
    double a = j0, b = (s / x - c) / x;  // j_1, free of cancellation for x > 1
    if (l == 1) return b;
    for (int k = 1; k < l; ++k) {
      const double n = (2.0 * k + 1.0) / x * b - a;
      a = b;
      b = n;
    }
    return b;
  }
  const int top = l + 30 + (int)ax + (int)std::sqrt(50.0 * (l + 1));
  double up = 0.0, cur = 1e-280, want = 0.0;
  for (int k = top; k >= 1; --k) {
    const double dn = (2.0 * k + 1.0) / x * cur - up;
    up = cur;
    cur = dn;
    if (k - 1 == l) want = cur;
    if (std::fabs(cur) > 1e200) {
      cur *= 1e-200;
      up *= 1e-200;
      want *= 1e-200;
    }
  }
  // normalise by the larger of j_0, j_1 (j_0 alone fails near the zeros of sin x)
  if (ax >= 1.0) {
    const double j1 = (s / x - c) / x;
    if (std::fabs(j1) > std::fabs(j0)) return want * (j1 / up);
  }
  return want * (j0 / cur);
}

static double sph_jl_deriv(int l, double x) {
  // j_l' = j_{l-1} - (l+1)/x j_l  (j_0' = -j_1)
  if (l == 0) return -sph_jl(1, x);
  return sph_jl(l - 1, x) - (l + 1.0) / x * sph_jl(l, x);
}

// The one zero of j_l inside (a, b), where j_l changes sign: Newton steps
// kept inside the shrinking bracket, bisection when a step leaves it.
static double zero_in(int l, double a, double b) {
  double fa = sph_jl(l, a);
  double x = 0.5 * (a + b);
  for (int it = 0; it < 200; ++it) {
    const double f = sph_jl(l, x);
    if (f == 0.0) return x;
    if ((f < 0.0) == (fa < 0.0)) {
      a = x;
      fa = f;
    } else {
      b = x;
    }
    const double d = sph_jl_deriv(l, x);
    double nx = (d != 0.0) ? x - f / d : 0.5 * (a + b);
    if (!(nx > a && nx < b)) nx = 0.5 * (a + b);
    if (std::fabs(nx - x) <= 2e-16 * std::fabs(x) || b - a <= 4e-16 * std::fabs(x)) return nx;
    x = nx;
  }
  return x;
}

std::vector<double> jl_zeros(int lmax, int nk) {
  // zeros of consecutive orders interlace: z_{l-1,n} < z_{l,n} < z_{l-1,n+1};
  // j_0's zeros are n*pi, so every zero is bracketed by two of the order below
  const double pi = std::acos(-1.0);
  const int width = nk + lmax + 1;
  std::vector<double> prev(width), cur(width), out((size_t)(lmax + 1) * nk);
  for (int n = 0; n < width; ++n) prev[n] = (n + 1) * pi;
  for (int n = 0; n < nk; ++n) out[n] = prev[n];
  for (int l = 1; l <= lmax; ++l) {
    const int cnt = width - l;
    for (int n = 0; n < cnt; ++n) cur[n] = zero_in(l, prev[n], prev[n + 1]);
    for (int n = 0; n < nk; ++n) out[(size_t)l * nk + n] = cur[n];
    std::swap(prev, cur);
  }
  return out;
}

}  // namespace kg
