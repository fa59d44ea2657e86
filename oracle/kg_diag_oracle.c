/*
 * kg_diag_oracle.c -- CPU FP64 restatement of the diagnostics that follow the
 * step in the reference loop (MoL/diagnostic_lib.f90:185-274,
 * tools/integral.f90).  TEST INFRASTRUCTURE ONLY: the checker for kg_amd's
 * diagnostic kernels.
 *
 * Arrays are the oracle's padded layout [kk][jj][ii]; a complex state is
 * passed as separate real and imaginary arrays (im may be NULL for a real
 * state).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "kg_oracle.h"

#define IDX(s, i, j, k) (((int64_t)(k) * (s)->nj + (j)) * (s)->ni + (i))
#define T2(s, i, j) ((int64_t)(j) * (s)->ni + (i))

static double cabs_(double re, double im) { return im == 0.0 ? fabs(re) : hypot(re, im); }

/* density (diagnostic_lib.f90:258-271):
 *   rho = 0.5*( |Pi| + g11 |psi_r|**2 + 2 g13 |psi_r| psi_phi + g22 |psi_th|**2
 *             + g33 |psi_phi|**2 + m**2 |phi| )
 * evaluated left to right; the sum is complex when psi_phi is, and rho (a
 * real array) keeps its real part. */
void kgd_density(const kgo_sim* s, const double* ure, const double* uim, double* rho) {
  const int64_t n = s->npts;
  const double m2 = s->m * s->m;
  for (int k = 0; k < s->nk; ++k)
    for (int j = 0; j < s->nj; ++j)
      for (int i = 0; i < s->ni; ++i) {
        const int64_t p = IDX(s, i, j, k), t = T2(s, i, j);
        double re[5], im[5];
        for (int f = 0; f < 5; ++f) {
          re[f] = ure[f * n + p];
          im[f] = uim ? uim[f * n + p] : 0.0;
        }
        const double a2 = cabs_(re[1], im[1]), a3 = cabs_(re[2], im[2]), a4 = cabs_(re[3], im[3]);
        double sum = cabs_(re[4], im[4]);
        sum = sum + s->g_rr[t] * (a2 * a2);
        sum = sum + ((2.0 * s->g_rp[t]) * a2) * re[3];
        sum = sum + s->g_tt[t] * (a3 * a3);
        sum = sum + s->g_pp[t] * (a4 * a4);
        sum = sum + m2 * cabs_(re[0], im[0]);
        rho[p] = 0.5 * sum;
      }
}

/* currents_ADM (diagnostic_lib.f90:214-224):
 *   j1 = -Pi (g11 psi_r + g13 psi_phi),  j2 = -Pi g22 psi_th,
 *   j3 = -Pi (g13 psi_r + g33 psi_phi)
 * complex products (a+ib)(c+id) = (ac - bd) + i(ad + bc). */
void kgd_currents(const kgo_sim* s, const double* ure, const double* uim, double* jre, double* jim) {
  const int64_t n = s->npts;
  for (int k = 0; k < s->nk; ++k)
    for (int j = 0; j < s->nj; ++j)
      for (int i = 0; i < s->ni; ++i) {
        const int64_t p = IDX(s, i, j, k), t = T2(s, i, j);
        const double pr = ure[4 * n + p], pi = uim ? uim[4 * n + p] : 0.0;
        const double rr = ure[1 * n + p], ri = uim ? uim[1 * n + p] : 0.0;
        const double tr = ure[2 * n + p], ti = uim ? uim[2 * n + p] : 0.0;
        const double fr = ure[3 * n + p], fi = uim ? uim[3 * n + p] : 0.0;
        const double g11 = s->g_rr[t], g13 = s->g_rp[t], g22 = s->g_tt[t], g33 = s->g_pp[t];
        /* c1 = g11*psi_r + g13*psi_phi */
        const double c1r = g11 * rr + g13 * fr, c1i = g11 * ri + g13 * fi;
        const double c3r = g13 * rr + g33 * fr, c3i = g13 * ri + g33 * fi;
        /* -(Pi*c) */
        jre[0 * n + p] = -(pr * c1r - pi * c1i);
        jre[2 * n + p] = -(pr * c3r - pi * c3i);
        /* j2 = -((Pi*g22)*psi_th) */
        const double q2r = pr * g22, q2i = pi * g22;
        jre[1 * n + p] = -(q2r * tr - q2i * ti);
        if (jim) {
          jim[0 * n + p] = -(pr * c1i + pi * c1r);
          jim[2 * n + p] = -(pr * c3i + pi * c3r);
          jim[1 * n + p] = -(q2r * ti + q2i * tr);
        }
      }
}

/* integrate_real (integral.f90:646-669): trapezium_real (:220-255) over
 * ii in [g, ni-1-g], jj in [g, nj-1-g], all kk, nested pair sums. */
double kgd_integrate(const kgo_sim* s, const double* f) {
  const int g = s->g;
  const int il = g, ir = s->ni - 1 - g, jl = g, jr = s->nj - 1 - g, kl = 0, kr = s->nk - 1;
  const int mi = ir - il + 1, mj = jr - jl + 1;
  double* f2d = (double*)calloc((size_t)mi * mj, sizeof(double));
  double* f1d = (double*)calloc((size_t)mi, sizeof(double));
  for (int k = kl; k < kr; ++k)
    for (int j = 0; j < mj; ++j)
      for (int i = 0; i < mi; ++i) {
        double* a = &f2d[(int64_t)j * mi + i];
        *a = *a + f[IDX(s, il + i, jl + j, k)] + f[IDX(s, il + i, jl + j, k + 1)];
      }
  for (int q = 0; q < mi * mj; ++q) f2d[q] = 0.50 * f2d[q] * s->dz;
  for (int j = 0; j < mj - 1; ++j)
    for (int i = 0; i < mi; ++i) f1d[i] = f1d[i] + f2d[(int64_t)j * mi + i] + f2d[(int64_t)(j + 1) * mi + i];
  for (int i = 0; i < mi; ++i) f1d[i] = 0.50 * f1d[i] * s->dy;
  double f0d = 0.0;
  for (int i = 0; i < mi - 1; ++i) f0d = f0d + f1d[i] + f1d[i + 1];
  f0d = 0.50 * f0d * s->dx;
  free(f2d);
  free(f1d);
  return f0d;
}

/* trapezium_2D_complex (integral.f90:330-362) of f[i][j] (i: theta, j: phi) */
void kgd_trapezium_2d_complex(int m, int n, double dx, double dy, const double* fre, const double* fim,
                              double* ore, double* oim) {
  double sr = 0.0, si = 0.0;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      const int c = (i == 0 || i == m - 1) + 2 * (j == 0 || j == n - 1);
      const double a = fre[(int64_t)i * n + j], b = fim ? fim[(int64_t)i * n + j] : 0.0;
      if (c == 0) {
        sr += a;
        si += b;
      } else if (c == 3) {
        sr += 0.25 * a;
        si += 0.25 * b;
      } else {
        sr += 0.5 * a;
        si += 0.5 * b;
      }
    }
  *ore = dx * dy * sr;
  *oim = dx * dy * si;
}
