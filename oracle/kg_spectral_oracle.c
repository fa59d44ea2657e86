/*
 * kg_spectral_oracle.c -- CPU FP64 restatement of the spectral projection of
 * FlavioRosales/KG (tools/FFT_sph.f90, tools/integral.f90).  TEST
 * INFRASTRUCTURE ONLY: the checker for kg_amd's spectral kernels.
 *
 * The reference evaluates the special functions with GSL (FFT_sph.f90:22-34,
 * 290-297: gsl_sf_legendre_Plm, gsl_sf_bessel_jl, gsl_sf_bessel_zero_Jnu;
 * linked as -lgsl from main/Makefile:8,27, no version pinned).  GSL is not
 * available here; this file restates the published algorithms:
 *   P_l^m(x)  -- upward recurrence in l from P_m^m = (-1)^m (2m-1)!! (1-x^2)^(m/2)
 *                (Condon-Shortley phase included, as GSL's Plm);
 *   j_l(x)    -- upward recurrence from j_0, j_1 for x > l, otherwise Miller's
 *                downward recurrence normalised by the larger of j_0, j_1
 *                (series for tiny x);
 *   zeros of J_{l+1/2} = zeros of j_l -- bracketed on a grid and bisected to
 *                full precision.
 * Parity of these three is pinned against closed forms and scipy in tests/.
 *
 * Layouts (C order): Y_lm [lm][theta j][phi k], lm = l*l + (m + l);
 * spectral field f[k][j][i] (r fastest, as the reference's f(Nx,Ny,Nz));
 * flm [i][lm]; Bessel basis jb [l][n][i]; radial spectrum ft [n][lm].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
double kgs_plm(int l, int m, double x) {
  if (m < 0 || m > l || fabs(x) > 1.0) return 0.0;
  double pmm = 1.0;
  if (m > 0) {
    const double s = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; ++i) {
      pmm *= -fact * s;
      fact += 2.0;
    }
  }
  if (l == m) return pmm;
  double pmmp1 = x * (2.0 * m + 1.0) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.0;
  for (int ll = m + 2; ll <= l; ++ll) {
    pll = ((2.0 * ll - 1.0) * x * pmmp1 - (ll + m - 1.0) * pmm) / (double)(ll - m);
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

double kgs_sph_jl(int l, double x) {
  if (x == 0.0) return l == 0 ? 1.0 : 0.0;
  const double ax = fabs(x);
  if (ax < 1e-3 * (l + 1)) {
    /* series x^l/(2l+1)!! (1 - x^2/(2(2l+3)) + x^4/(8(2l+3)(2l+5))) */
    double df = 1.0;
    for (int i = 1; i <= l; ++i) df *= (2.0 * i + 1.0);
    const double x2 = x * x;
    return pow(x, l) / df * (1.0 - x2 / (2.0 * (2 * l + 3)) + x2 * x2 / (8.0 * (2 * l + 3) * (2 * l + 5)));
  }
  const double j0 = sin(x) / x;
  if (l == 0) return j0;
  if (ax > l) {
    /* upward recurrence is stable above x = l; j_1 has no cancellation there */
    const double j1 = sin(x) / (x * x) - cos(x) / x;
    if (l == 1) return j1;
    double jm = j0, jc = j1;
    for (int k = 1; k < l; ++k) {
      const double jp = (2.0 * k + 1.0) / x * jc - jm;
      jm = jc;
      jc = jp;
    }
    return jc;
  }
  /* Miller: start well above l and recur downward */
  int start = l + 20 + (int)(ax) + (int)sqrt(40.0 * (l + 1));
  double jp = 0.0, jc = 1e-300, out = 0.0;
  for (int k = start; k >= 1; --k) {
    const double jm = (2.0 * k + 1.0) / x * jc - jp;
    jp = jc;
    jc = jm;
    if (k - 1 == l) out = jc;
    if (fabs(jc) > 1e250) { /* rescale */
      jc *= 1e-250;
      jp *= 1e-250;
      out *= 1e-250;
    }
  }
  /* normalise by whichever of j_0, j_1 is larger: near a zero of sin x the
   * ratio j_0/jc alone loses every digit */
  if (ax >= 1.0) {
    const double j1 = sin(x) / (x * x) - cos(x) / x;
    if (fabs(j1) > fabs(j0)) return out * (j1 / jp);
  }
  return out * (j0 / jc);
}

double kgs_jl_zero(int l, int n) {
  /* n-th positive zero of j_l: scan with step 0.05 from just above 0 */
  const double h = 0.05;
  double a = (l == 0) ? 1e-3 : (double)l * 0.5 + 1e-3;
  double fa = kgs_sph_jl(l, a);
  int found = 0;
  for (int it = 0; it < 10000000; ++it) {
    const double b = a + h;
    const double fb = kgs_sph_jl(l, b);
    if ((fa < 0.0 && fb >= 0.0) || (fa > 0.0 && fb <= 0.0)) {
      if (++found == n) {
        double lo = a, hi = b, flo = fa;
        for (int k = 0; k < 200; ++k) {
          const double mid = 0.5 * (lo + hi);
          if (mid <= lo || mid >= hi) break;
          const double fm = kgs_sph_jl(l, mid);
          if ((flo < 0.0) == (fm < 0.0)) {
            lo = mid;
            flo = fm;
          } else {
            hi = mid;
          }
        }
        return 0.5 * (lo + hi);
      }
    }
    a = b;
    fa = fb;
  }
  return NAN;
}

/* ------------------------------------------------------------------------
 * sph_harmonics (FFT_sph.f90:282-336): Y_lm = norm * Plm * e^{i m phi};
 * |cos theta| > 0.9999 uses Plm := fact for m = 0 and 0 otherwise;
 * Y_{l,-m} = (-1)^m conj(Y_lm).
 * ------------------------------------------------------------------------ */
void kgs_sph_harmonics(int lmax, int nth, int nph, const double* theta, const double* phi, double* yre,
                       double* yim) {
  const double pi = acos(-1.0);
  const int64_t sl = (int64_t)nth * nph;
  memset(yre, 0, sizeof(double) * (size_t)((lmax + 1) * (lmax + 1) * sl));
  memset(yim, 0, sizeof(double) * (size_t)((lmax + 1) * (lmax + 1) * sl));
  for (int l = 0; l <= lmax; ++l) {
    const double fact = sqrt((2.0 * l + 1.0) / (4.0 * pi));
    for (int m = 0; m <= l; ++m) {
      double norm = fact;
      if (m > 0)
        for (int i = 1; i <= 2 * m; ++i) norm = norm / sqrt((double)(l - m + i));
      for (int i = 0; i < nth; ++i) {
        const double x = cos(theta[i]);
        double plm;
        if (fabs(x) > 0.9999)
          plm = (m == 0) ? fact : 0.0;
        else
          plm = kgs_plm(l, m, x);
        for (int j = 0; j < nph; ++j) {
          const double pr = cos(m * phi[j]), pim = sin(m * phi[j]);
          const double a = norm * plm;
          const double re = a * pr, im = a * pim;
          const int64_t p = (int64_t)(l * l + m + l) * sl + (int64_t)i * nph + j;
          yre[p] = re;
          yim[p] = im;
          if (m > 0) {
            const double sgn = (m % 2) ? -1.0 : 1.0;
            const int64_t q = (int64_t)(l * l - m + l) * sl + (int64_t)i * nph + j;
            yre[q] = sgn * re;
            yim[q] = sgn * (-im);
          }
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------
 * project_sh (FFT_sph.f90:247-279) with trapezium_2D_complex
 * (integral.f90:330-362) on one radius: f[j][k] complex (re, im arrays).
 * ------------------------------------------------------------------------ */
void kgs_project_sh(int nth, int nph, double dth, double dph, const double* theta, const double* fre,
                    const double* fim, int lmax, const double* yre, const double* yim, double* ore, double* oim) {
  const int64_t sl = (int64_t)nth * nph;
  for (int l = 0; l <= lmax; ++l)
    for (int m = -l; m <= l; ++m) {
      const int lm = l * l + m + l;
      const double* yr = yre + (int64_t)lm * sl;
      const double* yi = yim + (int64_t)lm * sl;
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < nth; ++i) {
        const double st = sin(theta[i]);
        for (int j = 0; j < nph; ++j) {
          /* fun = f * conj(Y) * sin(theta) */
          const double a = fre[i * nph + j], b = fim[i * nph + j];
          const double c = yr[i * nph + j], d = -yi[i * nph + j];
          double fr = a * c - b * d, fi = a * d + b * c;
          fr = fr * st;
          fi = fi * st;
          const int edge = (i == 0 || i == nth - 1) + 2 * (j == 0 || j == nph - 1);
          if (edge == 0) {
            sr += fr;
            si += fi;
          } else if (edge == 3) {
            sr += 0.25 * fr;
            si += 0.25 * fi;
          } else {
            sr += 0.5 * fr;
            si += 0.5 * fi;
          }
        }
      }
      ore[lm] = dth * dph * sr;
      oim[lm] = dth * dph * si;
    }
}

/* trapezium_1D_complex (integral.f90:190-212) */
void kgs_trapezium_1d_complex(int n, double dx, const double* fre, const double* fim, double* ore, double* oim) {
  double sr = 0.5 * (fre[0] + fre[n - 1]), si = 0.5 * (fim[0] + fim[n - 1]);
  for (int i = 1; i < n - 1; ++i) {
    sr += fre[i];
    si += fim[i];
  }
  *ore = dx * sr;
  *oim = dx * si;
}

/* ------------------------------------------------------------------------
 * Bessel basis (FFT_sph.f90:13-95).  kind 0: k_{l,n} = z_n(J_{l+1/2})/R_max,
 * j_l(k r_i) normalised by sqrt(trapz(j_l^2 r^2, dr)); kind 1: k_n =
 * kmax*n/nk for every l, not normalised.  r[0..nr) uniform, dr = |r1-r0|.
 * ------------------------------------------------------------------------ */
void kgs_bessel_basis(int lmax, int nk, int nr, const double* r, double R_max, int kind, double kmax,
                      double* kv, double* jb) {
  const double dr = fabs(r[1] - r[0]);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)nr);
  for (int l = 0; l <= lmax; ++l)
    for (int n = 1; n <= nk; ++n) {
      const double k = kind == 0 ? kgs_jl_zero(l, n) / R_max : kmax * (double)n / (double)nk;
      kv[l * nk + n - 1] = k;
      double* row = jb + ((int64_t)l * nk + (n - 1)) * nr;
      for (int i = 0; i < nr; ++i) row[i] = kgs_sph_jl(l, k * r[i]);
      if (kind == 0) {
        for (int i = 0; i < nr; ++i) tmp[i] = row[i] * row[i] * (r[i] * r[i]);
        /* trapezium_1D_real (integral.f90:176-188) */
        double s = 0.0;
        for (int i = 1; i < nr - 1; ++i) s += tmp[i];
        double integ = dr * (0.5 * tmp[0] + s + 0.5 * tmp[nr - 1]);
        integ = sqrt(integ);
        if (integ > 0.0)
          for (int i = 0; i < nr; ++i) row[i] = row[i] / integ;
      }
    }
  free(tmp);
}

/* radial transform: ft(n,l,m) = trapz_i(flm(r_i,l,m) * (jb(l,n,i) * r_i^2), dr)
 * (the contraction SphFT_discrete_radial defines, FFT_sph.f90:231-234) */
void kgs_radial_project(int lmax, int nk, int nr, const double* r, double dr, const double* jb, const double* fre,
                        const double* fim, double* ore, double* oim) {
  const int nlm = (lmax + 1) * (lmax + 1);
  double* tr = (double*)malloc(sizeof(double) * (size_t)nr);
  double* ti = (double*)malloc(sizeof(double) * (size_t)nr);
  for (int l = 0; l <= lmax; ++l)
    for (int m = -l; m <= l; ++m) {
      const int lm = l * l + m + l;
      for (int n = 0; n < nk; ++n) {
        const double* row = jb + ((int64_t)l * nk + n) * nr;
        for (int i = 0; i < nr; ++i) {
          const double w = row[i] * (r[i] * r[i]);
          tr[i] = fre[(int64_t)i * nlm + lm] * w;
          ti[i] = fim[(int64_t)i * nlm + lm] * w;
        }
        kgs_trapezium_1d_complex(nr, dr, tr, ti, &ore[(int64_t)n * nlm + lm], &oim[(int64_t)n * nlm + lm]);
      }
    }
  free(tr);
  free(ti);
}

/* ------------------------------------------------------------------------
 * domain_to_sph (FFT_sph.f90:525-651): trilinear interpolation of a padded
 * field fd[kk][jj][ii] (ii = i+g, jj = j+g, kk = k) onto the spectral grid
 * r_i = xmin + i dx (i < Nx), theta_j = j dy (j < Ny), phi_k = k dz (k < Nz).
 * ------------------------------------------------------------------------ */
void kgs_domain_to_sph(int Nx, int Ny, int Nz, int g, double xmin, double dx, double dy, double dz, const double* fd,
                       double* f) {
  const int ni = Nx + 2 * g + 1, nj = Ny + 2 * g + 1;
  const int ilo = -g, ihi = Nx + g, jlo = -g, jhi = Ny + g, klo = 0, khi = Nz;
#define XD(i) (xmin + (double)((i) + g) * dx)
#define YD(j) (-dy / 2.0 + (double)(j)*dy)
#define ZD(k) ((double)(k)*dz)
#define FD(i, j, k) fd[((int64_t)(k)*nj + ((j) + g)) * ni + ((i) + g)]
  const double rmin = XD(ilo), dr = XD(ilo + 1) - XD(ilo);
  const double thetamin = YD(jlo), dth = fabs(YD(jlo + 1) - YD(jlo));
  const double phimin = ZD(klo), dph = fabs(ZD(klo + 1) - ZD(klo));
  for (int k = 0; k < Nz; ++k)
    for (int j = 0; j < Ny; ++j)
      for (int i = 0; i < Nx; ++i) {
        const double rr = xmin + (double)i * dx, tt = (double)j * dy, pp = (double)k * dz;
        int ir = (int)((rr - rmin) / dr) + ilo;
        int ith = (int)((tt - thetamin) / dth) + jlo;
        int iph = (int)((pp - phimin) / dph) + klo;
        if (ir < ilo) ir = ilo;
        if (ir > ihi - 1) ir = ihi - 1;
        if (ith < jlo) ith = jlo;
        if (ith > jhi - 1) ith = jhi - 1;
        if (iph < klo) iph = klo;
        if (iph > khi - 1) iph = khi - 1;
        while (ir < ihi - 1 && rr >= XD(ir + 1)) ++ir;
        while (ir > ilo && rr < XD(ir)) --ir;
        while (ith < jhi - 1 && tt >= YD(ith + 1)) ++ith;
        while (ith > jlo && tt < YD(ith)) --ith;
        while (iph < khi - 1 && pp >= ZD(iph + 1)) ++iph;
        while (iph > klo && pp < ZD(iph)) --iph;
        double* out = &f[((int64_t)k * Ny + j) * Nx + i];
        if (tt <= thetamin) {
          *out = FD(ir, jlo, iph);
          continue;
        } else if (tt >= thetamin + dth * (double)(jhi - jlo)) {
          *out = FD(ir, jhi, iph);
          continue;
        }
        const int iph1 = iph + 1;
        if (iph1 > khi) continue;
        const double f000 = FD(ir, ith, iph), f001 = FD(ir, ith, iph1);
        const double f010 = FD(ir, ith + 1, iph), f011 = FD(ir, ith + 1, iph1);
        const double f100 = FD(ir + 1, ith, iph), f101 = FD(ir + 1, ith, iph1);
        const double f110 = FD(ir + 1, ith + 1, iph), f111 = FD(ir + 1, ith + 1, iph1);
        const double wr = (rr - XD(ir)) / dr, wth = (tt - YD(ith)) / dth, wph = (pp - ZD(iph)) / dph;
        const double f00 = (1.0 - wph) * f000 + wph * f001;
        const double f01 = (1.0 - wph) * f010 + wph * f011;
        const double f10 = (1.0 - wph) * f100 + wph * f101;
        const double f11 = (1.0 - wph) * f110 + wph * f111;
        const double f0 = (1.0 - wth) * f00 + wth * f01;
        const double f1 = (1.0 - wth) * f10 + wth * f11;
        *out = (1.0 - wr) * f0 + wr * f1;
      }
#undef XD
#undef YD
#undef ZD
#undef FD
}
