/*
 * kg_ic_oracle.c -- CPU FP64 restatement of the initial-data generators of
 * the reference (main/initial_conditions.f90:23-134, tools/random.f90:17-156,
 * tools/fft_lib.f90:141-272).  TEST INFRASTRUCTURE ONLY: the checker for
 * kg_amd's initial-data kernels.
 *
 * The reference draws its random phases unseeded (random.f90:44-45); here the
 * uniform numbers are an input, [k][j][i] with i fastest (the Fortran order of
 * random_value(i,j,k)).  A complex state is a pair of real arrays.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "kg_oracle.h"

#define IDX(s, i, j, k) (((int64_t)(k) * (s)->nj + (j)) * (s)->ni + (i))
#define T2(s, i, j) ((int64_t)(j) * (s)->ni + (i))

/* Uniform numbers in [0, 1) for the random phases (the reference calls an
 * unseeded random_number, random.f90:47-48): SplitMix64 (Steele, Lea and
 * Flood, OOPSLA 2014) on seed + (p+1) * 0x9E3779B97F4A7C15, top 53 bits. */
void kgi_uniforms(uint64_t seed, int64_t count, double* out) {
  for (int64_t p = 0; p < count; ++p) {
    uint64_t z = seed + (uint64_t)(p + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    out[p] = (double)(z >> 11) * 0x1.0p-53;
  }
}

/* ---------------------------------------------------------------- FFT
 * fft_1D (fft_lib.f90:141-176): recursive radix-2 decimation in time with
 * the + sign, exp(2 pi i k/N); DFT for odd lengths (:119-138). */
static void fft1(int n, const double* xr, const double* xi, int stride, double* yr, double* yi) {
  if (n % 2 != 0) {
    const double pii = acos(-1.0);
    for (int k = 0; k < n; ++k) {
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < n; ++j) {
        const double th = 2.0 * pii * (double)(j * k) / (double)n;
        const double c = cos(th), s = sin(th);
        const double ar = xr[(int64_t)j * stride], ai = xi[(int64_t)j * stride];
        sr = sr + (ar * c - ai * s);
        si = si + (ar * s + ai * c);
      }
      yr[k] = sr;
      yi[k] = si;
    }
    return;
  }
  const int h = n / 2;
  double* er = (double*)malloc(sizeof(double) * (size_t)h * 4);
  double *ei = er + h, *odr = er + 2 * h, *odi = er + 3 * h;
  fft1(h, xr, xi, 2 * stride, er, ei);
  fft1(h, xr + stride, xi + stride, 2 * stride, odr, odi);
  const double pii = acos(-1.0);
  for (int k = 0; k < h; ++k) {
    const double th = 2.0 * pii * (double)k / (double)n;
    const double c = cos(th), s = sin(th);
    const double tr = c * odr[k] - s * odi[k], ti = c * odi[k] + s * odr[k];
    yr[k] = er[k] + tr;
    yi[k] = ei[k] + ti;
    yr[k + h] = er[k] - tr;
    yi[k + h] = ei[k] - ti;
  }
  free(er);
}

/* fft_3D (fft_lib.f90:196-226) along i, then j, then k, on [k][j][i] arrays */
static void fft3(int n, double* re, double* im) {
  double* tr = (double*)malloc(sizeof(double) * (size_t)n * 4);
  double *ti = tr + n, *yr = tr + 2 * n, *yi = tr + 3 * n;
  const int64_t n2 = (int64_t)n * n;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j) { /* columns: fft over i */
      double* br = re + k * n2 + (int64_t)j * n;
      double* bi = im + k * n2 + (int64_t)j * n;
      fft1(n, br, bi, 1, yr, yi);
      memcpy(br, yr, sizeof(double) * n);
      memcpy(bi, yi, sizeof(double) * n);
    }
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < n; ++i) { /* rows: fft over j */
      for (int j = 0; j < n; ++j) {
        tr[j] = re[k * n2 + (int64_t)j * n + i];
        ti[j] = im[k * n2 + (int64_t)j * n + i];
      }
      fft1(n, tr, ti, 1, yr, yi);
      for (int j = 0; j < n; ++j) {
        re[k * n2 + (int64_t)j * n + i] = yr[j];
        im[k * n2 + (int64_t)j * n + i] = yi[j];
      }
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) { /* fft over k */
      for (int k = 0; k < n; ++k) {
        tr[k] = re[k * n2 + (int64_t)j * n + i];
        ti[k] = im[k * n2 + (int64_t)j * n + i];
      }
      fft1(n, tr, ti, 1, yr, yi);
      for (int k = 0; k < n; ++k) {
        re[k * n2 + (int64_t)j * n + i] = yr[k];
        im[k * n2 + (int64_t)j * n + i] = yi[k];
      }
    }
  free(tr);
}

/* function_cartesian (random.f90:17-59): psi_x = ifft_3D(psi_p),
 * psi_p = exp(-0.5/p0^2 |omega|^2) exp(2 pi i U), ifft_3D(f) =
 * conj(fft_3D(conj f))/N^3 (fft_lib.f90:252-264); omega in the shifted order
 * (shift_3D swaps the halves of each axis). */
void kgi_random_field(int n, double rmax, double p0, const double* U, double* re, double* im) {
  const double pii = acos(-1.0);
  const double R = rmax + 1.0;
  const double dx = (R - (-R)) / (double)n;
  const int64_t n3 = (int64_t)n * n * n;
  double* om = (double*)malloc(sizeof(double) * (size_t)n);
  for (int p = 0; p < n; ++p) {
    const int q = (p + n / 2) % n; /* shifted position p holds unshifted index q */
    om[p] = 2.0 * pii * (double)(q - n / 2) / ((double)n * dx);
  }
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const int64_t p = ((int64_t)k * n + j) * n + i;
        const double w2 = om[i] * om[i] + om[j] * om[j] + om[k] * om[k];
        const double amp = exp(-0.5 / (p0 * p0) * w2);
        const double th = 2.0 * pii * U[p];
        /* conj(psi_p) */
        re[p] = amp * cos(th);
        im[p] = -(amp * sin(th));
      }
  fft3(n, re, im);
  const double nn = (double)n * (double)n * (double)n;
  for (int64_t p = 0; p < n3; ++p) {
    re[p] = re[p] / nn;
    im[p] = -im[p] / nn;
  }
  free(om);
}

/* convert_to_spherical (random.f90:128-156) with trilinear_interpolation
 * (:72-126): u(1) at every padded point from the Cartesian field.  Outside
 * the box the reference's search runs past its array; here the cell is
 * clamped (the weights are clamped to [0, 1] as in :96-99). */
void kgi_sample(const kgo_sim* s, int n, double rmax, const double* fre, const double* fim, double* pre,
                double* pim) {
  const double R = rmax + 1.0;
  const double xmin = -R, dx = (R - (-R)) / (double)n;
  const int64_t n2 = (int64_t)n * n;
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj)
      for (int ii = 0; ii < s->ni; ++ii) {
        const double r = s->x[ii], th = s->y[jj], ph = s->z[kk];
        const double xc[3] = {r * sin(th) * cos(ph), r * sin(th) * sin(ph), r * cos(th)};
        int c[3];
        double w[3];
        for (int a = 0; a < 3; ++a) {
          int i = 0;
          for (i = 0; i < n - 1; ++i)
            if (xmin + (double)i * dx <= xc[a] && xmin + (double)(i + 1) * dx > xc[a]) break;
          if (i > n - 2) i = n - 2;
          const double x0 = xmin + (double)i * dx, x1 = xmin + (double)(i + 1) * dx;
          double d = (xc[a] - x0) / (x1 - x0);
          d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
          c[a] = i;
          w[a] = d;
        }
        const double xd = w[0], yd = w[1], zd = w[2];
#define F(arr, a, b, cc) arr[(int64_t)(cc) * n2 + (int64_t)(b) * n + (a)]
        double out[2];
        for (int part = 0; part < 2; ++part) {
          const double* f = part ? fim : fre;
          const int i = c[0], j = c[1], k = c[2];
          double v = F(f, i, j, k) * (1 - xd) * (1 - yd) * (1 - zd);
          v = v + F(f, i + 1, j, k) * xd * (1 - yd) * (1 - zd);
          v = v + F(f, i, j + 1, k) * (1 - xd) * yd * (1 - zd);
          v = v + F(f, i, j, k + 1) * (1 - xd) * (1 - yd) * zd;
          v = v + F(f, i + 1, j + 1, k) * xd * yd * (1 - zd);
          v = v + F(f, i + 1, j, k + 1) * xd * (1 - yd) * zd;
          v = v + F(f, i, j + 1, k + 1) * (1 - xd) * yd * zd;
          v = v + F(f, i + 1, j + 1, k + 1) * xd * yd * zd;
          out[part] = v;
        }
#undef F
        pre[IDX(s, ii, jj, kk)] = out[0];
        pim[IDX(s, ii, jj, kk)] = out[1];
      }
}

/* fills of one phase on the 5 fields of a real array: 0 = boundary_xL then
 * xR, 1 = yL then yR, 2 = zL then zR (lines_method_lib.f90:507-597) */
static void fill_phase(kgo_sim* s, double* u, int phase) {
  const int ni = s->ni, nj = s->nj, g = s->g, Ny = s->Ny, Nz = s->Nz;
  const int dis = Nz / 2;
  for (int c = 0; c < 5; ++c) {
    double* f = u + (int64_t)c * s->npts;
    if (phase == 0) {
      for (int kk = 0; kk < s->nk; ++kk)
        for (int jj = 0; jj < nj; ++jj) {
          double* a = f + IDX(s, 0, jj, kk);
          a[0] = 3.0 * a[1] - 3.0 * a[2] + a[3];
          /* inflow copies u_p, which the reference has not written yet at
           * this point: taken as zero */
          if (s->bc_rmax != KGO_BC_SOMMERFELD)  /* Sommerfeld keeps the outer column's own value */
            a[ni - 1] = s->bc_rmax == KGO_BC_OUTFLOW ? a[ni - 2] : 0.0;
        }
    } else if (phase == 1) {
      for (int k = 0; k <= Nz; ++k) {
        const int kt = (k + dis) % (Nz - 1);
        for (int gg = g; gg >= 0; --gg)
          memcpy(f + IDX(s, 0, -gg + g, k), f + IDX(s, 0, gg + 1 + g, kt), sizeof(double) * (size_t)ni);
      }
      for (int k = 0; k <= Nz; ++k) {
        const int kt = (k + dis) % (Nz - 1);
        for (int gg = g; gg >= 0; --gg)
          memcpy(f + IDX(s, 0, Ny + gg + g, k), f + IDX(s, 0, Ny - gg - 1 + g, kt), sizeof(double) * (size_t)ni);
      }
    } else {
      const int64_t pl = (int64_t)ni * nj;
      memcpy(f, f + (int64_t)(Nz - 1) * pl, sizeof(double) * (size_t)pl);
      memcpy(f + (int64_t)Nz * pl, f + pl, sizeof(double) * (size_t)pl);
    }
  }
}

double kgd_integrate(const kgo_sim* s, const double* f);
void kgd_density(const kgo_sim* s, const double* ure, const double* uim, double* rho);

/* initial_conditions (initial_conditions.f90:23-134), non-MKS branches.
 * type 0 'spherical_shell', 1 'random', 2 'random_shell'.  ure/uim: the five
 * fields (uim NULL for the real spherical shell).  For the random types fre /
 * fim hold psi_x of kgi_random_field.  Returns Amp. */
double kgi_initial(kgo_sim* s, int type, double mu, double sigma, int n, double rmax, const double* fre,
                   const double* fim, double* ure, double* uim) {
  const int64_t np = s->npts;
  double* tmp = (double*)malloc(sizeof(double) * (size_t)np);
  memset(ure, 0, sizeof(double) * (size_t)np * 5);
  if (uim) memset(uim, 0, sizeof(double) * (size_t)np * 5);
  double amp;
  if (type == 0) {
    /* u1 = 1/x exp(-(x-mu)^2/sigma^2); u2 stays 0 (its line is commented
     * out); u3 = dx2(u1), u4 = dx3(u1); dtphi = -2(x-mu)/(x sigma^2) exp(..);
     * u5 = 1/alpha (dtphi - beta u2) */
    for (int kk = 0; kk < s->nk; ++kk)
      for (int jj = 0; jj < s->nj; ++jj)
        for (int ii = 0; ii < s->ni; ++ii) {
          const double x = s->x[ii];
          ure[IDX(s, ii, jj, kk)] = 1.0 / x * exp(-((x - mu) * (x - mu)) / (sigma * sigma));
        }
    kgo_dx2(s, ure, ure + 2 * np);
    kgo_dx3(s, ure, ure + 3 * np);
    for (int kk = 0; kk < s->nk; ++kk)
      for (int jj = 0; jj < s->nj; ++jj)
        for (int ii = 0; ii < s->ni; ++ii) {
          const int64_t p = IDX(s, ii, jj, kk), t = T2(s, ii, jj);
          const double x = s->x[ii];
          const double dtphi = -2.0 * (x - mu) / (x * (sigma * sigma)) * exp(-((x - mu) * (x - mu)) / (sigma * sigma));
          ure[4 * np + p] = 1.0 / s->alpha[t] * (dtphi - s->beta_r[t] * ure[np + p]);
        }
    kgd_density(s, ure, NULL, tmp);
    for (int kk = 0; kk < s->nk; ++kk)
      for (int jj = 0; jj < s->nj; ++jj)
        for (int ii = 0; ii < s->ni; ++ii) {
          const int64_t p = IDX(s, ii, jj, kk), t = T2(s, ii, jj);
          tmp[p] = tmp[p] * s->alpha[t] * s->sqrtg[t];
        }
    amp = 1.0 / sqrt(kgd_integrate(s, tmp));
  } else {
    kgi_sample(s, n, rmax, fre, fim, ure, uim);
    if (type == 1) { /* boundary_xL, boundary_xR before the derivatives */
      fill_phase(s, ure, 0);
      fill_phase(s, uim, 0);
    } else { /* u1 = u1 * 1/x * exp(-(x-mu)^2/sigma^2) */
      for (int kk = 0; kk < s->nk; ++kk)
        for (int jj = 0; jj < s->nj; ++jj)
          for (int ii = 0; ii < s->ni; ++ii) {
            const int64_t p = IDX(s, ii, jj, kk);
            const double x = s->x[ii], e = exp(-((x - mu) * (x - mu)) / (sigma * sigma));
            ure[p] = ure[p] * 1.0 / x * e;
            uim[p] = uim[p] * 1.0 / x * e;
          }
    }
    for (int part = 0; part < 2; ++part) {
      double* u = part ? uim : ure;
      kgo_dx1(s, u, u + np);
      kgo_dx2(s, u, u + 2 * np);
      kgo_dx3(s, u, u + 3 * np);
      for (int kk = 0; kk < s->nk; ++kk)
        for (int jj = 0; jj < s->nj; ++jj)
          for (int ii = 0; ii < s->ni; ++ii) {
            const int64_t p = IDX(s, ii, jj, kk), t = T2(s, ii, jj);
            const double dtphi = type == 1 ? 0.0 : u[np + p];
            u[4 * np + p] = 1.0 / s->alpha[t] * (dtphi - s->beta_r[t] * u[np + p]);
          }
    }
    if (type == 2) { /* xL, xR, zL, zR, yL, yR */
      for (int part = 0; part < 2; ++part) {
        double* u = part ? uim : ure;
        fill_phase(s, u, 0);
        fill_phase(s, u, 2);
        fill_phase(s, u, 1);
      }
    }
    /* Amp = 1/sqrt(integrate(m |u1|^2 sqrtg)) */
    for (int kk = 0; kk < s->nk; ++kk)
      for (int jj = 0; jj < s->nj; ++jj)
        for (int ii = 0; ii < s->ni; ++ii) {
          const int64_t p = IDX(s, ii, jj, kk), t = T2(s, ii, jj);
          const double a = hypot(ure[p], uim[p]);
          tmp[p] = s->m * (a * a) * s->sqrtg[t];
        }
    amp = 1.0 / sqrt(kgd_integrate(s, tmp));
  }
  for (int64_t p = 0; p < 5 * np; ++p) {
    ure[p] = amp * ure[p];
    if (uim) uim[p] = amp * uim[p];
  }
  free(tmp);
  return amp;
}
