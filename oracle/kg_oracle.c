/*
 * kg_oracle.c -- CPU FP64 restatement of the Klein-Gordon method-of-lines
 * step of FlavioRosales/KG.  TEST INFRASTRUCTURE ONLY (see kg_oracle.h).
 *
 * Every routine mirrors one Fortran routine of the reference and cites it.
 * Whole-array Fortran statements become loops over the padded array with the
 * same temporaries and the same left-to-right association, so that the
 * rounding of each intermediate matches the reference (real part; the
 * reference's complex/real division goes through __divdc3 and may differ from
 * a real division in the last bit).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math).
 */
#include "kg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(s, ii, jj, kk) (((int64_t)(kk) * (s)->nj + (jj)) * (s)->ni + (ii))
#define IDX2(s, ii, jj) ((int64_t)(jj) * (s)->ni + (ii))

/* Loops over independent points run on every host thread (OpenMP); each
 * point's arithmetic is the same on any thread, so results do not depend on
 * the thread count. */
#ifdef _OPENMP
#include <omp.h>
#define KGO_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define KGO_PARALLEL_FOR
#endif

int kgo_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static double* alloc0(int64_t n) {
  double* p = (double*)calloc((size_t)n, sizeof(double));
  return p;
}

/* --------------------------------------------------------------------------
 * Grid: mesh_refinement_lib.f90:100-147 with nproc = 1, rank = 0.
 * -------------------------------------------------------------------------- */
static void grid_build(kgo_sim* s) {
  /* mesh_refinement_boundary: dx = (xmax - xmin)/nproc; xmin' = xmin + 0*dx;
   * xmax' = xmin + 1*dx  (mesh_refinement_lib.f90:109-112) */
  double dxp = (s->xmax - s->xmin) / 1.0;
  double xmin = s->xmin + 0.0 * dxp;
  double xmax = s->xmin + 1.0 * dxp;
  s->dx = (xmax - xmin) / (double)s->Nx;               /* :115 */
  s->dy = acos(-1.0) / (double)(s->Ny - 1);            /* :116 */
  s->dz = 2.0 * acos(-1.0) / (double)(s->Nz - 1);      /* :117 */
  for (int ii = 0; ii < s->ni; ++ii) {
    int i = ii - s->g;
    s->x[ii] = xmin + (double)(i + s->g) * s->dx;      /* :130 */
  }
  for (int jj = 0; jj < s->nj; ++jj) {
    int j = jj - s->g;
    s->y[jj] = -s->dy / 2.0 + (double)j * s->dy;       /* :131 */
  }
  for (int kk = 0; kk < s->nk; ++kk) s->z[kk] = (double)kk * s->dz;  /* :132 */
}

double kgo_reference_dt(int Nx, int Ny, int Nz, int g, double xmin, double xmax, double CFL) {
  /* diagnostic_lib.f90:101-104: dxx = x(iL-g+1) - x(iL-g);
   * dt = CFL*min(dxx, exp(xmin)*dy, exp(xmin)*sin(dy/2)*dz) */
  double dxp = (xmax - xmin) / 1.0;
  double xm = xmin + 0.0 * dxp, xM = xmin + 1.0 * dxp;
  double dx = (xM - xm) / (double)Nx;
  double dy = acos(-1.0) / (double)(Ny - 1);
  double dz = 2.0 * acos(-1.0) / (double)(Nz - 1);
  double x0 = xm + (double)(0) * dx;        /* i = -g */
  double x1 = xm + (double)(1) * dx;        /* i = -g+1 */
  (void)g;
  double dxx = x1 - x0;
  double e = exp(xmin);
  double c1 = e * dy;
  double c2 = e * sin(dy / 2.0) * dz;
  double mn = dxx;
  if (c1 < mn) mn = c1;
  if (c2 < mn) mn = c2;
  return CFL * mn;
}

/* --------------------------------------------------------------------------
 * Metric: geometry.f90:92-352 (load_metric, boundary_theta_dn/up,
 * load_inverse).  All formulas depend on (x, y) only.
 * -------------------------------------------------------------------------- */
#define GDN(mu, nu) gdn[(mu) * 4 + (nu)]
#define GUP(mu, nu) gup[(mu) * 4 + (nu)]

static void metric_point(const kgo_sim* s, double x, double y, double* gdn) {
  const double s2_13 = (s->plant == KGO_PLANT_GDN13_NO_SIN2) ? 1.0 : sin(y) * sin(y);
  const double M = s->M_bh, a = s->a;
  for (int q = 0; q < 16; ++q) gdn[q] = 0.0;
  if (s->metric == KGO_METRIC_MINKOWSKI) {           /* geometry.f90:104-110 */
    GDN(0, 0) = -1.0;
    GDN(1, 1) = 1.0;
    GDN(2, 2) = x * x;
    double sy = sin(y);
    GDN(3, 3) = x * x * (sy * sy);
  } else if (s->metric == KGO_METRIC_KERR_SCHILD) {  /* geometry.f90:112-160 */
    double cy = cos(y), sy = sin(y);
    double c2 = cy * cy, s2 = sy * sy;
    double a2 = a * a;
    double rho2 = x * x + a2 * c2;
    GDN(0, 0) = -1.0 + 2.0 * M * x / (x * x + a2 * c2);
    GDN(0, 1) = (2.0 * M * x / (x * x + a2 * c2));
    GDN(0, 2) = 0.0;
    GDN(0, 3) = -2.0 * a * M * x * s2 / (x * x + a2 * c2);
    GDN(1, 0) = GDN(0, 1);
    GDN(1, 1) = (1.0 + 2.0 * M * x / (x * x + a2 * c2));
    GDN(1, 2) = 0.0;
    GDN(1, 3) = (-a * (1.0 + 2.0 * M * x / (x * x + a2 * c2)) * s2_13);
    GDN(2, 2) = (x * x + a2 * c2);
    GDN(3, 0) = GDN(0, 3);
    GDN(3, 1) = GDN(1, 3);
    double ax = a2 + x * x;
    GDN(3, 3) = ((ax * ax) - a2 * (-2.0 * M * x + a2 + x * x) * s2) * s2 / (x * x + a2 * c2);
    (void)rho2;
  } else if (s->metric == KGO_METRIC_MKS) {          /* geometry.f90:162-212 */
    const double h_mks = 0.0;
    double r = exp(x);
    double th = y + 0.5 * h_mks * sin(2.0 * y);
    double dth = 1.0 + h_mks * cos(2.0 * y);
    double ct = cos(th), st = sin(th);
    double c2 = ct * ct, s2 = st * st, a2 = a * a;
    GDN(0, 0) = -1.0 + 2.0 * M * r / (r * r + a2 * c2);
    GDN(0, 1) = (2.0 * M * r / (r * r + a2 * c2)) * r;
    GDN(0, 3) = -2.0 * a * M * r * s2 / (r * r + a2 * c2);
    GDN(1, 0) = GDN(0, 1);
    GDN(1, 1) = (1.0 + 2.0 * M * r / (r * r + a2 * c2)) * (r * r);
    GDN(1, 3) = (-a * (1.0 + 2.0 * M * r / (r * r + a2 * c2)) * s2) * r;
    GDN(2, 2) = (r * r + a2 * c2) * (dth * dth);
    GDN(3, 0) = GDN(0, 3);
    GDN(3, 1) = GDN(1, 3);
    double ar = a2 + r * r;
    GDN(3, 3) = ((ar * ar) - a2 * (-2.0 * M * r + a2 + r * r) * s2) * s2 / (r * r + a2 * c2);
  } else {                                           /* Schw_EF, geometry.f90:213-246 */
    GDN(0, 0) = -(1.0 - 2.0 * M / x);
    GDN(0, 1) = 2.0 * M / x;
    GDN(1, 0) = GDN(0, 1);
    GDN(1, 1) = 1.0 + 2.0 * M / x;
    GDN(2, 2) = x * x;
    double sy = sin(y);
    GDN(3, 3) = x * x * (sy * sy);
  }
}

static void inverse_point(const double* gdn, double* gup) {
  /* geometry.f90:286-320 (g**2 evaluated as g*g) */
  double g00 = GDN(0, 0), g01 = GDN(0, 1), g03 = GDN(0, 3), g11 = GDN(1, 1), g13 = GDN(1, 3),
         g22 = GDN(2, 2), g33 = GDN(3, 3);
  double idetg = 1.0 / (((g33 * g11 - g13 * g13) * g00 - g33 * (g01 * g01) + 2.0 * g13 * g03 * g01 -
                         g11 * (g03 * g03)) * g22);
  for (int q = 0; q < 16; ++q) gup[q] = 0.0;
  GUP(0, 0) = g22 * (g11 * g33 - g13 * g13) * idetg;
  GUP(0, 1) = g22 * (g13 * g03 - g01 * g33) * idetg;
  GUP(0, 2) = 0.0;
  GUP(0, 3) = g22 * (g13 * g01 - g11 * g03) * idetg;
  GUP(1, 1) = g22 * (g33 * g00 - g03 * g03) * idetg;
  GUP(1, 2) = 0.0;
  GUP(1, 3) = g22 * (-g13 * g00 + g03 * g01) * idetg;
  GUP(2, 2) = ((g33 * g11 - g13 * g13) * g00 - g33 * (g01 * g01) + 2.0 * g13 * g03 * g01 -
               g11 * (g03 * g03)) * idetg;
  GUP(2, 3) = 0.0;
  GUP(3, 3) = g22 * (g11 * g00 - g01 * g01) * idetg;
}

static void dx_beta_build(kgo_sim* s);

static void metric_build(kgo_sim* s) {
  const int ni = s->ni, nj = s->nj;
  const int64_t n2 = (int64_t)ni * nj;
  double gdn[16], gup[16];
  /* g_dn on all rows (geometry.f90:104-247) */
  for (int jj = 0; jj < nj; ++jj)
    for (int ii = 0; ii < ni; ++ii) {
      metric_point(s, s->x[ii], s->y[jj], gdn);
      for (int q = 0; q < 16; ++q) s->g_dn[q * n2 + IDX2(s, ii, jj)] = gdn[q];
    }
  /* boundary_theta_dn (geometry.f90:371-377): row j = Ny+g copied from
   * row j = Ny+g-1 at k_target; components (0,2),(1,2) negated.  The copy
   * source plane k_target carries the same (k-independent) values. */
  {
    int jd = nj - 1, js = nj - 2;
    for (int ii = 0; ii < ni; ++ii) {
      for (int q = 0; q < 16; ++q) s->g_dn[q * n2 + IDX2(s, ii, jd)] = s->g_dn[q * n2 + IDX2(s, ii, js)];
      s->g_dn[(0 * 4 + 2) * n2 + IDX2(s, ii, jd)] = -s->g_dn[(0 * 4 + 2) * n2 + IDX2(s, ii, js)];
      s->g_dn[(1 * 4 + 2) * n2 + IDX2(s, ii, jd)] = -s->g_dn[(1 * 4 + 2) * n2 + IDX2(s, ii, js)];
    }
  }
  /* symmetric copies (geometry.f90:260-265) and sqrtg (:267) */
  for (int64_t p = 0; p < n2; ++p) {
    for (int q = 0; q < 16; ++q) gdn[q] = s->g_dn[q * n2 + p];
    gdn[1 * 4 + 0] = gdn[0 * 4 + 1];
    gdn[2 * 4 + 0] = gdn[0 * 4 + 2];
    gdn[3 * 4 + 0] = gdn[0 * 4 + 3];
    gdn[2 * 4 + 1] = gdn[1 * 4 + 2];
    gdn[3 * 4 + 1] = gdn[1 * 4 + 3];
    gdn[3 * 4 + 2] = gdn[2 * 4 + 3];
    for (int q = 0; q < 16; ++q) s->g_dn[q * n2 + p] = gdn[q];
    s->sqrtg[p] = sqrt(GDN(2, 2) * (GDN(1, 1) * GDN(3, 3) - GDN(1, 3) * GDN(1, 3)));
  }
  /* inverse (geometry.f90:286-320) */
  for (int64_t p = 0; p < n2; ++p) {
    for (int q = 0; q < 16; ++q) gdn[q] = s->g_dn[q * n2 + p];
    inverse_point(gdn, gup);
    for (int q = 0; q < 16; ++q) s->g_up[q * n2 + p] = gup[q];
  }
  /* boundary_theta_up (geometry.f90:398-404) */
  {
    int jd = nj - 1, js = nj - 2;
    for (int ii = 0; ii < ni; ++ii) {
      for (int q = 0; q < 16; ++q) s->g_up[q * n2 + IDX2(s, ii, jd)] = s->g_up[q * n2 + IDX2(s, ii, js)];
      s->g_up[(0 * 4 + 2) * n2 + IDX2(s, ii, jd)] = -s->g_up[(0 * 4 + 2) * n2 + IDX2(s, ii, js)];
      s->g_up[(1 * 4 + 2) * n2 + IDX2(s, ii, jd)] = -s->g_up[(1 * 4 + 2) * n2 + IDX2(s, ii, js)];
    }
  }
  /* symmetric copies (:324-329), lapse, shift, spatial inverse (:332-342) */
  for (int64_t p = 0; p < n2; ++p) {
    for (int q = 0; q < 16; ++q) gup[q] = s->g_up[q * n2 + p];
    gup[1 * 4 + 0] = gup[0 * 4 + 1];
    gup[2 * 4 + 0] = gup[0 * 4 + 2];
    gup[3 * 4 + 0] = gup[0 * 4 + 3];
    gup[2 * 4 + 1] = gup[1 * 4 + 2];
    gup[3 * 4 + 1] = gup[1 * 4 + 3];
    gup[3 * 4 + 2] = gup[2 * 4 + 3];
    for (int q = 0; q < 16; ++q) s->g_up[q * n2 + p] = gup[q];
    double alpha = sqrt(-1.0 / GUP(0, 0));
    double a2 = alpha * alpha;
    double beta[4];
    beta[1] = a2 * GUP(0, 1);
    beta[2] = a2 * GUP(0, 2);
    beta[3] = a2 * GUP(0, 3);
    double gam[4][4];
    for (int i = 1; i <= 3; ++i)
      for (int j = 1; j <= 3; ++j) gam[i][j] = GUP(i, j) + beta[i] * beta[j] / a2;
    s->alpha[p] = alpha;
    s->beta_r[p] = beta[1];
    s->g_rr[p] = gam[1][1];
    s->g_rp[p] = gam[1][3];
    s->g_tt[p] = gam[2][2];
    s->g_pp[p] = gam[3][3];
  }
  /* dx_beta = Re dx1(beta(1)) (geometry.f90:344); k-independent, so one
   * plane of the 3-D stencil. */
  dx_beta_build(s);
}

/* --------------------------------------------------------------------------
 * Stencils: tools/finite_differences.f90 (2nd-order, the live set).
 * res = 0.5/h.  Bounds: x in [0, ni), y in [0, nj), z in [0, nk) with
 * bounds(3,2) = Nz = nk-1.
 * -------------------------------------------------------------------------- */
void kgo_dx1_raw(int ni, int nj, int nk, double h, const double* f, double* df) {
  /* first_derivative_x_2_uniform, finite_differences.f90:36-42 */
  double res = 0.50 / h;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < nk; ++kk)
    for (int jj = 0; jj < nj; ++jj) {
      const double* a = f + ((int64_t)kk * nj + jj) * ni;
      double* d = df + ((int64_t)kk * nj + jj) * ni;
      int lo = 0, hi = ni - 1;
      d[lo] = (-3.0 * a[lo] + 4.0 * a[lo + 1] - a[lo + 2]) * res;
      for (int i = lo + 1; i <= hi - 1; ++i) d[i] = (-a[i - 1] + a[i + 1]) * res;
      d[hi] = (+3.0 * a[hi] - 4.0 * a[hi - 1] + a[hi - 2]) * res;
    }
}

void kgo_dx2_raw(int ni, int nj, int nk, double h, const double* f, double* df) {
  /* first_derivative_y_2 with inpar = .false., finite_differences.f90:214-230.
   * Bounds(3,2) = nk-1 = Nz; dis = Nz/2; k_target = mod(k+dis, Nz-1). */
  double res = 0.50 / h;
  int Nz = nk - 1;
  int dis = Nz / 2;
  const int64_t pl = (int64_t)ni * nj;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < nk; ++kk) {
    int kt = (kk + dis) % (Nz - 1);
    const double* a = f + kk * pl;
    const double* at = f + kt * pl;
    double* d = df + kk * pl;
    int lo = 0, hi = nj - 1;
    for (int i = 0; i < ni; ++i)  /* :217, overwritten by :223 for every k */
      d[lo * ni + i] = (-3.0 * a[lo * ni + i] + 4.0 * a[(lo + 1) * ni + i] - a[(lo + 2) * ni + i]) * res;
    for (int i = 0; i < ni; ++i) d[lo * ni + i] = (-a[(lo + 1) * ni + i] + at[(lo + 2) * ni + i]) * res;
    for (int j = lo + 1; j <= hi - 1; ++j)
      for (int i = 0; i < ni; ++i) d[j * ni + i] = (-a[(j - 1) * ni + i] + a[(j + 1) * ni + i]) * res;
    for (int i = 0; i < ni; ++i)
      d[hi * ni + i] = (+3.0 * a[hi * ni + i] - 4.0 * a[(hi - 1) * ni + i] + a[(hi - 2) * ni + i]) * res;
  }
}

void kgo_dx3_raw(int ni, int nj, int nk, double h, const double* f, double* df) {
  /* first_derivative_z_2, finite_differences.f90:328-335 (k = Nz row uses
   * the pair (Nz-1, 2) as written) */
  double res = 0.50 / h;
  const int64_t pl = (int64_t)ni * nj;
  int lo = 0, hi = nk - 1;
  KGO_PARALLEL_FOR
  for (int64_t p = 0; p < pl; ++p) df[lo * pl + p] = (-f[(hi - 2) * pl + p] + f[(lo + 1) * pl + p]) * res;
  KGO_PARALLEL_FOR
  for (int k = lo + 1; k <= hi - 1; ++k)
    for (int64_t p = 0; p < pl; ++p) df[k * pl + p] = (-f[(k - 1) * pl + p] + f[(k + 1) * pl + p]) * res;
  KGO_PARALLEL_FOR
  for (int64_t p = 0; p < pl; ++p) df[hi * pl + p] = (-f[(lo + 2) * pl + p] + f[(hi - 1) * pl + p]) * res;
}

void kgo_advec_x_raw(int ni, int nj, int nk, double h, const double* f, const double* beta, double* df) {
  /* advec_x_uniform, finite_differences.f90:779-799; beta is a full 3-D array */
  double res = 0.5 / h;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < nk; ++kk)
    for (int jj = 0; jj < nj; ++jj) {
      const int64_t o = ((int64_t)kk * nj + jj) * ni;
      const double* a = f + o;
      const double* b = beta + o;
      double* d = df + o;
      int lo = 0, hi = ni - 1;
      d[lo] = (-3.0 * a[lo] + 4.0 * a[lo + 1] - a[lo + 2]) * res;
      d[lo + 1] = (-3.0 * a[lo + 1] + 4.0 * a[lo + 2] - a[lo + 3]) * res;
      for (int i = lo + 2; i <= hi - 2; ++i) {
        if (b[i] > 0.0)
          d[i] = (-3.0 * a[i] + 4.0 * a[i + 1] - a[i + 2]) * res;
        else if (b[i] < 0.0)
          d[i] = (3.0 * a[i] - 4.0 * a[i - 1] + a[i - 2]) * res;
        else
          d[i] = (a[i + 1] - a[i - 1]) * res;
      }
      d[hi - 1] = (a[hi] - a[hi - 2]) * res;
      d[hi] = (3.0 * a[hi] - 4.0 * a[hi - 1] + a[hi - 2]) * res;
    }
}

/* --------------------------------------------------------------------------
 * The fd4 set (finite_differences.f90, real arrays with lower bound 0; N is
 * the last index along the differentiated axis).  Each row formula is
 * evaluated in the reference's order; the whole-array statements run one
 * after another, so a later row wins where rows coincide (not for N >= 4).
 * -------------------------------------------------------------------------- */
#define FD4_ROWS(D, A, N, RES)                                                                          \
  do {                                                                                                  \
    D(0) = (-25.0 * A(0) + 48.0 * A(1) - 36.0 * A(2) + 16.0 * A(3) - 3.0 * A(4)) * (RES);              \
    D(1) = (-3.0 * A(0) - 10.0 * A(1) + 18.0 * A(2) - 6.0 * A(3) + 1.0 * A(4)) * (RES);                \
    for (int q_ = 2; q_ <= (N)-2; ++q_) D(q_) = (A(q_ - 2) - 8.0 * A(q_ - 1) + 8.0 * A(q_ + 1) - A(q_ + 2)) * (RES); \
    D((N)-1) = (+3.0 * A(N) + 10.0 * A((N)-1) - 18.0 * A((N)-2) + 6.0 * A((N)-3) - 1.0 * A((N)-4)) * (RES); \
    D(N) = (+25.0 * A(N) - 48.0 * A((N)-1) + 36.0 * A((N)-2) - 16.0 * A((N)-3) + 3.0 * A((N)-4)) * (RES); \
  } while (0)

void kgo_dx1_4_raw(int ni, int nj, int nk, double h, const double* f, double* df) {
  /* first_derivative_x_4 (:124-151): res = 1/(12h), N = size(f,1)-1 */
  const double res = 1.0 / (12.0 * h);
  const int N = ni - 1;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < nk; ++kk)
    for (int jj = 0; jj < nj; ++jj) {
      const double* a = f + ((int64_t)kk * nj + jj) * ni;
      double* d = df + ((int64_t)kk * nj + jj) * ni;
#define A_(q) a[q]
#define D_(q) d[q]
      FD4_ROWS(D_, A_, N, res);
#undef A_
#undef D_
    }
}

void kgo_dx2_4_raw(int ni, int nj, int nk, double h, const double* f, double* df) {
  /* first_derivative_y_4 (:241-268): N = size(f,2)-1; no pole rows */
  const double res = 1.0 / (12.0 * h);
  const int N = nj - 1;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < nk; ++kk)
    for (int ii = 0; ii < ni; ++ii) {
      const double* a = f + (int64_t)kk * nj * ni + ii;
      double* d = df + (int64_t)kk * nj * ni + ii;
#define A_(q) a[(int64_t)(q) * ni]
#define D_(q) d[(int64_t)(q) * ni]
      FD4_ROWS(D_, A_, N, res);
#undef A_
#undef D_
    }
}

void kgo_dx3_4_raw(int ni, int nj, int nk, double h, const double* f, double* df) {
  /* first_derivative_z_4 (:340-367): N = size(f,3)-1; one-sided at both ends */
  const double res = 1.0 / (12.0 * h);
  const int N = nk - 1;
  const int64_t pl = (int64_t)ni * nj;
  KGO_PARALLEL_FOR
  for (int64_t p = 0; p < pl; ++p) {
    const double* a = f + p;
    double* d = df + p;
#define A_(q) a[(int64_t)(q) * pl]
#define D_(q) d[(int64_t)(q) * pl]
    FD4_ROWS(D_, A_, N, res);
#undef A_
#undef D_
  }
}

/* advec_x_4 on one r-row a[0..N] with the row's beta b[0..N] */
static void advec4_row(const double* a, const double* b, double* d, int N, double res) {
  d[0] = (-25.0 * a[0] + 48.0 * a[1] - 36.0 * a[2] + 16.0 * a[3] - 3.0 * a[4]) * res / 12.0;
  d[1] = (-3.0 * a[0] - 10.0 * a[1] + 18.0 * a[2] - 6.0 * a[3] + a[4]) * res / 12.0;
  d[2] = (a[0] - 8.0 * a[1] + 8.0 * a[3] - a[4]) * res / 12.0;
  for (int i = 3; i <= N - 3; ++i) {
    if (b[i] < 0.0)
      d[i] = (3.0 * a[i + 1] + 10.0 * a[i] - 18.0 * a[i - 1] + 6.0 * a[i - 2] - a[i - 3]) * res / 12.0;
    else if (b[i] > 0.0)
      d[i] = (-3.0 * a[i - 1] - 10.0 * a[i] + 18.0 * a[i + 1] - 6.0 * a[i + 2] + a[i + 3]) * res / 12.0;
    else
      d[i] = (a[i - 2] - 8.0 * a[i - 1] + 8.0 * a[i + 1] - a[i + 2]) * res / 12.0;
  }
  d[N - 2] = (a[N - 4] - 8.0 * a[N - 3] + 8.0 * a[N - 1] - a[N]) * res / 12.0;
  d[N - 1] = (3.0 * a[N] + 10.0 * a[N - 1] - 18.0 * a[N - 2] + 6.0 * a[N - 3] - a[N - 4]) * res / 12.0;
  d[N] = (25.0 * a[N] - 48.0 * a[N - 1] + 36.0 * a[N - 2] - 16.0 * a[N - 3] + 3.0 * a[N - 4]) * res / 12.0;
}

void kgo_advec_x_4_raw(int ni, int nj, int nk, double h, const double* f, const double* beta, double* df) {
  /* advec_x_4 (:922-957): res = 1/dx, every row (...)*res/12; rows 0, 1, 2,
   * N-2, N-1, N fixed; 3..N-3 upwinded by the sign of beta (beta < 0 first) */
  const double res = 1.0 / h;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < nk; ++kk)
    for (int jj = 0; jj < nj; ++jj) {
      const int64_t o = ((int64_t)kk * nj + jj) * ni;
      advec4_row(f + o, beta + o, df + o, ni - 1, res);
    }
}

/* d_r beta^r of the metric with the operator set's D_r (geometry.f90:344
 * uses dx1; the fd4 set takes x_4, so the coefficient is as accurate as the
 * derivatives it multiplies) */
static void dx_beta_build(kgo_sim* s) {
  const int ni = s->ni, nj = s->nj;
  if (s->order == 4) {
    kgo_dx1_4_raw(ni, nj, 1, s->dx, s->beta_r, s->dx_beta);
    return;
  }
  double res = 0.50 / s->dx;
  for (int jj = 0; jj < nj; ++jj) {
    const double* f = s->beta_r + IDX2(s, 0, jj);
    double* df = s->dx_beta + IDX2(s, 0, jj);
    int lo = 0, hi = ni - 1;
    df[lo] = (-3.0 * f[lo] + 4.0 * f[lo + 1] - f[lo + 2]) * res;
    for (int ii = lo + 1; ii <= hi - 1; ++ii) df[ii] = (-f[ii - 1] + f[ii + 1]) * res;
    df[hi] = (+3.0 * f[hi] - 4.0 * f[hi - 1] + f[hi - 2]) * res;
  }
}

int kgo_plant(kgo_sim* s, int code) {
  if (code < 0 || code >= KGO_PLANT_NUM) return -1;
  s->plant = code;
  metric_build(s);
  return 0;
}

int kgo_set_order(kgo_sim* s, int order) {
  if (order != 2 && order != 4) return -1;
  s->order = order;
  dx_beta_build(s);
  return 0;
}

/* the RHS's operators under the selected set */
static void op_dr(const kgo_sim* s, const double* f, double* df) {
  if (s->order == 4)
    kgo_dx1_4_raw(s->ni, s->nj, s->nk, s->dx, f, df);
  else
    kgo_dx1_raw(s->ni, s->nj, s->nk, s->dx, f, df);
}
static void op_dth(const kgo_sim* s, const double* f, double* df) {
  if (s->order == 4)
    kgo_dx2_4_raw(s->ni, s->nj, s->nk, s->dy, f, df);
  else
    kgo_dx2_raw(s->ni, s->nj, s->nk, s->dy, f, df);
}
static void op_dph(const kgo_sim* s, const double* f, double* df) {
  if (s->order == 4)
    kgo_dx3_4_raw(s->ni, s->nj, s->nk, s->dz, f, df);
  else
    kgo_dx3_raw(s->ni, s->nj, s->nk, s->dz, f, df);
}
void kgo_advec_x(const kgo_sim* s, const double* f, const double* beta2d, double* df);
static void op_adv(const kgo_sim* s, const double* f, double* df) {
  if (s->order != 4) {
    kgo_advec_x(s, f, s->beta_r, df);
    return;
  }
  /* advec_x_4 with beta(i,j,k) = beta2d(i,j) */
  const double res = 1.0 / s->dx;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj) {
      const int64_t o = ((int64_t)kk * s->nj + jj) * s->ni;
      advec4_row(f + o, s->beta_r + IDX2(s, 0, jj), df + o, s->ni - 1, res);
    }
}

void kgo_dx1(const kgo_sim* s, const double* f, double* df) { kgo_dx1_raw(s->ni, s->nj, s->nk, s->dx, f, df); }
void kgo_dx2(const kgo_sim* s, const double* f, double* df) { kgo_dx2_raw(s->ni, s->nj, s->nk, s->dy, f, df); }
void kgo_dx3(const kgo_sim* s, const double* f, double* df) { kgo_dx3_raw(s->ni, s->nj, s->nk, s->dz, f, df); }

void kgo_advec_x(const kgo_sim* s, const double* f, const double* beta2d, double* df) {
  /* same as advec_x_uniform with beta(i,j,k) = beta2d(i,j) */
  double res = 0.5 / s->dx;
  const int ni = s->ni, nj = s->nj, nk = s->nk;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < nk; ++kk)
    for (int jj = 0; jj < nj; ++jj) {
      const int64_t o = ((int64_t)kk * nj + jj) * ni;
      const double* a = f + o;
      const double* b = beta2d + (int64_t)jj * ni;
      double* d = df + o;
      int lo = 0, hi = ni - 1;
      d[lo] = (-3.0 * a[lo] + 4.0 * a[lo + 1] - a[lo + 2]) * res;
      d[lo + 1] = (-3.0 * a[lo + 1] + 4.0 * a[lo + 2] - a[lo + 3]) * res;
      for (int i = lo + 2; i <= hi - 2; ++i) {
        if (b[i] > 0.0)
          d[i] = (-3.0 * a[i] + 4.0 * a[i + 1] - a[i + 2]) * res;
        else if (b[i] < 0.0)
          d[i] = (3.0 * a[i] - 4.0 * a[i - 1] + a[i - 2]) * res;
        else
          d[i] = (a[i + 1] - a[i - 1]) * res;
      }
      d[hi - 1] = (a[hi] - a[hi - 2]) * res;
      d[hi] = (3.0 * a[hi] - 4.0 * a[hi - 1] + a[hi - 2]) * res;
    }
}

/* --------------------------------------------------------------------------
 * RHS: lines_method_lib.f90:671-753.  vars is a 5-field state.
 * -------------------------------------------------------------------------- */
#define FIELD(v, c) ((v) + (int64_t)(c) * s->npts)
#define T2(tab) (tab)[p2]

void kgo_rhs_phi(kgo_sim* s, double* rhs, const double* vars) {
  /* :675  rhs = alpha*vars(5) + beta(1)*vars(2) */
  const double *v2 = FIELD(vars, 1), *v5 = FIELD(vars, 4);
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj)
      for (int ii = 0; ii < s->ni; ++ii) {
        int64_t p = IDX(s, ii, jj, kk), p2 = IDX2(s, ii, jj);
        rhs[p] = T2(s->alpha) * v5[p] + T2(s->beta_r) * v2[p];
      }
}

void kgo_rhs_psi1(kgo_sim* s, double* rhs, const double* vars) {
  /* :687-689  rhs = dx1(alpha*vars(5)) + beta*ad_1(vars(2),beta) + dx_beta*vars(2) */
  const double *v2 = FIELD(vars, 1), *v5 = FIELD(vars, 4);
  double *t0 = s->tmp[0], *t1 = s->tmp[1], *t2 = s->tmp[2];
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj)
      for (int ii = 0; ii < s->ni; ++ii) {
        int64_t p = IDX(s, ii, jj, kk), p2 = IDX2(s, ii, jj);
        t0[p] = T2(s->alpha) * v5[p];
      }
  op_dr(s, t0, t1);
  op_adv(s, v2, t2);
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj)
      for (int ii = 0; ii < s->ni; ++ii) {
        int64_t p = IDX(s, ii, jj, kk), p2 = IDX2(s, ii, jj);
        rhs[p] = (s->plant == KGO_PLANT_DROP_PSI1_DXBETA) ? t1[p] + T2(s->beta_r) * t2[p]
                                                          : t1[p] + T2(s->beta_r) * t2[p] + T2(s->dx_beta) * v2[p];
      }
}

static void flux_F(kgo_sim* s, const double* vars, double* F) {
  /* alpha*vars(5) + beta(1)*vars(2)  (:700, :711) */
  const double *v2 = FIELD(vars, 1), *v5 = FIELD(vars, 4);
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj)
      for (int ii = 0; ii < s->ni; ++ii) {
        int64_t p = IDX(s, ii, jj, kk), p2 = IDX2(s, ii, jj);
        F[p] = T2(s->alpha) * v5[p] + T2(s->beta_r) * v2[p];
      }
}

void kgo_rhs_psi2(kgo_sim* s, double* rhs, const double* vars) {
  /* :700  rhs = dx2(alpha*vars(5) + beta*vars(2)) */
  flux_F(s, vars, s->tmp[0]);
  if (s->plant == KGO_PLANT_PSI2_ALPHA_PI) {
    const double* v5 = FIELD(vars, 4);
    for (int64_t p = 0; p < s->npts; ++p) s->tmp[0][p] = s->alpha[p % ((int64_t)s->ni * s->nj)] * v5[p];
  }
  op_dth(s, s->tmp[0], rhs);
}

void kgo_rhs_psi3(kgo_sim* s, double* rhs, const double* vars) {
  /* :711  rhs = dx3(alpha*vars(5) + beta*vars(2)) */
  flux_F(s, vars, s->tmp[0]);
  op_dph(s, s->tmp[0], rhs);
}

void kgo_rhs_pi(kgo_sim* s, double* rhs, const double* vars) {
  /* Kerr_Schild branch :728-736; Schw_EF branch :738-743.  Other metrics stop
   * in the reference (:744-748); here they fall back to the Kerr_Schild form
   * only when the caller asks for them explicitly (tests use KS and EF). */
  const double *v1 = FIELD(vars, 0), *v2 = FIELD(vars, 1), *v3 = FIELD(vars, 2), *v4 = FIELD(vars, 3),
               *v5 = FIELD(vars, 4);
  double *fr = s->tmp[0], *ft = s->tmp[1], *fp = s->tmp[2], *ps = s->tmp[3];
  double *d1 = s->tmp[4], *d2 = s->tmp[5], *d3 = s->tmp[6], *ad = s->tmp[7];
  const int ef = (s->metric == KGO_METRIC_SCHW_EF);
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj)
      for (int ii = 0; ii < s->ni; ++ii) {
        int64_t p = IDX(s, ii, jj, kk), p2 = IDX2(s, ii, jj);
        double al = T2(s->alpha), sg = T2(s->sqrtg);
        if (!ef) {
          fr[p] = al * sg * (T2(s->g_rr) * v2[p] + T2(s->g_rp) * v4[p]);
          fp[p] = (s->plant == KGO_PLANT_SWAP_PHI_FLUX) ? al * sg * (T2(s->g_pp) * v2[p] + T2(s->g_rp) * v4[p])
                                                        : al * sg * (T2(s->g_rp) * v2[p] + T2(s->g_pp) * v4[p]);
        } else {
          fr[p] = al * sg * T2(s->g_rr) * v2[p];
          fp[p] = al * T2(s->g_pp) * v4[p];
        }
        ft[p] = al * sg * T2(s->g_tt) * v3[p];
        ps[p] = (s->plant == KGO_PLANT_ADVECT_PI) ? v5[p] : v5[p] * sg;
      }
  op_dr(s, fr, d1);
  op_dth(s, ft, d2);
  op_dph(s, fp, d3);
  op_adv(s, ps, ad);
  const double m2 = s->m * s->m;
  KGO_PARALLEL_FOR
  for (int kk = 0; kk < s->nk; ++kk)
    for (int jj = 0; jj < s->nj; ++jj)
      for (int ii = 0; ii < s->ni; ++ii) {
        int64_t p = IDX(s, ii, jj, kk), p2 = IDX2(s, ii, jj);
        double sg = T2(s->sqrtg);
        double t3 = ef ? d3[p] : d3[p] / sg;
        const double pdb = (s->plant == KGO_PLANT_DROP_PI_DXBETA) ? 0.0 : v5[p] * T2(s->dx_beta);
        const double mass = (s->plant == KGO_PLANT_MASS_SIGN) ? -(T2(s->alpha) * m2 * v1[p]) : T2(s->alpha) * m2 * v1[p];
        rhs[p] = d1[p] / sg + d2[p] / sg + t3 + pdb + T2(s->beta_r) * ad[p] / sg - mass;
      }
  if (s->bc_rmax == KGO_BC_SOMMERFELD) {
    /* the radiative condition at the outer column (no reference counterpart) */
    const int N = s->ni - 1;
    const double r = s->x[N];
    for (int kk = 0; kk < s->nk; ++kk)
      for (int jj = 0; jj < s->nj; ++jj) {
        const int64_t p = IDX(s, N, jj, kk), p2 = IDX2(s, N, jj);
        const double* a = v5 + p;  /* a[-m] = Pi at column N-m */
        double d;
        if (s->order == 4)
          d = (+25.0 * a[0] - 48.0 * a[-1] + 36.0 * a[-2] - 16.0 * a[-3] + 3.0 * a[-4]) * (1.0 / (12.0 * s->dx));
        else
          d = (+3.0 * a[0] - 4.0 * a[-1] + a[-2]) * (0.50 / s->dx);
        const double v = s->alpha[p2] * sqrt(s->g_rr[p2]) - s->beta_r[p2];
        rhs[p] = -(v * (d + a[0] / r));
      }
  }
}

void kgo_rhs_all(kgo_sim* s, double* rhs5, const double* vars) {
  kgo_rhs_phi(s, FIELD(rhs5, 0), vars);
  kgo_rhs_psi1(s, FIELD(rhs5, 1), vars);
  kgo_rhs_psi2(s, FIELD(rhs5, 2), vars);
  kgo_rhs_psi3(s, FIELD(rhs5, 3), vars);
  kgo_rhs_pi(s, FIELD(rhs5, 4), vars);
}

/* --------------------------------------------------------------------------
 * Ghost fill: boundary() = xL, xR, yL, yR, zL, zR in this order
 * (lines_method_lib.f90:313-330, rank 0 = nproc-1).
 * -------------------------------------------------------------------------- */
void kgo_fill(kgo_sim* s, double* u, const double* u_p) {
  const int ni = s->ni, nj = s->nj, nk = s->nk, g = s->g;
  const int Ny = s->Ny, Nz = s->Nz;
  for (int c = 0; c < 5; ++c) {
    double* f = FIELD(u, c);
    const double* fp = FIELD(u_p, c);
    /* boundary_xL :516-517  u(-g) = 3u(-g+1) - 3u(-g+2) + u(-g+3) */
    for (int kk = 0; kk < nk; ++kk)
      for (int jj = 0; jj < nj; ++jj) {
        double* a = f + IDX(s, 0, jj, kk);
        a[0] = 3.0 * a[1] - 3.0 * a[2] + a[3];
      }
    /* boundary_xR :525-528 (none with Sommerfeld: the outer column is evolved) */
    for (int kk = 0; kk < nk && s->bc_rmax != KGO_BC_SOMMERFELD; ++kk)
      for (int jj = 0; jj < nj; ++jj) {
        int64_t o = IDX(s, ni - 1, jj, kk);
        if (s->bc_rmax == KGO_BC_OUTFLOW)
          f[o] = f[o - 1];
        else
          f[o] = fp[o];
      }
    /* boundary_yL :542-555  u(:,:,-gg,k) = u(:,:,gg+1,kt(k)), gg = g..0 */
    int dis = Nz / 2;
    for (int k = 0; k <= Nz; ++k) {
      int kt = (k + dis) % (Nz - 1);
      for (int gg = g; gg >= 0; --gg) {
        double* dst = f + IDX(s, 0, -gg + g, k);
        const double* src = f + IDX(s, 0, gg + 1 + g, kt);
        memcpy(dst, src, sizeof(double) * (size_t)ni);
      }
    }
    /* boundary_yR :565-578  u(:,:,Ny+gg,k) = u(:,:,Ny-gg-1,kt(k)) */
    for (int k = 0; k <= Nz; ++k) {
      int kt = (k + dis) % (Nz - 1);
      for (int gg = g; gg >= 0; --gg) {
        double* dst = f + IDX(s, 0, Ny + gg + g, k);
        const double* src = f + IDX(s, 0, Ny - gg - 1 + g, kt);
        memcpy(dst, src, sizeof(double) * (size_t)ni);
      }
    }
    /* boundary_zL :587 and boundary_zR :596 */
    const int64_t pl = (int64_t)ni * nj;
    memcpy(f + 0 * pl, f + (int64_t)(Nz - 1) * pl, sizeof(double) * (size_t)pl);
    memcpy(f + (int64_t)Nz * pl, f + 1 * pl, sizeof(double) * (size_t)pl);
  }
}

void kgo_fill_source_map(kgo_sim* s, int64_t* src) {
  /* Runs the fill on index tags: every point starts holding its own flat
   * index (a double, exact below 2^53) and the start-of-step state holds -1.
   * Afterwards a point holds q when its value is a copy of the evolved value
   * at q.  The xL extrapolation 3(q+1)-3(q+2)+(q+3) of consecutive tags
   * returns q itself, so a tag whose i index is -g marks a value made by the
   * extrapolation at that row; -1 marks an inflow value. */
  const int64_t n = s->npts;
  double* t = (double*)malloc(sizeof(double) * (size_t)n * 5);
  double* tp = (double*)malloc(sizeof(double) * (size_t)n * 5);
  for (int c = 0; c < 5; ++c)
    for (int64_t p = 0; p < n; ++p) {
      t[c * n + p] = (double)p;
      tp[c * n + p] = -1.0;
    }
  kgo_fill(s, t, tp);
  for (int64_t p = 0; p < n; ++p) src[p] = (int64_t)t[p];
  free(t);
  free(tp);
}

/* --------------------------------------------------------------------------
 * Integrators over an operator table (lines_method_lib.f90:211-474).
 * -------------------------------------------------------------------------- */
static void pirk_stage1(const kgo_ops* ops, kgo_bufs* b, double dt, double* st) {
  /* :278-289 (pirk2) == :340-351 (pirk3):
   *   rhs_u(1..4) = R(u); u_stage(c) = u(c) + dt*rhs_u(c), c = 1..4
   *   rhs_u(5) = R_pi(u_stage)   [u_stage(5) is the previous step's value]
   *   u_stage(5) = u(5) + dt*rhs_u(5) */
  const int64_t n = ops->n;
  ops->rhs14(ops->ctx, b->u, b->rhs_u);
  for (int c = 0; c < 4; ++c)
    KGO_PARALLEL_FOR
    for (int64_t p = 0; p < n; ++p) st[c * n + p] = b->u[c * n + p] + dt * b->rhs_u[c * n + p];
  ops->rhs5(ops->ctx, st, b->rhs_u + 4 * n);
  KGO_PARALLEL_FOR
  for (int64_t p = 0; p < n; ++p) st[4 * n + p] = b->u[4 * n + p] + dt * b->rhs_u[4 * n + p];
}

void kgo_pirk2_generic(const kgo_ops* ops, kgo_bufs* b, double dt) {
  const int64_t n = ops->n, n5 = 5 * n;
  memcpy(b->u_p, b->u, sizeof(double) * (size_t)n5);                 /* :275 */
  pirk_stage1(ops, b, dt, b->u_stage);                                 /* :278-289 */
  if (ops->fill) ops->fill(ops->ctx, b->u, b->u_p);                    /* :291 */
  ops->rhs14(ops->ctx, b->u_stage, b->rhs_u);                          /* :294-297 */
  ops->rhs5(ops->ctx, b->u_stage, b->rhs_u + 4 * n);                   /* :298 */
  KGO_PARALLEL_FOR
  for (int64_t p = 0; p < n5; ++p)                                     /* :300-304 */
    b->u[p] = 0.5 * (b->u_p[p] + b->u_stage[p] + dt * b->rhs_u[p]);
  if (ops->fill) ops->fill(ops->ctx, b->u, b->u_p);                    /* :306 */
}

void kgo_pirk3_generic(const kgo_ops* ops, kgo_bufs* b, double dt) {
  const int64_t n = ops->n, n5 = 5 * n;
  memcpy(b->u_p, b->u, sizeof(double) * (size_t)n5);                 /* :337 */
  pirk_stage1(ops, b, dt, b->u_stage);                                 /* :340-351 */
  if (ops->fill) ops->fill(ops->ctx, b->u, b->u_p);                    /* :353 */
  ops->rhs14(ops->ctx, b->u_stage, b->rhs_u);                          /* :356-359 */
  for (int c = 0; c < 4; ++c)                                          /* :361-364 */
    KGO_PARALLEL_FOR
    for (int64_t p = 0; p < n; ++p)
      b->u_stage_2[c * n + p] =
          0.75 * b->u_p[c * n + p] + 0.25 * (b->u_stage[c * n + p] + dt * b->rhs_u[c * n + p]);
  ops->rhs5(ops->ctx, b->u_stage_2, b->rhs_u + 4 * n);                 /* :366 */
  KGO_PARALLEL_FOR
  for (int64_t p = 0; p < n; ++p)                                      /* :367 */
    b->u_stage_2[4 * n + p] = 0.75 * b->u_p[4 * n + p] + 0.25 * (b->u_stage[4 * n + p] + dt * b->rhs_u[4 * n + p]);
  if (ops->fill) ops->fill(ops->ctx, b->u, b->u_p);                    /* :369 */
  ops->rhs14(ops->ctx, b->u_stage_2, b->rhs_u);                        /* :372-375 */
  ops->rhs5(ops->ctx, b->u_stage_2, b->rhs_u + 4 * n);                 /* :376 */
  const double third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
  KGO_PARALLEL_FOR
  for (int64_t p = 0; p < n5; ++p)                                     /* :378-382 */
    b->u[p] = third * b->u_p[p] + two_thirds * (b->u_stage_2[p] + dt * b->rhs_u[p]);
  if (ops->fill) ops->fill(ops->ctx, b->u, b->u_p);                    /* :384 */
}

void kgo_rk3_generic(const kgo_ops* ops, kgo_bufs* b, double dt) {
  const int64_t n = ops->n, n5 = 5 * n;
  memcpy(b->u_p, b->u, sizeof(double) * (size_t)n5);                 /* :223 */
  for (int rk = 1; rk <= 3; ++rk) {
    ops->rhs14(ops->ctx, b->u, b->rhs_u);                              /* rhs(1) */
    ops->rhs5(ops->ctx, b->u, b->rhs_u + 4 * n);
    if (rk == 1) {                                                     /* :229-237 */
      KGO_PARALLEL_FOR
      for (int64_t p = 0; p < n5; ++p) b->u[p] = b->u_p[p] + b->rhs_u[p] * dt;
    } else if (rk == 2) {
      KGO_PARALLEL_FOR
      for (int64_t p = 0; p < n5; ++p) b->u[p] = 0.750 * b->u_p[p] + 0.250 * (b->u[p] + b->rhs_u[p] * dt);
    } else {
      KGO_PARALLEL_FOR
      for (int64_t p = 0; p < n5; ++p)
        b->u[p] = b->u_p[p] / 3.0 + 2.0 * (b->u[p] + b->rhs_u[p] * dt) / 3.0;
    }
    if (ops->fill) ops->fill(ops->ctx, b->u, b->u_p);                  /* :239 */
  }
}

void kgo_rk4_generic(const kgo_ops* ops, kgo_bufs* b, double dt) {
  const int64_t n = ops->n, n5 = 5 * n;
  memcpy(b->u_p, b->u, sizeof(double) * (size_t)n5);                 /* :424 */
  for (int rk = 1; rk <= 4; ++rk) {
    ops->rhs14(ops->ctx, b->u, b->rhs_u);                              /* rhs(1) */
    ops->rhs5(ops->ctx, b->u, b->rhs_u + 4 * n);
    if (rk == 1) {                                                     /* :429-441 */
      KGO_PARALLEL_FOR
      for (int64_t p = 0; p < n5; ++p) {
        b->k1[p] = b->rhs_u[p] * dt;
        b->u[p] = b->u_p[p] + 0.50 * b->k1[p];
      }
    } else if (rk == 2) {
      KGO_PARALLEL_FOR
      for (int64_t p = 0; p < n5; ++p) {
        b->k2[p] = b->rhs_u[p] * dt;
        b->u[p] = b->u_p[p] + 0.50 * b->k2[p];
      }
    } else if (rk == 3) {
      KGO_PARALLEL_FOR
      for (int64_t p = 0; p < n5; ++p) {
        b->k3[p] = b->rhs_u[p] * dt;
        b->u[p] = b->u_p[p] + b->k3[p];
      }
    } else {
      KGO_PARALLEL_FOR
      for (int64_t p = 0; p < n5; ++p) {
        b->k4[p] = b->rhs_u[p] * dt;
        b->u[p] = b->u_p[p] + (b->k1[p] + 2.0 * b->k2[p] + 2.0 * b->k3[p] + b->k4[p]) / 6.0;
      }
    }
    if (ops->fill) ops->fill(ops->ctx, b->u, b->u_p);                  /* :443 */
  }
}

/* --------------------------------------------------------------------------
 * Klein-Gordon operator table and the simulation object.
 * -------------------------------------------------------------------------- */
static void kg_rhs14(void* ctx, const double* vars, double* out5) {
  kgo_sim* s = (kgo_sim*)ctx;
  kgo_rhs_phi(s, FIELD(out5, 0), vars);
  kgo_rhs_psi1(s, FIELD(out5, 1), vars);
  kgo_rhs_psi2(s, FIELD(out5, 2), vars);
  kgo_rhs_psi3(s, FIELD(out5, 3), vars);
}
static void kg_rhs5(void* ctx, const double* vars, double* out1) { kgo_rhs_pi((kgo_sim*)ctx, out1, vars); }
static void kg_fill(void* ctx, double* u, const double* u_p) { kgo_fill((kgo_sim*)ctx, u, u_p); }

static void sim_ops(kgo_sim* s, kgo_ops* ops, kgo_bufs* b) {
  ops->rhs14 = kg_rhs14;
  ops->rhs5 = kg_rhs5;
  ops->fill = kg_fill;
  ops->ctx = s;
  ops->n = s->npts;
  b->u = s->u;
  b->u_p = s->u_p;
  b->u_stage = s->u_stage;
  b->u_stage_2 = s->u_stage_2;
  b->rhs_u = s->rhs_u;
  b->k1 = s->k1;
  b->k2 = s->k2;
  b->k3 = s->k3;
  b->k4 = s->k4;
}

void kgo_step(kgo_sim* s, int scheme, int nsteps) {
  kgo_ops ops;
  kgo_bufs b;
  sim_ops(s, &ops, &b);
  for (int it = 0; it < nsteps; ++it) {
    switch (scheme) {
      case KGO_PIRK2: kgo_pirk2_generic(&ops, &b, s->dt); break;
      case KGO_PIRK3: kgo_pirk3_generic(&ops, &b, s->dt); break;
      case KGO_RK3: kgo_rk3_generic(&ops, &b, s->dt); break;
      default: kgo_rk4_generic(&ops, &b, s->dt); break;
    }
    s->t = s->t + s->dt;  /* :243, :308, :385, :446 */
  }
}
void kgo_pirk2(kgo_sim* s, int nsteps) { kgo_step(s, KGO_PIRK2, nsteps); }
void kgo_pirk3(kgo_sim* s, int nsteps) { kgo_step(s, KGO_PIRK3, nsteps); }
void kgo_rk3(kgo_sim* s, int nsteps) { kgo_step(s, KGO_RK3, nsteps); }
void kgo_rk4(kgo_sim* s, int nsteps) { kgo_step(s, KGO_RK4, nsteps); }

kgo_sim* kgo_create(int Nx, int Ny, int Nz, int g, double xmin, double xmax, int metric, double a, double M_bh,
                    double m, int bc_rmax, double CFL, double dt) {
  if (Nx < 4 || Ny < 4 || Nz < 4 || g < 1 || !(xmax > xmin)) return NULL;
  kgo_sim* s = (kgo_sim*)calloc(1, sizeof(kgo_sim));
  s->Nx = Nx;
  s->Ny = Ny;
  s->Nz = Nz;
  s->g = g;
  s->ni = Nx + 2 * g + 1;
  s->nj = Ny + 2 * g + 1;
  s->nk = Nz + 1;
  s->npts = (int64_t)s->ni * s->nj * s->nk;
  s->xmin = xmin;
  s->xmax = xmax;
  s->metric = metric;
  s->a = a;
  s->M_bh = M_bh;
  s->m = m;
  s->bc_rmax = bc_rmax;
  s->order = 2;
  s->CFL = CFL;
  s->x = alloc0(s->ni);
  s->y = alloc0(s->nj);
  s->z = alloc0(s->nk);
  grid_build(s);
  const int64_t n2 = (int64_t)s->ni * s->nj;
  s->alpha = alloc0(n2);
  s->beta_r = alloc0(n2);
  s->dx_beta = alloc0(n2);
  s->sqrtg = alloc0(n2);
  s->g_rr = alloc0(n2);
  s->g_rp = alloc0(n2);
  s->g_tt = alloc0(n2);
  s->g_pp = alloc0(n2);
  s->g_dn = alloc0(16 * n2);
  s->g_up = alloc0(16 * n2);
  metric_build(s);
  s->dt = dt > 0.0 ? dt : kgo_reference_dt(Nx, Ny, Nz, g, xmin, xmax, CFL);
  s->t = 0.0;
  const int64_t n5 = 5 * s->npts;
  s->u = alloc0(n5);
  s->u_p = alloc0(n5);
  s->u_stage = alloc0(n5);   /* zero: the reference leaves it unset; see :146-155 */
  s->u_stage_2 = alloc0(n5);
  s->rhs_u = alloc0(n5);
  s->k1 = alloc0(n5);
  s->k2 = alloc0(n5);
  s->k3 = alloc0(n5);
  s->k4 = alloc0(n5);
  for (int q = 0; q < 8; ++q) s->tmp[q] = alloc0(s->npts);
  return s;
}

void kgo_destroy(kgo_sim* s) {
  if (!s) return;
  double* ptrs[] = {s->x, s->y, s->z, s->alpha, s->beta_r, s->dx_beta, s->sqrtg, s->g_rr, s->g_rp, s->g_tt,
                    s->g_pp, s->g_dn, s->g_up, s->u, s->u_p, s->u_stage, s->u_stage_2, s->rhs_u, s->k1,
                    s->k2, s->k3, s->k4};
  for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) free(ptrs[q]);
  for (int q = 0; q < 8; ++q) free(s->tmp[q]);
  free(s);
}

double* kgo_field(kgo_sim* s, int which) {
  switch (which) {
    case 0: return s->u;
    case 1: return s->u_p;
    case 2: return s->u_stage;
    case 3: return s->u_stage_2;
    case 4: return s->rhs_u;
    default: return NULL;
  }
}

double* kgo_table(kgo_sim* s, int which) {
  switch (which) {
    case 0: return s->alpha;
    case 1: return s->beta_r;
    case 2: return s->dx_beta;
    case 3: return s->sqrtg;
    case 4: return s->g_rr;
    case 5: return s->g_rp;
    case 6: return s->g_tt;
    case 7: return s->g_pp;
    case 8: return s->g_dn;
    case 9: return s->g_up;
    case 10: return s->x;
    case 11: return s->y;
    case 12: return s->z;
    default: return NULL;
  }
}

void kgo_info(const kgo_sim* s, int64_t* o, double* d) {
  o[0] = s->Nx; o[1] = s->Ny; o[2] = s->Nz; o[3] = s->g;
  o[4] = s->ni; o[5] = s->nj; o[6] = s->nk; o[7] = s->npts;
  d[0] = s->dx; d[1] = s->dy; d[2] = s->dz; d[3] = s->dt;
  d[4] = s->t; d[5] = s->xmin; d[6] = s->xmax; d[7] = s->m;
}
