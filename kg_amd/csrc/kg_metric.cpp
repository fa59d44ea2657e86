// kg_metric.cpp -- grid and metric coefficient tables, built once on the host
// at kg_create (the reference builds them once in its constructor:
// MoL/lines_method_lib.f90:89-127).
//
// Grid: MoL/mesh_refinement_lib.f90:100-132 with nproc = 1.
// Metric: MoL/geometry.f90:92-352 -- g_dn per metric, the theta-ghost row
// copy (boundary_theta_dn/up, :371-377, :398-404), sqrt(gamma) (:267), the
// closed-form inverse (:286-320), lapse, shift and spatial inverse
// (:332-342), d_r beta^r (:344).  Every expression keeps the reference's
// association; x**2 is x*x as Fortran evaluates integer powers.
#include <cmath>
#include <cstring>
#include <vector>

#include "kg_internal.h"

namespace kg {

int build_grid(int Nx, int Ny, int Nz, int g, double xmin, double xmax, double CFL, double dt_override,
               HostGrid* o) {
  o->Nx = Nx;
  o->Ny = Ny;
  o->Nz = Nz;
  o->g = g;
  o->ni = Nx + 2 * g + 1;
  o->nj = Ny + 2 * g + 1;
  o->nk = Nz + 1;
  o->ld = (o->ni + 15) / 16 * 16;
  // mesh_refinement_boundary, rank 0 of 1 (:109-117)
  const double span = (xmax - xmin) / 1.0;
  const double x_lo = xmin + 0.0 * span, x_hi = xmin + 1.0 * span;
  o->xmin = x_lo;
  o->xmax = x_hi;
  o->dx = (x_hi - x_lo) / (double)Nx;
  o->dy = std::acos(-1.0) / (double)(Ny - 1);
  o->dz = 2.0 * std::acos(-1.0) / (double)(Nz - 1);
  if (dt_override > 0.0) {
    o->dt = dt_override;
  } else {
    // diagnostic_lib.f90:101-104 (dxx is the coordinate difference
    // x(-g+1) - x(-g); exp(xmin) multiplies the angular spacings)
    const double x0 = x_lo + 0.0 * o->dx, x1 = x_lo + 1.0 * o->dx;
    const double dxx = x1 - x0;
    const double e = std::exp(xmin);
    double mn = dxx;
    const double c1 = e * o->dy;
    const double c2 = e * std::sin(o->dy / 2.0) * o->dz;
    if (c1 < mn) mn = c1;
    if (c2 < mn) mn = c2;
    o->dt = CFL * mn;
  }
  return 0;
}

namespace {

struct G4 {
  double v[4][4];
};

G4 metric_dn(int metric, double a, double M, double x, double y) {
  G4 g{};
  if (metric == 0) {  // Minkowski (geometry.f90:104-110)
    const double sy = std::sin(y);
    g.v[0][0] = -1.0;
    g.v[1][1] = 1.0;
    g.v[2][2] = x * x;
    g.v[3][3] = x * x * (sy * sy);
  } else if (metric == 1) {  // Kerr_Schild (geometry.f90:119-160)
    const double cy = std::cos(y), sy = std::sin(y);
    const double c2 = cy * cy, s2 = sy * sy, a2 = a * a;
    g.v[0][0] = -1.0 + 2.0 * M * x / (x * x + a2 * c2);
    g.v[0][1] = 2.0 * M * x / (x * x + a2 * c2);
    g.v[0][3] = -2.0 * a * M * x * s2 / (x * x + a2 * c2);
    g.v[1][1] = 1.0 + 2.0 * M * x / (x * x + a2 * c2);
    g.v[1][3] = -a * (1.0 + 2.0 * M * x / (x * x + a2 * c2)) * s2;
    g.v[2][2] = x * x + a2 * c2;
    const double q = a2 + x * x;
    g.v[3][3] = ((q * q) - a2 * (-2.0 * M * x + a2 + x * x) * s2) * s2 / (x * x + a2 * c2);
  } else if (metric == 2) {  // MKS with h = 0 (geometry.f90:162-212)
    const double h = 0.0;
    const double r = std::exp(x);
    const double th = y + 0.5 * h * std::sin(2.0 * y);
    const double dth = 1.0 + h * std::cos(2.0 * y);
    const double ct = std::cos(th), st = std::sin(th);
    const double c2 = ct * ct, s2 = st * st, a2 = a * a;
    g.v[0][0] = -1.0 + 2.0 * M * r / (r * r + a2 * c2);
    g.v[0][1] = (2.0 * M * r / (r * r + a2 * c2)) * r;
    g.v[0][3] = -2.0 * a * M * r * s2 / (r * r + a2 * c2);
    g.v[1][1] = (1.0 + 2.0 * M * r / (r * r + a2 * c2)) * (r * r);
    g.v[1][3] = (-a * (1.0 + 2.0 * M * r / (r * r + a2 * c2)) * s2) * r;
    g.v[2][2] = (r * r + a2 * c2) * (dth * dth);
    const double q = a2 + r * r;
    g.v[3][3] = ((q * q) - a2 * (-2.0 * M * r + a2 + r * r) * s2) * s2 / (r * r + a2 * c2);
  } else {  // Schw_EF (geometry.f90:213-246)
    const double sy = std::sin(y);
    g.v[0][0] = -(1.0 - 2.0 * M / x);
    g.v[0][1] = 2.0 * M / x;
    g.v[1][1] = 1.0 + 2.0 * M / x;
    g.v[2][2] = x * x;
    g.v[3][3] = x * x * (sy * sy);
  }
  // symmetric lower triangle (geometry.f90:260-265)
  g.v[1][0] = g.v[0][1];
  g.v[2][0] = g.v[0][2];
  g.v[3][0] = g.v[0][3];
  g.v[2][1] = g.v[1][2];
  g.v[3][1] = g.v[1][3];
  g.v[3][2] = g.v[2][3];
  return g;
}

G4 inverse(const G4& d) {
  // geometry.f90:286-320
  const double g00 = d.v[0][0], g01 = d.v[0][1], g03 = d.v[0][3], g11 = d.v[1][1], g13 = d.v[1][3],
               g22 = d.v[2][2], g33 = d.v[3][3];
  const double det3 = (g33 * g11 - g13 * g13) * g00 - g33 * (g01 * g01) + 2.0 * g13 * g03 * g01 - g11 * (g03 * g03);
  const double idetg = 1.0 / (det3 * g22);
  G4 u{};
  u.v[0][0] = g22 * (g11 * g33 - g13 * g13) * idetg;
  u.v[0][1] = g22 * (g13 * g03 - g01 * g33) * idetg;
  u.v[0][3] = g22 * (g13 * g01 - g11 * g03) * idetg;
  u.v[1][1] = g22 * (g33 * g00 - g03 * g03) * idetg;
  u.v[1][3] = g22 * (-g13 * g00 + g03 * g01) * idetg;
  u.v[2][2] = ((g33 * g11 - g13 * g13) * g00 - g33 * (g01 * g01) + 2.0 * g13 * g03 * g01 - g11 * (g03 * g03)) * idetg;
  u.v[3][3] = g22 * (g11 * g00 - g01 * g01) * idetg;
  u.v[1][0] = u.v[0][1];
  u.v[3][0] = u.v[0][3];
  u.v[3][1] = u.v[1][3];
  return u;
}

}  // namespace

int build_tables(const HostGrid& G, int metric, double a, double M_bh, double m, int fd4, double* tab) {
  const int ni = G.ni, nj = G.nj, ld = G.ld, g = G.g;
  const size_t n2 = (size_t)nj * ld;
  std::memset(tab, 0, sizeof(double) * kNumTables * n2);
  std::vector<double> xs(ni), ys(nj);
  for (int ii = 0; ii < ni; ++ii) xs[ii] = G.xmin + (double)((ii - g) + g) * G.dx;  // :130
  for (int jj = 0; jj < nj; ++jj) ys[jj] = -G.dy / 2.0 + (double)(jj - g) * G.dy;  // :131
  std::vector<double> beta_row(ni);
  for (int jj = 0; jj < nj; ++jj) {
    // boundary_theta_dn/up: the last theta-ghost row j = Ny+g takes the
    // values of row Ny+g-1 (the negated (0,2),(1,2) entries are zero and
    // do not enter any table), for g_dn and again for g_up.
    const int jsrc = (jj == nj - 1) ? nj - 2 : jj;
    for (int ii = 0; ii < ni; ++ii) {
      const G4 dn = metric_dn(metric, a, M_bh, xs[ii], ys[jsrc]);
      const G4 up = inverse(dn);
      const double sqrtg = std::sqrt(dn.v[2][2] * (dn.v[1][1] * dn.v[3][3] - dn.v[1][3] * dn.v[1][3]));
      const double alpha = std::sqrt(-1.0 / up.v[0][0]);
      const double a2 = alpha * alpha;
      double beta[4] = {0.0, a2 * up.v[0][1], a2 * up.v[0][2], a2 * up.v[0][3]};
      const double grr = up.v[1][1] + beta[1] * beta[1] / a2;
      const double grp = up.v[1][3] + beta[1] * beta[3] / a2;
      const double gtt = up.v[2][2] + beta[2] * beta[2] / a2;
      const double gpp = up.v[3][3] + beta[3] * beta[3] / a2;
      const size_t p = (size_t)jj * ld + ii;
      const double asg = alpha * sqrtg;
      tab[T_ALPHA * n2 + p] = alpha;
      tab[T_BETA * n2 + p] = beta[1];
      tab[T_SQRTG * n2 + p] = sqrtg;
      tab[T_ASG * n2 + p] = asg;
      tab[T_GRR * n2 + p] = grr;
      tab[T_GRP * n2 + p] = grp;
      tab[T_ASGTT * n2 + p] = asg * gtt;
      tab[T_GPP * n2 + p] = gpp;
      tab[T_AM2 * n2 + p] = alpha * (m * m);
      tab[T_GTT * n2 + p] = gtt;
      tab[T_VOUT * n2 + p] = alpha * std::sqrt(grr) - beta[1];
      beta_row[ii] = beta[1];
    }
    // dx_beta = Re dx1(beta^r) (geometry.f90:344; stencil finite_differences.f90:36-42);
    // with the fd4 set, x_4 (:124-151), as accurate as the derivatives it multiplies
    double* d = tab + T_DXBETA * n2 + (size_t)jj * ld;
    const double* f = beta_row.data();
    if (fd4) {
      const double res = 1.0 / (12.0 * G.dx);
      const int N = ni - 1;
      d[0] = (-25.0 * f[0] + 48.0 * f[1] - 36.0 * f[2] + 16.0 * f[3] - 3.0 * f[4]) * res;
      d[1] = (-3.0 * f[0] - 10.0 * f[1] + 18.0 * f[2] - 6.0 * f[3] + 1.0 * f[4]) * res;
      for (int ii = 2; ii <= N - 2; ++ii) d[ii] = (f[ii - 2] - 8.0 * f[ii - 1] + 8.0 * f[ii + 1] - f[ii + 2]) * res;
      d[N - 1] = (+3.0 * f[N] + 10.0 * f[N - 1] - 18.0 * f[N - 2] + 6.0 * f[N - 3] - 1.0 * f[N - 4]) * res;
      d[N] = (+25.0 * f[N] - 48.0 * f[N - 1] + 36.0 * f[N - 2] - 16.0 * f[N - 3] + 3.0 * f[N - 4]) * res;
    } else {
      const double res = 0.50 / G.dx;
      const int lo = 0, hi = ni - 1;
      d[lo] = (-3.0 * f[lo] + 4.0 * f[lo + 1] - f[lo + 2]) * res;
      for (int ii = lo + 1; ii <= hi - 1; ++ii) d[ii] = (-f[ii - 1] + f[ii + 1]) * res;
      d[hi] = (+3.0 * f[hi] - 4.0 * f[hi - 1] + f[hi - 2]) * res;
    }
  }
  return 0;
}

}  // namespace kg
