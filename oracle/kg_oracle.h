/*
 * kg_oracle.h -- CPU FP64 restatement of the Klein-Gordon method-of-lines
 * step of FlavioRosales/KG.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker for the HIP engine in kg_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * It follows the reference's whole-array Fortran statements one by one
 * (temporaries, operand order and left-to-right association kept), on real
 * FP64 arrays.  A complex reference state is handled as two real systems
 * (real and imaginary parts), which is exact because every coefficient on
 * the path is real (MoL/lines_method_lib.f90:671-753).
 *
 * Array layout: dense padded 3-D arrays, r fastest:
 *   a[(kk*nj + jj)*ni + ii],  ii = i+g in [0,ni), jj = j+g in [0,nj), kk = k in [0,nk)
 * with ni = Nx+2g+1, nj = Ny+2g+1, nk = Nz+1 (MoL/lines_method_lib.f90:135-138).
 * A 5-field state is 5 such arrays back to back (field 0 = phi, 1 = psi_r,
 * 2 = psi_theta, 3 = psi_phi, 4 = Pi).
 *
 * Metric coefficients are stored as 2-D (ni x nj) tables: every metric the
 * reference implements depends on x (r) and y (theta) only
 * (MoL/geometry.f90:104-247), so the k index of the reference's 3-D arrays
 * carries no information.
 *
 * Parity status per function (pins live in tests/):
 *   kgo_dx1/kgo_dx2/kgo_dx3/kgo_advec_x: bit-exact against the reference's own
 *     tools/finite_differences.f90 compiled from source (oracle/_ref), plus
 *     polynomial exactness.
 *   kgo_metric_build: closed-form Schwarzschild/Kerr-Schild identities and the
 *     reference's own inverse-metric self-check (geometry.f90:409-492).
 *   kgo_fill: index-tag maps checked against the closed-form reflection maps.
 *   kgo_rk3/kgo_rk4/kgo_pirk2/kgo_pirk3: amplification polynomials on linear
 *     scalar problems through the same stage code (kgo_*_generic), temporal
 *     convergence order on the full operator, and exact symmetry invariants.
 */
#ifndef KG_ORACLE_H
#define KG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { KGO_METRIC_MINKOWSKI = 0, KGO_METRIC_KERR_SCHILD = 1, KGO_METRIC_MKS = 2, KGO_METRIC_SCHW_EF = 3 };
/* KGO_BC_SOMMERFELD has no counterpart in the reference (which has 'outflow'
 * and 'inflow' only, lines_method_lib.f90:522-535): the outer column
 * ii = ni-1 is not filled, and R_Pi there is the radiative condition
 * -(v * (D_r Pi + Pi / r)), v = alpha*sqrt(gamma^rr) - beta^r, D_r the
 * operator set's last-row one-sided stencil, r = x(ni-1); the other fields
 * keep their right-hand sides (edge rows) there. */
enum { KGO_BC_OUTFLOW = 0, KGO_BC_INFLOW = 1, KGO_BC_SOMMERFELD = 2 };
enum { KGO_PIRK2 = 0, KGO_PIRK3 = 1, KGO_RK3 = 2, KGO_RK4 = 3 };

typedef struct kgo_sim {
  /* grid (mesh_refinement_lib.f90:100-147, nproc = 1) */
  int Nx, Ny, Nz, g;
  int ni, nj, nk;
  int64_t npts;        /* ni*nj*nk */
  double xmin, xmax, dx, dy, dz;
  double *x, *y, *z;   /* x[ii], y[jj], z[kk] */
  /* metric (geometry.f90:92-352), 2-D tables [jj*ni + ii] */
  int metric;
  double a, M_bh, m;
  double *alpha, *beta_r, *dx_beta, *sqrtg;
  double *g_rr, *g_rp, *g_tt, *g_pp;  /* gamma_up(1,1),(1,3),(2,2),(3,3) */
  /* full 4x4 g_dn and g_up, 2-D tables, index [(mu*4+nu)*ni*nj + jj*ni + ii] */
  double *g_dn, *g_up;
  /* time stepping */
  int bc_rmax;
  int order;           /* operator set: 2 = the reference's live stencils, 4 = its fd4 set */
  int plant;           /* KGO_PLANT_*: a deliberate mistake for the tests' mutation checks (0: none) */
  double t, dt, CFL;
  /* state: 5 fields each, npts doubles per field */
  double *u, *u_p, *u_stage, *u_stage_2, *rhs_u, *k1, *k2, *k3, *k4;
  /* scratch */
  double *tmp[8];
} kgo_sim;

/* Creates a simulation.  dt <= 0 selects the reference formula
 * (diagnostic_lib.f90:101-104).  Returns NULL on invalid input. */
kgo_sim* kgo_create(int Nx, int Ny, int Nz, int g, double xmin, double xmax,
                    int metric, double a, double M_bh, double m,
                    int bc_rmax, double CFL, double dt);
void kgo_destroy(kgo_sim* s);

/* accessors for ctypes */
double* kgo_field(kgo_sim* s, int which);   /* 0=u 1=u_p 2=u_stage 3=u_stage_2 4=rhs_u */
double* kgo_table(kgo_sim* s, int which);   /* 0 alpha 1 beta_r 2 dx_beta 3 sqrtg 4 g_rr 5 g_rp 6 g_tt 7 g_pp 8 g_dn 9 g_up 10 x 11 y 12 z */
void kgo_info(const kgo_sim* s, int64_t* out8, double* outd8);

/* host threads the independent-point loops run on (OpenMP; 1 without it) */
int kgo_threads(void);

/* reference-formula dt (diagnostic_lib.f90:101-104) */
double kgo_reference_dt(int Nx, int Ny, int Nz, int g, double xmin, double xmax, double CFL);

/* Stencils on full padded arrays (finite_differences.f90). res = 0.5/h. */
void kgo_dx1(const kgo_sim* s, const double* f, double* df);
void kgo_dx2(const kgo_sim* s, const double* f, double* df);
void kgo_dx3(const kgo_sim* s, const double* f, double* df);
void kgo_advec_x(const kgo_sim* s, const double* f, const double* beta2d, double* df);
/* Same stencils on caller-described arrays (for the reference cross-check). */
void kgo_dx1_raw(int ni, int nj, int nk, double h, const double* f, double* df);
void kgo_dx2_raw(int ni, int nj, int nk, double h, const double* f, double* df);
void kgo_dx3_raw(int ni, int nj, int nk, double h, const double* f, double* df);
void kgo_advec_x_raw(int ni, int nj, int nk, double h, const double* f, const double* beta3d, double* df);

/* The 4th-order set of finite_differences.f90 (x_4 :124-151, y_4 :241-268,
 * z_4 :340-367, advec_x_4 :922-957) on dense real arrays. */
void kgo_dx1_4_raw(int ni, int nj, int nk, double h, const double* f, double* df);
void kgo_dx2_4_raw(int ni, int nj, int nk, double h, const double* f, double* df);
void kgo_dx3_4_raw(int ni, int nj, int nk, double h, const double* f, double* df);
void kgo_advec_x_4_raw(int ni, int nj, int nk, double h, const double* f, const double* beta3d, double* df);
/* Selects the operator set of the RHS (2 or 4; default 2).  With 4 every
 * first derivative of lines_method_lib.f90:671-753 uses the fd4 set, and so
 * does the metric's d_r beta^r.  Returns 0, or -1 for another order. */
int kgo_set_order(kgo_sim* s, int order);

/* Deliberate mistakes, planted so the tests can show that their pins catch
 * them (mutation checks; 0 = the restatement as written).  Each is a
 * plausible slip in transcribing lines_method_lib.f90:671-753 or
 * geometry.f90:112-160. */
enum {
  KGO_PLANT_NONE = 0,
  KGO_PLANT_DROP_PI_DXBETA = 1,    /* R_Pi without vars(5)*dx_beta (:735) */
  KGO_PLANT_SWAP_PHI_FLUX = 2,     /* phi flux with gamma^{r phi} and gamma^{phi phi} swapped (:731) */
  KGO_PLANT_PSI2_ALPHA_PI = 3,     /* R_psi_theta = D_theta(alpha Pi), beta psi_r dropped (:700) */
  KGO_PLANT_DROP_PSI1_DXBETA = 4,  /* R_psi_r without dx_beta*vars(2) (:689) */
  KGO_PLANT_GDN13_NO_SIN2 = 5,     /* g_dn(1,3) without its sin^2(theta) factor (geometry.f90:130) */
  KGO_PLANT_ADVECT_PI = 6,         /* A_r of Pi instead of Pi*sqrtg in R_Pi (:733) */
  KGO_PLANT_MASS_SIGN = 7,         /* + alpha m^2 phi in R_Pi (:736) */
  KGO_PLANT_NUM = 8
};
/* Selects a planted mistake (rebuilding the metric tables); returns 0, or -1
 * for an unknown code. */
int kgo_plant(kgo_sim* s, int code);

/* RHS pieces (lines_method_lib.f90:671-753). vars = 5-field state. */
void kgo_rhs_phi(kgo_sim* s, double* rhs, const double* vars);
void kgo_rhs_psi1(kgo_sim* s, double* rhs, const double* vars);
void kgo_rhs_psi2(kgo_sim* s, double* rhs, const double* vars);
void kgo_rhs_psi3(kgo_sim* s, double* rhs, const double* vars);
void kgo_rhs_pi(kgo_sim* s, double* rhs, const double* vars);
void kgo_rhs_all(kgo_sim* s, double* rhs5, const double* vars);

/* ghost fill of a 5-field state (lines_method_lib.f90:313-330,507-597) */
void kgo_fill(kgo_sim* s, double* u, const double* u_p);

/* integrators (lines_method_lib.f90:211-474); advance s->u by nsteps */
void kgo_pirk2(kgo_sim* s, int nsteps);
void kgo_pirk3(kgo_sim* s, int nsteps);
void kgo_rk3(kgo_sim* s, int nsteps);
void kgo_rk4(kgo_sim* s, int nsteps);
void kgo_step(kgo_sim* s, int scheme, int nsteps);

/* The integrators are written once over an operator table, so the stage
 * arithmetic can be pinned on small linear systems as well as run with the
 * Klein-Gordon operator.  n = values per field; a state is 5*n doubles.
 *   rhs14: writes fields 1..4 (phi, psi_r, psi_theta, psi_phi) of R(vars)
 *   rhs5:  writes field 5 (Pi) of R(vars) into out (n doubles)
 *   fill:  ghost fill of u (u_p is the start-of-step state), may be NULL */
typedef struct kgo_ops {
  void (*rhs14)(void* ctx, const double* vars, double* out5);
  void (*rhs5)(void* ctx, const double* vars, double* out1);
  void (*fill)(void* ctx, double* u, const double* u_p);
  void* ctx;
  int64_t n;
} kgo_ops;
typedef struct kgo_bufs {
  double *u, *u_p, *u_stage, *u_stage_2, *rhs_u, *k1, *k2, *k3, *k4;
} kgo_bufs;
void kgo_pirk2_generic(const kgo_ops* ops, kgo_bufs* b, double dt);
void kgo_pirk3_generic(const kgo_ops* ops, kgo_bufs* b, double dt);
void kgo_rk3_generic(const kgo_ops* ops, kgo_bufs* b, double dt);
void kgo_rk4_generic(const kgo_ops* ops, kgo_bufs* b, double dt);

/* Index-tag probe of the fill maps: writes, for every padded point of one
 * field, the flat index of the evolved point its value is copied from after
 * kgo_fill.  A source whose i index is -g stands for the xL extrapolation at
 * that row; -1 stands for an inflow value. */
void kgo_fill_source_map(kgo_sim* s, int64_t* src);

#ifdef __cplusplus
}
#endif
#endif
