// kg_internal.h -- shared definitions of the HIP engine (host and device).
//
// Device layout of one real system: 5 fields, each [nk][nj][ld] doubles,
// r fastest, ld = ni rounded up to 16 doubles (128-byte rows).  The padded
// index ranges follow the reference: ii = i+g in [0,ni), jj = j+g in [0,nj),
// kk = k in [0,nk) (MoL/lines_method_lib.f90:135-138).
//
// Metric coefficients are 2-D tables [nj][ld]: every metric of the reference
// depends on (r, theta) only (MoL/geometry.f90:104-247).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

struct kg_handle;

namespace kg {

constexpr int kNumFields = 5;  // phi, psi_r, psi_theta, psi_phi, Pi
constexpr int kNumTables = 12;

// table slots
enum Table : int {
  T_ALPHA = 0,   // alpha
  T_BETA = 1,    // beta^r
  T_DXBETA = 2,  // d_r beta^r = Re dx1(beta)         (geometry.f90:344)
  T_SQRTG = 3,   // sqrt(gamma)                       (geometry.f90:267)
  T_ASG = 4,     // alpha*sqrtg (the reference's product, rounded once)
  T_GRR = 5,     // gamma^{rr}
  T_GRP = 6,     // gamma^{r phi}
  T_ASGTT = 7,   // (alpha*sqrtg)*gamma^{theta theta}
  T_GPP = 8,     // gamma^{phi phi}
  T_AM2 = 9,     // alpha*m**2
  T_GTT = 10,    // gamma^{theta theta} (diagnostics: currents_ADM, density)
  T_VOUT = 11    // alpha*sqrt(gamma^{rr}) - beta^r: outgoing radial characteristic speed
                 // (the Sommerfeld outer boundary)
};

struct Geom {
  int ni, nj, nk, ld;     // padded extents, row pitch (doubles)
  int Nx, Ny, Nz, g;
  int64_t plane;          // ld*nj
  int64_t fstride;        // plane*nk rounded up: size of one field of one system
  int64_t sys_stride;     // offset between systems inside a field buffer
  double rr, rt, rp;      // 0.5/dx, 0.5/dy, 0.5/dz
  double r4r, r4t, r4p;   // fd4 set: 1/(12 dx), 1/(12 dy), 1/(12 dz)
  double rar;             // fd4 set: 1/dx of advec_x_4
  double dt;
  int metric_ef;          // 1: Schw_EF form of rhs_pi (lines_method_lib.f90:738-743)
  int bc_inflow;          // 1: 'inflow' at r_max
  int fd4;                // 1: the 4th-order operator set
  int bc_somm;            // 1: Sommerfeld at r_max (the outer column evolves; no xR fill)
  double r_out;           // x of the outer column i = Nx+g (Sommerfeld's 1/r)
};

// Host-side table builder (kg_metric.cpp): fills 10 tables [nj][ld] and the
// grid spacings.  Returns 0 on success.
struct HostGrid {
  int Nx, Ny, Nz, g, ni, nj, nk, ld;
  double xmin, xmax, dx, dy, dz, dt;
};
int build_grid(int Nx, int Ny, int Nz, int g, double xmin, double xmax, double CFL, double dt_override,
               HostGrid* out);
int build_tables(const HostGrid& G, int metric, double a, double M_bh, double m, int fd4,
                 double* tables /*kNumTables*nj*ld*/);

// Integrator stage kinds (one kernel launch each).
enum StageMode : int {
  // RK4 keeps the increments k_s = R*dt and forms each stage state
  // u = u_p + c*k_s (and its ghost fill) where it reads it
  M_RK4_S1 = 0,   // k1 = R(P)dt
  M_RK4_S2 = 1,   // k2 = R(P + 0.5k1)dt, stored as Y = P + 0.5k2 and C = k1 + 2k2
  M_RK4_S3 = 2,   // k3 = R(Y)dt, Y = P + 0.5k2 stored (and ghost-filled) by stage 2
  M_RK4_S4 = 3,   // X = P + ((C + 2k3) + R(P + k3)dt)/6, C = k1 + 2k2
  M_RK3_S1 = 4,   // U1 = P + R(P)dt
  M_RK3_S2 = 5,   // U2 = 0.75P + 0.25(U1 + R(U1)dt)
  M_RK3_S3 = 6,   // P = P/3 + 2(U2 + R(U2)dt)/3
  M_PIRK_1A = 7,  // S_c = u_c + dt R_c(u), c = 1..4
  M_PIRK_1B = 8,  // S5' = u5 + dt R_pi(S1..4, S5)
  M_PIRK2_2 = 9,  // u = 0.5((u + S) + dt R(S))
  M_PIRK3_2A = 10,  // S2_c = 0.75u_c + 0.25(S_c + dt R_c(S)), c = 1..4
  M_PIRK3_2B = 11,  // S2_5' = 0.75u5 + 0.25(S5 + dt R_pi(S2_1..4, S2_5))
  M_PIRK3_3 = 12,   // u = (1/3)u + (2/3)(S2 + dt R(S2))
  M_NUM_MODES = 13
};

// Roles of the buffers a stage reads and writes.  Every state buffer holds
// 5 (or 6) field slots `per` doubles apart, so a role is its base address;
// field 4 (Pi) may sit in slot 4 or in the extra slot 5 (the PIRK stage
// arrays keep two Pi slots).
enum Role : int {
  R_IN = 0,    // fields the RHS is evaluated on (RK4 stages 2 and 4: u_p; stage 3: Y = u_p + 0.5*k2)
  R_KIN = 1,   // RK4 stages 2 and 4: increment k the stage state u_p + c*k is formed from
  R_BASE = 2,  // u_p / P / u of the combine
  R_AUX = 3,   // RK4 k1 + 2*k2 (written by stage 2, read by stage 4), PIRK3 stage state (stage 2)
  R_OUT = 4,   // stage output
  R_NUM = 5
};

struct StageArgs {
  Geom G;
  const double* tab;        // 10 tables, each nj*ld
  double* rb[R_NUM];        // base of each role's buffer (slot 0)
  int ri4[R_NUM];           // slot of field 4 in each role (4 or 5)
  int64_t per;              // doubles between two field slots
  int i_lo, i_hi, j_lo, j_hi, k_lo, k_hi;  // computed box, [lo, hi)
  int kc;                   // phi planes per workgroup
  int tj;                   // theta rows per workgroup (4 or 8)
  int nti, ntj, ntk;        // tile counts (set by launch_stage)
  int nts;                  // theta strips of the fd4 set's last r columns (0: none; launch_stage)
  int nsys;                 // independent real systems sharing the grid
};

struct FillArgs {
  Geom G;
  double* u[5];
  const double* u_p[5];
};

// Device buffers hold fields as [field][system][nk][nj][ld]; a field pointer
// addresses system 0 and system s sits sys_stride doubles further.

// launchers (kg_kernels.hip)
hipError_t launch_stage(int mode, const StageArgs& a, hipStream_t s);
// Sommerfeld outer column of a stage (G.bc_somm): every stage but PIRK 1A/2A
// (whose outer column the stage kernel computes with its edge rows) runs it
// after the stage kernel with the same arguments.
bool somm_stage(int mode);
int64_t somm_items(const StageArgs& a);
int somm_block();  // work items per workgroup of the Sommerfeld column kernels
hipError_t launch_somm(int mode, const StageArgs& a, hipStream_t s);
// batched: grid q owns blocks [bstart[q], bstart[q+1]) of somm_block() items
hipError_t launch_somm_batch(int mode, const StageArgs* descs, const int* bstart, int ngrids, int nblocks, int fd4,
                             hipStream_t s);
hipError_t launch_fill(const FillArgs& a, hipStream_t s, int phase /*0 x,1 y,2 z*/, int nsys);
// Batched launches: one launch covers every grid of descs[0..ngrids).
// Stage: grid q owns workgroups [tstart[q], tstart[q+1]), its tiles times its
// systems (finalize_tiles sets the tile counts); any stage of any scheme.
void finalize_tiles(StageArgs& a);
// theta rows per workgroup a 2nd-order stage runs with for the tuned `tj`
// (RK4 stages 2 and 4 and the PIRK stages of Pi alone always 4)
int stage_tj(int mode, int tj);
// nwg > 0: a persistent launch of nwg workgroups (a multiple of 8) that march
// the `total` tiles in turn; otherwise one workgroup per tile
hipError_t launch_stage_batch(int mode, const StageArgs* descs, const int* tstart, int ngrids, int total, int nwg,
                              int tj, int ef, int fd4, hipStream_t s);
// Fill pass `phase`: grid q owns blocks [bstart[q], ...) of 256 work items, nitems[q] items.
int64_t fill_work_items(const Geom& G, int nsys, int phase);
// wide = 1 when some grid's pass has 2^32 items or more (64-bit index arithmetic)
hipError_t launch_fill_batch(const FillArgs* descs, const int* bstart, int ngrids, const int64_t* nitems,
                             int nblocks, int phase, int wide, hipStream_t s);
// dense [sys][5][nk][nj][ni] (device) <-> pitched fields; to_dev = 1: dense -> fields
// fortran = 1: the host array is the reference's u(5, i, j, k) [parts fastest]
hipError_t launch_pack(const Geom& G, double* dense, double* const fields[5], int nsys, int to_dev, hipStream_t s,
                       int fortran = 0);

// What the spectral projector needs from a handle: its device, stream, grid
// and the current state's field buffers (these move between steps for RK4).
struct HandleView {
  const ::kg_handle* handle;
  int device;
  hipStream_t stream;
  Geom G;
  HostGrid hg;
  int nsys;
  double* u[5];
  const double* tab;  // kNumTables metric tables [nj][ld]
  double m;           // field mass
  int metric;         // KG_METRIC_*
};
int handle_view(const ::kg_handle* h, HandleView* v);
// Device scratch owned by the handle (grown on demand, freed by kg_destroy);
// work on the handle's stream only.
hipError_t handle_scratch(const ::kg_handle* h, size_t bytes, void** out);

// Volume integral (integral.f90:646-669) of an integrand of the state, on the
// view's stream, synchronous.  quantity 0..4: that field; 5: density*sqrtg;
// 6: (density*alpha)*sqrtg; 7: (m*|phi|**2)*sqrtg.  out = real, imaginary.
hipError_t diag_integrate(const HandleView& v, int quantity, double out[2]);

// Special functions for the spectral tables (kg_special.cpp, host).
double ylm_norm_plm(int l, int m, double x);  // N_lm P_l^m(x), Condon-Shortley phase
double sph_jl(int l, double x);
std::vector<double> jl_zeros(int lmax, int nk);  // [l][n]: the first nk positive zeros of j_l

}  // namespace kg
