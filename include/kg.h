/*
 * kg.h -- C ABI of the MI355X engine for the Klein-Gordon method-of-lines
 * step of FlavioRosales/KG (library: kg_amd/libkg_hip.so).
 *
 * The reference path sits behind the Fortran type-bound interface of
 * `type(lines_method)` (MoL/lines_method_lib.f90:19-44): the driver calls
 * `MoL%pirk2()` once per step (main/evolve.f90:25) on the state
 * `MoL%u(nvars, iL-g:iR+g, jL-g:jR+g, kL:kR)`, complex(8), column-major.
 * These entry points replace that interface one for one; a Fortran driver
 * binds them through ISO_C_BINDING (kg_amd/fortran/kg_hip_mod.f90).
 *
 * Conventions
 *  - Every function returns a kg_status; KG_OK = 0.  On error the message is
 *    available from kg_last_error(h) (or kg_last_error(NULL) when no handle
 *    exists yet).  The reference prints and stops on errors
 *    (lines_method_lib.f90:529-533,744-748); the Fortran shim keeps that by
 *    stopping on a nonzero status.
 *  - One handle = one grid (the reference's one-rank object).  Calls on one
 *    handle must come from one thread at a time.  kg_step and the kg_<scheme>
 *    calls are asynchronous on the handle's HIP stream; downloads and
 *    kg_sync synchronise.
 *  - Host arrays are owned by the caller and only read or written inside
 *    the call.  Device memory is owned by the handle.
 *
 * Reference sizes: Nx radial cells, Ny = Ntheta+1, Nz = Nphi+1, g = g_pts.
 * Padded extents: ni = Nx+2g+1, nj = Ny+2g+1, nk = Nz+1
 * (lines_method_lib.f90:135-138 with nproc = 1).
 */
#ifndef KG_H
#define KG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t kg_status;
enum {
  KG_OK = 0,
  KG_ERR_INVALID = 1,      /* bad argument or descriptor */
  KG_ERR_UNSUPPORTED = 2,  /* a branch the reference stops on */
  KG_ERR_HIP = 3,          /* HIP runtime error */
  KG_ERR_NOMEM = 4,        /* device allocation failed */
  KG_ERR_NO_DEVICE = 5     /* no GPU visible */
};

/* time integrators: lines_method_lib.f90 pirk2 :271-332, pirk3 :334-408,
 * rk3 :211-269, rk4 :411-474 */
enum { KG_PIRK2 = 0, KG_PIRK3 = 1, KG_RK3 = 2, KG_RK4 = 3 };
/* metrics: geometry.f90:104-247.  The RHS of Pi exists only for Kerr_Schild
 * and Schw_EF (lines_method_lib.f90:727-748); the others are rejected at
 * kg_create with KG_ERR_UNSUPPORTED, as the reference stops there. */
enum { KG_METRIC_MINKOWSKI = 0, KG_METRIC_KERR_SCHILD = 1, KG_METRIC_MKS = 2, KG_METRIC_SCHW_EF = 3 };
/* boundary_type_rmax: 'outflow' / 'inflow' (lines_method_lib.f90:522-535),
 * and KG_BC_SOMMERFELD, which the reference does not have: the outer column
 * i = Nx+g is evolved instead of filled, its phi and psi with their own
 * right-hand sides (the operator set's one-sided r rows) and its Pi with the
 * radiative condition d_t Pi = -v (d_r Pi + Pi/r), v = alpha sqrt(gamma^rr) -
 * beta^r the outgoing radial characteristic speed, d_r the operator set's
 * one-sided edge row; exact for outgoing waves f(t - r)/r in flat space. */
enum { KG_BC_OUTFLOW = 0, KG_BC_INFLOW = 1, KG_BC_SOMMERFELD = 2 };
/* host array layouts accepted by upload/download */
enum {
  KG_LAYOUT_FORTRAN = 0, /* u(nvars,ni,nj,nk), variable fastest: complex(8) pairs
                            when is_complex, else real(8) -- MoL%u as is */
  KG_LAYOUT_SOA = 1      /* [part][nvars][nk][nj][ni] real(8), r fastest; part =
                            real, imaginary (imaginary only when is_complex) */
};

/* Operator set of the right-hand side.  KG_OPS_REF2: the reference's live
 * 2nd-order stencils (finite_differences.f90:17-44,193-239,310-338,767-801).
 * KG_OPS_FD4: the 4th-order set the same file defines (x_4 :124-151, y_4
 * :241-268, z_4 :340-367, advec_x_4 :922-957) in every first derivative of
 * lines_method_lib.f90:671-753 and in the metric's d_r beta^r; for RK3 and
 * RK4 (the PIRK schemes return KG_ERR_UNSUPPORTED). */
enum { KG_OPS_REF2 = 0, KG_OPS_FD4 = 1 };

/* Descriptor: the arguments of lines_method_constructor
 * (lines_method_lib.f90:49-51) plus the scheme and the value type. */
typedef struct kg_desc {
  int32_t nvars;      /* must be 5: phi, psi_r, psi_theta, psi_phi, Pi */
  int32_t g_pts;      /* ghost points in r and theta (reference: 2) */
  int32_t Nx, Ny, Nz; /* reference sizes (Ny = Ntheta+1, Nz = Nphi+1) */
  int32_t scheme;     /* KG_PIRK2 ... KG_RK4: what kg_step advances with */
  int32_t metric;     /* KG_METRIC_* */
  int32_t bc_rmax;    /* KG_BC_* */
  int32_t is_complex; /* 1: complex(8) state (two real systems on the device) */
  int32_t op_set;     /* KG_OPS_REF2 (0) or KG_OPS_FD4 (1) */
  double xmin, xmax;  /* radial domain */
  double tmin;        /* initial time */
  double CFL;         /* used when dt <= 0 */
  double dt;          /* > 0 overrides dt = CFL*min(...) (diagnostic_lib.f90:101-104) */
  double a, M_bh;     /* black-hole spin and mass */
  double m;           /* field mass */
} kg_desc;

typedef struct kg_handle kg_handle;

/* lines_method_constructor (lines_method_lib.f90:49-129): grid, metric
 * tables, dt; state and stage buffers zeroed. */
kg_status kg_create(const kg_desc* desc, kg_handle** out);
/* releases device memory (no reference counterpart: Fortran deallocation) */
kg_status kg_destroy(kg_handle* h);
/* message of the last failed call on h (h = NULL: last failure without a handle) */
const char* kg_last_error(const kg_handle* h);

/* dims[0..7] = Nx, Ny, Nz, g, ni, nj, nk, npts(padded) */
kg_status kg_get_dims(const kg_handle* h, int64_t dims[8]);
/* this%t and this%dt (lines_method_lib.f90:24,112) */
kg_status kg_get_time(const kg_handle* h, double* t, double* dt);
kg_status kg_set_time(kg_handle* h, double t);

/* MoL%u <-> device (whole padded array, ghosts included) */
kg_status kg_upload_state(kg_handle* h, const double* u, int32_t layout);
kg_status kg_download_state(kg_handle* h, double* u, int32_t layout);
/* The persistent PIRK stage arrays MoL%u_stage (which = 0) and
 * MoL%u_stage_2 (which = 1).  Only their Pi component carries over between
 * steps (lines_method_lib.f90:288-289,366-367); the full array is exchanged
 * so a run can be resumed exactly. */
kg_status kg_upload_aux(kg_handle* h, int32_t which, const double* u, int32_t layout);
kg_status kg_download_aux(kg_handle* h, int32_t which, double* u, int32_t layout);
/* Device-to-device state exchange in KG_LAYOUT_SOA order (for callers that
 * keep data in HBM).  dir = 0: src -> handle; dir = 1: handle -> dst. */
kg_status kg_copy_state_device(kg_handle* h, void* dev_ptr, int32_t dir);

/* check_point / read_from_check_point (lines_method_lib.f90:599-667): the
 * state, its time and, for PIRK2/PIRK3, the stage arrays whose Pi carries
 * over between steps, in a binary file of this engine (the reference writes
 * list-directed text and its reader deletes the file before reading it).
 * Loading requires a handle created with the same descriptor; a restarted run
 * continues bit for bit as if it had not stopped. */
kg_status kg_checkpoint_save(kg_handle* h, const char* path);
kg_status kg_checkpoint_load(kg_handle* h, const char* path);

/* boundary() of the integrators (lines_method_lib.f90:313-330): xL, xR, yL,
 * yR, zL, zR on MoL%u, as the initial-data code calls them
 * (initial_conditions.f90:111-121). */
kg_status kg_fill_ghosts(kg_handle* h);

/* One step of a given scheme: MoL%pirk2() (lines_method_lib.f90:271),
 * MoL%pirk3() (:334), MoL%rk3(rhs) (:211), MoL%rk4(rhs) (:411).  The rk3/rk4
 * rhs callback is the Klein-Gordon right-hand side (all five rhs_* on u). */
kg_status kg_pirk2(kg_handle* h);
kg_status kg_pirk3(kg_handle* h);
kg_status kg_rk3(kg_handle* h);
kg_status kg_rk4(kg_handle* h);
/* nsteps steps of desc.scheme: the `do l=1,Nt` loop body of evolve.f90:22-26
 * without diagnostics, replayed from a captured HIP graph. */
kg_status kg_step(kg_handle* h, int32_t nsteps);
/* Batched stepping (no reference counterpart: the reference runs one grid
 * per process).  Advances every handle of hs[0..n) by nsteps steps of their
 * common desc.scheme (RK4, RK3, PIRK2 or PIRK3), each stage kernel and each
 * ghost-fill pass one launch over the tiles of all grids (grids may differ in
 * size, real or complex state, mass, black-hole parameters, initial data and
 * buffer or stale-Pi parity; all must share the scheme, the metric form and
 * the operator set).  Each member marches its own phi chunk (kg_set_tuning
 * kc); theta rows per workgroup are hs[0]'s.  Work runs on hs[0]'s stream;
 * later calls on any member are ordered after it.  Each member's result,
 * including its PIRK stage arrays, equals stepping it alone.
 * KG_ERR_UNSUPPORTED for mixed schemes, operator sets or metric forms. */
kg_status kg_batch_step(kg_handle* const* hs, int32_t n, int32_t nsteps);
/* kg_batch_step without allocation.  No counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and ordering as kg_batch_step, but the launch descriptors go to a
 * device array of the library itself, 4 MiB per GPU (held until the process
 * ends; its address is resolved at kg_create), instead of a buffer the lead
 * handle grows; the call makes no hipMalloc/hipFree and creates no event.
 * The descriptors take 2 x (launches per step) x n x 504 bytes plus small
 * prefix arrays (PIRK2: 4 launches per step, about 1020 members at most;
 * RK4, RK3 and PIRK3: 6, about 680; with KG_BC_SOMMERFELD a stage adds its
 * outer-column launch: RK4 about 410): a call that needs more
 * returns KG_ERR_INVALID.
 * Every member's stage buffers must exist (kg_create makes those of its
 * desc.scheme).  Kernel timing (kg_profile_enable) is not recorded.  All
 * calls on one GPU share the array and the last call's descriptors stay in
 * it: calls on one GPU that overlap (from other threads, or on other
 * streams than an earlier call's still-running lead stream) race. */
kg_status kg_batch_step_noalloc(kg_handle* const* hs, int32_t n, int32_t nsteps);
/* Batched stage launches led by handle h (it is hs[0]): wg_per_cu = 0 (the
 * default) launches one workgroup per tile; wg_per_cu > 0 makes each stage
 * one persistent launch of wg_per_cu workgroups per compute unit that march
 * the tiles of all members in turn.  Results do not depend on it.  No
 * reference counterpart. */
kg_status kg_set_batch_persistent(kg_handle* h, int32_t wg_per_cu);
/* waits for the handle's stream */
kg_status kg_sync(kg_handle* h);

/* Launch shape of the stage kernel (no reference counterpart): kc = phi
 * planes each workgroup marches over at most (the ceil(Nz/kc) chunks of a
 * tile column are balanced: 257 planes at kc = 32 march as 9 chunks of 29),
 * tj = theta rows per workgroup (4 or 8; the fd4 stages always run 8).
 * Results do not depend on it; autotuning callers sweep it.  Setting the
 * current shape again is a no-op. */
kg_status kg_set_tuning(kg_handle* h, int32_t kc, int32_t tj);

/* Kernel timing with HIP events on the handle's stream (no reference
 * counterpart; the reference only prints MPI_Wtime, save_gnuplot.f90:154).
 * kind: 0 = RHS+stage kernels, 1 = ghost-fill kernels, 2 = Sommerfeld
 * outer-column kernels.  While enabled, kg_step launches eagerly with an
 * event pair around every kernel. */
kg_status kg_profile_enable(kg_handle* h, int32_t on);
kg_status kg_profile_read(kg_handle* h, int32_t kind, double* total_ms, int64_t* launches);

/* ------------------------------------------------------------------------
 * Spherical-Bessel spectral projection (tools/FFT_sph.f90), on the grid of a
 * handle.  The spectral grid is the reference's node-centred one
 * (mesh_refinement_lib.f90:135-143): r_i = xmin + i dx (i < Nx),
 * theta_j = j dtheta (j < Ny), phi_k = k dphi (k < Nz).  Indices lm =
 * l*l + l + m for l = 0..lmax, m = -l..l.  Complex outputs are interleaved
 * (re, im) doubles.  Outputs with on_device = 1 are device pointers filled
 * asynchronously on the handle's stream; otherwise host arrays, filled
 * before the call returns.
 * ------------------------------------------------------------------------ */
typedef struct kg_spec kg_spec;
/* radial basis: KG_BASIS_ZEROS = generate_full_bessel_basis (FFT_sph.f90:13-95):
 * k_{l,n} = z_n(J_{l+1/2}) / R_max, j_l(k r) normalised to unit trapezoid
 * norm with weight r^2;  KG_BASIS_UNIFORM: k_n = kmax n / nk for every l,
 * j_l(k r) unnormalised. */
enum { KG_BASIS_ZEROS = 0, KG_BASIS_UNIFORM = 1 };
/* arithmetic of the radial contraction: FP64, or BF16 matrix cores with
 * split operands (five 8-bit slices per operand, exact FP32 accumulation of
 * each slice product over 32-term chunks, chunks combined in FP64) */
enum { KG_SPEC_FP64 = 0, KG_SPEC_SPLIT_BF16 = 1 };
typedef struct kg_spec_desc {
  int32_t lmax;  /* l_max (1..16) */
  int32_t nk;    /* Nk: radial wave numbers n = 1..nk */
  int32_t basis; /* KG_BASIS_* */
  int32_t reserved;
  double R_max;  /* KG_BASIS_ZEROS: normalisation radius (the reference's r_max) */
  double kmax;   /* KG_BASIS_UNIFORM: largest wave number */
} kg_spec_desc;

/* scalar_field_constructor's spectral set-up (diagnostic_lib.f90:144-149):
 * Y_lm tables (sph_harmonics, FFT_sph.f90:282-336) and the Bessel basis, on
 * h's device.  The projector reads h's current state at each call. */
kg_status kg_spec_create(kg_handle* h, const kg_spec_desc* desc, kg_spec** out);
kg_status kg_spec_destroy(kg_spec* p);
const char* kg_spec_last_error(const kg_spec* p);
/* Host copies of the set-up tables; any pointer may be NULL.
 * ylm: complex [nlm][Ny][Nz] (sph_harmonics' Ylm(l,m,j,k));  kv: [lmax+1][nk]
 * (k_values);  jb: [lmax+1][nk][Nx] (j_basis). */
kg_status kg_spec_get_tables(kg_spec* p, double* ylm, double* kv, double* jb);
/* domain_to_sph (FFT_sph.f90:525-651) of state field `field` (0..4):
 * out = real [nsys][Nz][Ny][Nx] (nsys = 2 for a complex state: real part,
 * then imaginary part). */
kg_status kg_spec_domain_to_sph(kg_spec* p, int32_t field, double* out, int32_t on_device);
/* project_sh (FFT_sph.f90:247-279) at every spectral radius, as
 * SphFT_discrete_radial applies it (:228-231): flm = complex [nlm][Nx]. */
kg_status kg_spec_project_sh(kg_spec* p, int32_t field, double* flm, int32_t on_device);
/* The radial contraction SphFT_discrete_radial defines (:233-236):
 * ft(n,l,m) = trapezoid_i( flm(r_i) j_l(k_n r_i) r_i^2, dr ), computed from
 * domain_to_sph + project_sh of `field`: ft = complex [nlm][nk]. */
kg_status kg_spec_radial(kg_spec* p, int32_t field, int32_t precision, double* ft, int32_t on_device);
/* kg_spec_radial of many projectors (grids of any size, one lmax class:
 * all lmax <= 8 or all > 8), one launch per pass (phi, theta, radial) over
 * all of them on ps[0]'s stream; fts[q] receives projector q's complex
 * [nlm][nk].  Results equal kg_spec_radial on each projector bit for bit.
 * No counterpart in the reference (one grid per process); its per-grid
 * routine is SphFT_discrete_radial (FFT_sph.f90:213-244). */
kg_status kg_spec_radial_batch(kg_spec* const* ps, int32_t n, int32_t field, int32_t precision, double* const* fts,
                               int32_t on_device);
/* The radial contraction alone, on caller-supplied host flm (complex
 * [nlm][Nx]) -> host ft (complex [nlm][nk]). */
kg_status kg_spec_radial_flm(kg_spec* p, const double* flm, int32_t precision, double* ft);
/* Detector diagnostics at spectral radii index_det[0..n_det) (the
 * reference's save-step diagnostics, diagnostic_lib.f90:197-199):
 * momentum = complex [n_det][3], the fluxes trapezium_2D(domain_to_sph(
 * j_i sqrtg)) of currents_ADM / currents_detectors (:214-224,250-274);
 * alm = complex [n_det][nlm], project_sh of domain_to_sph(phi)
 * (proyector_detectors, :231-248).  The reference never assigns index_det;
 * callers pass the radius indices.  Either output may be NULL. */
kg_status kg_spec_detectors(kg_spec* p, int32_t n_det, const int32_t* index_det, double* momentum, double* alm);

/* ------------------------------------------------------------------------
 * Volume integral of the state (integrate, integral.f90:646-669, with
 * trapezium_real :220-255 over ii in [g, ni-1-g], jj in [g, nj-1-g], all
 * phi planes).  quantity 0..4: that field (complex for a complex state);
 * KG_DIAG_DENSITY: density() (diagnostic_lib.f90:258-271) times sqrt(gamma),
 * the integrand of Mass_scalar (:207).  out[0] = real part, out[1] =
 * imaginary part.  Synchronous.
 * ------------------------------------------------------------------------ */
enum { KG_DIAG_DENSITY = 5 };
kg_status kg_diag_integrate(kg_handle* h, int32_t quantity, double out[2]);
const char* kg_diag_last_error(void);
/* The arrays the reference writes to HDF5 every save step
 * (save_gnuplot.f90:121-144), formed on the GPU over its slab i = -g..Nx,
 * j = 0..Ny, k = 0..Nz-1: quantity 0 |phi|, 1 density(), 2 |psi_r|,
 * 3 |psi_theta|, 4 |psi_phi|, 5 |Pi|, 6 Re phi, 7 Im phi.  out = host array
 * [Nz][Ny+1][Nx+g+1] (i fastest).  Synchronous. */
kg_status kg_diag_output(kg_handle* h, int32_t quantity, double* out);

/* ------------------------------------------------------------------------
 * Initial data built on the GPU: the reference's `subroutine
 * initial_conditions` (main/initial_conditions.f90:1-220), selected there by
 * initial_conditions_type with the globals mu and sigma.
 *  KG_IC_SPHERICAL_SHELL  'spherical_shell' (:23-88, every metric but MKS):
 *      phi = 1/x exp(-(x-mu)^2/sigma^2), psi_r = 0 (its assignment is
 *      commented out there), psi_theta, psi_phi its stencil derivatives,
 *      Pi = 1/alpha (dtphi - beta^r psi_r); Amp = 1/sqrt(integrate(rho alpha
 *      sqrt(gamma))).  The MKS branch scales by an unset local (:26) and is
 *      refused with KG_ERR_UNSUPPORTED.
 *  KG_IC_RANDOM           'random' (:90-144): the Cartesian random field
 *      (tools/random.f90:17-59) sampled onto the grid, xL/xR fills,
 *      derivatives, Pi with dtphi = 0; Amp = 1/sqrt(integrate(m |phi|^2
 *      sqrt(gamma))).
 *  KG_IC_RANDOM_SHELL     'random_shell' (:146-188): the sampled field times
 *      the shell envelope, derivatives, Pi with dtphi = psi_r, then the xL,
 *      xR, zL, zR, yL, yR fills; Amp as for 'random'.
 * The random types need a complex handle (is_complex = 1), as the reference's
 * state is complex.  The random field lives on ncart^3 points over
 * [-(r_max+1), r_max+1)^3 with r_max = x(iR+g) (the reference fixes
 * ncart = 256); every grid point must fall inside the box's cells, i.e.
 * r_max + 1 < ncart/2.  The random phases U in [0,1) (random.f90:47-48, an
 * unseeded random_number there) come from `uniforms` when it is not NULL
 * (host array of ncart^3 values, i fastest: Fortran's random_value(i,j,k))
 * and otherwise from a SplitMix64 stream keyed by `seed`.  An 'inflow' xR
 * fill reads the start-of-step state, which the reference has not written at
 * this point; it is taken as zero.  *amp_out (may be NULL) receives Amp.
 * The time is left unchanged.  Synchronous.
 * ------------------------------------------------------------------------ */
enum { KG_IC_SPHERICAL_SHELL = 0, KG_IC_RANDOM = 1, KG_IC_RANDOM_SHELL = 2 };
typedef struct kg_ic_desc {
  int32_t type;            /* KG_IC_* */
  int32_t ncart;           /* random types: Cartesian points per axis, a power of two in [4, 1024] */
  double mu, sigma;        /* shell centre and width (global_numbers) */
  double p0;               /* random types: momentum width (the reference passes 1) */
  uint64_t seed;           /* random types: phase stream when uniforms is NULL */
  const double* uniforms;  /* random types: optional host array [ncart^3] */
} kg_ic_desc;
kg_status kg_initial_data(kg_handle* h, const kg_ic_desc* d, double* amp_out);
/* The Cartesian field psi_x = ifft_3D(psi_p) of random.f90:17-59 alone, on
 * device `device`: out = host complex(8) array psi_x(ncart,ncart,ncart) in
 * Fortran order (re, im interleaved, i fastest); r_max as the reference
 * passes it (the box is [-(r_max+1), r_max+1)).  Synchronous. */
kg_status kg_ic_random_field(int32_t device, int32_t ncart, double r_max, double p0, uint64_t seed,
                             const double* uniforms, double* out);
const char* kg_ic_last_error(void);

/* ------------------------------------------------------------------------
 * Two fixed refinement levels (no reference counterpart: the reference's
 * mesh_refinement_lib.f90:16-147 is a slab decomposition of one uniform
 * grid).  A base grid (`base`, any scheme and operator set) and a 2:1
 * radially refined level on [xmin, r_patch] with the same angular grid, dr/2
 * and dt/2, advanced Berger-Oliger style: per base step the base level steps,
 * then the fine level takes two steps, each after its last `buffer` columns
 * (beyond r_patch) are refilled from the base level -- in r by 6-point
 * interpolation (3, -25, 150, 150, -25, 3)/256 between base points and
 * copies at coincident ones, in time through the last three base times
 * (weights -1/8, 3/4, 3/8 at the half step; 1/2, 1/2 while only two are
 * known) -- with its outer column held at its start-of-step value; then the
 * base columns 0 .. I-3 (I: the base column at r_patch) take the fine values
 * at the same radii.  The caller sets both levels' initial data through the
 * handles kg_refined_levels returns (kg_upload_state, kg_fill_ghosts, ...)
 * and reads results the same way; the levels must not be stepped directly.
 * ------------------------------------------------------------------------ */
typedef struct kg_refined kg_refined;
typedef struct kg_refine_desc {
  double r_patch;   /* outer radius of the refined patch (rounded to a base grid point) */
  int32_t buffer;   /* fine buffer columns beyond r_patch (even, 8..64; 0: 16) */
  int32_t reserved;
} kg_refine_desc;
kg_status kg_refined_create(const kg_desc* base, const kg_refine_desc* desc, kg_refined** out);
kg_status kg_refined_levels(kg_refined* r, kg_handle** base, kg_handle** fine);
/* nsteps base steps (each with two fine steps) */
kg_status kg_refined_step(kg_refined* r, int32_t nsteps);
kg_status kg_refined_destroy(kg_refined* r);
const char* kg_refined_last_error(const kg_refined* r);

/* ------------------------------------------------------------------------
 * HBM bandwidth probe (measurement only; no reference counterpart): the
 * average duration of `reps` launches of a streaming kernel over FP64 arrays
 * of n doubles (n even) on `device`, `blocks` workgroups of 256 threads.
 * kind 0: read (n x 8 B); 1: write (n x 8 B); 2: copy (read + write);
 * 3: d = a + b + c (three reads, one write); 4, 5: copy and that mix with
 * non-temporal loads and stores.  Allocates and frees its four arrays.
 * ------------------------------------------------------------------------ */
kg_status kg_hbm_probe(int32_t device, int32_t kind, int64_t n, int32_t blocks, int32_t reps, double* ms_per_launch);
const char* kg_probe_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* KG_H */
