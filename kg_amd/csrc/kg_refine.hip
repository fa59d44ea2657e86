// kg_refine.hip -- two fixed refinement levels with Berger-Oliger time
// sub-stepping: a base grid and a 2:1 radially refined patch on
// [xmin, r_patch].  The reference has no refinement levels (its
// mesh_refinement_lib.f90:16-147 cuts one uniform grid into MPI slabs); this
// is new here, built on the engine's own stepping.
//
// Levels.  The base level is a kg_handle as the caller describes it.  The
// fine level is a kg_handle on the same angular grid with dr/2 and dt/2; its
// padded columns ii = 0 .. 2I + B (x = xmin + ii dr/2), where base column I
// sits at r_patch and the last B fine columns are a buffer zone beyond it.
// The fine level's outer column is held at its start-of-step value during a
// step ('inflow'); the errors that this and the one-sided rows make reach at
// most (stages x stencil reach) columns inward per step, which B = 16 covers
// for every scheme and operator set.
//
// One base step t -> t + dt:
//   1. the base level steps (its history of the buffer columns keeps the
//      last three base times);
//   2. the fine buffer is refilled from the base level at t (even fine
//      columns copy base columns; odd ones take the 6-point interpolation
//      (3, -25, 150, 150, -25, 3)/256 of the base columns around them) and
//      the fine level steps to t + dt/2;
//   3. the buffer is refilled from the base level at t + dt/2, interpolated
//      in time through the last three base times (-1/8, 3/4, 3/8; with two
//      known times, 1/2, 1/2), and the fine level steps to t + dt;
//   4. base columns 0 .. I-3 take the fine values at the same radii
//      (injection: fine column 2i).  The interpolation in 2-3 reads base
//      columns >= I-2 only, so it never reads restricted values.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/kg.h"
#include "kg_internal.h"

using namespace kg;

namespace {

thread_local std::string g_ref_err;

typedef __attribute__((address_space(1))) double gdouble;

// history of the base level's buffer columns [c_lo, c_lo + ncol): [slot][sys][field][k][j][col]
__global__ __launch_bounds__(256) void save_cols_kernel(const Geom G, const double* u0, const double* u1,
                                                        const double* u2, const double* u3, const double* u4,
                                                        double* __restrict__ slot, int c_lo, int ncol, int nsys) {
  const int64_t n = (int64_t)nsys * 5 * G.nk * G.nj * ncol;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const int c = (int)(e % ncol);
  int64_t r = e / ncol;
  const int jj = (int)(r % G.nj);
  r /= G.nj;
  const int kk = (int)(r % G.nk);
  r /= G.nk;
  const int f = (int)(r % 5);
  const int s = (int)(r / 5);
  const double* u = f == 0 ? u0 : f == 1 ? u1 : f == 2 ? u2 : f == 3 ? u3 : u4;
  slot[e] = ((const gdouble*)u)[(int64_t)s * G.sys_stride + (int64_t)kk * G.plane + (int64_t)jj * G.ld + c_lo + c];
}

struct ProlongArgs {
  Geom Gf;                 // fine level
  double* u[5];            // fine state fields
  const double* slot[3];   // base history at the times being combined
  double w[3];             // their weights
  int nterms;              // 1 (copy), 2 (linear), 3 (quadratic)
  int c_lo, ncol;          // base columns held in the history
  int f_lo;                // first fine buffer column
  int nbuf;                // buffer columns
  int nsys;
};

__global__ __launch_bounds__(256) void prolong_kernel(const ProlongArgs a) {
  const Geom& G = a.Gf;
  const int64_t n = (int64_t)a.nsys * 5 * G.nk * G.nj * a.nbuf;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const int b = (int)(e % a.nbuf);
  int64_t r = e / a.nbuf;
  const int jj = (int)(r % G.nj);
  r /= G.nj;
  const int kk = (int)(r % G.nk);
  r /= G.nk;
  const int f = (int)(r % 5);
  const int s = (int)(r / 5);
  const int64_t row = (((int64_t)s * 5 + f) * G.nk + kk) * G.nj + jj;  // history row of this point
  // base value at history column c, interpolated in time
  auto V = [&](int c) -> double {
    const int64_t o = row * a.ncol + (c - a.c_lo);
    const double v0 = ((const gdouble*)a.slot[0])[o];
    if (a.nterms == 1) return v0;
    const double v1 = ((const gdouble*)a.slot[1])[o];
    if (a.nterms == 2) return a.w[0] * v0 + a.w[1] * v1;
    const double v2 = ((const gdouble*)a.slot[2])[o];
    return a.w[0] * v0 + a.w[1] * v1 + a.w[2] * v2;
  };
  const int fi = a.f_lo + b;  // fine column
  double v;
  if ((fi & 1) == 0) {
    v = V(fi / 2);
  } else {
    const int c = (fi - 1) / 2;  // fine column fi lies midway between base columns c and c+1
    v = (3.0 * V(c - 2) - 25.0 * V(c - 1) + 150.0 * V(c) + 150.0 * V(c + 1) - 25.0 * V(c + 2) + 3.0 * V(c + 3)) *
        0.00390625;
  }
  ((gdouble*)a.u[f])[(int64_t)s * G.sys_stride + (int64_t)kk * G.plane + (int64_t)jj * G.ld + fi] = v;
}

// base columns 0 .. ncol-1 <- fine columns 0, 2, .., 2(ncol-1)
__global__ __launch_bounds__(256) void restrict_kernel(const Geom Gc, const Geom Gf, double* c0, double* c1,
                                                       double* c2, double* c3, double* c4, const double* f0,
                                                       const double* f1, const double* f2, const double* f3,
                                                       const double* f4, int ncol, int nsys) {
  const int64_t n = (int64_t)nsys * 5 * Gc.nk * Gc.nj * ncol;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const int c = (int)(e % ncol);
  int64_t r = e / ncol;
  const int jj = (int)(r % Gc.nj);
  r /= Gc.nj;
  const int kk = (int)(r % Gc.nk);
  r /= Gc.nk;
  const int f = (int)(r % 5);
  const int s = (int)(r / 5);
  double* cu = f == 0 ? c0 : f == 1 ? c1 : f == 2 ? c2 : f == 3 ? c3 : c4;
  const double* fu = f == 0 ? f0 : f == 1 ? f1 : f == 2 ? f2 : f == 3 ? f3 : f4;
  ((gdouble*)cu)[(int64_t)s * Gc.sys_stride + (int64_t)kk * Gc.plane + (int64_t)jj * Gc.ld + c] =
      ((const gdouble*)fu)[(int64_t)s * Gf.sys_stride + (int64_t)kk * Gf.plane + (int64_t)jj * Gf.ld + 2 * c];
}

unsigned nblocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct kg_refined {
  kg_handle* base = nullptr;
  kg_handle* fine = nullptr;
  int I = 0;           // base column at r_patch
  int B = 16;          // fine buffer columns
  int c_lo = 0, ncol = 0;
  double* hist = nullptr;  // 3 slots of the base buffer columns
  size_t slot_len = 0;
  int slot_t = 0, slot_m = 1, slot_n = 2;  // slots of base times t, t - dt, t + dt
  int nlev = 0;                            // base times held (0..3)
  hipEvent_t base_done = nullptr, fine_done = nullptr;
  std::string err;
};

namespace {

kg_status rfail(kg_refined* r, kg_status st, const std::string& msg) {
  if (r) r->err = msg;
  g_ref_err = msg;
  return st;
}

kg_status save_slot(kg_refined* r, int slot) {
  HandleView v;
  handle_view(r->base, &v);
  const int64_t n = (int64_t)v.nsys * 5 * v.G.nk * v.G.nj * r->ncol;
  hipLaunchKernelGGL(save_cols_kernel, dim3(nblocks(n)), dim3(256), 0, v.stream, v.G, v.u[0], v.u[1], v.u[2],
                     v.u[3], v.u[4], r->hist + (size_t)slot * r->slot_len, r->c_lo, r->ncol, v.nsys);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? KG_OK : rfail(r, KG_ERR_HIP, std::string("save: ") + hipGetErrorString(e));
}

kg_status prolong(kg_refined* r, int nterms, const int* slots, const double* w) {
  HandleView v;
  handle_view(r->fine, &v);
  ProlongArgs a{};
  a.Gf = v.G;
  for (int f = 0; f < 5; ++f) a.u[f] = v.u[f];
  for (int q = 0; q < nterms; ++q) {
    a.slot[q] = r->hist + (size_t)slots[q] * r->slot_len;
    a.w[q] = w[q];
  }
  a.nterms = nterms;
  a.c_lo = r->c_lo;
  a.ncol = r->ncol;
  a.f_lo = v.G.ni - r->B;
  a.nbuf = r->B;
  a.nsys = v.nsys;
  const int64_t n = (int64_t)v.nsys * 5 * v.G.nk * v.G.nj * r->B;
  hipLaunchKernelGGL(prolong_kernel, dim3(nblocks(n)), dim3(256), 0, v.stream, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? KG_OK : rfail(r, KG_ERR_HIP, std::string("prolong: ") + hipGetErrorString(e));
}

kg_status restrict_to_base(kg_refined* r) {
  HandleView vc, vf;
  handle_view(r->base, &vc);
  handle_view(r->fine, &vf);
  const int ncol = r->I - 2;  // base columns 0 .. I-3
  const int64_t n = (int64_t)vc.nsys * 5 * vc.G.nk * vc.G.nj * ncol;
  hipLaunchKernelGGL(restrict_kernel, dim3(nblocks(n)), dim3(256), 0, vf.stream, vc.G, vf.G, vc.u[0], vc.u[1],
                     vc.u[2], vc.u[3], vc.u[4], vf.u[0], vf.u[1], vf.u[2], vf.u[3], vf.u[4], ncol, vc.nsys);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? KG_OK : rfail(r, KG_ERR_HIP, std::string("restrict: ") + hipGetErrorString(e));
}

#define KR_TRY(expr)            \
  do {                          \
    kg_status s_ = (expr);      \
    if (s_ != KG_OK) return s_; \
  } while (0)
#define KR_HIP(r, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return rfail((r), KG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

kg_status level_step(kg_refined* r, kg_handle* h) {
  const kg_status st = kg_step(h, 1);
  if (st != KG_OK) return rfail(r, st, std::string("level step: ") + kg_last_error(h));
  return KG_OK;
}

}  // namespace

extern "C" {

const char* kg_refined_last_error(const kg_refined* r) { return r ? r->err.c_str() : g_ref_err.c_str(); }

kg_status kg_refined_create(const kg_desc* base, const kg_refine_desc* rd, kg_refined** out) {
  if (!base || !rd || !out) return rfail(nullptr, KG_ERR_INVALID, "null argument");
  *out = nullptr;
  const int B = rd->buffer > 0 ? rd->buffer : 16;
  if (B < 8 || B > 64 || (B & 1)) return rfail(nullptr, KG_ERR_INVALID, "buffer must be even, in [8, 64]");
  kg_refined* r = new kg_refined();
  r->B = B;
  kg_status st = kg_create(base, &r->base);
  if (st != KG_OK) {
    const std::string m = kg_last_error(nullptr);
    delete r;
    return rfail(nullptr, st, "base level: " + m);
  }
  HandleView vb;
  handle_view(r->base, &vb);
  const double dx0 = vb.hg.dx, xmin = vb.hg.xmin;
  const int I = (int)std::lround((rd->r_patch - xmin) / dx0);
  const int g = base->g_pts;
  // the history reads base columns I-2 .. I+B/2+3, inside the base level's computed columns
  if (I < 2 * g + 4 || I + B / 2 + 3 > vb.G.ni - 2) {
    kg_destroy(r->base);
    delete r;
    return rfail(nullptr, KG_ERR_INVALID, "r_patch must leave room for the buffer inside the base grid");
  }
  r->I = I;
  r->c_lo = I - 2;
  r->ncol = B / 2 + 6;
  kg_desc fd = *base;
  fd.Nx = 2 * I + B - 2 * g;
  fd.xmax = xmin + (double)fd.Nx * (dx0 / 2.0);
  fd.bc_rmax = KG_BC_INFLOW;  // the fine outer column holds its start-of-step value
  fd.dt = vb.hg.dt / 2.0;
  st = kg_create(&fd, &r->fine);
  if (st != KG_OK) {
    const std::string m = kg_last_error(nullptr);
    kg_destroy(r->base);
    delete r;
    return rfail(nullptr, st, "fine level: " + m);
  }
  r->slot_len = (size_t)vb.nsys * 5 * vb.G.nk * vb.G.nj * r->ncol;
  hipError_t e = hipMalloc(&r->hist, 3 * r->slot_len * sizeof(double));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&r->base_done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&r->fine_done, hipEventDisableTiming);
  if (e != hipSuccess) {
    kg_refined_destroy(r);
    return rfail(nullptr, KG_ERR_NOMEM, std::string("refinement buffers: ") + hipGetErrorString(e));
  }
  *out = r;
  return KG_OK;
}

kg_status kg_refined_levels(kg_refined* r, kg_handle** base, kg_handle** fine) {
  if (!r) return rfail(nullptr, KG_ERR_INVALID, "null refinement");
  if (base) *base = r->base;
  if (fine) *fine = r->fine;
  return KG_OK;
}

kg_status kg_refined_step(kg_refined* r, int32_t nsteps) {
  if (!r) return rfail(nullptr, KG_ERR_INVALID, "null refinement");
  if (nsteps < 0) return rfail(r, KG_ERR_INVALID, "nsteps < 0");
  HandleView vb, vf;
  handle_view(r->base, &vb);
  handle_view(r->fine, &vf);
  for (int it = 0; it < nsteps; ++it) {
    if (r->nlev == 0) {  // the base level's state at the first step's start
      KR_TRY(save_slot(r, r->slot_t));
      r->nlev = 1;
    }
    // 1. base step, then its buffer columns at t + dt
    KR_TRY(level_step(r, r->base));
    KR_TRY(save_slot(r, r->slot_n));
    KR_HIP(r, hipEventRecord(r->base_done, vb.stream));
    KR_HIP(r, hipStreamWaitEvent(vf.stream, r->base_done, 0));
    // 2. fine step from t: buffer from the base level at t
    {
      const int s1[1] = {r->slot_t};
      const double w1[1] = {1.0};
      KR_TRY(prolong(r, 1, s1, w1));
    }
    KR_TRY(level_step(r, r->fine));
    // 3. fine step from t + dt/2: buffer interpolated in time
    if (r->nlev >= 2) {
      const int s3[3] = {r->slot_m, r->slot_t, r->slot_n};
      const double w3[3] = {-0.125, 0.75, 0.375};
      KR_TRY(prolong(r, 3, s3, w3));
    } else {
      const int s2[2] = {r->slot_t, r->slot_n};
      const double w2[2] = {0.5, 0.5};
      KR_TRY(prolong(r, 2, s2, w2));
    }
    KR_TRY(level_step(r, r->fine));
    // 4. restriction
    KR_TRY(restrict_to_base(r));
    KR_HIP(r, hipEventRecord(r->fine_done, vf.stream));
    KR_HIP(r, hipStreamWaitEvent(vb.stream, r->fine_done, 0));
    // t + dt becomes t
    const int old_m = r->slot_m;
    r->slot_m = r->slot_t;
    r->slot_t = r->slot_n;
    r->slot_n = old_m;
    if (r->nlev < 3) ++r->nlev;
  }
  return KG_OK;
}

kg_status kg_refined_destroy(kg_refined* r) {
  if (!r) return KG_OK;
  if (r->base) (void)kg_sync(r->base);
  if (r->fine) (void)kg_sync(r->fine);
  if (r->hist) (void)hipFree(r->hist);
  if (r->base_done) (void)hipEventDestroy(r->base_done);
  if (r->fine_done) (void)hipEventDestroy(r->fine_done);
  if (r->fine) kg_destroy(r->fine);
  if (r->base) kg_destroy(r->base);
  delete r;
  return KG_OK;
}

}  // extern "C"
