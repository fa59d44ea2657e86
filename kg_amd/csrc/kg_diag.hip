// kg_diag.hip -- on-device diagnostics of the state (MoL/diagnostic_lib.f90
// :185-274, tools/integral.f90:220-255,646-704): the volume integral of a
// field or of the energy density, computed where the state lives.
//
// integrate (integral.f90:646-669) takes ii in [g, ni-1-g], jj in [g, nj-1-g]
// and every phi plane (the pole ghost rows jj = g, g+Ny and both periodic
// duplicate planes included) and applies the trapezoid rule along each axis
// (trapezium_real, :220-255: weight 1/2 at both ends).  Here one workgroup
// reduces one (theta, phi) row with those weights; a second pass adds the row
// sums in a fixed order, so the result is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/kg.h"
#include "kg_internal.h"

using namespace kg;

namespace {

struct IntArgs {
  const double* u[5];
  const double* tab;
  int64_t sys_stride;
  int64_t plane;
  int ld, nj, g, Nx, Ny, Nz, nsys;
  int quantity;  // 0..4 field, 5 density*sqrtg, 6 (density*alpha)*sqrtg, 7 (m*|phi|**2)*sqrtg
  double m2;     // m**2
  double m;
};

__device__ __forceinline__ double mag(double re, double im, bool cplx) { return cplx ? hypot(re, im) : fabs(re); }

// density (diagnostic_lib.f90:258-271), left to right; the real part of the
// sum (rho is a real array): 0.5*( |Pi| + g11|psi_r|^2 + 2 g13|psi_r| psi_phi
// + g22|psi_th|^2 + g33|psi_phi|^2 + m^2|phi| )
__device__ double density(const IntArgs& a, int64_t o, uint32_t t, bool cplx) {
  double re[5], im[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    re[f] = a.u[f][o];
    im[f] = cplx ? a.u[f][o + a.sys_stride] : 0.0;
  }
  const int64_t n2 = (int64_t)a.nj * a.ld;
  const double g11 = a.tab[T_GRR * n2 + t], g13 = a.tab[T_GRP * n2 + t];
  const double g22 = a.tab[T_GTT * n2 + t], g33 = a.tab[T_GPP * n2 + t];
  const double a2 = mag(re[1], im[1], cplx), a3 = mag(re[2], im[2], cplx), a4 = mag(re[3], im[3], cplx);
  double sum = mag(re[4], im[4], cplx);
  sum = sum + g11 * (a2 * a2);
  sum = sum + ((2.0 * g13) * a2) * re[3];
  sum = sum + g22 * (a3 * a3);
  sum = sum + g33 * (a4 * a4);
  sum = sum + a.m2 * mag(re[0], im[0], cplx);
  return 0.5 * sum;
}

// one block per (jj, kk) row of the region; partial[row] = w_j w_k sum_i w_i q
// (two partials per row -- real, imaginary -- for a complex field)
__global__ __launch_bounds__(256) void integrate_rows(const IntArgs a, double* __restrict__ partial) {
  __shared__ double red[2][256];
  const int nrow = a.Ny + 1;
  const int row = blockIdx.x;
  const int jr = row % nrow, kk = row / nrow;  // jj = g + jr, kk in [0, Nz]
  const int jj = a.g + jr;
  const bool cplx = a.nsys == 2;
  double sr = 0.0, si = 0.0;
  for (int ir = threadIdx.x; ir <= a.Nx; ir += 256) {
    const int ii = a.g + ir;
    const uint32_t t = (uint32_t)(jj * a.ld + ii);
    const int64_t o = (int64_t)kk * a.plane + t;
    const double w = (ir == 0 || ir == a.Nx) ? 0.5 : 1.0;
    const int64_t n2 = (int64_t)a.nj * a.ld;
    if (a.quantity == 5) {
      sr += w * (density(a, o, t, cplx) * a.tab[T_SQRTG * n2 + t]);
    } else if (a.quantity == 6) {
      // MoL%rho*MoL%alpha*MoL%sqrtg (initial_conditions.f90:86)
      sr += w * (density(a, o, t, cplx) * a.tab[T_ALPHA * n2 + t] * a.tab[T_SQRTG * n2 + t]);
    } else if (a.quantity == 7) {
      // MoL%m*abs(MoL%u(1,:,:,:))**2 *MoL%sqrtg (initial_conditions.f90:142)
      const double p = mag(a.u[0][o], cplx ? a.u[0][o + a.sys_stride] : 0.0, cplx);
      sr += w * (a.m * (p * p) * a.tab[T_SQRTG * n2 + t]);
    } else {
      sr += w * a.u[a.quantity][o];
      if (cplx) si += w * a.u[a.quantity][o + a.sys_stride];
    }
  }
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double wj = (jr == 0 || jr == a.Ny) ? 0.5 : 1.0;
    const double wk = (kk == 0 || kk == a.Nz) ? 0.5 : 1.0;
    partial[2 * row] = (wj * wk) * red[0][0];
    partial[2 * row + 1] = (wj * wk) * red[1][0];
  }
}

// one block: out = dx dy dz * sum of the row partials, in row order per thread
__global__ __launch_bounds__(256) void integrate_sum(const double* __restrict__ partial, int nrows, double scale,
                                                     double* __restrict__ out) {
  __shared__ double red[2][256];
  double sr = 0.0, si = 0.0;
  for (int r = threadIdx.x; r < nrows; r += 256) {
    sr += partial[2 * r];
    si += partial[2 * r + 1];
  }
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = scale * red[0][0];
    out[1] = scale * red[1][0];
  }
}

// Output quantities over the slab the reference writes (save_gnuplot.f90:
// 121-144: ii in [0, Nx+g] (i = -g..Nx), jj in [g, g+Ny], kk in [0, Nz-1]):
// 0 |phi|, 1 density(), 2 |psi_r|, 3 |psi_theta|, 4 |psi_phi|, 5 |Pi|,
// 6 Re phi, 7 Im phi.  out[k][j][i], i fastest (a Fortran-ordered
// (i, j, k) array, as the reference's HDF5 datasets hold it).
__global__ __launch_bounds__(256) void output_slab(const IntArgs a, int quantity, double* __restrict__ out) {
  const int ni_o = a.Nx + a.g + 1, nj_o = a.Ny + 1;
  const int ir = blockIdx.x * 256 + threadIdx.x;
  const int jr = blockIdx.y, kk = blockIdx.z;
  if (ir >= ni_o) return;
  const int jj = a.g + jr, ii = ir;
  const uint32_t t = (uint32_t)(jj * a.ld + ii);
  const int64_t o = (int64_t)kk * a.plane + t;
  const bool cplx = a.nsys == 2;
  double v;
  if (quantity == 1) {
    v = density(a, o, t, cplx);
  } else if (quantity == 6) {
    v = a.u[0][o];
  } else if (quantity == 7) {
    v = cplx ? a.u[0][o + a.sys_stride] : 0.0;
  } else {
    const int f = quantity == 0 ? 0 : quantity - 1;
    v = mag(a.u[f][o], cplx ? a.u[f][o + a.sys_stride] : 0.0, cplx);
  }
  out[((int64_t)kk * nj_o + jr) * ni_o + ir] = v;
}

thread_local std::string g_diag_error;

kg_status dfail(kg_status st, const char* msg, hipError_t e = hipSuccess) {
  char buf[400];
  snprintf(buf, sizeof buf, "%s%s%s", msg, e != hipSuccess ? ": " : "", e != hipSuccess ? hipGetErrorString(e) : "");
  g_diag_error = buf;
  return st;
}

IntArgs make_args(const HandleView& v, int quantity) {
  IntArgs a;
  for (int f = 0; f < 5; ++f) a.u[f] = v.u[f];
  a.tab = v.tab;
  a.sys_stride = v.G.sys_stride;
  a.plane = v.G.plane;
  a.ld = v.G.ld;
  a.nj = v.G.nj;
  a.g = v.G.g;
  a.Nx = v.G.Nx;
  a.Ny = v.G.Ny;
  a.Nz = v.G.Nz;
  a.nsys = v.nsys;
  a.quantity = quantity;
  a.m2 = v.m * v.m;
  a.m = v.m;
  return a;
}

}  // namespace

const char* kg_diag_last_error(void) { return g_diag_error.c_str(); }

namespace kg {
hipError_t diag_integrate(const HandleView& v, int quantity, double out[2]) {
  const IntArgs a = make_args(v, quantity);
  const int nrows = (a.Ny + 1) * (a.Nz + 1);
  double* buf = nullptr;
  hipError_t e = handle_scratch(v.handle, sizeof(double) * (2 * (size_t)nrows + 2), (void**)&buf);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(integrate_rows, dim3(nrows), dim3(256), 0, v.stream, a, buf + 2);
  e = hipGetLastError();
  if (e == hipSuccess) {
    const double scale = v.hg.dx * v.hg.dy * v.hg.dz;
    hipLaunchKernelGGL(integrate_sum, dim3(1), dim3(256), 0, v.stream, buf + 2, nrows, scale, buf);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, buf, 2 * sizeof(double), hipMemcpyDeviceToHost, v.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
  return e;
}
}  // namespace kg

kg_status kg_diag_integrate(kg_handle* h, int32_t quantity, double out[2]) {
  if (!h || !out) return dfail(KG_ERR_INVALID, "null argument");
  if (quantity < 0 || quantity > KG_DIAG_DENSITY) return dfail(KG_ERR_INVALID, "quantity must be 0..4 or KG_DIAG_DENSITY");
  HandleView v;
  if (handle_view(h, &v) != 0) return dfail(KG_ERR_INVALID, "bad handle");
  hipError_t e = hipSetDevice(v.device);
  if (e != hipSuccess) return dfail(KG_ERR_HIP, "hipSetDevice", e);
  e = diag_integrate(v, quantity, out);
  if (e == hipErrorOutOfMemory) return dfail(KG_ERR_NOMEM, "hipMalloc", e);
  if (e != hipSuccess) return dfail(KG_ERR_HIP, "integrate", e);
  return KG_OK;
}

kg_status kg_diag_output(kg_handle* h, int32_t quantity, double* out) {
  if (!h || !out) return dfail(KG_ERR_INVALID, "null argument");
  if (quantity < 0 || quantity > 7) return dfail(KG_ERR_INVALID, "quantity must be 0..7");
  HandleView v;
  if (handle_view(h, &v) != 0) return dfail(KG_ERR_INVALID, "bad handle");
  hipError_t e = hipSetDevice(v.device);
  if (e != hipSuccess) return dfail(KG_ERR_HIP, "hipSetDevice", e);
  const IntArgs a = make_args(v, quantity);
  const int ni_o = a.Nx + a.g + 1, nj_o = a.Ny + 1, nk_o = a.Nz;
  const size_t n = (size_t)ni_o * nj_o * nk_o;
  double* buf = nullptr;
  e = hipMalloc((void**)&buf, n * sizeof(double));
  if (e != hipSuccess) return dfail(KG_ERR_NOMEM, "hipMalloc", e);
  hipLaunchKernelGGL(output_slab, dim3((ni_o + 255) / 256, nj_o, nk_o), dim3(256), 0, v.stream, a, quantity, buf);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, buf, n * sizeof(double), hipMemcpyDeviceToHost, v.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
  (void)hipFree(buf);
  if (e != hipSuccess) return dfail(KG_ERR_HIP, "output", e);
  return KG_OK;
}
