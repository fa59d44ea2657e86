// kg_kernels.hip -- gfx950 kernels of the Klein-Gordon method-of-lines step.
//
// One launch per integrator stage: every thread owns one (r, theta) column
// point (ii, jj) of a 64 x 4 tile and marches over a chunk of phi planes.  At
// each point it evaluates the right-hand side of the five fields
// (MoL/lines_method_lib.f90:671-753) with the reference's 2nd-order stencils
// and edge rows (tools/finite_differences.f90:17-44,193-239,310-338,767-801)
// and applies the stage combine of the integrator in the same expression.
// The metric tables are k-independent, so a thread loads its coefficients
// once for the whole phi march.
//
// Floating-point contraction is off for this file (build flag
// -ffp-contract=off): every product and sum is rounded on its own, in the
// reference's left-to-right order, so the results match a CPU evaluation of
// the Fortran expressions bit for bit.
#include <hip/hip_runtime.h>

#include "kg_internal.h"

namespace kg {
namespace {

constexpr int TI = 64;  // tile width along r (one wavefront)
constexpr int TJ = 4;   // tile height along theta

struct View {
  const double* __restrict__ tab;
  int64_t n2;  // nj*ld
  int ld;
  int64_t plane;
  __device__ __forceinline__ double t(int which, int ii, int jj) const {
    return tab[which * n2 + (int64_t)jj * ld + ii];
  }
  __device__ __forceinline__ double f(const double* __restrict__ p, int ii, int jj, int kk) const {
    return p[(int64_t)kk * plane + (int64_t)jj * ld + ii];
  }
};

// D_r (first_derivative_x_2_uniform, finite_differences.f90:36-42)
template <class Q>
__device__ __forceinline__ double d_r(const Q& q, int ii, int ni, double res) {
  if (ii == 0) return (-3.0 * q(0) + 4.0 * q(1) - q(2)) * res;
  if (ii == ni - 1) return (+3.0 * q(ni - 1) - 4.0 * q(ni - 2) + q(ni - 3)) * res;
  return (-q(ii - 1) + q(ii + 1)) * res;
}

// A_r (advec_x_uniform, finite_differences.f90:779-799)
template <class Q>
__device__ __forceinline__ double a_r(const Q& q, int ii, int ni, double beta, double res) {
  if (ii <= 1) return (-3.0 * q(ii) + 4.0 * q(ii + 1) - q(ii + 2)) * res;
  if (ii == ni - 2) return (q(ni - 1) - q(ni - 3)) * res;
  if (ii == ni - 1) return (3.0 * q(ni - 1) - 4.0 * q(ni - 2) + q(ni - 3)) * res;
  if (beta > 0.0) return (-3.0 * q(ii) + 4.0 * q(ii + 1) - q(ii + 2)) * res;
  if (beta < 0.0) return (3.0 * q(ii) - 4.0 * q(ii - 1) + q(ii - 2)) * res;
  return (q(ii + 1) - q(ii - 1)) * res;
}

// D_theta (first_derivative_y_2 with inpar = .false., finite_differences.f90:214-230)
template <class Q>
__device__ __forceinline__ double d_th(const Q& q, int jj, int kk, int nj, int kt, double res) {
  if (jj == 0) return (-q(1, kk) + q(2, kt)) * res;
  if (jj == nj - 1) return (+3.0 * q(nj - 1, kk) - 4.0 * q(nj - 2, kk) + q(nj - 3, kk)) * res;
  return (-q(jj - 1, kk) + q(jj + 1, kk)) * res;
}

// D_phi (first_derivative_z_2, finite_differences.f90:328-335)
template <class Q>
__device__ __forceinline__ double d_ph(const Q& q, int kk, int Nz, double res) {
  if (kk == 0) return (-q(Nz - 2) + q(1)) * res;
  if (kk == Nz) return (-q(2) + q(Nz - 1)) * res;
  return (-q(kk - 1) + q(kk + 1)) * res;
}

template <int MODE>
struct ModeTraits {
  static constexpr bool need14 = !(MODE == M_PIRK_1B || MODE == M_PIRK3_2B);
  static constexpr bool need5 = !(MODE == M_PIRK_1A || MODE == M_PIRK3_2A);
};

template <int MODE, int EF>
__global__ __launch_bounds__(TI* TJ) void stage_kernel(const StageArgs a) {
  const Geom& G = a.G;
  const int b = blockIdx.x;
  const int ti = b % a.nti;
  const int rem = b / a.nti;
  const int tj = rem % a.ntj;
  const int tk = rem / a.ntj;
  const int ii = a.i_lo + ti * TI + (int)threadIdx.x;
  const int jj = a.j_lo + tj * TJ + (int)threadIdx.y;
  if (ii >= a.i_hi || jj >= a.j_hi) return;
  const int k0 = a.k_lo + tk * a.kc;
  const int k1 = min(k0 + a.kc, a.k_hi);
  const int64_t sys = (int64_t)blockIdx.y * G.sys_stride;  // system (complex: real, imaginary)

  const View V{a.tab, (int64_t)G.nj * G.ld, G.ld, G.plane};
  const int ni = G.ni, nj = G.nj, Nz = G.Nz;
  const double dt = G.dt;

  const double* __restrict__ f_phi = a.in[0] + sys;
  const double* __restrict__ f_psr = a.in[1] + sys;
  const double* __restrict__ f_pst = a.in[2] + sys;
  const double* __restrict__ f_psp = a.in[3] + sys;
  const double* __restrict__ f_pi = a.in[4] + sys;

  // coefficients of this (r, theta) point
  const double al = V.t(T_ALPHA, ii, jj);
  const double be = V.t(T_BETA, ii, jj);
  const double dxb = V.t(T_DXBETA, ii, jj);
  const double sg = V.t(T_SQRTG, ii, jj);
  const double asg = V.t(T_ASG, ii, jj);
  const double grp = V.t(T_GRP, ii, jj);
  const double gpp = V.t(T_GPP, ii, jj);
  const double am2 = V.t(T_AM2, ii, jj);

  const int dis = Nz / 2;
  for (int kk = k0; kk < k1; ++kk) {
    const int kt = (kk + dis) % (Nz - 1);
    double R[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double Pi = V.f(f_pi, ii, jj, kk);
    const double psr = V.f(f_psr, ii, jj, kk);
    if constexpr (ModeTraits<MODE>::need14) {
      // rhs_phi (:675)
      R[0] = al * Pi + be * psr;
      // rhs_psi1 (:687-689)
      auto aPi = [&](int i) { return V.t(T_ALPHA, i, jj) * V.f(f_pi, i, jj, kk); };
      auto psr_r = [&](int i) { return V.f(f_psr, i, jj, kk); };
      R[1] = d_r(aPi, ii, ni, G.rr) + be * a_r(psr_r, ii, ni, be, G.rr) + dxb * psr;
      // rhs_psi2 (:700) and rhs_psi3 (:711): D(alpha*Pi + beta*psi_r)
      auto Fth = [&](int j, int k) {
        return V.t(T_ALPHA, ii, j) * V.f(f_pi, ii, j, k) + V.t(T_BETA, ii, j) * V.f(f_psr, ii, j, k);
      };
      auto Fph = [&](int k) { return al * V.f(f_pi, ii, jj, k) + be * V.f(f_psr, ii, jj, k); };
      R[2] = d_th(Fth, jj, kk, nj, kt, G.rt);
      R[3] = d_ph(Fph, kk, Nz, G.rp);
    }
    if constexpr (ModeTraits<MODE>::need5) {
      // rhs_pi (:728-736 Kerr_Schild; :738-743 Schw_EF)
      auto fr = [&](int i) {
        if constexpr (EF)
          return V.t(T_ASG, i, jj) * V.t(T_GRR, i, jj) * V.f(f_psr, i, jj, kk);
        else
          return V.t(T_ASG, i, jj) *
                 (V.t(T_GRR, i, jj) * V.f(f_psr, i, jj, kk) + V.t(T_GRP, i, jj) * V.f(f_psp, i, jj, kk));
      };
      auto ft = [&](int j, int k) { return V.t(T_ASGTT, ii, j) * V.f(f_pst, ii, j, k); };
      auto fp = [&](int k) {
        if constexpr (EF)
          return (al * gpp) * V.f(f_psp, ii, jj, k);
        else
          return asg * (grp * V.f(f_psr, ii, jj, k) + gpp * V.f(f_psp, ii, jj, k));
      };
      auto ps = [&](int i) { return V.f(f_pi, i, jj, kk) * V.t(T_SQRTG, i, jj); };
      const double t1 = d_r(fr, ii, ni, G.rr) / sg;
      const double t2 = d_th(ft, jj, kk, nj, kt, G.rt) / sg;
      const double t3 = EF ? d_ph(fp, kk, Nz, G.rp) : d_ph(fp, kk, Nz, G.rp) / sg;
      const double t4 = Pi * dxb;
      const double t5 = be * a_r(ps, ii, ni, be, G.rr) / sg;
      const double t6 = am2 * V.f(f_phi, ii, jj, kk);
      R[4] = t1 + t2 + t3 + t4 + t5 - t6;
    }

    const int64_t p = (int64_t)kk * G.plane + (int64_t)jj * G.ld + ii + sys;
    if constexpr (MODE == M_RK4_S1 || MODE == M_RK4_S2 || MODE == M_RK4_S3) {
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const double k = R[c] * dt;
        double acc;
        if constexpr (MODE == M_RK4_S1)
          acc = k;
        else
          acc = a.aux[c][p] + 2.0 * k;
        a.out2[c][p] = acc;
        if constexpr (MODE == M_RK4_S3)
          a.out[c][p] = a.base[c][p] + k;
        else
          a.out[c][p] = a.base[c][p] + 0.50 * k;
      }
    } else if constexpr (MODE == M_RK4_S4) {
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const double k = R[c] * dt;
        a.out[c][p] = a.base[c][p] + (a.aux[c][p] + k) / 6.0;
      }
    } else if constexpr (MODE == M_RK3_S1) {
#pragma unroll
      for (int c = 0; c < 5; ++c) a.out[c][p] = a.base[c][p] + R[c] * dt;
    } else if constexpr (MODE == M_RK3_S2) {
#pragma unroll
      for (int c = 0; c < 5; ++c) a.out[c][p] = 0.750 * a.base[c][p] + 0.250 * (a.in[c][p] + R[c] * dt);
    } else if constexpr (MODE == M_RK3_S3) {
#pragma unroll
      for (int c = 0; c < 5; ++c) a.out[c][p] = a.base[c][p] / 3.0 + 2.0 * (a.in[c][p] + R[c] * dt) / 3.0;
    } else if constexpr (MODE == M_PIRK_1A) {
#pragma unroll
      for (int c = 0; c < 4; ++c) a.out[c][p] = a.base[c][p] + dt * R[c];
    } else if constexpr (MODE == M_PIRK_1B) {
      a.out[4][p] = a.base[4][p] + dt * R[4];
    } else if constexpr (MODE == M_PIRK2_2) {
#pragma unroll
      for (int c = 0; c < 5; ++c) a.out[c][p] = 0.5 * (a.base[c][p] + a.in[c][p] + dt * R[c]);
    } else if constexpr (MODE == M_PIRK3_2A) {
#pragma unroll
      for (int c = 0; c < 4; ++c) a.out[c][p] = 0.75 * a.base[c][p] + 0.25 * (a.aux[c][p] + dt * R[c]);
    } else if constexpr (MODE == M_PIRK3_2B) {
      a.out[4][p] = 0.75 * a.base[4][p] + 0.25 * (a.aux[4][p] + dt * R[4]);
    } else if constexpr (MODE == M_PIRK3_3) {
      const double third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
#pragma unroll
      for (int c = 0; c < 5; ++c) a.out[c][p] = third * a.base[c][p] + two_thirds * (a.in[c][p] + dt * R[c]);
    }
  }
}

// ---------------------------------------------------------------------------
// Ghost fill (lines_method_lib.f90:313-330,507-597), three ordered passes.
// Each pass only reads points it does not write.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fill_x_kernel(const FillArgs a) {
  const Geom& G = a.G;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.z % 5;
  const int64_t sys = (int64_t)(blockIdx.z / 5) * G.sys_stride;
  if (t >= (int64_t)G.nj * G.nk) return;
  const int jj = (int)(t % G.nj), kk = (int)(t / G.nj);
  double* u = a.u[c] + sys + (int64_t)kk * G.plane + (int64_t)jj * G.ld;
  // boundary_xL (:516-517)
  u[0] = 3.0 * u[1] - 3.0 * u[2] + u[3];
  // boundary_xR (:525-528)
  if (G.bc_inflow) {
    const double* up = a.u_p[c] + sys + (int64_t)kk * G.plane + (int64_t)jj * G.ld;
    u[G.ni - 1] = up[G.ni - 1];
  } else {
    u[G.ni - 1] = u[G.ni - 2];
  }
}

__global__ __launch_bounds__(256) void fill_y_kernel(const FillArgs a) {
  // boundary_yL (:542-555) and boundary_yR (:565-578): whole r-lines,
  // rows -gg <- gg+1 and Ny+gg <- Ny-gg-1, taken at k_target = mod(k+Nz/2, Nz-1)
  const Geom& G = a.G;
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii >= G.ni) return;
  const int nrow = G.g + 1;
  const int row = blockIdx.y % (2 * nrow);
  const int kk = blockIdx.y / (2 * nrow);
  const int c = blockIdx.z % 5;
  const int64_t sys = (int64_t)(blockIdx.z / 5) * G.sys_stride;
  const int kt = (kk + G.Nz / 2) % (G.Nz - 1);
  int jd, js;
  if (row < nrow) {
    const int gg = row;
    jd = G.g - gg;          // j = -gg
    js = G.g + gg + 1;      // j = gg+1
  } else {
    const int gg = row - nrow;
    jd = G.g + G.Ny + gg;       // j = Ny+gg
    js = G.g + G.Ny - gg - 1;   // j = Ny-gg-1
  }
  double* u = a.u[c] + sys;
  u[(int64_t)kk * G.plane + (int64_t)jd * G.ld + ii] = u[(int64_t)kt * G.plane + (int64_t)js * G.ld + ii];
}

__global__ __launch_bounds__(256) void fill_z_kernel(const FillArgs a) {
  // boundary_zL (:587) and boundary_zR (:596)
  const Geom& G = a.G;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.plane) return;
  const int c = blockIdx.z % 5;
  const int64_t sys = (int64_t)(blockIdx.z / 5) * G.sys_stride;
  double* u = a.u[c] + sys;
  if (blockIdx.y == 0)
    u[t] = u[(int64_t)(G.Nz - 1) * G.plane + t];
  else
    u[(int64_t)G.Nz * G.plane + t] = u[G.plane + t];
}

// dense [sys][5][nk][nj][ni] <-> pitched fields
__global__ __launch_bounds__(256) void pack_kernel(const Geom G, double* __restrict__ src, double* d0,
                                                   double* d1, double* d2, double* d3, double* d4, int to_dev) {
  const int nbx = (G.ni + 63) / 64;
  const int ii = (blockIdx.x % nbx) * 64 + threadIdx.x;
  const int jj = blockIdx.x / nbx;
  const int kk = blockIdx.y;
  if (ii >= G.ni) return;
  const int c = blockIdx.z % 5;
  const int s = blockIdx.z / 5;
  double* dst[5] = {d0, d1, d2, d3, d4};
  const int64_t dense = ((((int64_t)s * 5 + c) * G.nk + kk) * G.nj + jj) * G.ni + ii;
  const int64_t pitched = (int64_t)s * G.sys_stride + (int64_t)kk * G.plane + (int64_t)jj * G.ld + ii;
  if (to_dev)
    dst[c][pitched] = src[dense];
  else
    src[dense] = dst[c][pitched];
}

template <int MODE>
hipError_t launch_mode(const StageArgs& a, hipStream_t s, int nsys) {
  dim3 grid((unsigned)(a.nti * a.ntj * a.ntk), (unsigned)nsys, 1);
  dim3 block(TI, TJ, 1);
  if (a.G.metric_ef)
    hipLaunchKernelGGL((stage_kernel<MODE, 1>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((stage_kernel<MODE, 0>), grid, block, 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_stage(int mode, const StageArgs& a0, hipStream_t s) {
  StageArgs a = a0;
  const int nsys = a.nsys;
  a.nti = (a.i_hi - a.i_lo + TI - 1) / TI;
  a.ntj = (a.j_hi - a.j_lo + TJ - 1) / TJ;
  a.ntk = (a.k_hi - a.k_lo + a.kc - 1) / a.kc;
  switch (mode) {
    case M_RK4_S1: return launch_mode<M_RK4_S1>(a, s, nsys);
    case M_RK4_S2: return launch_mode<M_RK4_S2>(a, s, nsys);
    case M_RK4_S3: return launch_mode<M_RK4_S3>(a, s, nsys);
    case M_RK4_S4: return launch_mode<M_RK4_S4>(a, s, nsys);
    case M_RK3_S1: return launch_mode<M_RK3_S1>(a, s, nsys);
    case M_RK3_S2: return launch_mode<M_RK3_S2>(a, s, nsys);
    case M_RK3_S3: return launch_mode<M_RK3_S3>(a, s, nsys);
    case M_PIRK_1A: return launch_mode<M_PIRK_1A>(a, s, nsys);
    case M_PIRK_1B: return launch_mode<M_PIRK_1B>(a, s, nsys);
    case M_PIRK2_2: return launch_mode<M_PIRK2_2>(a, s, nsys);
    case M_PIRK3_2A: return launch_mode<M_PIRK3_2A>(a, s, nsys);
    case M_PIRK3_2B: return launch_mode<M_PIRK3_2B>(a, s, nsys);
    case M_PIRK3_3: return launch_mode<M_PIRK3_3>(a, s, nsys);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fill(const FillArgs& a, hipStream_t s, int phase, int nsys) {
  const Geom& G = a.G;
  if (phase == 0) {
    const int64_t n = (int64_t)G.nj * G.nk;
    hipLaunchKernelGGL(fill_x_kernel, dim3((unsigned)((n + 255) / 256), 1, 5 * nsys), dim3(256), 0, s, a);
  } else if (phase == 1) {
    hipLaunchKernelGGL(fill_y_kernel, dim3((unsigned)((G.ni + 63) / 64), 2 * (G.g + 1) * G.nk, 5 * nsys),
                       dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL(fill_z_kernel, dim3((unsigned)((G.plane + 255) / 256), 2, 5 * nsys), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_pack(const Geom& G, double* src, double* const d[5], int nsys, int to_dev, hipStream_t s) {
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((G.ni + 63) / 64 * G.nj), (unsigned)G.nk, 5 * nsys), dim3(64),
                     0, s, G, src, d[0], d[1], d[2], d[3], d[4], to_dev);
  return hipGetLastError();
}

}  // namespace kg
