// kg_initial.hip -- initial data generated where the state lives: the
// reference's initial_conditions (main/initial_conditions.f90:23-186) for the
// types 'spherical_shell', 'random' and 'random_shell', including the
// Cartesian random field of tools/random.f90:17-59 (a 3-D FFT written here,
// following tools/fft_lib.f90:141-264) and its trilinear sampling onto the
// spherical grid (random.f90:72-156).
//
// Every value that depends on one coordinate only (radii, the shell envelope,
// sin/cos of the grid angles, FFT twiddles, Cartesian frequencies) is computed
// on the host with the same expressions the reference evaluates and uploaded,
// so the spherical shell, the stencils, the cell search and the FFT butterflies
// repeat the reference's operations bit for bit.  The per-point transcendental
// functions of the random spectrum (exp of the Gaussian, cos/sin of the random
// phase) use the device library and agree to rounding.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kg.h"
#include "kg_internal.h"

using namespace kg;

namespace {

constexpr int kFftLds = 4096;  // complex values per FFT workgroup (64 KiB of LDS)

thread_local std::string g_ic_error;

kg_status ifail(kg_status st, const char* msg, hipError_t e = hipSuccess) {
  char buf[400];
  snprintf(buf, sizeof buf, "%s%s%s", msg, e != hipSuccess ? ": " : "", e != hipSuccess ? hipGetErrorString(e) : "");
  g_ic_error = buf;
  return st;
}

// ---------------------------------------------------------------- random field

// Uniform numbers in [0, 1) when the caller supplies none: SplitMix64
// (Steele, Lea and Flood, OOPSLA 2014) on the counter seed + (p+1) * golden
// gamma, top 53 bits.  The reference draws from an unseeded random_number.
__host__ __device__ inline double uniform01(uint64_t seed, uint64_t p) {
  uint64_t z = seed + (p + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * 0x1.0p-53;
}

// conj(psi_p), psi_p = exp(-0.5/p0**2 * (wx**2 + wy**2 + wz**2)) exp(i 2 pi U)
// (random.f90:53), at [k][j][i]; om = the shifted frequencies of one axis
__global__ __launch_bounds__(256) void psi_p_kernel(double2* __restrict__ out, const double* __restrict__ om, int log2n,
                                                    double q, double twopi, const double* __restrict__ uni,
                                                    uint64_t seed) {
  const int64_t n3 = (int64_t)1 << (3 * log2n);
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n3) return;
  const int n = 1 << log2n;
  const int i = (int)(p & (n - 1)), j = (int)((p >> log2n) & (n - 1)), k = (int)(p >> (2 * log2n));
  double w2 = om[i] * om[i] + om[j] * om[j];
  w2 = w2 + om[k] * om[k];
  const double amp = exp(q * w2);
  const double u = uni ? uni[p] : uniform01(seed, (uint64_t)p);
  const double th = twopi * u;
  out[p] = make_double2(amp * cos(th), -(amp * sin(th)));
}

__device__ __forceinline__ int bitrev(int p, int bits) { return (int)(__brev((unsigned)p) >> (32 - bits)); }

// One workgroup transforms `nl` lines of length n = 2**log2n along one axis of
// an [n][n][n] array: fft_1D (fft_lib.f90:141-176) unrolled into its
// butterfly levels.  The recursion splits f into f(1::2), f(2::2) down to
// length 1, where DFT (:114-138) returns 0 + f*exp(0); level `len` then forms
// E + w O and E - w O with w = exp(2 pi i k/len).  Loading the line in
// bit-reversed order puts every sub-transform in a contiguous block, so the
// levels run in place in LDS.  Axis 0 (i, contiguous) lines are whole rows;
// for axes 1 and 2 a workgroup takes nl consecutive i at one (k) or one (j),
// so every load and store moves runs of nl contiguous values.  LAST: the
// final axis also applies ifft_3D's conj(.)/n**3 (:252-264).
template <bool LAST>
__global__ __launch_bounds__(256) void fft_lines_kernel(double2* __restrict__ a, const double2* __restrict__ tw,
                                                        int log2n, int nl, int axis, double nn) {
  __shared__ double2 s[kFftLds];
  const int n = 1 << log2n;
  const int64_t n2 = (int64_t)n * n;
  int64_t base, sl, sp;
  if (axis == 0) {
    base = (int64_t)blockIdx.x * nl * n;
    sl = n;
    sp = 1;
  } else {
    const int nb = n / nl;
    const int ib = (int)(blockIdx.x % nb), outer = (int)(blockIdx.x / nb);
    base = (axis == 1 ? (int64_t)outer * n2 : (int64_t)outer * n) + (int64_t)ib * nl;
    sl = 1;
    sp = axis == 1 ? n : n2;
  }
  const int tot = nl * n;
  for (int e = threadIdx.x; e < tot; e += 256) {
    int l, p;
    if (axis == 0) {
      l = e >> log2n;
      p = e & (n - 1);
    } else {
      l = e % nl;
      p = e / nl;
    }
    const double2 x = a[base + l * sl + p * sp];
    double2 y;
    y.x = 0.0 + (x.x * 1.0 - x.y * 0.0);
    y.y = 0.0 + (x.x * 0.0 + x.y * 1.0);
    s[l * n + bitrev(p, log2n)] = y;
  }
  __syncthreads();
  for (int lev = 1; lev <= log2n; ++lev) {
    const int half = 1 << (lev - 1);
    const int tws = n >> lev;
    for (int e = threadIdx.x; e < tot / 2; e += 256) {
      const int l = e >> (log2n - 1);
      const int b = e & (n / 2 - 1);
      const int k = b & (half - 1);
      const int i0 = ((b >> (lev - 1)) << lev) + k;
      const int i1 = i0 + half;
      const double2 w = tw[k * tws];
      const double2 E = s[l * n + i0], O = s[l * n + i1];
      const double tr = w.x * O.x - w.y * O.y;
      const double ti = w.x * O.y + w.y * O.x;
      s[l * n + i0] = make_double2(E.x + tr, E.y + ti);
      s[l * n + i1] = make_double2(E.x - tr, E.y - ti);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < tot; e += 256) {
    int l, p;
    if (axis == 0) {
      l = e >> log2n;
      p = e & (n - 1);
    } else {
      l = e % nl;
      p = e / nl;
    }
    double2 v = s[l * n + p];
    if (LAST) v = make_double2(v.x / nn, (-v.y) / nn);
    a[base + l * sl + p * sp] = v;
  }
}

// ---------------------------------------------------------------- spherical grid

struct IcTables {
  const double* x;      // [ni] radius x(i)
  const double* env;    // [ni] exp(-(x-mu)**2/sigma**2)
  const double* shell;  // [ni] 1/x * env
  const double* dtphi;  // [ni] -2 (x-mu)/(x sigma**2) * env
  const double* st;     // [nj] sin(theta_j)
  const double* ct;     // [nj] cos(theta_j)
  const double* cp;     // [nk] cos(phi_k)
  const double* sp;     // [nk] sin(phi_k)
};

struct IcArgs {
  Geom G;
  int nsys;
  double* u[5];
  const double* tab;
  IcTables T;
};

// the first cell i in [0, n-2] with node(i) <= v < node(i+1), node(i) = c0 +
// i*h (the linear search of random.f90:80-88), found from an estimate; the
// host has checked that every sampled point lies inside [node(0), node(n-1))
__device__ inline int find_cell(double v, double c0, double h, int n) {
  const double t = (v - c0) / h;
  int i = t >= 0.0 ? (t < (double)(n - 2) ? (int)t : n - 2) : 0;
  while (i > 0 && !(c0 + (double)i * h <= v)) --i;
  while (i < n - 2 && !(c0 + (double)(i + 1) * h > v)) ++i;
  return i;
}

__device__ inline double clamp01(double d) { return d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d); }

// u(1) at every padded point: convert_to_spherical (random.f90:128-156) with
// trilinear_interpolation (:72-126); random_shell (initial_conditions.f90:158)
// then multiplies by 1/x * exp(-(x-mu)**2/sigma**2)
__global__ __launch_bounds__(256) void sample_kernel(const IcArgs a, const double2* __restrict__ psi, int log2n,
                                                     double c0, double h, int envelope) {
  const Geom& G = a.G;
  const int ii = blockIdx.x * 256 + threadIdx.x;
  const int jj = blockIdx.y, kk = blockIdx.z;
  if (ii >= G.ni) return;
  const int n = 1 << log2n;
  const double r = a.T.x[ii];
  const double xc = r * a.T.st[jj] * a.T.cp[kk];
  const double yc = r * a.T.st[jj] * a.T.sp[kk];
  const double zc = r * a.T.ct[jj];
  const int i = find_cell(xc, c0, h, n), j = find_cell(yc, c0, h, n), k = find_cell(zc, c0, h, n);
  const double xd = clamp01((xc - (c0 + (double)i * h)) / ((c0 + (double)(i + 1) * h) - (c0 + (double)i * h)));
  const double yd = clamp01((yc - (c0 + (double)j * h)) / ((c0 + (double)(j + 1) * h) - (c0 + (double)j * h)));
  const double zd = clamp01((zc - (c0 + (double)k * h)) / ((c0 + (double)(k + 1) * h) - (c0 + (double)k * h)));
  const int64_t sn = n, sn2 = (int64_t)n * n;
  const int64_t b = (int64_t)k * sn2 + (int64_t)j * sn + i;
  const double2 c000 = psi[b], c100 = psi[b + 1], c010 = psi[b + sn], c001 = psi[b + sn2];
  const double2 c110 = psi[b + sn + 1], c101 = psi[b + sn2 + 1], c011 = psi[b + sn2 + sn];
  const double2 c111 = psi[b + sn2 + sn + 1];
  double v[2];
#pragma unroll
  for (int part = 0; part < 2; ++part) {
    const double f000 = part ? c000.y : c000.x, f100 = part ? c100.y : c100.x;
    const double f010 = part ? c010.y : c010.x, f001 = part ? c001.y : c001.x;
    const double f110 = part ? c110.y : c110.x, f101 = part ? c101.y : c101.x;
    const double f011 = part ? c011.y : c011.x, f111 = part ? c111.y : c111.x;
    double s = f000 * (1.0 - xd) * (1.0 - yd) * (1.0 - zd);
    s = s + f100 * xd * (1.0 - yd) * (1.0 - zd);
    s = s + f010 * (1.0 - xd) * yd * (1.0 - zd);
    s = s + f001 * (1.0 - xd) * (1.0 - yd) * zd;
    s = s + f110 * xd * yd * (1.0 - zd);
    s = s + f101 * xd * (1.0 - yd) * zd;
    s = s + f011 * (1.0 - xd) * yd * zd;
    s = s + f111 * xd * yd * zd;
    if (envelope) s = s * 1.0 / r * a.T.env[ii];
    v[part] = s;
  }
  const int64_t o = (int64_t)kk * G.plane + (int64_t)jj * G.ld + ii;
  a.u[0][o] = v[0];
  a.u[0][o + G.sys_stride] = v[1];
}

// spherical_shell (initial_conditions.f90:59-63): u(1) = 1/x exp(..), u(2) = 0
// (its assignment is commented out there)
__global__ __launch_bounds__(256) void shell_kernel(const IcArgs a) {
  const Geom& G = a.G;
  const int ii = blockIdx.x * 256 + threadIdx.x;
  const int jj = blockIdx.y, kk = blockIdx.z;
  if (ii >= G.ni) return;
  const int64_t o = (int64_t)kk * G.plane + (int64_t)jj * G.ld + ii;
  a.u[0][o] = a.T.shell[ii];
  a.u[1][o] = 0.0;
  if (a.nsys == 2) {
    a.u[0][o + G.sys_stride] = 0.0;
    a.u[1][o + G.sys_stride] = 0.0;
  }
}

// u(2), u(3), u(4) = dx1, dx2, dx3 of u(1) on the whole padded array
// (finite_differences.f90:36-42, :214-230 with inpar = .false., :328-335),
// the bits of `mask` selecting which
__global__ __launch_bounds__(256) void deriv_kernel(const IcArgs a, int mask) {
  const Geom& G = a.G;
  const int ii = blockIdx.x * 256 + threadIdx.x;
  const int jj = blockIdx.y, kk = blockIdx.z;
  if (ii >= G.ni) return;
  for (int sy = 0; sy < a.nsys; ++sy) {
    const double* f = a.u[0] + (int64_t)sy * G.sys_stride;
    const int64_t o = (int64_t)kk * G.plane + (int64_t)jj * G.ld + ii;
    if (mask & 1) {
      const double* r = f + (int64_t)kk * G.plane + (int64_t)jj * G.ld;
      double d;
      if (ii == 0)
        d = (-3.0 * r[0] + 4.0 * r[1] - r[2]) * G.rr;
      else if (ii == G.ni - 1)
        d = (+3.0 * r[ii] - 4.0 * r[ii - 1] + r[ii - 2]) * G.rr;
      else
        d = (-r[ii - 1] + r[ii + 1]) * G.rr;
      a.u[1][o + (int64_t)sy * G.sys_stride] = d;
    }
    if (mask & 2) {
      const double* c = f + (int64_t)kk * G.plane + ii;  // column at (ii, kk), rows ld apart
      double d;
      if (jj == 0) {
        const int kt = (kk + G.Nz / 2) % (G.Nz - 1);
        d = (-c[(int64_t)1 * G.ld] + f[(int64_t)kt * G.plane + (int64_t)2 * G.ld + ii]) * G.rt;
      } else if (jj == G.nj - 1) {
        d = (+3.0 * c[(int64_t)jj * G.ld] - 4.0 * c[(int64_t)(jj - 1) * G.ld] + c[(int64_t)(jj - 2) * G.ld]) * G.rt;
      } else {
        d = (-c[(int64_t)(jj - 1) * G.ld] + c[(int64_t)(jj + 1) * G.ld]) * G.rt;
      }
      a.u[2][o + (int64_t)sy * G.sys_stride] = d;
    }
    if (mask & 4) {
      const double* c = f + (int64_t)jj * G.ld + ii;  // (ii, jj), planes `plane` apart
      const int hi = G.nk - 1;
      double d;
      if (kk == 0)
        d = (-c[(int64_t)(hi - 2) * G.plane] + c[G.plane]) * G.rp;
      else if (kk == hi)
        d = (-c[(int64_t)2 * G.plane] + c[(int64_t)(hi - 1) * G.plane]) * G.rp;
      else
        d = (-c[(int64_t)(kk - 1) * G.plane] + c[(int64_t)(kk + 1) * G.plane]) * G.rp;
      a.u[3][o + (int64_t)sy * G.sys_stride] = d;
    }
  }
}

// u(5) = 1/alpha * (partial_t_phi - beta^r u(2)) (initial_conditions.f90:81,
// :136, :165); mode 0: partial_t_phi = the shell's dtphi (imaginary part 0),
// 1: 0, 2: u(2)
__global__ __launch_bounds__(256) void pi_kernel(const IcArgs a, int mode) {
  const Geom& G = a.G;
  const int ii = blockIdx.x * 256 + threadIdx.x;
  const int jj = blockIdx.y, kk = blockIdx.z;
  if (ii >= G.ni) return;
  const int64_t n2 = (int64_t)G.nj * G.ld;
  const int64_t t = (int64_t)jj * G.ld + ii;
  const double alpha = a.tab[T_ALPHA * n2 + t], beta = a.tab[T_BETA * n2 + t];
  for (int sy = 0; sy < a.nsys; ++sy) {
    const int64_t o = (int64_t)kk * G.plane + t + (int64_t)sy * G.sys_stride;
    const double u2 = a.u[1][o];
    const double dtphi = mode == 0 ? (sy == 0 ? a.T.dtphi[ii] : 0.0) : (mode == 1 ? 0.0 : u2);
    a.u[4][o] = 1.0 / alpha * (dtphi - beta * u2);
  }
}

// MoL%u = Amp * MoL%u over every padded point of the five fields
__global__ __launch_bounds__(256) void scale_kernel(const IcArgs a, double amp) {
  const Geom& G = a.G;
  const int ii = blockIdx.x * 256 + threadIdx.x;
  const int jj = blockIdx.y, kk = blockIdx.z;
  if (ii >= G.ni) return;
  const int64_t o = (int64_t)kk * G.plane + (int64_t)jj * G.ld + ii;
  for (int f = 0; f < 5; ++f)
    for (int sy = 0; sy < a.nsys; ++sy) a.u[f][o + (int64_t)sy * G.sys_stride] = amp * a.u[f][o + (int64_t)sy * G.sys_stride];
}

// ---------------------------------------------------------------- host side

int ilog2_exact(int n) {
  int b = 0;
  while ((1 << b) < n) ++b;
  return (1 << b) == n ? b : -1;
}

// RAII device buffers
struct DevBuf {
  std::vector<void*> p;
  ~DevBuf() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t alloc(T** out, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e == hipSuccess) {
      p.push_back(q);
      *out = (T*)q;
    }
    return e;
  }
};

// Cartesian box of random.f90:25-47: R = r_max + 1, nodes -R + i dx, dx = 2R/n
struct CartBox {
  int n = 0, log2n = 0, nl = 0;
  double R = 0.0, c0 = 0.0, h = 0.0;
};

kg_status cart_box(int n, double r_max, CartBox* b) {
  const int l2 = ilog2_exact(n);
  if (l2 < 2 || n > 1024) return ifail(KG_ERR_INVALID, "ncart must be a power of two in [4, 1024]");
  if (!std::isfinite(r_max) || !(r_max > 0.0)) return ifail(KG_ERR_INVALID, "r_max must be positive");
  b->n = n;
  b->log2n = l2;
  b->nl = n < kFftLds / n ? n : kFftLds / n;
  b->R = r_max + 1.0;
  const double xmin = -b->R, xmax = b->R;
  b->c0 = xmin;
  b->h = (xmax - xmin) / (double)n;
  return KG_OK;
}

// conj(fft_3D(conj(psi_p)))/n**3 into d_psi ([k][j][i] complex), on stream s
kg_status random_field(const CartBox& B, double p0, uint64_t seed, const double* host_uniforms, double2* d_psi,
                       hipStream_t s, DevBuf& mem) {
  if (!std::isfinite(p0) || !(p0 > 0.0)) return ifail(KG_ERR_INVALID, "p0 must be positive");
  const int n = B.n;
  const double pii = std::acos(-1.0);
  // omega(i) = 2 pi (i-1-n/2)/(n dx) (random.f90:42-44), shifted (shift_3D,
  // fft_lib.f90:329-349): position p holds the unshifted index (p + n/2) mod n
  std::vector<double> om(n);
  for (int p = 0; p < n; ++p) {
    const int qi = (p + n / 2) % n;
    om[p] = 2.0 * pii * (double)(qi - n / 2) / ((double)n * B.h);
  }
  // twiddles exp(2 pi i k/n), k < n/2 (fft_lib.f90:163-167); level len uses
  // entry k n/len, which holds the same double as 2 pi k/len
  std::vector<double2> tw(n / 2);
  for (int k = 0; k < n / 2; ++k) {
    const double th = 2.0 * pii * (double)k / (double)n;
    tw[k] = make_double2(std::cos(th), std::sin(th));
  }
  double* d_om = nullptr;
  double2* d_tw = nullptr;
  double* d_uni = nullptr;
  const int64_t n3 = (int64_t)n * n * n;
  hipError_t e = mem.alloc(&d_om, (size_t)n);
  if (e == hipSuccess) e = mem.alloc(&d_tw, (size_t)n / 2);
  if (e == hipSuccess && host_uniforms) e = mem.alloc(&d_uni, (size_t)n3);
  if (e != hipSuccess) return ifail(KG_ERR_NOMEM, "hipMalloc", e);
  e = hipMemcpyAsync(d_om, om.data(), n * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_tw, tw.data(), (n / 2) * sizeof(double2), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && host_uniforms)
    e = hipMemcpyAsync(d_uni, host_uniforms, (size_t)n3 * sizeof(double), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return ifail(KG_ERR_HIP, "upload", e);
  const double q = -0.5 / (p0 * p0);
  hipLaunchKernelGGL(psi_p_kernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, d_psi, d_om, B.log2n, q,
                     2.0 * pii, d_uni, seed);
  if ((e = hipGetLastError()) != hipSuccess) return ifail(KG_ERR_HIP, "psi_p", e);
  const unsigned ngrp = (unsigned)((int64_t)n * n / B.nl);
  const double nn = (double)n * (double)n * (double)n;
  for (int axis = 0; axis < 3; ++axis) {
    if (axis < 2)
      hipLaunchKernelGGL(fft_lines_kernel<false>, dim3(ngrp), dim3(256), 0, s, d_psi, d_tw, B.log2n, B.nl, axis, nn);
    else
      hipLaunchKernelGGL(fft_lines_kernel<true>, dim3(ngrp), dim3(256), 0, s, d_psi, d_tw, B.log2n, B.nl, axis, nn);
    if ((e = hipGetLastError()) != hipSuccess) return ifail(KG_ERR_HIP, "fft", e);
  }
  return KG_OK;
}

}  // namespace

extern "C" {

const char* kg_ic_last_error(void) { return g_ic_error.c_str(); }

kg_status kg_ic_random_field(int32_t device, int32_t ncart, double r_max, double p0, uint64_t seed,
                             const double* uniforms, double* out) {
  if (!out) return ifail(KG_ERR_INVALID, "null output");
  CartBox B;
  kg_status st = cart_box(ncart, r_max, &B);
  if (st != KG_OK) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return ifail(KG_ERR_NO_DEVICE, "no GPU visible");
  if (device < 0 || device >= ndev) return ifail(KG_ERR_INVALID, "device out of range");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return ifail(KG_ERR_HIP, "hipSetDevice", e);
  hipStream_t s = nullptr;
  if ((e = hipStreamCreate(&s)) != hipSuccess) return ifail(KG_ERR_HIP, "hipStreamCreate", e);
  const int64_t n3 = (int64_t)ncart * ncart * ncart;
  {
    DevBuf mem;
    double2* d_psi = nullptr;
    e = mem.alloc(&d_psi, (size_t)n3);
    if (e != hipSuccess) {
      st = ifail(KG_ERR_NOMEM, "hipMalloc", e);
    } else {
      st = random_field(B, p0, seed, uniforms, d_psi, s, mem);
      if (st == KG_OK) {
        e = hipMemcpyAsync(out, d_psi, (size_t)n3 * sizeof(double2), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = ifail(KG_ERR_HIP, "download", e);
      } else {
        (void)hipStreamSynchronize(s);
      }
    }
  }
  (void)hipStreamDestroy(s);
  return st;
}

kg_status kg_initial_data(kg_handle* h, const kg_ic_desc* d, double* amp_out) {
  if (!h || !d) return ifail(KG_ERR_INVALID, "null argument");
  if (d->type < KG_IC_SPHERICAL_SHELL || d->type > KG_IC_RANDOM_SHELL) return ifail(KG_ERR_INVALID, "unknown initial-data type");
  HandleView v;
  if (handle_view(h, &v) != 0) return ifail(KG_ERR_INVALID, "bad handle");
  const Geom& G = v.G;
  const bool uses_shell = d->type != KG_IC_RANDOM;
  if (uses_shell && (!std::isfinite(d->mu) || !std::isfinite(d->sigma) || d->sigma == 0.0))
    return ifail(KG_ERR_INVALID, "mu must be finite and sigma finite and nonzero");
  if (d->type == KG_IC_SPHERICAL_SHELL && v.metric == KG_METRIC_MKS)
    return ifail(KG_ERR_UNSUPPORTED,
                 "the MKS branch of spherical_shell scales by an unset amplitude (initial_conditions.f90:26)");
  if (d->type != KG_IC_SPHERICAL_SHELL && v.nsys != 2)
    return ifail(KG_ERR_INVALID, "random initial data is complex: create the handle with is_complex = 1");
  hipError_t e = hipSetDevice(v.device);
  if (e != hipSuccess) return ifail(KG_ERR_HIP, "hipSetDevice", e);
  hipStream_t s = v.stream;
  const HostGrid& hg = v.hg;

  // one-coordinate tables (mesh_refinement_lib.f90:126-132 for x, y, z)
  std::vector<double> tabs(4 * (size_t)G.ni + 2 * (size_t)G.nj + 2 * (size_t)G.nk);
  double* hx = tabs.data();
  double* henv = hx + G.ni;
  double* hshell = henv + G.ni;
  double* hdt = hshell + G.ni;
  double* hst = hdt + G.ni;
  double* hct = hst + G.nj;
  double* hcp = hct + G.nj;
  double* hsp = hcp + G.nk;
  const double mu = d->mu, sg = d->sigma;
  for (int ii = 0; ii < G.ni; ++ii) {
    const double x = hg.xmin + (double)ii * hg.dx;
    hx[ii] = x;
    henv[ii] = std::exp(-((x - mu) * (x - mu)) / (sg * sg));
    hshell[ii] = 1.0 / x * henv[ii];
    hdt[ii] = -2.0 * (x - mu) / (x * (sg * sg)) * henv[ii];
  }
  for (int jj = 0; jj < G.nj; ++jj) {
    const double th = -hg.dy / 2.0 + (double)(jj - G.g) * hg.dy;
    hst[jj] = std::sin(th);
    hct[jj] = std::cos(th);
  }
  for (int kk = 0; kk < G.nk; ++kk) {
    const double ph = (double)kk * hg.dz;
    hcp[kk] = std::cos(ph);
    hsp[kk] = std::sin(ph);
  }

  CartBox B;
  if (d->type != KG_IC_SPHERICAL_SHELL) {
    // r_max = x(iR+g) (diagnostic_lib.f90:90-92); the box must hold every
    // sampled point for the cell search of random.f90:80-88 to succeed
    kg_status st = cart_box(d->ncart, hx[G.ni - 1], &B);
    if (st != KG_OK) return st;
    const double rpad = hx[G.ni - 1];
    const double top = B.c0 + (double)(B.n - 1) * B.h;
    if (!(-rpad >= B.c0) || !(rpad < top))
      return ifail(KG_ERR_INVALID, "the Cartesian box [-(r_max+1), r_max+1) with ncart points leaves its last cell "
                                   "below r_max: raise ncart (needs r_max + 1 < ncart/2)");
  }

  DevBuf mem;
  double* d_tabs = nullptr;
  e = mem.alloc(&d_tabs, tabs.size());
  if (e != hipSuccess) return ifail(KG_ERR_NOMEM, "hipMalloc", e);
  e = hipMemcpyAsync(d_tabs, tabs.data(), tabs.size() * sizeof(double), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return ifail(KG_ERR_HIP, "upload", e);
  IcArgs a;
  a.G = G;
  a.nsys = v.nsys;
  for (int f = 0; f < 5; ++f) a.u[f] = v.u[f];
  a.tab = v.tab;
  a.T.x = d_tabs;
  a.T.env = a.T.x + G.ni;
  a.T.shell = a.T.env + G.ni;
  a.T.dtphi = a.T.shell + G.ni;
  a.T.st = a.T.dtphi + G.ni;
  a.T.ct = a.T.st + G.nj;
  a.T.cp = a.T.ct + G.nj;
  a.T.sp = a.T.cp + G.nk;
  const dim3 grid((unsigned)((G.ni + 255) / 256), (unsigned)G.nj, (unsigned)G.nk), blk(256);

  // start-of-step state read by an 'inflow' xR fill: the reference's u_p has
  // not been written when the initial data is built; it is taken as zero
  FillArgs fa;
  fa.G = G;
  double* d_zero = nullptr;
  if (G.bc_inflow && d->type != KG_IC_SPHERICAL_SHELL) {
    const size_t per = (size_t)G.sys_stride * v.nsys;
    if ((e = mem.alloc(&d_zero, per)) != hipSuccess) return ifail(KG_ERR_NOMEM, "hipMalloc", e);
    if ((e = hipMemsetAsync(d_zero, 0, per * sizeof(double), s)) != hipSuccess) return ifail(KG_ERR_HIP, "memset", e);
  }
  for (int f = 0; f < 5; ++f) {
    fa.u[f] = v.u[f];
    fa.u_p[f] = d_zero ? d_zero : v.u[f];
  }

  int quantity;
  if (d->type == KG_IC_SPHERICAL_SHELL) {
    hipLaunchKernelGGL(shell_kernel, grid, blk, 0, s, a);
    hipLaunchKernelGGL(deriv_kernel, grid, blk, 0, s, a, 2 | 4);
    hipLaunchKernelGGL(pi_kernel, grid, blk, 0, s, a, 0);
    quantity = 6;
  } else {
    double2* d_psi = nullptr;
    const int64_t n3 = (int64_t)B.n * B.n * B.n;
    if ((e = mem.alloc(&d_psi, (size_t)n3)) != hipSuccess) return ifail(KG_ERR_NOMEM, "hipMalloc", e);
    kg_status st = random_field(B, d->p0, d->seed, d->uniforms, d_psi, s, mem);
    if (st != KG_OK) {
      (void)hipStreamSynchronize(s);
      return st;
    }
    hipLaunchKernelGGL(sample_kernel, grid, blk, 0, s, a, d_psi, B.log2n, B.c0, B.h,
                       d->type == KG_IC_RANDOM_SHELL ? 1 : 0);
    if (d->type == KG_IC_RANDOM) {  // boundary_xL, boundary_xR (initial_conditions.f90:107-117)
      if ((e = launch_fill(fa, s, 0, v.nsys)) != hipSuccess) return ifail(KG_ERR_HIP, "fill", e);
    }
    hipLaunchKernelGGL(deriv_kernel, grid, blk, 0, s, a, 1 | 2 | 4);
    hipLaunchKernelGGL(pi_kernel, grid, blk, 0, s, a, d->type == KG_IC_RANDOM ? 1 : 2);
    if (d->type == KG_IC_RANDOM_SHELL) {  // xL, xR, zL, zR, yL, yR (:168-182)
      const int order[3] = {0, 2, 1};
      for (int ph : order)
        if ((e = launch_fill(fa, s, ph, v.nsys)) != hipSuccess) return ifail(KG_ERR_HIP, "fill", e);
    }
    quantity = 7;
  }
  if ((e = hipGetLastError()) != hipSuccess) return ifail(KG_ERR_HIP, "initial data", e);
  // Amp = 1/sqrt(integrate(...)) (:86, :142, :186); the integral synchronises
  double I[2];
  if ((e = diag_integrate(v, quantity, I)) != hipSuccess) return ifail(KG_ERR_HIP, "integrate", e);
  const double amp = 1.0 / std::sqrt(I[0]);
  hipLaunchKernelGGL(scale_kernel, grid, blk, 0, s, a, amp);
  if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return ifail(KG_ERR_HIP, "scale", e);
  if (amp_out) *amp_out = amp;
  return KG_OK;
}

}  // extern "C"
