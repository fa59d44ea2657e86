// kg_spectral.hip -- spherical-Bessel spectral projection of a handle's state
// (tools/FFT_sph.f90) and the kg_spec_* C ABI of include/kg.h.
//
// Pipeline for one field (FFT_sph.f90:213-244 as the contraction it defines):
//   1. domain_to_sph (:525-651): trilinear interpolation from the padded,
//      theta-cell-centred evolution grid to the node-centred spectral grid.
//      Per-axis bracket indices and weights are found once on the host with
//      the reference's search (initial guess, clamp, correction loops) and the
//      per-point arithmetic is the reference's, in its order, so the values
//      equal the oracle's bit for bit.
//   2. project_sh (:247-279) at every radius.  Y_lm = N_lm P_l^m(cos t) e^{imp}
//      and the trapezoid weights (integral.f90:330-362) are products of a
//      theta factor and a phi factor, so the double sum splits: a phi pass
//      forms C_m = sum_k w_k f cos(m p_k), S_m = sum_k w_k f sin(m p_k) for
//      m <= lmax while it interpolates (phi_kernel: one read of the field,
//      HBM-bound), and a theta pass contracts them with
//      L_lm(t_j) = dt dp w_j sin(t_j) N_lm P_l^m(t_j) (theta_kernel).
//   3. the radial contraction ft(n,lm) = sum_i W_l(n,i) flm(i,lm) with
//      W_l(n,i) = j_l(k_n r_i) r_i^2 w_i dr, a dense real product per l:
//      FP64 (radial_f64_kernel) or BF16 matrix cores with split operands
//      (radial_mfma_kernel).
//
// Split-BF16 contraction.  Within a chunk of 32 radii every operand row is
// scaled by a power of two below its largest magnitude and cut into S = 5
// slices x = sum_p t_p 2^{-8(p+1)} with integer |t_p| <= 256 (each exact in
// BF16).  Slice products t_p u_q are integers below 2^16 in units of
// 2^{-8(p+q+2)}; the products of one diagonal p + q = d (at most 5 pairs)
// summed over the 32 radii stay below 2^24 units, so the FP32 accumulation of
// the matrix cores (one v_mfma_f32_16x16x32_bf16 per slice pair and chunk) is
// exact.  Diagonals d <= 4 are kept (15 slice products);
// each chunk's FP32 sums are scaled back and added in FP64.  The dropped
// terms are below ~6 * 2^-40 of sum_i |W||flm| per output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kg.h"
#include "kg_internal.h"

using namespace kg;

namespace {

constexpr int kSlices = 5;   // BF16 slices per operand
constexpr int kChunk = 32;   // radii per exact FP32 accumulation

struct InterpArgs {
  const int* ir;      // [Nr] padded column below r_i
  const double* wr;   // [Nr]
  const int* ith;     // [Nt] padded row below theta_j
  const double* wth;  // [Nt]
  const int* tpole;   // [Nt] 0: interpolate, 1: copy row 0, 2: copy row nj-1
  const int* iph;     // [Np] padded plane below phi_k
  const double* wph;  // [Np]
  const int* pskip;   // [Np] 1: no upper plane (left at zero)
  int ni, nj, ld;
  int64_t plane;
  int Nr, Nt, Np;
};

// The projection reads its tables and the field through address-space
// qualified pointers: the field and the per-radius tables as global memory,
// the per-theta and per-phi tables and the phi factors (read-only for the
// whole launch, the same address across a wave) as constant memory, which
// lets those loads go through the scalar unit instead of 64 vector lanes.
typedef __attribute__((address_space(1))) const double gcd;
typedef __attribute__((address_space(1))) const int gci;
typedef __attribute__((address_space(4))) const double kcd;
typedef __attribute__((address_space(4))) const int kci;

__device__ __forceinline__ double interp(const InterpArgs& a, const double* __restrict__ F, int i, int j, int k) {
  const int ir = ((gci*)a.ir)[i];
  const int tp = ((kci*)a.tpole)[j];
  const int kp = ((kci*)a.iph)[k];
  auto at = [&](int ii, int jj, int kk) { return ((gcd*)F)[(int64_t)kk * a.plane + (int64_t)jj * a.ld + ii]; };
  if (tp == 1) return at(ir, 0, kp);
  if (tp == 2) return at(ir, a.nj - 1, kp);
  if (((kci*)a.pskip)[k]) return 0.0;
  const int jt = ((kci*)a.ith)[j];
  const int k1 = kp + 1;
  const double f000 = at(ir, jt, kp), f001 = at(ir, jt, k1);
  const double f010 = at(ir, jt + 1, kp), f011 = at(ir, jt + 1, k1);
  const double f100 = at(ir + 1, jt, kp), f101 = at(ir + 1, jt, k1);
  const double f110 = at(ir + 1, jt + 1, kp), f111 = at(ir + 1, jt + 1, k1);
  const double w_r = ((gcd*)a.wr)[i], w_th = ((kcd*)a.wth)[j], w_ph = ((kcd*)a.wph)[k];
  const double f00 = (1.0 - w_ph) * f000 + w_ph * f001;
  const double f01 = (1.0 - w_ph) * f010 + w_ph * f011;
  const double f10 = (1.0 - w_ph) * f100 + w_ph * f101;
  const double f11 = (1.0 - w_ph) * f110 + w_ph * f111;
  const double f0 = (1.0 - w_th) * f00 + w_th * f01;
  const double f1 = (1.0 - w_th) * f10 + w_th * f11;
  return (1.0 - w_r) * f0 + w_r * f1;
}

// out[s][k][j][i]; grid (ceil(Nr/64), ceil(Nt/4), Np*nsys), block 256
__global__ __launch_bounds__(256) void d2s_kernel(const InterpArgs a, const double* __restrict__ F,
                                                  int64_t sys_stride, double* __restrict__ out) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = blockIdx.z % a.Np;
  const int s = blockIdx.z / a.Np;
  if (i >= a.Nr || j >= a.Nt) return;
  out[(((int64_t)s * a.Np + k) * a.Nt + j) * a.Nr + i] = interp(a, F + s * sys_stride, i, j, k);
}

// cs[s][j][q][i], q = 2m (cos) / 2m+1 (sin), m < MC; trig[k][m][2] carries
// the phi trapezoid weight.  Blocks (bx, by, s) of a grid (ceil(Nr/64),
// ceil(Nt/4), nsys), 256 threads
template <int MC>
__device__ __forceinline__ void phi_body(const InterpArgs& a, const double* __restrict__ F, int64_t sys_stride,
                                         const double* __restrict__ trig, double* __restrict__ cs, int bx, int by,
                                         int s) {
  const int i = bx * 64 + (threadIdx.x & 63);
  const int j = by * 4 + (threadIdx.x >> 6);
  if (i >= a.Nr || j >= a.Nt) return;
  const double* Fs = F + s * sys_stride;
  double acc[2 * MC];
#pragma unroll
  for (int q = 0; q < 2 * MC; ++q) acc[q] = 0.0;
  for (int k = 0; k < a.Np; ++k) {
    const double f = interp(a, Fs, i, j, k);
    kcd* tk = (kcd*)trig + (int64_t)k * 2 * MC;
#pragma unroll
    for (int q = 0; q < 2 * MC; ++q) acc[q] += tk[q] * f;
  }
  __attribute__((address_space(1))) double* o =
      (__attribute__((address_space(1))) double*)cs + (((int64_t)s * a.Nt + j) * 2 * MC) * a.Nr + i;
#pragma unroll
  for (int q = 0; q < 2 * MC; ++q) o[(int64_t)q * a.Nr] = acc[q];
}

template <int MC>
__global__ __launch_bounds__(256) void phi_kernel(const InterpArgs a, const double* __restrict__ F,
                                                  int64_t sys_stride, const double* __restrict__ trig,
                                                  double* __restrict__ cs) {
  phi_body<MC>(a, F, sys_stride, trig, cs, blockIdx.x, blockIdx.y, blockIdx.z);
}

// flm[lm][part][i] from cs; one thread per (i, l, m >= 0) writes (l, m) and
// (l, -m).  leg[j][pair], pair = l(l+1)/2 + m.  grid (ceil(Nr/64),
// ceil(npair/4)), block 256
__device__ __forceinline__ void theta_body(const double* __restrict__ cs, const double* __restrict__ leg,
                                           double* __restrict__ flm, int Nr, int Nt, int MC, int lmax, int nsys,
                                           int bx, int by) {
  const int i = bx * 64 + (threadIdx.x & 63);
  const int pair = by * 4 + (threadIdx.x >> 6);
  const int npair = (lmax + 1) * (lmax + 2) / 2;
  if (i >= Nr || pair >= npair) return;
  int l = 0;
  while ((l + 1) * (l + 2) / 2 <= pair) ++l;
  const int m = pair - l * (l + 1) / 2;
  const int64_t jstride = (int64_t)2 * MC * Nr;
  gcd* ca = (gcd*)cs + (int64_t)(2 * m) * Nr + i;
  gcd* sa = ca + Nr;
  kcd* lg = (kcd*)leg;
  double lca = 0.0, lsa = 0.0, lcb = 0.0, lsb = 0.0;
  if (nsys == 1) {
    for (int j = 0; j < Nt; ++j) {
      const double L = lg[(int64_t)j * npair + pair];
      lca += L * ca[j * jstride];
      lsa += L * sa[j * jstride];
    }
  } else {
    const int64_t s1 = (int64_t)Nt * jstride;
    for (int j = 0; j < Nt; ++j) {
      const double L = lg[(int64_t)j * npair + pair];
      lca += L * ca[j * jstride];
      lsa += L * sa[j * jstride];
      lcb += L * ca[s1 + j * jstride];
      lsb += L * sa[s1 + j * jstride];
    }
  }
  // f conj(Y_lm):   m >= 0: (Ca + Sb) + i (Cb - Sa)
  //                 m <  0: (-1)^m [(Ca - Sb) + i (Cb + Sa)]   (Y_{l,-m} = (-1)^m conj Y_lm)
  const int lm = l * l + l + m;
  __attribute__((address_space(1))) double* out = (__attribute__((address_space(1))) double*)flm;
  out[((int64_t)lm * 2 + 0) * Nr + i] = lca + lsb;
  out[((int64_t)lm * 2 + 1) * Nr + i] = lcb - lsa;
  if (m > 0) {
    const double sg = (m & 1) ? -1.0 : 1.0;
    const int ln = l * l + l - m;
    out[((int64_t)ln * 2 + 0) * Nr + i] = sg * (lca - lsb);
    out[((int64_t)ln * 2 + 1) * Nr + i] = sg * (lcb + lsa);
  }
}

__global__ __launch_bounds__(256) void theta_kernel(const double* __restrict__ cs, const double* __restrict__ leg,
                                                    double* __restrict__ flm, int Nr, int Nt, int MC, int lmax,
                                                    int nsys) {
  theta_body(cs, leg, flm, Nr, Nt, MC, lmax, nsys, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------- j_l on the device
__device__ double sph_jl_dev(int l, double x) {
  if (x == 0.0) return l == 0 ? 1.0 : 0.0;
  const double ax = fabs(x);
  if (ax < 1e-4) {
    double df = 1.0;
    for (int i = 3; i <= 2 * l + 1; i += 2) df *= i;
    double xl = 1.0;
    for (int i = 0; i < l; ++i) xl *= x;
    return xl / df * (1.0 - x * x / (2.0 * (2 * l + 3)));
  }
  double s, c;
  sincos(x, &s, &c);
  const double j0 = s / x;
  if (l == 0) return j0;
  if (ax > l) {
    double a = j0, b = (s / x - c) / x;
    for (int k = 1; k < l; ++k) {
      const double n = (2.0 * k + 1.0) / x * b - a;
      a = b;
      b = n;
    }
    return b;
  }
  const int top = l + 30 + (int)ax + (int)sqrt(50.0 * (l + 1));
  double up = 0.0, cur = 1e-280, want = 0.0;
  for (int k = top; k >= 1; --k) {
    const double dn = (2.0 * k + 1.0) / x * cur - up;
    up = cur;
    cur = dn;
    if (k - 1 == l) want = cur;
    if (fabs(cur) > 1e200) {
      cur *= 1e-200;
      up *= 1e-200;
      want *= 1e-200;
    }
  }
  if (ax >= 1.0) {
    const double j1 = (s / x - c) / x;
    if (fabs(j1) > fabs(j0)) return want * (j1 / up);
  }
  return want * (j0 / cur);
}

// One block per (l, n): j_l(k r_i) for all i, normalised for KG_BASIS_ZEROS
// by sqrt(trapezoid(j^2 r^2, dr)) (FFT_sph.f90:56-69).  mode 0 writes the
// contraction weights W = j r^2 w_i dr, mode 1 the basis itself.
__global__ __launch_bounds__(256) void bessel_kernel(const double* __restrict__ kv, const double* __restrict__ r,
                                                     int Nr, int normalise, double dr, int lmax_unused, int nk,
                                                     int mode, double* __restrict__ out) {
  __shared__ double red[256];
  const int ln = blockIdx.x;
  const int l = ln / nk;
  const double k = kv[ln];
  double* row = out + (int64_t)ln * Nr;
  double part = 0.0;
  for (int i = threadIdx.x; i < Nr; i += 256) {
    const double v = sph_jl_dev(l, k * r[i]);
    row[i] = v;
    const double t = v * v * (r[i] * r[i]);
    part += (i == 0 || i == Nr - 1) ? 0.5 * t : t;
  }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  double norm = 1.0;
  if (normalise) {
    const double integ = sqrt(dr * red[0]);
    if (integ > 0.0) norm = integ;
  }
  for (int i = threadIdx.x; i < Nr; i += 256) {
    const double v = row[i] / norm;
    const double w = (i == 0 || i == Nr - 1) ? 0.5 : 1.0;
    row[i] = mode == 0 ? v * (r[i] * r[i]) * w * dr : v;
  }
}

// ---------------------------------------------------------------- radial FP64
// Block = (l, 64 wave numbers); 4 column groups of CPT columns each.
template <int CPT>
__device__ __forceinline__ void radial_f64_body(const int2* __restrict__ tiles, const double* __restrict__ W,
                                                const double* __restrict__ flm, double* __restrict__ ft, int Nr,
                                                int nk, int tile) {
  __shared__ double sW[64][65];
  __shared__ double sB[4 * CPT][64];
  const int2 t = tiles[tile];
  const int l = t.x, n0 = t.y;
  const int C = 2 * (2 * l + 1);
  const int lane = threadIdx.x & 63, cg = threadIdx.x >> 6;
  const double* Wl = W + (int64_t)l * nk * Nr;
  const double* B = flm + (int64_t)2 * l * l * Nr;
  double acc[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) acc[q] = 0.0;
  for (int i0 = 0; i0 < Nr; i0 += 64) {
    const int ii = i0 + lane;
    for (int rr = cg; rr < 64; rr += 4) {
      const int n = n0 + rr;
      sW[rr][lane] = (n < nk && ii < Nr) ? Wl[(int64_t)n * Nr + ii] : 0.0;
    }
    for (int c = cg; c < 4 * CPT; c += 4) sB[c][lane] = (c < C && ii < Nr) ? B[(int64_t)c * Nr + ii] : 0.0;
    __syncthreads();
    for (int x = 0; x < 64; ++x) {
      const double w = sW[lane][x];
#pragma unroll
      for (int q = 0; q < CPT; ++q) acc[q] += w * sB[cg + 4 * q][x];
    }
    __syncthreads();
  }
  const int n = n0 + lane;
  if (n >= nk) return;
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    const int c = cg + 4 * q;
    if (c < C) ft[((int64_t)(l * l + (c >> 1)) * nk + n) * 2 + (c & 1)] = acc[q];
  }
}

template <int CPT>
__global__ __launch_bounds__(256) void radial_f64_kernel(const int2* __restrict__ tiles, const double* __restrict__ W,
                                                         const double* __restrict__ flm, double* __restrict__ ft,
                                                         int Nr, int nk) {
  radial_f64_body<CPT>(tiles, W, flm, ft, Nr, nk, blockIdx.x);
}

// ---------------------------------------------------------------- radial split BF16
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_exact(double t, float unit) {
  const float v = (float)t * unit;  // t integer, |t| <= 256: exact
  return (short)(__float_as_uint(v) >> 16);
}

// Cut 8 values (already scaled into (-1, 1)) into kSlices BF16 slices; slice p
// of value e is element e of sl[p].
__device__ __forceinline__ void slice8(const double* v, v8s sl[kSlices]) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    double y = v[e];
#pragma unroll
    for (int p = 0; p < kSlices; ++p) {
      y *= 256.0;
      const double t = rint(y);
      y -= t;
      sl[p][e] = bf16_exact(t, __uint_as_float((uint32_t)(127 - 8 * (p + 1)) << 23));
    }
  }
}

// One wave per {l, 16 wave numbers n0..n0+15}, covering every 16-column tile
// of that l (the W rows are loaded and sliced once per chunk for all of
// them); one chunk of 32 radii per 16x16x32 BF16 MFMA (gfx950).  Operand layout: lane L holds
// A[L%16][8(L/16)+e], B[8(L/16)+e][L%16], e = 0..7, and the result
// D[4(L/16)+e][L%16], e = 0..3.  A = W_l rows (wave numbers), B = flm columns.
// Every partial sum of a diagonal is an integer below 2^24 units, so the FP32
// accumulation is exact in any order and the result does not depend on how
// the matrix core groups the 32 products.
constexpr int kMaxCT = 5;  // 16-column tiles of one l: C = 2(2l+1) <= 66 for l <= 16

__device__ __forceinline__ void radial_mfma_body(const int4* __restrict__ tiles, int ntiles,
                                                 const double* __restrict__ W, const double* __restrict__ flm,
                                                 double* __restrict__ ft, int Nr, int nk, int block) {
  const int tile = block * 4 + (threadIdx.x >> 6);
  if (tile >= ntiles) return;  // whole wave
  const int4 t = tiles[tile];
  const int l = t.x, n0 = t.y;
  const int C = 2 * (2 * l + 1);
  const int nct = (C + 15) / 16;  // column tiles, all handled by this wave: W is sliced once for them
  const int L = threadIdx.x & 63;
  const int q = L >> 4, rl = L & 15;
  const int n = n0 + rl;
  const bool vn = n < nk;
  const double* Arow = W + ((int64_t)l * nk + (vn ? n : 0)) * Nr;
  double out[kMaxCT][4];
#pragma unroll
  for (int ct = 0; ct < kMaxCT; ++ct)
#pragma unroll
    for (int e = 0; e < 4; ++e) out[ct][e] = 0.0;
  for (int i0 = 0; i0 < Nr; i0 += kChunk) {
    double a[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int i = i0 + 8 * q + e;
      a[e] = (vn && i < Nr) ? Arow[i] : 0.0;
    }
    double am = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) am = fmax(am, fabs(a[e]));
    am = fmax(am, __shfl_xor(am, 16));
    am = fmax(am, __shfl_xor(am, 32));
    int ea = 0;
    if (am > 0.0) frexp(am, &ea);  // am < 2^ea
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = ldexp(a[e], -ea);
    v8s as[kSlices];
    slice8(a, as);
    // row 4q+e of the result needs that row's exponent (held by lane 4q+e)
    int er[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) er[e] = __shfl(ea, 4 * q + e);
#pragma unroll
    for (int ct = 0; ct < kMaxCT; ++ct) {
      if (ct >= nct) break;  // wave-uniform
      const int c = 16 * ct + rl;
      const bool vc = c < C;
      const double* Bcol = flm + ((int64_t)2 * l * l + (vc ? c : 0)) * Nr;
      double b[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int i = i0 + 8 * q + e;
        b[e] = (vc && i < Nr) ? Bcol[i] : 0.0;
      }
      double bm = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) bm = fmax(bm, fabs(b[e]));
      bm = fmax(bm, __shfl_xor(bm, 16));
      bm = fmax(bm, __shfl_xor(bm, 32));
      int eb = 0;
      if (bm > 0.0) frexp(bm, &eb);
#pragma unroll
      for (int e = 0; e < 8; ++e) b[e] = ldexp(b[e], -eb);
      v8s bs[kSlices];
      slice8(b, bs);
      v4f acc[kSlices];
#pragma unroll
      for (int d = 0; d < kSlices; ++d) acc[d] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int d = 0; d < kSlices; ++d) {
#pragma unroll
        for (int p = 0; p <= d; ++p) {
          const int r = d - p;
          acc[d] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, as[p]),
                                                           __builtin_bit_cast(v8bf, bs[r]), acc[d], 0, 0, 0);
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // the slices carry their 2^{-8(p+1)} scale, so acc[d] is already in units of
        // the scaled operands; only the two row/column exponents remain
        double sum = 0.0;
#pragma unroll
        for (int d = 0; d < kSlices; ++d) sum += (double)acc[d][e];
        out[ct][e] += ldexp(sum, er[e] + eb);
      }
    }
  }
#pragma unroll
  for (int ct = 0; ct < kMaxCT; ++ct) {
    if (ct >= nct) break;
    const int c = 16 * ct + rl;
    if (c >= C) continue;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int nn = n0 + 4 * q + e;
      if (nn < nk) ft[((int64_t)(l * l + (c >> 1)) * nk + nn) * 2 + (c & 1)] = out[ct][e];
    }
  }
}

__global__ __launch_bounds__(256) void radial_mfma_kernel(const int4* __restrict__ tiles, int ntiles,
                                                          const double* __restrict__ W,
                                                          const double* __restrict__ flm, double* __restrict__ ft,
                                                          int Nr, int nk) {
  radial_mfma_body(tiles, ntiles, W, flm, ft, Nr, nk, blockIdx.x);
}

// ---------------------------------------------------------------- many projectors
// One launch per pass over the projectors of many grids: member q owns blocks
// [start[q], start[q+1]) of the pass and runs the single-grid body on them, so
// every member's arithmetic is that of its own launch.
struct SpecMember {
  InterpArgs ia;
  const double* F;  // the projected field (system 0)
  int64_t sys_stride;
  const double* trig;
  const double* leg;
  const double* W;
  double* cs;
  double* flm;
  double* ft;
  const int2* tiles64;
  const int4* tiles16;
  int Nr, Nt, nk, lmax, mc, nsys, npair, ntiles64, ntiles16;
};

__device__ __forceinline__ int spec_member(const int* __restrict__ start, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int MC>
__global__ __launch_bounds__(256) void phi_kernel_batch(const SpecMember* __restrict__ ms,
                                                        const int* __restrict__ start, int n) {
  const int q = spec_member(start, n, blockIdx.x);
  const SpecMember& m = ms[q];
  const int b = blockIdx.x - start[q];
  const int gx = (m.Nr + 63) / 64, gy = (m.Nt + 3) / 4;
  phi_body<MC>(m.ia, m.F, m.sys_stride, m.trig, m.cs, b % gx, (b / gx) % gy, b / (gx * gy));
}

__global__ __launch_bounds__(256) void theta_kernel_batch(const SpecMember* __restrict__ ms,
                                                          const int* __restrict__ start, int n) {
  const int q = spec_member(start, n, blockIdx.x);
  const SpecMember& m = ms[q];
  const int b = blockIdx.x - start[q];
  const int gx = (m.Nr + 63) / 64;
  theta_body(m.cs, m.leg, m.flm, m.Nr, m.Nt, m.mc, m.lmax, m.nsys, b % gx, b / gx);
}

template <int CPT>
__global__ __launch_bounds__(256) void radial_f64_kernel_batch(const SpecMember* __restrict__ ms,
                                                               const int* __restrict__ start, int n) {
  const int q = spec_member(start, n, blockIdx.x);
  const SpecMember& m = ms[q];
  radial_f64_body<CPT>(m.tiles64, m.W, m.flm, m.ft, m.Nr, m.nk, blockIdx.x - start[q]);
}

__global__ __launch_bounds__(256) void radial_mfma_kernel_batch(const SpecMember* __restrict__ ms,
                                                                const int* __restrict__ start, int n) {
  const int q = spec_member(start, n, blockIdx.x);
  const SpecMember& m = ms[q];
  radial_mfma_body(m.tiles16, m.ntiles16, m.W, m.flm, m.ft, m.Nr, m.nk, blockIdx.x - start[q]);
}

// ---------------------------------------------------------------- detectors
// currents_ADM (diagnostic_lib.f90:214-224) times sqrt(gamma) at a padded
// point: j1 = -Pi (g11 psi_r + g13 psi_phi), j2 = -Pi g22 psi_th,
// j3 = -Pi (g13 psi_r + g33 psi_phi), complex products (ac - bd, ad + bc).
struct CurArgs {
  const double* u[5];
  const double* tab;
  int64_t sys_stride, plane, n2;
  int ld, nsys;
};

__device__ void current_at(const CurArgs& c, int ii, int jj, int kk, double* v /* re0 re1 re2 im0 im1 im2 */) {
  const int64_t o = (int64_t)kk * c.plane + (int64_t)jj * c.ld + ii;
  const int64_t t = (int64_t)jj * c.ld + ii;
  const bool cp = c.nsys == 2;
  const double pr = c.u[4][o], pi = cp ? c.u[4][o + c.sys_stride] : 0.0;
  const double rr = c.u[1][o], ri = cp ? c.u[1][o + c.sys_stride] : 0.0;
  const double tr = c.u[2][o], ti = cp ? c.u[2][o + c.sys_stride] : 0.0;
  const double fr = c.u[3][o], fi = cp ? c.u[3][o + c.sys_stride] : 0.0;
  const double g11 = c.tab[T_GRR * c.n2 + t], g13 = c.tab[T_GRP * c.n2 + t];
  const double g22 = c.tab[T_GTT * c.n2 + t], g33 = c.tab[T_GPP * c.n2 + t];
  const double sg = c.tab[T_SQRTG * c.n2 + t];
  const double c1r = g11 * rr + g13 * fr, c1i = g11 * ri + g13 * fi;
  const double c3r = g13 * rr + g33 * fr, c3i = g13 * ri + g33 * fi;
  const double q2r = pr * g22, q2i = pi * g22;
  const double j1r = -(pr * c1r - pi * c1i), j1i = -(pr * c1i + pi * c1r);
  const double j2r = -(q2r * tr - q2i * ti), j2i = -(q2r * ti + q2i * tr);
  const double j3r = -(pr * c3r - pi * c3i), j3i = -(pr * c3i + pi * c3r);
  // j_i(c) * sqrtg: complex times real
  v[0] = j1r * sg;
  v[1] = j2r * sg;
  v[2] = j3r * sg;
  v[3] = j1i * sg;
  v[4] = j2i * sg;
  v[5] = j3i * sg;
}

// domain_to_sph of the six components at spectral point (i, j, k): the
// arithmetic of interp() on values formed at the corners
__device__ void interp_current(const InterpArgs& a, const CurArgs& c, int i, int j, int k, double* out) {
  const int ir = a.ir[i];
  const int tp = a.tpole[j];
  const int kp = a.iph[k];
  if (tp == 1 || tp == 2) {
    current_at(c, ir, tp == 1 ? 0 : a.nj - 1, kp, out);
    return;
  }
  if (a.pskip[k]) {
#pragma unroll
    for (int q = 0; q < 6; ++q) out[q] = 0.0;
    return;
  }
  const int jt = a.ith[j];
  const int k1 = kp + 1;
  double f000[6], f001[6], f010[6], f011[6], f100[6], f101[6], f110[6], f111[6];
  current_at(c, ir, jt, kp, f000);
  current_at(c, ir, jt, k1, f001);
  current_at(c, ir, jt + 1, kp, f010);
  current_at(c, ir, jt + 1, k1, f011);
  current_at(c, ir + 1, jt, kp, f100);
  current_at(c, ir + 1, jt, k1, f101);
  current_at(c, ir + 1, jt + 1, kp, f110);
  current_at(c, ir + 1, jt + 1, k1, f111);
  const double w_r = a.wr[i], w_th = a.wth[j], w_ph = a.wph[k];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const double f00 = (1.0 - w_ph) * f000[q] + w_ph * f001[q];
    const double f01 = (1.0 - w_ph) * f010[q] + w_ph * f011[q];
    const double f10 = (1.0 - w_ph) * f100[q] + w_ph * f101[q];
    const double f11 = (1.0 - w_ph) * f110[q] + w_ph * f111[q];
    const double f0 = (1.0 - w_th) * f00 + w_th * f01;
    const double f1 = (1.0 - w_th) * f10 + w_th * f11;
    out[q] = (1.0 - w_r) * f0 + w_r * f1;
  }
}

// currents_detectors (diagnostic_lib.f90:250-274): one block per detector
// radius, trapezium_2D (integral.f90:330-362) of the interpolated j_i sqrtg
// over the sphere; mom[d][c] complex.
__global__ __launch_bounds__(256) void detector_flux_kernel(const InterpArgs a, const CurArgs c,
                                                            const int* __restrict__ idx, double dth, double dph,
                                                            double* __restrict__ mom) {
  __shared__ double red[6][256];
  const int d = blockIdx.x;
  const int i = idx[d];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int n = a.Nt * a.Np;
  for (int e = threadIdx.x; e < n; e += 256) {
    const int j = e / a.Np, k = e % a.Np;
    double v[6];
    interp_current(a, c, i, j, k, v);
    const int edge = (j == 0 || j == a.Nt - 1) + 2 * (k == 0 || k == a.Np - 1);
    const double w = edge == 0 ? 1.0 : (edge == 3 ? 0.25 : 0.5);
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] += w * v[q];
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) {
#pragma unroll
      for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s2];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      mom[(d * 3 + cc) * 2 + 0] = dth * dph * red[cc][0];
      mom[(d * 3 + cc) * 2 + 1] = dth * dph * red[3 + cc][0];
    }
  }
}

// alm[d][lm] (complex) = flm at detector radius idx[d]
__global__ void gather_alm_kernel(const double* __restrict__ flm, const int* __restrict__ idx, int ndet, int nlm,
                                  int Nr, double* __restrict__ alm) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= ndet * nlm) return;
  const int d = e / nlm, lm = e % nlm;
  alm[2 * e] = flm[((int64_t)lm * 2 + 0) * Nr + idx[d]];
  alm[2 * e + 1] = flm[((int64_t)lm * 2 + 1) * Nr + idx[d]];
}

}  // namespace

// ===================================================================== host
struct kg_spec {
  kg_handle* h = nullptr;
  int device = 0;
  int lmax = 0, nk = 0, basis = 0;
  double R_max = 0.0, kmax = 0.0;
  int Nr = 0, Nt = 0, Np = 0, nsys = 1;
  int mc = 0;        // m columns of the phi pass (template width)
  int nlm = 0, npair = 0;
  double dr = 0.0;
  std::vector<double> r, kv, ylm_re_im;  // host copies for kg_spec_get_tables
  void* d_interp = nullptr;              // one allocation for the interpolation tables
  InterpArgs ia{};
  double* d_trig = nullptr;
  double* d_leg = nullptr;
  double* d_kv = nullptr;
  double* d_r = nullptr;
  double* d_W = nullptr;
  double* d_cs = nullptr;
  double* d_flm = nullptr;   // [lm][part][i]
  double* d_flmi = nullptr;  // complex [lm][i], interleaved
  double* d_ft = nullptr;
  double* d_f = nullptr;
  int2* d_tiles64 = nullptr;
  int ntiles64 = 0;
  int4* d_tiles16 = nullptr;
  int ntiles16 = 0;
  void* d_batch = nullptr;  // member descriptors of kg_spec_radial_batch (this projector leading)
  size_t batch_bytes = 0;
  hipEvent_t batch_done = nullptr;
  std::string err;
};

namespace {

thread_local std::string g_spec_error;

kg_status sfail(kg_spec* p, kg_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (p) p->err = buf;
  g_spec_error = buf;
  return st;
}

#define SP_HIP(p, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return sfail((p), KG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

template <typename T>
kg_status dalloc(kg_spec* p, T** ptr, size_t count) {
  if (hipMalloc((void**)ptr, count * sizeof(T) + 16) != hipSuccess) {
    *ptr = nullptr;
    return sfail(p, KG_ERR_NOMEM, "device allocation of %zu bytes failed", count * sizeof(T));
  }
  return KG_OK;
}

void free_all(kg_spec* p) {
  void* ptrs[] = {p->d_interp, p->d_trig, p->d_leg, p->d_kv, p->d_r,       p->d_W,       p->d_cs,
                  p->d_flm,    p->d_flmi, p->d_ft,  p->d_f,  p->d_tiles64, p->d_tiles16, p->d_batch};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p->batch_done) (void)hipEventDestroy(p->batch_done);
}

// Bracket search of domain_to_sph (FFT_sph.f90:555-611) along one axis:
// coord(i) of the evolution grid for i in [lo, hi], a query value v.
template <typename Coord>
int bracket(Coord coord, int lo, int hi, double vmin, double dv, double v) {
  int i = (int)((v - vmin) / dv) + lo;
  if (i < lo) i = lo;
  if (i > hi - 1) i = hi - 1;
  while (i < hi - 1 && v >= coord(i + 1)) ++i;
  while (i > lo && v < coord(i)) --i;
  return i;
}

kg_status build_interp(kg_spec* p, const HandleView& v) {
  const HostGrid& g = v.hg;
  const int G = g.g;
  const int ilo = -G, ihi = g.Nx + G, jlo = -G, jhi = g.Ny + G, klo = 0, khi = g.Nz;
  // mesh_refinement_create (mesh_refinement_lib.f90:126-143)
  auto xd = [&](int i) { return g.xmin + (double)(i + G) * g.dx; };
  auto yd = [&](int j) { return -g.dy / 2.0 + (double)j * g.dy; };
  auto zd = [&](int k) { return (double)k * g.dz; };
  const double rmin = xd(ilo), drr = xd(ilo + 1) - xd(ilo);
  const double thetamin = yd(jlo), dth = std::fabs(yd(jlo + 1) - yd(jlo));
  const double phimin = zd(klo), dph = std::fabs(zd(klo + 1) - zd(klo));
  const int Nr = p->Nr, Nt = p->Nt, Np = p->Np;
  std::vector<int> ir(Nr), ith(Nt), tpole(Nt), iph(Np), pskip(Np);
  std::vector<double> wr(Nr), wth(Nt), wph(Np);
  for (int i = 0; i < Nr; ++i) {
    const double rr = g.xmin + (double)i * g.dx;
    const int a = bracket(xd, ilo, ihi, rmin, drr, rr);
    ir[i] = a + G;
    wr[i] = (rr - xd(a)) / drr;
  }
  for (int j = 0; j < Nt; ++j) {
    const double tt = (double)j * g.dy;
    const int a = bracket(yd, jlo, jhi, thetamin, dth, tt);
    ith[j] = a + G;
    wth[j] = (tt - yd(a)) / dth;
    tpole[j] = tt <= thetamin ? 1 : (tt >= thetamin + dth * (double)(jhi - jlo) ? 2 : 0);
  }
  for (int k = 0; k < Np; ++k) {
    const double pp = (double)k * g.dz;
    const int a = bracket(zd, klo, khi, phimin, dph, pp);
    iph[k] = a;
    wph[k] = (pp - zd(a)) / dph;
    pskip[k] = (a + 1 > khi) ? 1 : 0;
  }
  // every access of interp() stays inside the padded field
  for (int i = 0; i < Nr; ++i)
    if (ir[i] < 0 || ir[i] + 1 >= g.ni) return sfail(p, KG_ERR_INVALID, "radial bracket out of range");
  for (int j = 0; j < Nt; ++j)
    if (ith[j] < 0 || ith[j] + 1 >= g.nj) return sfail(p, KG_ERR_INVALID, "theta bracket out of range");
  for (int k = 0; k < Np; ++k)
    if (iph[k] < 0 || iph[k] + (pskip[k] ? 0 : 1) >= g.nk) return sfail(p, KG_ERR_INVALID, "phi bracket out of range");

  const size_t ints = (size_t)Nr + 2 * Nt + 2 * Np;
  const size_t dbls = (size_t)Nr + Nt + Np;
  const size_t bytes = dbls * 8 + ints * 4 + 64;
  std::vector<char> hostbuf(bytes, 0);
  char* base = hostbuf.data();
  size_t off = 0;
  auto put = [&](const void* src, size_t n) {
    const size_t o = off;
    std::memcpy(base + off, src, n);
    off += (n + 7) & ~(size_t)7;
    return o;
  };
  const size_t o_wr = put(wr.data(), Nr * 8), o_wth = put(wth.data(), Nt * 8), o_wph = put(wph.data(), Np * 8);
  const size_t o_ir = put(ir.data(), Nr * 4), o_ith = put(ith.data(), Nt * 4), o_tp = put(tpole.data(), Nt * 4);
  const size_t o_iph = put(iph.data(), Np * 4), o_ps = put(pskip.data(), Np * 4);
  char* d = nullptr;
  if (dalloc(p, &d, off) != KG_OK) return KG_ERR_NOMEM;
  p->d_interp = d;
  SP_HIP(p, hipMemcpy(d, base, off, hipMemcpyHostToDevice));
  InterpArgs& A = p->ia;
  A.wr = (const double*)(d + o_wr);
  A.wth = (const double*)(d + o_wth);
  A.wph = (const double*)(d + o_wph);
  A.ir = (const int*)(d + o_ir);
  A.ith = (const int*)(d + o_ith);
  A.tpole = (const int*)(d + o_tp);
  A.iph = (const int*)(d + o_iph);
  A.pskip = (const int*)(d + o_ps);
  A.ni = v.G.ni;
  A.nj = v.G.nj;
  A.ld = v.G.ld;
  A.plane = v.G.plane;
  A.Nr = Nr;
  A.Nt = Nt;
  A.Np = Np;
  return KG_OK;
}

kg_status build_angular(kg_spec* p, const HandleView& v) {
  const HostGrid& g = v.hg;
  const int Nt = p->Nt, Np = p->Np, lmax = p->lmax, MC = p->mc;
  const double pi = std::acos(-1.0);
  // phi factors with the trapezoid weight (integral.f90:349-356)
  std::vector<double> trig((size_t)Np * 2 * MC, 0.0);
  for (int k = 0; k < Np; ++k) {
    const double phi = (double)k * g.dz;
    const double w = (k == 0 || k == Np - 1) ? 0.5 : 1.0;
    for (int m = 0; m <= lmax; ++m) {
      trig[((size_t)k * MC + m) * 2 + 0] = w * std::cos((double)m * phi);
      trig[((size_t)k * MC + m) * 2 + 1] = w * std::sin((double)m * phi);
    }
  }
  // theta factors: N_lm P_l^m with the pole rule of sph_harmonics
  // (FFT_sph.f90:307-315: |cos t| > 0.9999 -> P := sqrt((2l+1)/4pi) for m = 0,
  // 0 otherwise), times sin t, the trapezoid weight and dtheta dphi
  const int npair = p->npair;
  std::vector<double> leg((size_t)Nt * npair, 0.0);
  std::vector<double> amp((size_t)Nt * npair, 0.0);  // N_lm P_l^m as in Y_lm
  for (int j = 0; j < Nt; ++j) {
    const double th = (double)j * g.dy;
    const double x = std::cos(th);
    const double w = (j == 0 || j == Nt - 1) ? 0.5 : 1.0;
    for (int l = 0; l <= lmax; ++l) {
      const double fact = std::sqrt((2.0 * l + 1.0) / (4.0 * pi));
      for (int m = 0; m <= l; ++m) {
        double a;
        if (std::fabs(x) > 0.9999)
          a = (m == 0) ? fact * fact : 0.0;
        else
          a = ylm_norm_plm(l, m, x);
        const int pr = l * (l + 1) / 2 + m;
        amp[(size_t)j * npair + pr] = a;
        leg[(size_t)j * npair + pr] = (g.dy * g.dz) * w * std::sin(th) * a;
      }
    }
  }
  // Y_lm for kg_spec_get_tables: complex [nlm][Nt][Np]
  p->ylm_re_im.assign((size_t)p->nlm * Nt * Np * 2, 0.0);
  for (int l = 0; l <= lmax; ++l)
    for (int m = 0; m <= l; ++m)
      for (int j = 0; j < Nt; ++j) {
        const double a = amp[(size_t)j * npair + l * (l + 1) / 2 + m];
        for (int k = 0; k < Np; ++k) {
          const double phi = (double)k * g.dz;
          const double re = a * std::cos((double)m * phi), im = a * std::sin((double)m * phi);
          const size_t o = (((size_t)(l * l + l + m) * Nt + j) * Np + k) * 2;
          p->ylm_re_im[o] = re;
          p->ylm_re_im[o + 1] = im;
          if (m > 0) {
            const double sg = (m & 1) ? -1.0 : 1.0;
            const size_t q = (((size_t)(l * l + l - m) * Nt + j) * Np + k) * 2;
            p->ylm_re_im[q] = sg * re;
            p->ylm_re_im[q + 1] = -sg * im;
          }
        }
      }
  if (dalloc(p, &p->d_trig, trig.size()) != KG_OK || dalloc(p, &p->d_leg, leg.size()) != KG_OK)
    return KG_ERR_NOMEM;
  SP_HIP(p, hipMemcpy(p->d_trig, trig.data(), trig.size() * 8, hipMemcpyHostToDevice));
  SP_HIP(p, hipMemcpy(p->d_leg, leg.data(), leg.size() * 8, hipMemcpyHostToDevice));
  return KG_OK;
}

kg_status build_radial(kg_spec* p, const HandleView& v) {
  const HostGrid& g = v.hg;
  const int Nr = p->Nr, nk = p->nk, lmax = p->lmax;
  p->r.resize(Nr);
  for (int i = 0; i < Nr; ++i) p->r[i] = g.xmin + (double)i * g.dx;  // spectral radii
  p->dr = std::fabs(p->r[1] - p->r[0]);
  p->kv.resize((size_t)(lmax + 1) * nk);
  const std::vector<double> z = p->basis == KG_BASIS_ZEROS ? jl_zeros(lmax, nk) : std::vector<double>();
  for (int l = 0; l <= lmax; ++l)
    for (int n = 1; n <= nk; ++n)
      p->kv[(size_t)l * nk + n - 1] =
          p->basis == KG_BASIS_ZEROS ? z[(size_t)l * nk + n - 1] / p->R_max : p->kmax * (double)n / (double)nk;
  if (dalloc(p, &p->d_r, Nr) != KG_OK || dalloc(p, &p->d_kv, p->kv.size()) != KG_OK ||
      dalloc(p, &p->d_W, (size_t)(lmax + 1) * nk * Nr) != KG_OK)
    return KG_ERR_NOMEM;
  SP_HIP(p, hipMemcpy(p->d_r, p->r.data(), Nr * 8, hipMemcpyHostToDevice));
  SP_HIP(p, hipMemcpy(p->d_kv, p->kv.data(), p->kv.size() * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(bessel_kernel, dim3((lmax + 1) * nk), dim3(256), 0, v.stream, p->d_kv, p->d_r, Nr,
                     p->basis == KG_BASIS_ZEROS ? 1 : 0, p->dr, lmax, nk, 0, p->d_W);
  SP_HIP(p, hipGetLastError());
  // tile lists of the two contraction kernels
  std::vector<int2> t64;
  std::vector<int4> t16;
  for (int l = 0; l <= lmax; ++l) {
    for (int n0 = 0; n0 < nk; n0 += 64) t64.push_back(int2{l, n0});
    for (int n0 = 0; n0 < nk; n0 += 16) t16.push_back(int4{l, n0, 0, 0});  // every column tile of l
  }
  p->ntiles64 = (int)t64.size();
  p->ntiles16 = (int)t16.size();
  if (dalloc(p, &p->d_tiles64, t64.size()) != KG_OK || dalloc(p, &p->d_tiles16, t16.size()) != KG_OK)
    return KG_ERR_NOMEM;
  SP_HIP(p, hipMemcpy(p->d_tiles64, t64.data(), t64.size() * sizeof(int2), hipMemcpyHostToDevice));
  SP_HIP(p, hipMemcpy(p->d_tiles16, t16.data(), t16.size() * sizeof(int4), hipMemcpyHostToDevice));
  SP_HIP(p, hipStreamSynchronize(v.stream));
  return KG_OK;
}

kg_status view(kg_spec* p, HandleView* v) {
  if (handle_view(p->h, v) != 0) return sfail(p, KG_ERR_INVALID, "projector has no handle");
  SP_HIP(p, hipSetDevice(v->device));
  return KG_OK;
}

kg_status check_field(kg_spec* p, int field) {
  if (field < 0 || field >= kNumFields) return sfail(p, KG_ERR_INVALID, "field %d out of range 0..4", field);
  return KG_OK;
}

kg_status enqueue_phi(kg_spec* p, const HandleView& v, int field) {
  const dim3 grid((p->Nr + 63) / 64, (p->Nt + 3) / 4, p->nsys);
  if (p->mc == 9)
    hipLaunchKernelGGL(phi_kernel<9>, grid, dim3(256), 0, v.stream, p->ia, v.u[field], v.G.sys_stride, p->d_trig,
                       p->d_cs);
  else
    hipLaunchKernelGGL(phi_kernel<17>, grid, dim3(256), 0, v.stream, p->ia, v.u[field], v.G.sys_stride, p->d_trig,
                       p->d_cs);
  SP_HIP(p, hipGetLastError());
  return KG_OK;
}

kg_status enqueue_theta(kg_spec* p, const HandleView& v) {
  const dim3 grid((p->Nr + 63) / 64, (p->npair + 3) / 4);
  hipLaunchKernelGGL(theta_kernel, grid, dim3(256), 0, v.stream, p->d_cs, p->d_leg, p->d_flm, p->Nr, p->Nt, p->mc,
                     p->lmax, p->nsys);
  SP_HIP(p, hipGetLastError());
  return KG_OK;
}

kg_status enqueue_radial(kg_spec* p, hipStream_t s, int precision) {
  if (precision == KG_SPEC_FP64) {
    if (p->lmax <= 8)
      hipLaunchKernelGGL(radial_f64_kernel<9>, dim3(p->ntiles64), dim3(256), 0, s, p->d_tiles64, p->d_W, p->d_flm,
                         p->d_ft, p->Nr, p->nk);
    else
      hipLaunchKernelGGL(radial_f64_kernel<17>, dim3(p->ntiles64), dim3(256), 0, s, p->d_tiles64, p->d_W,
                         p->d_flm, p->d_ft, p->Nr, p->nk);
  } else if (precision == KG_SPEC_SPLIT_BF16) {
    hipLaunchKernelGGL(radial_mfma_kernel, dim3((p->ntiles16 + 3) / 4), dim3(256), 0, s, p->d_tiles16, p->ntiles16,
                       p->d_W, p->d_flm, p->d_ft, p->Nr, p->nk);
  } else {
    return sfail(p, KG_ERR_INVALID, "unknown precision %d", precision);
  }
  SP_HIP(p, hipGetLastError());
  return KG_OK;
}

kg_status deliver(kg_spec* p, hipStream_t s, const double* src, double* dst, size_t count, int on_device) {
  SP_HIP(p, hipMemcpyAsync(dst, src, count * 8, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  if (!on_device) SP_HIP(p, hipStreamSynchronize(s));
  return KG_OK;
}

// flm [lm][part][i] (device) <-> complex [lm][i] interleaved
__global__ void interleave_kernel(double* __restrict__ planar, double* __restrict__ inter, int Nr, int nlm,
                                  int to_inter) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = (int64_t)nlm * Nr;
  if (e >= n) return;
  const int64_t lm = e / Nr, i = e % Nr;
  const int64_t re = (lm * 2) * Nr + i, im = re + Nr;
  if (to_inter) {
    inter[2 * e] = planar[re];
    inter[2 * e + 1] = planar[im];
  } else {
    planar[re] = inter[2 * e];
    planar[im] = inter[2 * e + 1];
  }
}

}  // namespace

// ===================================================================== C ABI
const char* kg_spec_last_error(const kg_spec* p) { return p ? p->err.c_str() : g_spec_error.c_str(); }

kg_status kg_spec_create(kg_handle* h, const kg_spec_desc* d, kg_spec** out) {
  if (!h || !d || !out) return sfail(nullptr, KG_ERR_INVALID, "null argument");
  *out = nullptr;
  if (d->lmax < 0 || d->lmax > 16) return sfail(nullptr, KG_ERR_INVALID, "lmax must be in 0..16");
  if (d->nk < 1) return sfail(nullptr, KG_ERR_INVALID, "nk must be >= 1");
  if (d->basis != KG_BASIS_ZEROS && d->basis != KG_BASIS_UNIFORM)
    return sfail(nullptr, KG_ERR_INVALID, "unknown basis %d", d->basis);
  if (d->basis == KG_BASIS_ZEROS && !(d->R_max > 0.0)) return sfail(nullptr, KG_ERR_INVALID, "R_max must be > 0");
  if (d->basis == KG_BASIS_UNIFORM && !(d->kmax > 0.0)) return sfail(nullptr, KG_ERR_INVALID, "kmax must be > 0");
  kg_spec* p = new kg_spec();
  p->h = h;
  p->lmax = d->lmax;
  p->nk = d->nk;
  p->basis = d->basis;
  p->R_max = d->R_max;
  p->kmax = d->kmax;
  HandleView v;
  kg_status st = view(p, &v);
  if (st != KG_OK) {
    g_spec_error = p->err;
    delete p;
    return st;
  }
  p->device = v.device;
  p->Nr = v.hg.Nx;
  p->Nt = v.hg.Ny;
  p->Np = v.hg.Nz;
  p->nsys = v.nsys;
  p->mc = p->lmax + 1 <= 9 ? 9 : 17;
  p->nlm = (p->lmax + 1) * (p->lmax + 1);
  p->npair = (p->lmax + 1) * (p->lmax + 2) / 2;
  if (p->Nr < 2) st = sfail(p, KG_ERR_INVALID, "need at least two radii");
  if (st == KG_OK) st = build_interp(p, v);
  if (st == KG_OK) st = build_angular(p, v);
  if (st == KG_OK) st = build_radial(p, v);
  if (st == KG_OK) st = dalloc(p, &p->d_cs, (size_t)p->nsys * p->Nt * 2 * p->mc * p->Nr);
  if (st == KG_OK) st = dalloc(p, &p->d_flm, (size_t)p->nlm * 2 * p->Nr);
  if (st == KG_OK) st = dalloc(p, &p->d_flmi, (size_t)p->nlm * 2 * p->Nr);
  if (st == KG_OK) st = dalloc(p, &p->d_ft, (size_t)p->nlm * p->nk * 2);
  if (st != KG_OK) {
    g_spec_error = p->err;
    free_all(p);
    delete p;
    return st;
  }
  *out = p;
  return KG_OK;
}

kg_status kg_spec_destroy(kg_spec* p) {
  if (!p) return KG_OK;
  (void)hipSetDevice(p->device);
  free_all(p);
  delete p;
  return KG_OK;
}

kg_status kg_spec_get_tables(kg_spec* p, double* ylm, double* kv, double* jb) {
  if (!p) return sfail(nullptr, KG_ERR_INVALID, "null projector");
  if (ylm) std::memcpy(ylm, p->ylm_re_im.data(), p->ylm_re_im.size() * 8);
  if (kv) std::memcpy(kv, p->kv.data(), p->kv.size() * 8);
  if (jb) {
    HandleView v;
    const kg_status st = view(p, &v);
    if (st != KG_OK) return st;
    const size_t n = (size_t)(p->lmax + 1) * p->nk * p->Nr;
    double* tmp = nullptr;
    if (dalloc(p, &tmp, n) != KG_OK) return KG_ERR_NOMEM;
    hipLaunchKernelGGL(bessel_kernel, dim3((p->lmax + 1) * p->nk), dim3(256), 0, v.stream, p->d_kv, p->d_r, p->Nr,
                       p->basis == KG_BASIS_ZEROS ? 1 : 0, p->dr, p->lmax, p->nk, 1, tmp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(jb, tmp, n * 8, hipMemcpyDeviceToHost, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return sfail(p, KG_ERR_HIP, "Bessel table: %s", hipGetErrorString(e));
  }
  return KG_OK;
}

kg_status kg_spec_domain_to_sph(kg_spec* p, int32_t field, double* out, int32_t on_device) {
  if (!p || !out) return sfail(p, KG_ERR_INVALID, "null argument");
  HandleView v;
  kg_status st = view(p, &v);
  if (st == KG_OK) st = check_field(p, field);
  if (st != KG_OK) return st;
  const size_t n = (size_t)p->nsys * p->Np * p->Nt * p->Nr;
  double* dst = out;
  if (!on_device) {
    if (!p->d_f && dalloc(p, &p->d_f, n) != KG_OK) return KG_ERR_NOMEM;
    dst = p->d_f;
  }
  const dim3 grid((p->Nr + 63) / 64, (p->Nt + 3) / 4, p->Np * p->nsys);
  hipLaunchKernelGGL(d2s_kernel, grid, dim3(256), 0, v.stream, p->ia, v.u[field], v.G.sys_stride, dst);
  SP_HIP(p, hipGetLastError());
  if (!on_device) return deliver(p, v.stream, p->d_f, out, n, 0);
  return KG_OK;
}

kg_status kg_spec_project_sh(kg_spec* p, int32_t field, double* flm, int32_t on_device) {
  if (!p || !flm) return sfail(p, KG_ERR_INVALID, "null argument");
  HandleView v;
  kg_status st = view(p, &v);
  if (st == KG_OK) st = check_field(p, field);
  if (st == KG_OK) st = enqueue_phi(p, v, field);
  if (st == KG_OK) st = enqueue_theta(p, v);
  if (st != KG_OK) return st;
  const size_t n = (size_t)p->nlm * p->Nr;
  double* inter = on_device ? flm : p->d_flmi;
  hipLaunchKernelGGL(interleave_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, v.stream, p->d_flm, inter,
                     p->Nr, p->nlm, 1);
  SP_HIP(p, hipGetLastError());
  if (!on_device) return deliver(p, v.stream, inter, flm, 2 * n, 0);
  return KG_OK;
}

kg_status kg_spec_radial(kg_spec* p, int32_t field, int32_t precision, double* ft, int32_t on_device) {
  if (!p || !ft) return sfail(p, KG_ERR_INVALID, "null argument");
  HandleView v;
  kg_status st = view(p, &v);
  if (st == KG_OK) st = check_field(p, field);
  if (st == KG_OK) st = enqueue_phi(p, v, field);
  if (st == KG_OK) st = enqueue_theta(p, v);
  if (st == KG_OK) st = enqueue_radial(p, v.stream, precision);
  if (st != KG_OK) return st;
  return deliver(p, v.stream, p->d_ft, ft, (size_t)p->nlm * p->nk * 2, on_device);
}

kg_status kg_spec_radial_batch(kg_spec* const* ps, int32_t n, int32_t field, int32_t precision, double* const* fts,
                               int32_t on_device) {
  if (!ps || !fts || n < 1) return sfail(nullptr, KG_ERR_INVALID, "need n >= 1 projectors and outputs");
  kg_spec* lead = ps[0];
  if (!lead) return sfail(nullptr, KG_ERR_INVALID, "batch: null projector 0");
  if (precision != KG_SPEC_FP64 && precision != KG_SPEC_SPLIT_BF16)
    return sfail(lead, KG_ERR_INVALID, "unknown precision %d", precision);
  HandleView lv;
  kg_status st = view(lead, &lv);
  if (st == KG_OK) st = check_field(lead, field);
  if (st != KG_OK) return st;
  std::vector<SpecMember> ms((size_t)n);
  std::vector<int> start((size_t)4 * (n + 1), 0);  // [pass][member]: phi, theta, radial FP64, radial BF16
  for (int q = 0; q < n; ++q) {
    kg_spec* p = ps[q];
    if (!p || !fts[q]) return sfail(lead, KG_ERR_INVALID, "batch: null projector or output %d", q);
    if (p->device != lead->device) return sfail(lead, KG_ERR_INVALID, "batch: projectors on different devices");
    if (p->mc != lead->mc)
      return sfail(lead, KG_ERR_UNSUPPORTED, "batch: lmax <= 8 and lmax > 8 projectors cannot share a launch");
    for (int o = 0; o < q; ++o)
      if (ps[o] == p) return sfail(lead, KG_ERR_INVALID, "batch: projector %d repeated", q);
    HandleView v;
    if (handle_view(p->h, &v) != 0) return sfail(lead, KG_ERR_INVALID, "batch: projector %d has no handle", q);
    // the member's pending work (its steps) completes before the batch reads its state
    if (v.stream != lv.stream) SP_HIP(lead, hipStreamSynchronize(v.stream));
    SpecMember& m = ms[(size_t)q];
    m.ia = p->ia;
    m.F = v.u[field];
    m.sys_stride = v.G.sys_stride;
    m.trig = p->d_trig;
    m.leg = p->d_leg;
    m.W = p->d_W;
    m.cs = p->d_cs;
    m.flm = p->d_flm;
    m.ft = p->d_ft;
    m.tiles64 = p->d_tiles64;
    m.tiles16 = p->d_tiles16;
    m.Nr = p->Nr;
    m.Nt = p->Nt;
    m.nk = p->nk;
    m.lmax = p->lmax;
    m.mc = p->mc;
    m.nsys = p->nsys;
    m.npair = p->npair;
    m.ntiles64 = p->ntiles64;
    m.ntiles16 = p->ntiles16;
    const int gx = (p->Nr + 63) / 64;
    const int64_t nb[4] = {(int64_t)gx * ((p->Nt + 3) / 4) * p->nsys, (int64_t)gx * ((p->npair + 3) / 4),
                           (int64_t)p->ntiles64, (int64_t)(p->ntiles16 + 3) / 4};
    for (int k = 0; k < 4; ++k) {
      const int64_t next = (int64_t)start[(size_t)k * (n + 1) + q] + nb[k];
      if (next > INT32_MAX) return sfail(lead, KG_ERR_INVALID, "batch: too many blocks");
      start[(size_t)k * (n + 1) + q + 1] = (int)next;
    }
  }
  const size_t mbytes = ((size_t)n * sizeof(SpecMember) + 255) & ~(size_t)255;
  const size_t bytes = mbytes + start.size() * sizeof(int);
  if (bytes > lead->batch_bytes) {
    SP_HIP(lead, hipStreamSynchronize(lv.stream));  // the old table may still be read
    if (lead->d_batch) (void)hipFree(lead->d_batch);
    lead->d_batch = nullptr;
    lead->batch_bytes = 0;
    char* d = nullptr;
    if (dalloc(lead, &d, bytes) != KG_OK) return KG_ERR_NOMEM;
    lead->d_batch = d;
    lead->batch_bytes = bytes;
  }
  if (!lead->batch_done) SP_HIP(lead, hipEventCreateWithFlags(&lead->batch_done, hipEventDisableTiming));
  std::vector<char> host(bytes, 0);
  std::memcpy(host.data(), ms.data(), (size_t)n * sizeof(SpecMember));
  std::memcpy(host.data() + mbytes, start.data(), start.size() * sizeof(int));
  // pageable source: the copy is staged before the call returns
  SP_HIP(lead, hipMemcpyAsync(lead->d_batch, host.data(), bytes, hipMemcpyHostToDevice, lv.stream));
  const SpecMember* dm = (const SpecMember*)lead->d_batch;
  const int* ds = (const int*)((const char*)lead->d_batch + mbytes);
  auto total = [&](int k) { return (unsigned)start[(size_t)k * (n + 1) + n]; };
  if (total(0) > 0) {
    if (lead->mc == 9)
      hipLaunchKernelGGL(phi_kernel_batch<9>, dim3(total(0)), dim3(256), 0, lv.stream, dm, ds, n);
    else
      hipLaunchKernelGGL(phi_kernel_batch<17>, dim3(total(0)), dim3(256), 0, lv.stream, dm, ds, n);
    SP_HIP(lead, hipGetLastError());
  }
  if (total(1) > 0) {
    hipLaunchKernelGGL(theta_kernel_batch, dim3(total(1)), dim3(256), 0, lv.stream, dm, ds + (n + 1), n);
    SP_HIP(lead, hipGetLastError());
  }
  if (precision == KG_SPEC_FP64 && total(2) > 0) {
    if (lead->lmax <= 8)
      hipLaunchKernelGGL(radial_f64_kernel_batch<9>, dim3(total(2)), dim3(256), 0, lv.stream, dm, ds + 2 * (n + 1), n);
    else
      hipLaunchKernelGGL(radial_f64_kernel_batch<17>, dim3(total(2)), dim3(256), 0, lv.stream, dm, ds + 2 * (n + 1),
                         n);
    SP_HIP(lead, hipGetLastError());
  } else if (precision == KG_SPEC_SPLIT_BF16 && total(3) > 0) {
    hipLaunchKernelGGL(radial_mfma_kernel_batch, dim3(total(3)), dim3(256), 0, lv.stream, dm, ds + 3 * (n + 1), n);
    SP_HIP(lead, hipGetLastError());
  }
  for (int q = 0; q < n; ++q) {
    kg_spec* p = ps[q];
    SP_HIP(lead, hipMemcpyAsync(fts[q], p->d_ft, (size_t)p->nlm * p->nk * 2 * 8,
                                on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, lv.stream));
  }
  // later work on any member (its steps, its own projections) waits for the batch
  SP_HIP(lead, hipEventRecord(lead->batch_done, lv.stream));
  for (int q = 1; q < n; ++q) {
    HandleView v;
    if (handle_view(ps[q]->h, &v) == 0 && v.stream != lv.stream)
      SP_HIP(lead, hipStreamWaitEvent(v.stream, lead->batch_done, 0));
  }
  if (!on_device) SP_HIP(lead, hipStreamSynchronize(lv.stream));
  return KG_OK;
}

kg_status kg_spec_radial_flm(kg_spec* p, const double* flm, int32_t precision, double* ft) {
  if (!p || !flm || !ft) return sfail(p, KG_ERR_INVALID, "null argument");
  HandleView v;
  kg_status st = view(p, &v);
  if (st != KG_OK) return st;
  const size_t n = (size_t)p->nlm * p->Nr;
  SP_HIP(p, hipMemcpyAsync(p->d_flmi, flm, 2 * n * 8, hipMemcpyHostToDevice, v.stream));
  hipLaunchKernelGGL(interleave_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, v.stream, p->d_flm,
                     p->d_flmi, p->Nr, p->nlm, 0);
  SP_HIP(p, hipGetLastError());
  st = enqueue_radial(p, v.stream, precision);
  if (st != KG_OK) return st;
  return deliver(p, v.stream, p->d_ft, ft, (size_t)p->nlm * p->nk * 2, 0);
}

kg_status kg_spec_detectors(kg_spec* p, int32_t n_det, const int32_t* index_det, double* momentum, double* alm) {
  if (!p || !index_det || n_det < 1) return sfail(p, KG_ERR_INVALID, "need a projector and n_det >= 1 indices");
  for (int d = 0; d < n_det; ++d)
    if (index_det[d] < 0 || index_det[d] >= p->Nr)
      return sfail(p, KG_ERR_INVALID, "index_det[%d] = %d outside 0..%d", d, index_det[d], p->Nr - 1);
  HandleView v;
  kg_status st = view(p, &v);
  if (st != KG_OK) return st;
  int* d_idx = nullptr;
  double* d_out = nullptr;
  const size_t nmom = (size_t)n_det * 3 * 2, nalm = (size_t)n_det * p->nlm * 2;
  if (dalloc(p, &d_idx, (size_t)n_det) != KG_OK) return KG_ERR_NOMEM;
  if (dalloc(p, &d_out, nmom + nalm) != KG_OK) {
    (void)hipFree(d_idx);
    return KG_ERR_NOMEM;
  }
  hipError_t e = hipMemcpyAsync(d_idx, index_det, sizeof(int) * n_det, hipMemcpyHostToDevice, v.stream);
  if (e == hipSuccess && momentum) {
    CurArgs c;
    for (int f = 0; f < 5; ++f) c.u[f] = v.u[f];
    c.tab = v.tab;
    c.sys_stride = v.G.sys_stride;
    c.plane = v.G.plane;
    c.n2 = (int64_t)v.G.nj * v.G.ld;
    c.ld = v.G.ld;
    c.nsys = v.nsys;
    hipLaunchKernelGGL(detector_flux_kernel, dim3(n_det), dim3(256), 0, v.stream, p->ia, c, d_idx, v.hg.dy, v.hg.dz,
                       d_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess && alm) {
    st = enqueue_phi(p, v, 0);
    if (st == KG_OK) st = enqueue_theta(p, v);
    if (st != KG_OK) {
      (void)hipStreamSynchronize(v.stream);
      (void)hipFree(d_idx);
      (void)hipFree(d_out);
      return st;
    }
    const int n = n_det * p->nlm;
    hipLaunchKernelGGL(gather_alm_kernel, dim3((n + 255) / 256), dim3(256), 0, v.stream, p->d_flm, d_idx, n_det,
                       p->nlm, p->Nr, d_out + nmom);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
  if (e == hipSuccess && momentum) e = hipMemcpy(momentum, d_out, nmom * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess && alm) e = hipMemcpy(alm, d_out + nmom, nalm * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d_idx);
  (void)hipFree(d_out);
  if (e != hipSuccess) return sfail(p, KG_ERR_HIP, "detectors: %s", hipGetErrorString(e));
  return KG_OK;
}
