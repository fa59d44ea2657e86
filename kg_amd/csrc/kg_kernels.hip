// kg_kernels.hip -- gfx950 kernels of the Klein-Gordon method-of-lines step.
//
// One launch per integrator stage: every thread owns one (r, theta) column
// point (ii, jj) of a 64 x TJ tile and marches over a chunk of phi planes.  At
// each point it evaluates the right-hand side of the five fields
// (MoL/lines_method_lib.f90:671-753) with the reference's 2nd-order stencils
// and edge rows (tools/finite_differences.f90:17-44,193-239,310-338,767-801)
// and applies the stage combine of the integrator in the same expression.
// The metric tables are k-independent, so a thread loads its coefficients
// once for the whole phi march.
//
// Per plane: products into LDS, loads of the next plane issued, barrier,
// derivatives from LDS, then the terms that need the next plane.  The loads
// are kept as raw operands (u_p and k of an RK4 stage state) and combined
// only after the barrier and the LDS work, so their latency overlaps both;
// scheduling barriers keep the compiler from moving that arithmetic ahead.
//
// Floating-point contraction is off for this file (build flag
// -ffp-contract=off): every product and sum is rounded on its own, in the
// reference's left-to-right order, so the results match a CPU evaluation of
// the Fortran expressions bit for bit.
#include <hip/hip_runtime.h>

#include "kg_internal.h"

namespace kg {
namespace {

// Every device buffer the kernels touch is hipMalloc'd global memory, and is
// accessed through address-space-1 pointers.  A pointer read from a
// descriptor table in memory (the batched kernels) is otherwise generic and
// compiles to flat loads and stores, which also count against the LDS
// counter: each wait for a field then drains the tile's LDS traffic as well
// (measured 7-15 % on a batched stage launch).
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const double gcdouble;
typedef double double2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) double2v gdouble2;
typedef __attribute__((address_space(1))) char gchar;
typedef __attribute__((address_space(1))) const char gcchar;

__device__ __forceinline__ gdouble* gptr(double* p) { return (gdouble*)p; }
__device__ __forceinline__ gcdouble* gptr(const double* p) { return (gcdouble*)p; }

// Global access as (uniform 64-bit base) + (32-bit byte offset): the base
// stays in SGPRs and a lane carries one 32-bit offset, which maps to the
// SGPR-base form of global_load/global_store.  Every field of one system is
// below 2 GiB (checked at kg_create), so offsets fit.
__device__ __forceinline__ double gld(const double* base, uint32_t off) {
  return *(gcdouble*)((gcchar*)base + off);
}
__device__ __forceinline__ void gst(double* base, uint32_t off, double v) {
  *(gdouble*)((gchar*)base + off) = v;
}

constexpr int TI = 64;  // tile width along r: one wavefront per theta row
// theta tiles per group of tile_in_run: the 2nd-order stages take the phi
// chunks of two theta tiles in turn, the fd4 stages one tile at a time
// (pairs measured 1.2 % slower there: 9.75 against 9.63 ms per RK4 step,
// and 8.64 against 8.52 with the theta tiles dealt over the XCDs)
constexpr int P2 = 2, P4 = 1;
constexpr int HALO_R = 2;  // r-neighbours needed on each side (A_r reaches i-2 and i+2)
constexpr int NXCD = 8;    // XCDs of an MI355X; workgroups are dealt round-robin over them

template <int MODE>
struct ModeTraits {
  // fields 1..4 (phi, psi_r, psi_theta, psi_phi) and/or field 5 (Pi) of the RHS
  static constexpr bool n14 = !(MODE == M_PIRK_1B || MODE == M_PIRK3_2B);
  static constexpr bool n5 = !(MODE == M_PIRK_1A || MODE == M_PIRK3_2A);
};

__device__ __forceinline__ int remap_xcd(int b, int nb) {
  // Workgroup b runs on XCD b % NXCD.  Give each XCD a contiguous run of
  // tiles so neighbouring tiles (which share halo planes and rows) meet in
  // the same L2.  Bijective on [0, nb).
  const int q = nb / NXCD, r = nb % NXCD;
  const int x = b % NXCD, idx = b / NXCD;
  return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

struct Tile {
  int ti, tj, tk;  // r tile, theta tile, phi chunk
};

// The order of the tiles in a run of workgroups: r tiles fastest (the r
// neighbours share halo columns in L2), then the phi chunks of a group of P
// theta tiles, taken in turn (chunk 0 of each tile of the group, then chunk
// 1, ...), then the groups; a last theta group of fewer tiles comes last.
__device__ __forceinline__ Tile tile_in_run(int t, int nti, int ntj, int ntk, int P) {
  Tile o;
  o.ti = t % nti;
  const int rest = t / nti;
  const int full = ntj / P;
  const int tiles_full = full * P * ntk;
  if (rest < tiles_full) {
    const int g2 = rest / (P * ntk), r2 = rest % (P * ntk);
    o.tj = g2 * P + r2 % P;
    o.tk = r2 / P;
  } else {
    const int r3 = rest - tiles_full, rem = ntj - full * P;
    o.tj = full * P + r3 % rem;
    o.tk = r3 / rem;
  }
  return o;
}

// Tile of workgroup b of a one-grid launch.  Workgroups are dealt
// round-robin over the XCDs, and all XCDs reach the same position of their
// share at about the same time.  The theta tiles are dealt the same way:
// with q = ntj / NXCD, XCD x runs theta tiles x, x + NXCD, ..., x +
// (q-1)*NXCD (in the order of tile_in_run), so the XCDs march
// theta-neighbouring tiles side by side and the halo rows a tile reads are
// its neighbours' own rows, read at about the same time and served from the
// memory-side cache rather than HBM (fd4 RK4 on 2048 x 128 x 256: 8.96 ->
// 8.53 ms per step).  The ntj % NXCD theta tiles left over follow, split
// into contiguous runs (remap_xcd).  Bijective on [0, nti*ntj*ntk).
__device__ __forceinline__ Tile tile_dealt(int b, int nti, int ntj, int ntk, int P) {
  const int q = ntj / NXCD;
  const int nA = NXCD * q * nti * ntk;
  if (b < nA) {
    Tile o = tile_in_run(b / NXCD, nti, q, ntk, P);
    o.tj = o.tj * NXCD + b % NXCD;
    return o;
  }
  const int rem = ntj - NXCD * q;
  Tile o = tile_in_run(remap_xcd(b - nA, rem * nti * ntk), nti, rem, ntk, P);
  o.tj += NXCD * q;
  return o;
}

// Stage kernel.  A workgroup owns a TI x TJ tile of (r, theta) columns and
// marches over a chunk of phi planes.  At each plane every thread loads its
// point's five fields, forms the products the stencils act on (alpha*Pi,
// psi_r, the r-flux, Pi*sqrtg, F = alpha*Pi + beta*psi_r, the theta-flux)
// with its own metric coefficients and shares them through LDS; r- and
// theta-neighbours are read back from LDS, phi-neighbours come from a
// register queue (the point's own F and phi-flux at k-1, k, k+1).  Halo
// columns (i0-2, i0-1, i0+TI, i0+TI+1) and rows (j0-1, j0+TJ) are formed by
// designated threads of the same workgroup.
template <int MODE>
struct Combine {
  // operands of the stage combine read at the point: base (u_p, P, u) and
  // aux (PIRK3 stage state; RK4 k1 and k2 in stage 4); fields [lo, hi)
  static constexpr int lo = (MODE == M_PIRK_1B || MODE == M_PIRK3_2B) ? 4 : 0;
  static constexpr int hi = (MODE == M_PIRK_1A || MODE == M_PIRK3_2A) ? 4 : 5;
  static constexpr bool base = !(MODE == M_RK4_S1 || MODE == M_RK4_S2 || MODE == M_RK4_S3);
  static constexpr bool aux = MODE == M_RK4_S4 || MODE == M_PIRK3_2A || MODE == M_PIRK3_2B;
  // RK4 stages 2 and 4 read the stage state u_p + c*k and its ghost fill;
  // stage 3 reads the state Y = u_p + k2/2 that stage 2 stored, ghost-filled
  static constexpr bool virt = MODE == M_RK4_S2 || MODE == M_RK4_S4;
  // stages 2 and 4 combine u_p and k at the point: the march holds both as
  // the raw operands of its stage state
  static constexpr bool keep = MODE == M_RK4_S2 || MODE == M_RK4_S4;
  // stages followed by a ghost fill compute only i in [1, ni-2],
  // j in [g+1, g+Ny-1], k in [0, Nz-1]: the edge rows of D_r at i = -g and
  // Nx+g, of D_theta at j = -g and Ny+g and of D_phi at k = Nz never occur
  static constexpr bool boxed = !(MODE == M_PIRK_1A || MODE == M_PIRK_1B || MODE == M_PIRK3_2A || MODE == M_PIRK3_2B);
};

// LDS of a 2nd-order tile of TIL x TJL points: 4 r-arrays and 2 theta-arrays
// with their halos
template <int TIL, int TJL>
constexpr int lds2_doubles() {
  return 4 * TJL * (TIL + 2 * HALO_R) + 2 * (TJL + 2) * TIL;
}
// the last r columns of a box as strips of SI2 x (TJ * TI / SI2) points
// (one strip covers 128 theta rows at TJ = 8, 64 at 4), as the fd4 set does;
// the RK stages only: on 2048 x 128 x 256 strips take the RK4 step from 7.84
// to 7.62 ms but the PIRK2 step from 3.61 to 3.68 (its stage 2 alone 1.62 ->
// 1.68 ms)
constexpr int SI2 = 4;
__host__ __device__ constexpr bool strips2(int mode) {
  return mode == M_RK4_S1 || mode == M_RK4_S2 || mode == M_RK4_S3 || mode == M_RK4_S4 || mode == M_RK3_S1 ||
         mode == M_RK3_S2 || mode == M_RK3_S3;
}

// LDS of a stage with strips: tiles and strips carve their arrays from one
// buffer (with the tiles' own arrays beside a strip buffer, RK4 stage 4
// takes 176 instead of 168 VGPRs and loses a wave per SIMD: 7.62 -> 7.90 ms
// per RK4 step)
template <int TJ>
constexpr int lds2_of() {
  return lds2_doubles<TI, TJ>() > lds2_doubles<SI2, TJ * TI / SI2>() ? lds2_doubles<TI, TJ>()
                                                                      : lds2_doubles<SI2, TJ * TI / SI2>();
}

// TIL x TJL: the tile's points in r and theta (TI x TJ, or a strip's SI2 x
// TJ*TI/SI2), on the same TI*TJ threads.  The stages with strips carve the
// LDS arrays from `lds`; the others declare their own (which leaves out the
// arrays a stage of phi and psi alone, or of Pi alone, does not use).
template <int MODE, int EF, int TJ, int TIL = TI, int TJL = TJ>
__device__ __forceinline__ void stage_body(const StageArgs& a, const Tile tile, const int sysidx, double* lds) {
  static_assert(TIL * TJL == TI * TJ, "a tile has one point per thread");
  constexpr bool N14 = ModeTraits<MODE>::n14;
  constexpr bool N5 = ModeTraits<MODE>::n5;
  constexpr int CLO = Combine<MODE>::lo, CHI = Combine<MODE>::hi;
  constexpr bool CBASE = Combine<MODE>::base, CAUX = Combine<MODE>::aux, VIRT = Combine<MODE>::virt;
  constexpr bool BOXED = Combine<MODE>::boxed;
  // stages 2 and 4 keep u_p and k of their march in registers for the
  // combine (stage 2 then needs ~164 VGPRs and runs 4 theta rows, three
  // workgroups to a CU: 8.59 -> 8.19 ms per RK4 step on 2048 x 128 x 256,
  // against reading u_p and k1 at the point again after the RHS)
  constexpr bool KEEP = Combine<MODE>::keep;

  constexpr int W = TIL + 2 * HALO_R;
  double(*sG)[W];    // alpha*Pi
  double(*sPR)[W];   // psi_r
  double(*sFR)[W];   // r-flux of rhs_pi
  double(*sPS)[W];   // Pi*sqrtg
  double(*sF)[TIL];  // alpha*Pi + beta*psi_r
  double(*sFT)[TIL];  // theta-flux of rhs_pi
  if constexpr (!strips2(MODE)) {
    __shared__ double aG[TJL][W], aPR[TJL][W], aFR[TJL][W], aPS[TJL][W];
    __shared__ double aF[TJL + 2][TIL], aFT[TJL + 2][TIL];
    sG = aG;
    sPR = aPR;
    sFR = aFR;
    sPS = aPS;
    sF = aF;
    sFT = aFT;
  } else {
    sG = reinterpret_cast<double(*)[W]>(lds);
    sPR = reinterpret_cast<double(*)[W]>(lds + TJL * W);
    sFR = reinterpret_cast<double(*)[W]>(lds + 2 * TJL * W);
    sPS = reinterpret_cast<double(*)[W]>(lds + 3 * TJL * W);
    sF = reinterpret_cast<double(*)[TIL]>(lds + 4 * TJL * W);
    sFT = reinterpret_cast<double(*)[TIL]>(lds + 4 * TJL * W + (TJL + 2) * TIL);
  }

  const Geom& G = a.G;
  const int ti = tile.ti, tj = tile.tj, tk = tile.tk;
  const int tid = (int)threadIdx.y * TI + (int)threadIdx.x;
  // (a strip's waves each hold one r column: uniform stencil rows per wave)
  const int lane = TIL == TI ? tid % TI : tid / TJL, row = TIL == TI ? tid / TI : tid % TJL;
  // (a strip's ti is the count of whole r tiles: it starts where they end)
  const int i0 = a.i_lo + ti * TI, j0 = a.j_lo + tj * TJL;
  const int ii = i0 + lane, jj = j0 + row;
  const int ni = G.ni, nj = G.nj, Nz = G.Nz, Ny = G.Ny, g = G.g, ld = G.ld;
  const uint32_t plane8 = (uint32_t)(G.plane * 8);
  const uint32_t n2b = (uint32_t)((int64_t)nj * ld * 8);  // bytes of one table
  const bool valid = ii < ni && jj < nj;
  const bool act = valid && ii < a.i_hi && jj < a.j_hi;
  const int k0 = a.k_lo + tk * a.kc;
  const int k1 = min(k0 + a.kc, a.k_hi);
  const int64_t sys = (int64_t)sysidx * G.sys_stride;
  const uint32_t sysb = (uint32_t)(sys * 8);  // system offset in bytes
  const double* tab = a.tab;
  auto T = [&](int which, uint32_t off) { return gld(tab, off + (uint32_t)which * n2b); };
  // address of field f of a role: slot f (field 4: slot ri4) of that role's buffer
  auto FP = [&](int r, int f) { return a.rb[r] + (int64_t)(f == 4 ? a.ri4[r] : f) * a.per; };
  auto OFF = [&](int i, int j, int k) { return (uint32_t)((j * ld + i) * 8) + (uint32_t)k * plane8; };
  const int dis = Nz / 2;

  // Value of field f of the stage input at padded point (i, j, k).  For the
  // RK4 stages that form u = u_p + c*k, points the ghost fill writes are
  // mapped to the evolved points they copy (lines_method_lib.f90:313-330,
  // 516-596): zL/zR, then yL/yR at k_target, then xL (extrapolation) / xR.
  // remap() resolves a point once; Ur() reads any field there.
  auto E = [&](int f, uint32_t o) {
    if constexpr (VIRT) {
      const double p = gld(FP(R_IN, f), o + sysb), k = gld(FP(R_KIN, f), o + sysb);
      if constexpr (MODE == M_RK4_S4)
        return p + k;  // u_p + k3
      else
        return p + 0.50 * k;  // u_p + 0.5*k1
    } else {
      return gld(FP(R_IN, f), o + sysb);  // u (or Y = u_p + 0.5*k2, stored by stage 2)
    }
  };
  struct Src {
    uint32_t o;
    int kind;  // 0: evolved value at o; 1: xL extrapolation from o, o+1, o+2; 2: u_p at o (inflow)
  };
  auto remap = [&](int i, int j, int k) -> Src {
    if constexpr (VIRT) {
      if (i < 1 || i > ni - 2 || j < g + 1 || j > g + Ny - 1 || k < 1 || k > Nz - 1) {
        if (k == 0)
          k = Nz - 1;
        else if (k == Nz)
          k = 1;
        if (j <= g) {
          j = 2 * g + 1 - j;
          k = (k + dis) % (Nz - 1);
        } else if (j >= g + Ny) {
          j = 2 * Ny + 2 * g - 1 - j;
          k = (k + dis) % (Nz - 1);
        }
        if (i == 0) return Src{OFF(1, j, k), 1};
        if (i == ni - 1 && !G.bc_somm) return G.bc_inflow ? Src{OFF(ni - 1, j, k), 2} : Src{OFF(ni - 2, j, k), 0};
        return Src{OFF(i, j, k), 0};
      }
    }
    return Src{OFF(i, j, k), 0};
  };
  auto Ur = [&](int f, const Src& q) -> double {
    if constexpr (VIRT) {
      if (q.kind == 1) return 3.0 * E(f, q.o) - 3.0 * E(f, q.o + 8) + E(f, q.o + 16);
      if (q.kind == 2) return gld(FP(R_IN, f), q.o + sysb);
    }
    return E(f, q.o);
  };
  auto U = [&](int f, int i, int j, int k) { return Ur(f, remap(i, j, k)); };
  // Deferred stage input.  fetch() issues the loads of field f at a resolved
  // point and keeps the raw operands (u_p and k for the RK4 stage states);
  // val() forms the value where it is used, so loads issued ahead of a
  // barrier stay in flight across it.  Points resolved through the r-edge
  // fills (xL extrapolation, inflow copy) are formed at once ("direct").
  struct Raw {
    double p, k;
  };
  // A zero the compiler cannot see through (dt * 0.0 is not folded without
  // fast-math): the direct branch's unused k must not be a constant, or the
  // stage arithmetic gets folded into the branches and the common path has
  // to wait for its loads there.
  const double opaque0 = G.dt * 0.0;
  auto fetch = [&](int f, const Src& q, Raw& r) {
    if constexpr (VIRT) {
      if (q.kind == 0) {
        r.p = gld(FP(R_IN, f), q.o + sysb);
        r.k = gld(FP(R_KIN, f), q.o + sysb);
      } else {
        r.p = Ur(f, q);
        r.k = opaque0;  // unused: val() selects p
      }
    } else {
      r.p = gld(FP(R_IN, f), q.o + sysb);
      r.k = opaque0;
    }
  };
  auto val = [&](const Raw& r, bool direct) -> double {
    if constexpr (VIRT) {
      const double v = (MODE == M_RK4_S4) ? r.p + r.k : r.p + 0.50 * r.k;  // as E() forms it
      return direct ? r.p : v;
    } else {
      return r.p;
    }
  };

  // own coefficients (the theta-ghost row of the tables is already the copy)
  const uint32_t own = valid ? (uint32_t)((jj * ld + ii) * 8) : 0u;
  const double al = T(T_ALPHA, own), be = T(T_BETA, own);
  const double dxb = T(T_DXBETA, own), sg = T(T_SQRTG, own);
  const double asg = T(T_ASG, own), grr = T(T_GRR, own);
  const double grp = T(T_GRP, own), asgtt = T(T_ASGTT, own);
  const double gpp = T(T_GPP, own), am2 = T(T_AM2, own);

  auto Fof = [&](double Pi, double psr) { return al * Pi + be * psr; };
  auto fpof = [&](double psr, double psp) {
    if constexpr (EF)
      return (al * gpp) * psp;
    else
      return asg * (grp * psr + gpp * psp);
  };

  // halo task of this thread: 1 = theta row, 2 = r column (two each side;
  // a strip, whose columns reach ni-1, the left two only), 0 = none
  constexpr int RH = TIL == TI ? 4 : 2;
  static_assert(2 * TIL + RH * TJL <= TI * TJ, "one halo task per thread at most");
  int hkind = 0, hi = 0, hj = 0, hrow = 0, hcol = 0;
  if (tid < 2 * TIL) {
    hkind = 1;
    hi = i0 + (tid % TIL);
    hj = tid < TIL ? j0 - 1 : j0 + TJL;
    hrow = tid < TIL ? 0 : TJL + 1;
    hcol = tid % TIL;
  } else if (tid < 2 * TIL + RH * TJL) {
    const int u = tid - 2 * TIL;
    const int cc = u % RH;
    hkind = 2;
    hj = j0 + u / RH;
    hrow = u / RH;
    hi = cc == 0 ? i0 - 2 : cc == 1 ? i0 - 1 : cc == 2 ? i0 + TIL : i0 + TIL + 1;
    hcol = hi - i0 + HALO_R;
  }
  if (hi < 0 || hi >= ni || hj < 0 || hj >= nj) hkind = 0;
  const uint32_t hown = hkind ? (uint32_t)((hj * ld + hi) * 8) : 0u;
  // halo coefficients: theta rows use alpha, beta, (alpha sqrtg) gamma^tt;
  // r columns use alpha, alpha*sqrtg, gamma^rr, gamma^rphi, sqrtg
  double h0 = 0, h1 = 0, h2 = 0, h3 = 0, h4 = 0;
  if (hkind == 1) {
    h0 = T(T_ALPHA, hown);
    h1 = T(T_BETA, hown);
    h2 = T(T_ASGTT, hown);
  } else if (hkind == 2) {
    h0 = T(T_ALPHA, hown);
    h1 = T(T_ASG, hown);
    h2 = T(T_GRR, hown);
    h3 = T(T_GRP, hown);
    h4 = T(T_SQRTG, hown);
  }
  // halo operands: Pi, psi_r and psi_theta (theta rows) or psi_phi (r columns)
  auto fetch_halo = [&](int k, Raw* x, bool& direct) {
    if (hkind == 0) return;
    const Src q = remap(hi, hj, k);
    direct = q.kind != 0;
    if (hkind == 1) {
      if constexpr (N14) {
        fetch(4, q, x[0]);
        fetch(1, q, x[1]);
      }
      if constexpr (N5) fetch(2, q, x[2]);
    } else {
      fetch(4, q, x[0]);
      fetch(1, q, x[1]);
      if constexpr (N5 && !EF) fetch(3, q, x[2]);
    }
  };

  // phi queue at the own point: F and the phi-flux at k-1 and k
  double Pi = 0, psr = 0, pst = 0, psp = 0;  // own values at plane kk
  double Fm = 0, fpm = 0, F0 = 0, fp0 = 0;
  Raw hr[3] = {{0, 0}, {0, 0}, {0, 0}};  // halo operands at plane kk
  bool hdir = false;
  Src qc{0u, 0};  // own point of plane kk, resolved
  Raw cr[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};  // raw own operands of plane kk (KEEP)
  if (valid) {
    const int km = k0 == 0 ? Nz - 2 : k0 - 1;  // D_phi at k = 0 pairs f(Nz-2) with f(1)
    const Src qm = remap(ii, jj, km);
    const double mPi = Ur(4, qm), mpr = Ur(1, qm);
    Fm = Fof(mPi, mpr);
    if constexpr (N5) fpm = fpof(mpr, Ur(3, qm));
    const Src q0 = remap(ii, jj, k0);
    qc = q0;
    const bool d0 = q0.kind != 0;
    fetch(4, q0, cr[0]);
    fetch(1, q0, cr[1]);
    Pi = val(cr[0], d0);
    psr = val(cr[1], d0);
    if constexpr (N5) {
      fetch(2, q0, cr[2]);
      fetch(3, q0, cr[3]);
      pst = val(cr[2], d0);
      psp = val(cr[3], d0);
      fp0 = fpof(psr, psp);
    }
    F0 = Fof(Pi, psr);
  }
  fetch_halo(k0, hr, hdir);

  const double rr = G.rr, rt = G.rt, rp = G.rp, dt = G.dt;
  const int c = lane + HALO_R;  // own column in the r-arrays

  for (int kk = k0; kk < k1; ++kk) {
    const uint32_t sk = (uint32_t)kk * plane8;
    const double hx0 = val(hr[0], hdir), hx1 = val(hr[1], hdir), hx2 = val(hr[2], hdir);
    // --- products of plane kk into LDS (from registers) ----------------------
    if (valid) {
      if constexpr (N14) {
        sG[row][c] = al * Pi;
        sPR[row][c] = psr;
        sF[row + 1][lane] = F0;
      }
      if constexpr (N5) {
        if constexpr (EF)
          sFR[row][c] = asg * grr * psr;
        else
          sFR[row][c] = asg * (grr * psr + grp * psp);
        sPS[row][c] = Pi * sg;
        sFT[row + 1][lane] = asgtt * pst;
      }
    }
    if (hkind == 1) {
      if constexpr (N14) sF[hrow][hcol] = h0 * hx0 + h1 * hx1;
      if constexpr (N5) sFT[hrow][hcol] = h2 * hx2;
    } else if (hkind == 2) {
      if constexpr (N14) {
        sG[hrow][hcol] = h0 * hx0;
        sPR[hrow][hcol] = hx1;
      }
      if constexpr (N5) {
        if constexpr (EF)
          sFR[hrow][hcol] = h1 * h2 * hx1;
        else
          sFR[hrow][hcol] = h1 * (h2 * hx1 + h3 * hx2);
        sPS[hrow][hcol] = hx0 * h4;
      }
    }
    // --- loads issued before the barrier: next plane, combine operands --------
    Raw nr[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};  // Pi, psi_r, psi_theta, psi_phi at kk+1
    bool ndir = false;
    Src qn{0u, 0};
    if (valid && kk < Nz) {
      qn = remap(ii, jj, kk + 1);
      ndir = qn.kind != 0;
      fetch(4, qn, nr[0]);
      fetch(1, qn, nr[1]);
      if constexpr (N5) {
        fetch(2, qn, nr[2]);
        fetch(3, qn, nr[3]);
      }
    }
    Raw nhr[3] = {{0, 0}, {0, 0}, {0, 0}};
    bool nhdir = false;
    if (kk + 1 < k1) fetch_halo(kk + 1, nhr, nhdir);
    // phi at the own point for the mass term of rhs_pi
    Raw vr{0, 0};
    if constexpr (N5) {
      if (act) fetch(0, qc, vr);
    }
    const uint32_t so = sk + sysb + own;
    __syncthreads();
    // keep the arithmetic on the operands above below the barrier
    __builtin_amdgcn_sched_barrier(0);
    double nPi = 0, npsr = 0, npst = 0, npsp = 0;  // plane kk+1, formed once its loads are needed
    auto next_values = [&]() {
      nPi = val(nr[0], ndir);
      npsr = val(nr[1], ndir);
      if constexpr (N5) {
        npst = val(nr[2], ndir);
        npsp = val(nr[3], ndir);
      }
    };

    // combine operands and RHS terms of plane kk; vb, va, vk: u_p / P / u,
    // the PIRK3 stage state or RK4 k1 + 2k2, RK4 k1 or k3 at the point
    double vb[5] = {0, 0, 0, 0, 0}, va[5] = {0, 0, 0, 0, 0}, vk[5] = {0, 0, 0, 0, 0};
    double R[5] = {0, 0, 0, 0, 0};
    double t1 = 0, t2 = 0, t5 = 0;
    double vphi = 0;
    const int kt = (kk + dis) % (Nz - 1);
    // D_r / A_r on an LDS r-array (finite_differences.f90:36-42, :779-799)
    auto Dr = [&](const double(*A)[W]) {
      if constexpr (!BOXED) {
        if (ii == 0) return (-3.0 * A[row][c] + 4.0 * A[row][c + 1] - A[row][c + 2]) * rr;
        if (ii == ni - 1) return (+3.0 * A[row][c] - 4.0 * A[row][c - 1] + A[row][c - 2]) * rr;
      }
      return (-A[row][c - 1] + A[row][c + 1]) * rr;
    };
    auto Ar = [&](const double(*A)[W]) {
      if (ii <= 1) return (-3.0 * A[row][c] + 4.0 * A[row][c + 1] - A[row][c + 2]) * rr;
      if (ii == ni - 2) return (A[row][c + 1] - A[row][c - 1]) * rr;
      if constexpr (!BOXED) {
        if (ii == ni - 1) return (3.0 * A[row][c] - 4.0 * A[row][c - 1] + A[row][c - 2]) * rr;
      }
      if (be > 0.0) return (-3.0 * A[row][c] + 4.0 * A[row][c + 1] - A[row][c + 2]) * rr;
      if (be < 0.0) return (3.0 * A[row][c] - 4.0 * A[row][c - 1] + A[row][c - 2]) * rr;
      return (A[row][c + 1] - A[row][c - 1]) * rr;
    };
    // D_theta on an LDS theta-array (finite_differences.f90:214-230); the
    // pole row reads plane k_target and the last row may reach below the
    // tile, both from global memory through `far`.
    auto Dth = [&](const double(*A)[TIL], auto far) {
      if constexpr (!BOXED) {
        if (jj == 0) return (-A[row + 2][lane] + far(2, kt)) * rt;
        if (jj == nj - 1) {
          const double m2 = row >= 1 ? A[row - 1][lane] : far(nj - 3, kk);
          return (+3.0 * A[row + 1][lane] - 4.0 * A[row][lane] + m2) * rt;
        }
      }
      return (-A[row][lane] + A[row + 2][lane]) * rt;
    };
    // D_phi (finite_differences.f90:328-335); k = Nz pairs f(Nz-1) with f(2)
    auto Dph = [&](double qm, double qp, auto at2) {
      if constexpr (!BOXED) {
        if (kk == Nz) return (-at2() + qm) * rp;
      }
      return (-qm + qp) * rp;
    };
    if (act) {
      // combine operands (issued ahead of the RHS arithmetic)
      // KEEP: u_p and k3 at the point are the raw operands of the march
      // (plane 0 resolves to plane Nz-1, so it reads plane 0 itself: the x and
      // y fills read it before the z fill overwrites it)
      const bool kept = KEEP && kk != 0;
      if constexpr (KEEP) {
        // u_p and k at the point (stage 2: k1, stage 4: k3)
        if (kept) {
          vb[0] = vr.p;
          vk[0] = vr.k;
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            vb[x == 0 ? 4 : x] = cr[x].p;
            vk[x == 0 ? 4 : x] = cr[x].k;
          }
        } else {
#pragma unroll
          for (int f = 0; f < 5; ++f) {
            vb[f] = gld(FP(R_IN, f), so);
            vk[f] = gld(FP(R_KIN, f), so);
          }
        }
      } else if constexpr (CBASE) {
#pragma unroll
        for (int f = CLO; f < CHI; ++f) vb[f] = gld(FP(R_BASE, f), so);
      }
      if constexpr (CAUX) {
#pragma unroll
        for (int f = CLO; f < CHI; ++f) va[f] = gld(FP(R_AUX, f), so);
      }
      // terms from LDS and the own registers first: they need none of the
      // loads issued before the barrier, which complete meanwhile
      if constexpr (N14) {
        R[0] = F0;  // rhs_phi = alpha*Pi + beta*psi_r (:675)
        R[1] = Dr(sG) + be * Ar(sPR) + dxb * psr;  // :687-689
        R[2] = Dth(sF, [&](int j, int k) {
          const uint32_t o = (uint32_t)((j * ld + ii) * 8);
          return T(T_ALPHA, o) * U(4, ii, j, k) + T(T_BETA, o) * U(1, ii, j, k);
        });
      }
      if constexpr (N5) {
        t1 = Dr(sFR) / sg;
        t2 = Dth(sFT, [&](int j, int k) { return T(T_ASGTT, (uint32_t)((j * ld + ii) * 8)) * U(2, ii, j, k); }) /
             sg;
        t5 = be * Ar(sPS) / sg;
      }
    }
    // the next plane's values, once the terms above have covered its loads
    __builtin_amdgcn_sched_barrier(0);
    next_values();
    const double Fp = Fof(nPi, npsr);
    double fpp = 0;
    if constexpr (N5) fpp = fpof(npsr, npsp);
    if (act) {
      if constexpr (N5) vphi = val(vr, qc.kind != 0);
      if constexpr (N14) R[3] = Dph(Fm, Fp, [&]() { return Fof(U(4, ii, jj, 2), U(1, ii, jj, 2)); });
      if constexpr (N5) {
        const double dp = Dph(fpm, fpp, [&]() { return fpof(U(1, ii, jj, 2), U(3, ii, jj, 2)); });
        const double t3 = EF ? dp : dp / sg;
        const double t4 = Pi * dxb;
        const double t6 = am2 * vphi;
        R[4] = t1 + t2 + t3 + t4 + t5 - t6;  // :728-736
      }
      // stage combine; `cur` is the RHS input at this point
      const double cur[5] = {vphi, psr, pst, psp, Pi};
      if constexpr (MODE == M_RK4_S1 || MODE == M_RK4_S3) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, R[f] * dt);  // k = rhs_u*dt (:430,436)
      } else if constexpr (MODE == M_RK4_S2) {
        // k2 = rhs_u*dt (:433); stored as the stage-3 state u_p + 0.5*k2 and
        // the first partial sum k1 + 2*k2 of the final combine
#pragma unroll
        for (int f = 0; f < 5; ++f) {
          const double k2 = R[f] * dt;
          gst(FP(R_OUT, f), so, vb[f] + 0.50 * k2);
          gst(FP(R_AUX, f), so, vk[f] + 2.0 * k2);
        }
      } else if constexpr (MODE == M_RK4_S4) {
        // :439-440 u = u_p + (k1 + 2k2 + 2k3 + k4)/6, k4 = rhs_u*dt; va = k1 + 2k2
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] + (va[f] + 2.0 * vk[f] + R[f] * dt) / 6.0);
      } else if constexpr (MODE == M_RK3_S1) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] + R[f] * dt);
      } else if constexpr (MODE == M_RK3_S2) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, 0.750 * vb[f] + 0.250 * (cur[f] + R[f] * dt));
      } else if constexpr (MODE == M_RK3_S3) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] / 3.0 + 2.0 * (cur[f] + R[f] * dt) / 3.0);
      } else if constexpr (MODE == M_PIRK_1A) {
#pragma unroll
        for (int f = 0; f < 4; ++f) gst(FP(R_OUT, f), so, vb[f] + dt * R[f]);
      } else if constexpr (MODE == M_PIRK_1B) {
        gst(FP(R_OUT, 4), so, vb[4] + dt * R[4]);
      } else if constexpr (MODE == M_PIRK2_2) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, 0.5 * (vb[f] + cur[f] + dt * R[f]));
      } else if constexpr (MODE == M_PIRK3_2A) {
#pragma unroll
        for (int f = 0; f < 4; ++f) gst(FP(R_OUT, f), so, 0.75 * vb[f] + 0.25 * (va[f] + dt * R[f]));
      } else if constexpr (MODE == M_PIRK3_2B) {
        gst(FP(R_OUT, 4), so, 0.75 * vb[4] + 0.25 * (va[4] + dt * R[4]));
      } else if constexpr (MODE == M_PIRK3_3) {
        const double third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, third * vb[f] + two_thirds * (cur[f] + dt * R[f]));
      }
    }
    Fm = F0;
    fpm = fp0;
    F0 = Fp;
    fp0 = fpp;
    __syncthreads();
    Pi = nPi;
    psr = npsr;
    pst = npst;
    psp = npsp;
    qc = qn;
    if constexpr (KEEP) {
#pragma unroll
      for (int x = 0; x < 4; ++x) cr[x] = nr[x];
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) hr[x] = nhr[x];
    hdir = nhdir;
  }
}

// ---------------------------------------------------------------------------
// The fd4 operator set: the reference's 4th-order stencils
// (tools/finite_differences.f90: x_4 :124-151, y_4 :241-268, z_4 :340-367,
// advec_x_4 :922-957) in the same right-hand side, for the integrators that
// fill the ghosts after every stage (RK3, RK4).  The march is stage_body's
// with wider reaches: 3 halo columns each side in r (the upwinded A_r reaches
// i-3 and i+3), 2 halo rows each side in theta, and a five-plane phi queue
// (the own F and phi-flux at k-2..k+2) fed by loads two planes ahead.  The
// computed box (i in [1, ni-2], j in [g+1, g+Ny-1], k in [0, Nz-1]) uses the
// one-sided r rows 1 and N-1 (N = ni-1), A_r's rows 1, 2, N-2 and N-1,
// central theta rows only, and the one-sided phi rows 0, 1 and Nz-1, whose
// planes beyond the queue are read from global memory.
// ---------------------------------------------------------------------------
constexpr int HR4 = 3;  // r halo of the fd4 set
// theta rows per workgroup of the fd4 stages (12 rows for the ~160-VGPR
// stages measured no faster on the full grid)
constexpr int TJ4 = 8;

// The last r columns of a box that is not a whole number of 64-point tiles
// wide (2051 = 32 x 64 + 3 on the 2048-point grid) run as strips of SI4 x
// SJ4 points instead of 64 x 8 tiles with a few active lanes: one strip
// covers 128 theta rows, where the 64-wide tiles of those columns took 16
// workgroups.  On 2048 x 128 x 256 those tiles cost 5.8 % of an RK4 step;
// the strips take 8.59 -> 8.41 ms per step back (a strip's march is several
// times slower than a tile's, and the 4096 tiles being 16 rounds of 256,
// its work puts one tile-march more on the launch's last round).  A strip
// needs no right r halo (its columns reach ni-1), so it is used when its r
// range holds at most SI4 - 1 columns of the box.
constexpr int SI4 = 4, SJ4 = 128;

// EDGE: the tile's r range reaches a fixed row of x_4 (1, N-1) or of
// advec_x_4 (1, 2, N-2, N-1), whose stencils then test the row per point;
// interior tiles run the march without those tests (~2 % of an RK4 step on
// 2048 x 128 x 256).  TIL x TJ: the tile's points in r and theta (64 x 8,
// or a strip's SI4 x SJ4), on the same 512 threads.
template <int TIL, int TJ>
constexpr int lds4_doubles() {  // LDS of a tile: 4 r-arrays with halos, 2 theta-arrays with halos
  return 4 * TJ * (TIL + 2 * HR4) + 2 * (TJ + 4) * TIL;
}
constexpr int LDS4 = lds4_doubles<TI, TJ4>() > lds4_doubles<SI4, SJ4>() ? lds4_doubles<TI, TJ4>()
                                                                        : lds4_doubles<SI4, SJ4>();

template <int MODE, int EF, bool EDGE, int TIL = TI, int TJ = TJ4>
__device__ __forceinline__ void stage4_body(const StageArgs& a, const Tile tile, const int sysidx, double* lds) {
  static_assert(TIL * TJ == TI * TJ4, "a tile has one point per thread");
  static_assert(Combine<MODE>::boxed && ModeTraits<MODE>::n14 && ModeTraits<MODE>::n5,
                "the fd4 set runs the stages that are followed by a ghost fill");
  constexpr bool CBASE = Combine<MODE>::base, CAUX = Combine<MODE>::aux, VIRT = Combine<MODE>::virt;
  // RK4 stages 2 and 4 combine u_p and k at the point: the march holds both
  // as the raw operands of its stage state (planes kk and kk+1 kept raw)
  constexpr bool KEEP = Combine<MODE>::keep;
  // stages that read one stored array per field load the point's planes
  // three ahead (the loads issued for plane kk+3 have a whole plane to
  // arrive: RK4 stage 1 1.59 -> 1.40 ms on the 2048 x 128 x 256 grid); the
  // stages that read the raw pairs (u_p, k) of a stage state keep two, which
  // measured ~3 % faster for them than three
  constexpr bool PF3 = !VIRT && !KEEP;

  // the tile's LDS arrays, carved from one workgroup buffer that every tile
  // shape shares
  constexpr int W = TIL + 2 * HR4;
  static_assert(lds4_doubles<TIL, TJ>() <= LDS4, "LDS buffer too small");
  double(*const sG)[W] = reinterpret_cast<double(*)[W]>(lds);               // alpha*Pi
  double(*const sPR)[W] = reinterpret_cast<double(*)[W]>(lds + TJ * W);     // psi_r
  double(*const sFR)[W] = reinterpret_cast<double(*)[W]>(lds + 2 * TJ * W);  // r-flux of rhs_pi
  double(*const sPS)[W] = reinterpret_cast<double(*)[W]>(lds + 3 * TJ * W);  // Pi*sqrtg
  double(*const sF)[TIL] = reinterpret_cast<double(*)[TIL]>(lds + 4 * TJ * W);  // alpha*Pi + beta*psi_r
  double(*const sFT)[TIL] =
      reinterpret_cast<double(*)[TIL]>(lds + 4 * TJ * W + (TJ + 4) * TIL);  // theta-flux of rhs_pi

  const Geom& G = a.G;
  const int ti = tile.ti, tj = tile.tj, tk = tile.tk;
  const int tid = (int)threadIdx.y * TI + (int)threadIdx.x;
  // (a strip's waves each hold one r column: uniform stencil rows per wave)
  const int lane = TIL == TI ? tid % TIL : tid / TJ, row = TIL == TI ? tid / TIL : tid % TJ;
  // (a strip's ti is the count of whole r tiles: it starts where they end)
  const int i0 = a.i_lo + ti * TI, j0 = a.j_lo + tj * TJ;
  const int ii = i0 + lane, jj = j0 + row;
  const int ni = G.ni, nj = G.nj, Nz = G.Nz, Ny = G.Ny, g = G.g, ld = G.ld;
  const int NR = ni - 1;  // last r index of the stencils
  const uint32_t plane8 = (uint32_t)(G.plane * 8);
  const uint32_t n2b = (uint32_t)((int64_t)nj * ld * 8);
  const bool valid = ii < ni && jj < nj;
  const bool act = valid && ii < a.i_hi && jj < a.j_hi;
  const int k0 = a.k_lo + tk * a.kc;
  const int k1 = min(k0 + a.kc, a.k_hi);
  const int64_t sys = (int64_t)sysidx * G.sys_stride;
  const uint32_t sysb = (uint32_t)(sys * 8);
  const double* tab = a.tab;
  auto T = [&](int which, uint32_t off) { return gld(tab, off + (uint32_t)which * n2b); };
  auto FP = [&](int r, int f) { return a.rb[r] + (int64_t)(f == 4 ? a.ri4[r] : f) * a.per; };
  auto OFF = [&](int i, int j, int k) { return (uint32_t)((j * ld + i) * 8) + (uint32_t)k * plane8; };
  const int dis = Nz / 2;

  // stage input at a padded point, with the ghost fill resolved (as stage_body)
  auto E = [&](int f, uint32_t o) {
    if constexpr (VIRT) {
      const double p = gld(FP(R_IN, f), o + sysb), k = gld(FP(R_KIN, f), o + sysb);
      if constexpr (MODE == M_RK4_S4)
        return p + k;
      else
        return p + 0.50 * k;
    } else {
      return gld(FP(R_IN, f), o + sysb);
    }
  };
  struct Src {
    uint32_t o;
    int kind;  // 0: evolved value at o; 1: xL extrapolation from o, o+1, o+2; 2: u_p at o (inflow)
  };
  auto remap = [&](int i, int j, int k) -> Src {
    if constexpr (VIRT) {
      if (i < 1 || i > ni - 2 || j < g + 1 || j > g + Ny - 1 || k < 1 || k > Nz - 1) {
        if (k == 0)
          k = Nz - 1;
        else if (k == Nz)
          k = 1;
        if (j <= g) {
          j = 2 * g + 1 - j;
          k = (k + dis) % (Nz - 1);
        } else if (j >= g + Ny) {
          j = 2 * Ny + 2 * g - 1 - j;
          k = (k + dis) % (Nz - 1);
        }
        if (i == 0) return Src{OFF(1, j, k), 1};
        if (i == ni - 1 && !G.bc_somm) return G.bc_inflow ? Src{OFF(ni - 1, j, k), 2} : Src{OFF(ni - 2, j, k), 0};
        return Src{OFF(i, j, k), 0};
      }
    }
    return Src{OFF(i, j, k), 0};
  };
  auto Ur = [&](int f, const Src& q) -> double {
    if constexpr (VIRT) {
      if (q.kind == 1) return 3.0 * E(f, q.o) - 3.0 * E(f, q.o + 8) + E(f, q.o + 16);
      if (q.kind == 2) return gld(FP(R_IN, f), q.o + sysb);
    }
    return E(f, q.o);
  };
  auto U = [&](int f, int i, int j, int k) { return Ur(f, remap(i, j, k)); };
  struct Raw {
    double p, k;
  };
  const double opaque0 = G.dt * 0.0;
  auto fetch = [&](int f, const Src& q, Raw& r) {
    if constexpr (VIRT) {
      if (q.kind == 0) {
        r.p = gld(FP(R_IN, f), q.o + sysb);
        r.k = gld(FP(R_KIN, f), q.o + sysb);
      } else {
        r.p = Ur(f, q);
        r.k = opaque0;
      }
    } else {
      r.p = gld(FP(R_IN, f), q.o + sysb);
      r.k = opaque0;
    }
  };
  auto val = [&](const Raw& r, bool direct) -> double {
    if constexpr (VIRT) {
      const double v = (MODE == M_RK4_S4) ? r.p + r.k : r.p + 0.50 * r.k;
      return direct ? r.p : v;
    } else {
      return r.p;
    }
  };

  const uint32_t own = valid ? (uint32_t)((jj * ld + ii) * 8) : 0u;
  const double al = T(T_ALPHA, own), be = T(T_BETA, own);
  const double dxb = T(T_DXBETA, own), sg = T(T_SQRTG, own);
  const double asg = T(T_ASG, own), grr = T(T_GRR, own);
  const double grp = T(T_GRP, own), asgtt = T(T_ASGTT, own);
  const double gpp = T(T_GPP, own), am2 = T(T_AM2, own);
  auto Fof = [&](double Pi, double psr) { return al * Pi + be * psr; };
  auto fpof = [&](double psr, double psp) {
    if constexpr (EF)
      return (al * gpp) * psp;
    else
      return asg * (grp * psr + gpp * psp);
  };

  // halo task: 1 = theta row (j0-2, j0-1, j0+TJ, j0+TJ+1), 2 = r column
  // (i0-3..i0-1, i0+TIL..i0+TIL+2; a strip has the left three only), 0 = none
  constexpr int RH = TIL == TI ? 6 : 3;  // r-halo columns per row
  static_assert(4 * TIL + RH * TJ <= TI * TJ4, "one halo task per thread at most");
  int hkind = 0, hi = 0, hj = 0, hrow = 0, hcol = 0;
  if (tid < 4 * TIL) {
    hkind = 1;
    const int r4 = tid / TIL;
    hi = i0 + (tid % TIL);
    hrow = r4 < 2 ? r4 : TJ + r4;
    hj = j0 - 2 + hrow;
    hcol = tid % TIL;
  } else if (tid < 4 * TIL + RH * TJ) {
    const int u = tid - 4 * TIL;
    const int cc = u % RH;
    hkind = 2;
    hrow = u / RH;
    hj = j0 + hrow;
    hi = cc < 3 ? i0 - 3 + cc : i0 + TIL + (cc - 3);
    hcol = hi - i0 + HR4;
  }
  if (hi < 0 || hi >= ni || hj < 0 || hj >= nj) hkind = 0;
  const uint32_t hown = hkind ? (uint32_t)((hj * ld + hi) * 8) : 0u;
  double h0 = 0, h1 = 0, h2 = 0, h3 = 0, h4 = 0;
  if (hkind == 1) {
    h0 = T(T_ALPHA, hown);
    h1 = T(T_BETA, hown);
    h2 = T(T_ASGTT, hown);
  } else if (hkind == 2) {
    h0 = T(T_ALPHA, hown);
    h1 = T(T_ASG, hown);
    h2 = T(T_GRR, hown);
    h3 = T(T_GRP, hown);
    h4 = T(T_SQRTG, hown);
  }
  auto fetch_halo = [&](int k, Raw* x, bool& direct) {
    if (hkind == 0) return;
    const Src q = remap(hi, hj, k);
    direct = q.kind != 0;
    fetch(4, q, x[0]);
    fetch(1, q, x[1]);
    if (hkind == 1)
      fetch(2, q, x[2]);
    else if constexpr (!EF)
      fetch(3, q, x[2]);
  };

  // own values at planes kk and kk+1; F and the phi-flux at kk-2..kk+1
  double Pi = 0, psr = 0, pst = 0, psp = 0, Pi1 = 0, psr1 = 0, pst1 = 0, psp1 = 0;
  double Fm2 = 0, Fm1 = 0, F0 = 0, Fp1 = 0, fpm2 = 0, fpm1 = 0, fp0 = 0, fpp1 = 0;
  Src qc{0u, 0}, qc1{0u, 0};
  Raw hr[3] = {{0, 0}, {0, 0}, {0, 0}};
  bool hdir = false;
  Raw cr[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}}, cr1[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};  // KEEP
  Raw pr[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};  // PF3: plane kk+2, loaded one iteration earlier
  bool pdir = false;
  Src qc2{0u, 0};
  if (valid) {
    for (int d = 2; d >= 1; --d) {
      if (k0 - d < 0) continue;
      const Src q = remap(ii, jj, k0 - d);
      const double mPi = Ur(4, q), mpr = Ur(1, q), mpp = Ur(3, q);
      if (d == 2) {
        Fm2 = Fof(mPi, mpr);
        fpm2 = fpof(mpr, mpp);
      } else {
        Fm1 = Fof(mPi, mpr);
        fpm1 = fpof(mpr, mpp);
      }
    }
    qc = remap(ii, jj, k0);
    qc1 = remap(ii, jj, k0 + 1);  // k0 + 1 <= Nz
    if constexpr (KEEP) {
      const int fl[4] = {4, 1, 2, 3};
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        fetch(fl[x], qc, cr[x]);
        fetch(fl[x], qc1, cr1[x]);
      }
      const bool d0 = qc.kind != 0, d1 = qc1.kind != 0;
      Pi = val(cr[0], d0);
      psr = val(cr[1], d0);
      pst = val(cr[2], d0);
      psp = val(cr[3], d0);
      Pi1 = val(cr1[0], d1);
      psr1 = val(cr1[1], d1);
      pst1 = val(cr1[2], d1);
      psp1 = val(cr1[3], d1);
    } else {
      Pi = Ur(4, qc);
      psr = Ur(1, qc);
      pst = Ur(2, qc);
      psp = Ur(3, qc);
      Pi1 = Ur(4, qc1);
      psr1 = Ur(1, qc1);
      pst1 = Ur(2, qc1);
      psp1 = Ur(3, qc1);
    }
    F0 = Fof(Pi, psr);
    fp0 = fpof(psr, psp);
    Fp1 = Fof(Pi1, psr1);
    fpp1 = fpof(psr1, psp1);
    if constexpr (PF3) {
      if (k0 + 2 <= Nz) {
        qc2 = remap(ii, jj, k0 + 2);
        pdir = qc2.kind != 0;
        fetch(4, qc2, pr[0]);
        fetch(1, qc2, pr[1]);
        fetch(2, qc2, pr[2]);
        fetch(3, qc2, pr[3]);
      }
    }
  }
  fetch_halo(k0, hr, hdir);

  const double r4r = G.r4r, r4t = G.r4t, r4p = G.r4p, rar = G.rar, dt = G.dt;
  const int c = lane + HR4;

  for (int kk = k0; kk < k1; ++kk) {
    const uint32_t sk = (uint32_t)kk * plane8;
    const double hx0 = val(hr[0], hdir), hx1 = val(hr[1], hdir), hx2 = val(hr[2], hdir);
    if (valid) {
      sG[row][c] = al * Pi;
      sPR[row][c] = psr;
      sF[row + 2][lane] = F0;
      if constexpr (EF)
        sFR[row][c] = asg * grr * psr;
      else
        sFR[row][c] = asg * (grr * psr + grp * psp);
      sPS[row][c] = Pi * sg;
      sFT[row + 2][lane] = asgtt * pst;
    }
    if (hkind == 1) {
      sF[hrow][hcol] = h0 * hx0 + h1 * hx1;
      sFT[hrow][hcol] = h2 * hx2;
    } else if (hkind == 2) {
      sG[hrow][hcol] = h0 * hx0;
      sPR[hrow][hcol] = hx1;
      if constexpr (EF)
        sFR[hrow][hcol] = h1 * h2 * hx1;
      else
        sFR[hrow][hcol] = h1 * (h2 * hx1 + h3 * hx2);
      sPS[hrow][hcol] = hx0 * h4;
    }
    // plane kk+2 (kk+3 with PF3) of the own point, the halo's next plane and
    // phi at the point
    Raw nr[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    bool ndir = false;
    Src qn{0u, 0};
    constexpr int AHEAD = PF3 ? 3 : 2;
    // (with PF3, plane kk+3 serves the next iteration only)
    if (valid && kk + AHEAD <= Nz && (!PF3 || kk + 1 < k1)) {
      qn = remap(ii, jj, kk + AHEAD);
      ndir = qn.kind != 0;
      fetch(4, qn, nr[0]);
      fetch(1, qn, nr[1]);
      fetch(2, qn, nr[2]);
      fetch(3, qn, nr[3]);
    }
    Raw nhr[3] = {{0, 0}, {0, 0}, {0, 0}};
    bool nhdir = false;
    if (kk + 1 < k1) fetch_halo(kk + 1, nhr, nhdir);
    Raw vr{0, 0};
    if (act) fetch(0, qc, vr);
    const uint32_t so = sk + sysb + own;
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);

    double vb[5] = {0, 0, 0, 0, 0}, va[5] = {0, 0, 0, 0, 0}, vk[5] = {0, 0, 0, 0, 0};
    double R[5] = {0, 0, 0, 0, 0};
    double t1 = 0, t2 = 0, t5 = 0;
    // x_4 on an LDS r-array: rows 1 and N-1 one-sided, central between
    auto Dr4 = [&](const double(*A)[W]) {
      if constexpr (EDGE) {
        if (ii == NR - 1)
          return (+3.0 * A[row][c + 1] + 10.0 * A[row][c] - 18.0 * A[row][c - 1] + 6.0 * A[row][c - 2] -
                  1.0 * A[row][c - 3]) * r4r;
        if (ii == 1)
          return (-3.0 * A[row][c - 1] - 10.0 * A[row][c] + 18.0 * A[row][c + 1] - 6.0 * A[row][c + 2] +
                  1.0 * A[row][c + 3]) * r4r;
      }
      return (A[row][c - 2] - 8.0 * A[row][c - 1] + 8.0 * A[row][c + 1] - A[row][c + 2]) * r4r;
    };
    // advec_x_4 at the point: its fixed rows 1, 2, N-2, N-1, upwinded between
    auto Ar4 = [&](const double(*A)[W]) {
      auto plus = [&]() {  // beta > 0
        return (-3.0 * A[row][c - 1] - 10.0 * A[row][c] + 18.0 * A[row][c + 1] - 6.0 * A[row][c + 2] +
                A[row][c + 3]) * rar / 12.0;
      };
      auto minus = [&]() {  // beta < 0
        return (3.0 * A[row][c + 1] + 10.0 * A[row][c] - 18.0 * A[row][c - 1] + 6.0 * A[row][c - 2] -
                A[row][c - 3]) * rar / 12.0;
      };
      if constexpr (EDGE) {
        if (ii == NR - 1) return minus();
        if (ii == NR - 2 || ii == 2)
          return (A[row][c - 2] - 8.0 * A[row][c - 1] + 8.0 * A[row][c + 1] - A[row][c + 2]) * rar / 12.0;
        if (ii == 1) return plus();
      }
      if (be < 0.0) return minus();
      if (be > 0.0) return plus();
      return (A[row][c - 2] - 8.0 * A[row][c - 1] + 8.0 * A[row][c + 1] - A[row][c + 2]) * rar / 12.0;
    };
    // y_4 away from its edge rows (the computed rows are all central)
    auto Dth4 = [&](const double(*A)[TIL]) {
      return (A[row][lane] - 8.0 * A[row + 1][lane] + 8.0 * A[row + 3][lane] - A[row + 4][lane]) * r4t;
    };
    // z_4: one-sided rows 0, 1 and Nz-1; planes beyond the queue from `far`
    auto Dph4 = [&](double qm2, double qm1, double q0, double qp1, double qp2, auto far) {
      if (kk == Nz - 1)
        return (+3.0 * qp1 + 10.0 * q0 - 18.0 * qm1 + 6.0 * qm2 - 1.0 * far(Nz - 4)) * r4p;
      if (kk == 0) return (-25.0 * q0 + 48.0 * qp1 - 36.0 * qp2 + 16.0 * far(3) - 3.0 * far(4)) * r4p;
      if (kk == 1) return (-3.0 * qm1 - 10.0 * q0 + 18.0 * qp1 - 6.0 * qp2 + 1.0 * far(4)) * r4p;
      return (qm2 - 8.0 * qm1 + 8.0 * qp1 - qp2) * r4p;
    };
    if (act) {
      // KEEP: u_p and k at the point are the raw operands of the march
      // (plane 0 resolves to plane Nz-1, so it reads plane 0 itself)
      const bool kept = KEEP && kk != 0;
      if constexpr (KEEP) {
        if (kept) {
          vb[0] = vr.p;
          vk[0] = vr.k;
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            vb[x == 0 ? 4 : x] = cr[x].p;
            vk[x == 0 ? 4 : x] = cr[x].k;
          }
        } else {
#pragma unroll
          for (int f = 0; f < 5; ++f) {
            vb[f] = gld(FP(R_IN, f), so);
            vk[f] = gld(FP(R_KIN, f), so);
          }
        }
      } else if constexpr (CBASE) {
#pragma unroll
        for (int f = 0; f < 5; ++f) vb[f] = gld(FP(R_BASE, f), so);
      }
      if constexpr (CAUX) {
#pragma unroll
        for (int f = 0; f < 5; ++f) va[f] = gld(FP(R_AUX, f), so);
      }
      R[0] = F0;
      R[1] = Dr4(sG) + be * Ar4(sPR) + dxb * psr;
      R[2] = Dth4(sF);
      t1 = Dr4(sFR) / sg;
      t2 = Dth4(sFT) / sg;
      t5 = be * Ar4(sPS) / sg;
    }
    __builtin_amdgcn_sched_barrier(0);
    const Raw* const p2 = PF3 ? pr : nr;  // plane kk+2
    const bool d2 = PF3 ? pdir : ndir;
    const double nPi = val(p2[0], d2), npsr = val(p2[1], d2);
    const double npst = val(p2[2], d2), npsp = val(p2[3], d2);
    const double Fp2 = Fof(nPi, npsr), fpp2 = fpof(npsr, npsp);
    if (act) {
      const double vphi = val(vr, qc.kind != 0);
      R[3] = Dph4(Fm2, Fm1, F0, Fp1, Fp2, [&](int k) { return Fof(U(4, ii, jj, k), U(1, ii, jj, k)); });
      const double dp = Dph4(fpm2, fpm1, fp0, fpp1, fpp2, [&](int k) { return fpof(U(1, ii, jj, k), U(3, ii, jj, k)); });
      const double t3 = EF ? dp : dp / sg;
      const double t4 = Pi * dxb;
      const double t6 = am2 * vphi;
      R[4] = t1 + t2 + t3 + t4 + t5 - t6;
      const double cur[5] = {vphi, psr, pst, psp, Pi};
      if constexpr (MODE == M_RK4_S1 || MODE == M_RK4_S3) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, R[f] * dt);
      } else if constexpr (MODE == M_RK4_S2) {
        // k2 stored as the stage-3 state u_p + 0.5*k2 and as k1 + 2*k2
#pragma unroll
        for (int f = 0; f < 5; ++f) {
          const double k2 = R[f] * dt;
          gst(FP(R_OUT, f), so, vb[f] + 0.50 * k2);
          gst(FP(R_AUX, f), so, vk[f] + 2.0 * k2);
        }
      } else if constexpr (MODE == M_RK4_S4) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] + (va[f] + 2.0 * vk[f] + R[f] * dt) / 6.0);
      } else if constexpr (MODE == M_RK3_S1) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] + R[f] * dt);
      } else if constexpr (MODE == M_RK3_S2) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, 0.750 * vb[f] + 0.250 * (cur[f] + R[f] * dt));
      } else if constexpr (MODE == M_RK3_S3) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] / 3.0 + 2.0 * (cur[f] + R[f] * dt) / 3.0);
      }
    }
    Fm2 = Fm1;
    Fm1 = F0;
    F0 = Fp1;
    Fp1 = Fp2;
    fpm2 = fpm1;
    fpm1 = fp0;
    fp0 = fpp1;
    fpp1 = fpp2;
    __syncthreads();
    Pi = Pi1;
    psr = psr1;
    pst = pst1;
    psp = psp1;
    Pi1 = nPi;
    psr1 = npsr;
    pst1 = npst;
    psp1 = npsp;
    qc = qc1;
    if constexpr (PF3) {
      qc1 = qc2;
      qc2 = qn;
#pragma unroll
      for (int x = 0; x < 4; ++x) pr[x] = nr[x];
      pdir = ndir;
    } else {
      qc1 = qn;
    }
    if constexpr (KEEP) {
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        cr[x] = cr1[x];
        cr1[x] = nr[x];
      }
    }
#pragma unroll
    for (int x = 0; x < 3; ++x) hr[x] = nhr[x];
    hdir = nhdir;
  }
}

// a workgroup-uniform choice between a strip and the edge and interior
// marches of the 64 x 8 tiles
template <int MODE, int EF>
__device__ __forceinline__ void stage4_run(const StageArgs& a, const Tile tile, const int sysidx) {
  __shared__ double lds[LDS4];
  if (tile.ti == a.nti) {
    stage4_body<MODE, EF, true, SI4, SJ4>(a, tile, sysidx, lds);
    return;
  }
  const int i0 = a.i_lo + tile.ti * TI, NR = a.G.ni - 1;
  if (i0 <= 2 || i0 + TI - 1 >= NR - 2)
    stage4_body<MODE, EF, true>(a, tile, sysidx, lds);
  else
    stage4_body<MODE, EF, false>(a, tile, sysidx, lds);
}

// One grid: workgroups [0, nts*ntk) run the strips (theta strip fastest;
// first, so their longer marches overlap the tiles'), the nti*ntj*ntk after
// them the 64 x 8 tiles, dealt over the XCDs.
template <int MODE, int EF>
__global__ __launch_bounds__(TI* TJ4) void stage4_kernel(const StageArgs a) {
  const int b = (int)blockIdx.x, ns = a.nts * a.ntk;
  const Tile t = b < ns ? Tile{a.nti, b % a.nts, b / a.nts} : tile_dealt(b - ns, a.nti, a.ntj, a.ntk, P4);
  stage4_run<MODE, EF>(a, t, (int)blockIdx.y);
}

// Many fd4 grids in one launch (as stage_kernel_batch).
template <int MODE, int EF>
__global__ __launch_bounds__(TI* TJ4) void stage4_kernel_batch(const StageArgs* __restrict__ descs,
                                                                const int* __restrict__ tstart, int ngrids, int total) {
  for (int w = (int)blockIdx.x; w < total; w += (int)gridDim.x) {  // as stage_kernel_batch
    const int b = remap_xcd(w, total);
    int lo = 0, hi = ngrids - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tstart[mid] <= b)
        lo = mid;
      else
        hi = mid - 1;
    }
    const StageArgs& a = descs[lo];
    const int local = b - tstart[lo];
    const int nt = a.nti * a.ntj * a.ntk;
    stage4_run<MODE, EF>(a, tile_in_run(local % nt, a.nti, a.ntj, a.ntk, P4), local / nt);
  }
}

// One grid: blockIdx.x = strip or tile, blockIdx.y = system.  The second
// launch bound is the minimum number of waves per SIMD.  RK4 stage 3 asks for
// 4 (128 VGPRs), stage 2 for 3 (~157-164 VGPRs with u_p and k1 kept for its
// stores) and the final stage for 2 (its natural count: 168 VGPRs now, which
// fits 3; at ~184 a bound of 3 or 4 spilled); stages 2 and 4 always run 4
// theta rows per workgroup, so several
// workgroups with their own barriers share a CU (stage 4: 3.25 -> 3.14 ms
// against one 8-row workgroup on the 2048 x 128 x 256 grid).
template <int MODE, int EF, int TJ>
__global__ __launch_bounds__(TI* TJ, (MODE == M_RK4_S4 ? 2 : MODE == M_RK4_S2 ? 3 : MODE == M_RK4_S3 ? 4 : 1)) void stage_kernel(const StageArgs a) {
  // workgroups [0, nts*ntk): the strips of the last r columns (RK stages);
  // then the tiles, dealt over the XCDs
  const int b = (int)blockIdx.x, ns = a.nts * a.ntk;
  if constexpr (strips2(MODE)) {
    __shared__ double lds[lds2_of<TJ>()];
    if (b < ns)
      stage_body<MODE, EF, TJ, SI2, TJ * TI / SI2>(a, Tile{a.nti, b % a.nts, b / a.nts}, (int)blockIdx.y, lds);
    else
      stage_body<MODE, EF, TJ>(a, tile_dealt(b - ns, a.nti, a.ntj, a.ntk, P2), (int)blockIdx.y, lds);
  } else {
    stage_body<MODE, EF, TJ>(a, tile_dealt(b, a.nti, a.ntj, a.ntk, P2), (int)blockIdx.y, nullptr);
  }
}

// Many grids in one launch: grid q owns the workgroups [tstart[q], tstart[q+1])
// (its tiles times its systems); the lookup is a uniform binary search.
__device__ __forceinline__ int find_grid(const int* __restrict__ start, int n, int b) {
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

// Launched with one workgroup per tile, or (persistent) with fewer: workgroup
// w then marches tiles w, w + gridDim.x, ... in turn (every march ends on a
// barrier, so the next tile's LDS writes follow the last reads).  gridDim.x is
// a multiple of 8, so each workgroup stays within its XCD's run of tiles.
template <int MODE, int EF, int TJ>
__global__ __launch_bounds__(TI* TJ, (MODE == M_RK4_S4 ? 2 : MODE == M_RK4_S2 ? 3 : MODE == M_RK4_S3 ? 4 : 1)) void stage_kernel_batch(
    const StageArgs* __restrict__ descs, const int* __restrict__ tstart, int ngrids, int total) {
  for (int w = (int)blockIdx.x; w < total; w += (int)gridDim.x) {
    const int b = remap_xcd(w, total);
    const int q = find_grid(tstart, ngrids, b);
    const StageArgs& a = descs[q];
    const int local = b - tstart[q];
    const int nt = a.nti * a.ntj * a.ntk;
    if constexpr (strips2(MODE)) {
      __shared__ double lds[lds2_doubles<TI, TJ>()];
      stage_body<MODE, EF, TJ>(a, tile_in_run(local % nt, a.nti, a.ntj, a.ntk, P2), local / nt, lds);
    } else {
      stage_body<MODE, EF, TJ>(a, tile_in_run(local % nt, a.nti, a.ntj, a.ntk, P2), local / nt, nullptr);
    }
  }
}

// ---------------------------------------------------------------------------
// Ghost fill (lines_method_lib.f90:313-330,507-597), three ordered passes.
// Each pass only reads points it does not write.
// ---------------------------------------------------------------------------
// Work items of each fill pass of one grid (all fields and systems).
__host__ __device__ inline int64_t fill_items(const Geom& G, int nsys, int phase) {
  if (phase == 0) return (int64_t)G.nj * G.nk * 5 * nsys;
  if (phase == 1) return (int64_t)((G.ni + 1) / 2) * 2 * (G.g + 1) * G.nk * 5 * nsys;
  return G.plane / 2 * 2 * 5 * nsys;
}

// A work item's index is split with the index type I: uint32_t whenever the
// pass has fewer than 2^32 items (every grid whose fields are below 2 GiB),
// int64_t otherwise; 64-bit division by a runtime divisor costs several times
// the 32-bit one, and made the theta pass bound by its index arithmetic.
template <int PHASE, typename I>
__device__ __forceinline__ void fill_item(const FillArgs& a, I e) {
  const Geom& G = a.G;
  if constexpr (PHASE == 0) {
    // boundary_xL (:516-517) and boundary_xR (:525-528)
    const int jj = (int)(e % (I)G.nj);
    I r = e / (I)G.nj;
    const int kk = (int)(r % (I)G.nk);
    r /= (I)G.nk;
    const int c = (int)(r % 5);
    const int64_t sys = (int64_t)(r / 5) * G.sys_stride;
    gdouble* u = gptr(a.u[c]) + sys + (int64_t)kk * G.plane + (int64_t)jj * G.ld;
    u[0] = 3.0 * u[1] - 3.0 * u[2] + u[3];
    if (G.bc_somm) {
      // Sommerfeld: the outer column is evolved, not filled
    } else if (G.bc_inflow) {
      gcdouble* up = gptr(a.u_p[c]) + sys + (int64_t)kk * G.plane + (int64_t)jj * G.ld;
      u[G.ni - 1] = up[G.ni - 1];
    } else {
      u[G.ni - 1] = u[G.ni - 2];
    }
  } else if constexpr (PHASE == 1) {
    // boundary_yL (:542-555) and boundary_yR (:565-578): whole r-lines,
    // rows -gg <- gg+1 and Ny+gg <- Ny-gg-1, taken at k_target = mod(k+Nz/2, Nz-1);
    // an item copies two neighbouring points of a line (rows start 128-byte
    // aligned: ld, the plane and the field stride are multiples of 16
    // doubles), the last odd point alone (the pass 69 -> 48 us at 2048 x 128
    // x 256)
    const int nrow = G.g + 1;
    const int hp = (G.ni + 1) / 2;
    const int ii = 2 * (int)(e % (I)hp);
    I r = e / (I)hp;
    const int row = (int)(r % (I)(2 * nrow));
    r /= (I)(2 * nrow);
    const int kk = (int)(r % (I)G.nk);
    r /= (I)G.nk;
    const int c = (int)(r % 5);
    const int64_t sys = (int64_t)(r / 5) * G.sys_stride;
    const int kt = (kk + G.Nz / 2) % (G.Nz - 1);
    int jd, js;
    if (row < nrow) {
      jd = G.g - row;          // j = -gg
      js = G.g + row + 1;      // j = gg+1
    } else {
      const int gg = row - nrow;
      jd = G.g + G.Ny + gg;       // j = Ny+gg
      js = G.g + G.Ny - gg - 1;   // j = Ny-gg-1
    }
    gdouble* u = gptr(a.u[c]) + sys;
    const int64_t dst = (int64_t)kk * G.plane + (int64_t)jd * G.ld + ii;
    const int64_t src = (int64_t)kt * G.plane + (int64_t)js * G.ld + ii;
    if (ii + 1 < G.ni)
      *(gdouble2*)(u + dst) = *(gdouble2*)(u + src);
    else
      u[dst] = u[src];
  } else {
    // boundary_zL (:587) and boundary_zR (:596); an item copies two
    // neighbouring points (the plane is a multiple of 16 doubles)
    const int64_t hp = G.plane / 2;
    const int64_t t = 2 * (int64_t)(e % (I)hp);
    I r = e / (I)hp;
    const int side = (int)(r % 2);
    r /= 2;
    const int c = (int)(r % 5);
    const int64_t sys = (int64_t)(r / 5) * G.sys_stride;
    gdouble* u = gptr(a.u[c]) + sys;
    if (side == 0)
      *(gdouble2*)(u + t) = *(gdouble2*)(u + (int64_t)(G.Nz - 1) * G.plane + t);
    else
      *(gdouble2*)(u + (int64_t)G.Nz * G.plane + t) = *(gdouble2*)(u + G.plane + t);
  }
}

// The three passes run in order (xL/xR, yL/yR, zL/zR; lines_method_lib.f90:313-330);
// each reads only points it does not write.
template <int PHASE, typename I>
__global__ __launch_bounds__(256) void fill_kernel(const FillArgs a, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) fill_item<PHASE, I>(a, (I)e);
}

template <int PHASE, typename I>
__global__ __launch_bounds__(256) void fill_kernel_batch(const FillArgs* __restrict__ descs,
                                                         const int* __restrict__ bstart, int ngrids,
                                                         const int64_t* __restrict__ nitems) {
  const int q = find_grid(bstart, ngrids, (int)blockIdx.x);
  const int64_t e = (int64_t)(blockIdx.x - bstart[q]) * 256 + threadIdx.x;
  if (e < nitems[q]) fill_item<PHASE, I>(descs[q], (I)e);
}

// ---------------------------------------------------------------------------
// Sommerfeld outer boundary (KG_BC_SOMMERFELD; the reference has 'outflow'
// and 'inflow' only, lines_method_lib.f90:522-535).  The outer column
// N = ni-1 is evolved instead of filled: phi and the psi fields with their own
// right-hand sides (lines_method_lib.f90:671-713) using the operator set's
// one-sided r rows (finite_differences.f90:36-42,779-799 / :124-151,922-957),
// Pi with the radiative condition
//     R_Pi = -(v * (D_r Pi + Pi / r)),   v = alpha*sqrt(gamma^rr) - beta^r,
// D_r the same one-sided edge row, r = x(N).  One work item per (system,
// plane, theta row) of the stage's box at column N; the stage combine is the
// stage kernel's.  PIRK stages 1A and 2A (phi and psi over the full box) are
// left to their stage kernel, which computes column N with the same edge rows.
// ---------------------------------------------------------------------------
template <int MODE, int FD4>
__device__ __forceinline__ void somm_item(const StageArgs& a, int64_t e) {
  constexpr bool N14 = ModeTraits<MODE>::n14;
  constexpr bool N5 = ModeTraits<MODE>::n5;
  constexpr int CLO = Combine<MODE>::lo, CHI = Combine<MODE>::hi;
  constexpr bool CBASE = Combine<MODE>::base, CAUX = Combine<MODE>::aux, VIRT = Combine<MODE>::virt;
  static_assert(!N14 || Combine<MODE>::boxed, "phi and psi of the outer column are computed for boxed stages only");
  const Geom& G = a.G;
  const int nbj = a.j_hi - a.j_lo, nbk = a.k_hi - a.k_lo;
  const int jj = a.j_lo + (int)(e % nbj);
  const int64_t rest = e / nbj;
  const int kk = a.k_lo + (int)(rest % nbk);
  const int sysidx = (int)(rest / nbk);
  const int ni = G.ni, nj = G.nj, Nz = G.Nz, Ny = G.Ny, g = G.g, ld = G.ld;
  const int N = ni - 1;
  const uint32_t plane8 = (uint32_t)(G.plane * 8);
  const uint32_t n2b = (uint32_t)((int64_t)nj * ld * 8);
  const uint32_t sysb = (uint32_t)((int64_t)sysidx * G.sys_stride * 8);
  const int dis = Nz / 2;
  auto T = [&](int which, int i, int j) { return gld(a.tab, (uint32_t)((j * ld + i) * 8) + (uint32_t)which * n2b); };
  auto FP = [&](int r, int f) { return a.rb[r] + (int64_t)(f == 4 ? a.ri4[r] : f) * a.per; };
  auto OFF = [&](int i, int j, int k) { return (uint32_t)((j * ld + i) * 8) + (uint32_t)k * plane8; };
  auto E = [&](int f, uint32_t o) {
    if constexpr (VIRT) {
      const double p = gld(FP(R_IN, f), o + sysb), k = gld(FP(R_KIN, f), o + sysb);
      if constexpr (MODE == M_RK4_S4)
        return p + k;
      else
        return p + 0.50 * k;
    } else {
      return gld(FP(R_IN, f), o + sysb);
    }
  };
  // stage input at (i, j, k) with the ghost fill resolved as the stage kernels
  // resolve it (no xR fill: the outer column holds its evolved value)
  auto U = [&](int f, int i, int j, int k) -> double {
    if constexpr (VIRT) {
      if (i < 1 || j < g + 1 || j > g + Ny - 1 || k < 1 || k > Nz - 1) {
        if (k == 0)
          k = Nz - 1;
        else if (k == Nz)
          k = 1;
        if (j <= g) {
          j = 2 * g + 1 - j;
          k = (k + dis) % (Nz - 1);
        } else if (j >= g + Ny) {
          j = 2 * Ny + 2 * g - 1 - j;
          k = (k + dis) % (Nz - 1);
        }
        if (i == 0) {
          const uint32_t o = OFF(1, j, k);
          return 3.0 * E(f, o) - 3.0 * E(f, o + 8) + E(f, o + 16);
        }
      }
    }
    return E(f, OFF(i, j, k));
  };
  const double rr = G.rr, rt = G.rt, rp = G.rp, dt = G.dt;
  const double r4r = G.r4r, r4t = G.r4t, r4p = G.r4p, rar = G.rar;
  double R[5] = {0, 0, 0, 0, 0};
  double cur[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int f = CLO; f < CHI; ++f) cur[f] = U(f, N, jj, kk);
  if constexpr (N14) {
    const double al = T(T_ALPHA, N, jj), be = T(T_BETA, N, jj), dxb = T(T_DXBETA, N, jj);
    auto F = [&](int j, int k) { return T(T_ALPHA, N, j) * U(4, N, j, k) + T(T_BETA, N, j) * U(1, N, j, k); };
    double A[5], P[5];
    constexpr int M = FD4 ? 5 : 3;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      A[m] = T(T_ALPHA, N - m, jj) * U(4, N - m, jj, kk);
      P[m] = U(1, N - m, jj, kk);
    }
    double dr, ar;
    if constexpr (FD4) {
      dr = (+25.0 * A[0] - 48.0 * A[1] + 36.0 * A[2] - 16.0 * A[3] + 3.0 * A[4]) * r4r;
      ar = (25.0 * P[0] - 48.0 * P[1] + 36.0 * P[2] - 16.0 * P[3] + 3.0 * P[4]) * rar / 12.0;
    } else {
      dr = (+3.0 * A[0] - 4.0 * A[1] + A[2]) * rr;
      ar = (3.0 * P[0] - 4.0 * P[1] + P[2]) * rr;
    }
    R[0] = al * cur[4] + be * cur[1];
    R[1] = dr + be * ar + dxb * cur[1];
    if constexpr (FD4) {
      R[2] = (F(jj - 2, kk) - 8.0 * F(jj - 1, kk) + 8.0 * F(jj + 1, kk) - F(jj + 2, kk)) * r4t;
      if (kk == 0)
        R[3] = (-25.0 * F(jj, 0) + 48.0 * F(jj, 1) - 36.0 * F(jj, 2) + 16.0 * F(jj, 3) - 3.0 * F(jj, 4)) * r4p;
      else if (kk == 1)
        R[3] = (-3.0 * F(jj, 0) - 10.0 * F(jj, 1) + 18.0 * F(jj, 2) - 6.0 * F(jj, 3) + 1.0 * F(jj, 4)) * r4p;
      else if (kk == Nz - 1)
        R[3] = (+3.0 * F(jj, Nz) + 10.0 * F(jj, Nz - 1) - 18.0 * F(jj, Nz - 2) + 6.0 * F(jj, Nz - 3) -
                1.0 * F(jj, Nz - 4)) * r4p;
      else
        R[3] = (F(jj, kk - 2) - 8.0 * F(jj, kk - 1) + 8.0 * F(jj, kk + 1) - F(jj, kk + 2)) * r4p;
    } else {
      R[2] = (-F(jj - 1, kk) + F(jj + 1, kk)) * rt;
      R[3] = (-F(jj, kk == 0 ? Nz - 2 : kk - 1) + F(jj, kk + 1)) * rp;
    }
  }
  if constexpr (N5) {
    double P[5];
    constexpr int M = FD4 ? 5 : 3;
#pragma unroll
    for (int m = 0; m < M; ++m) P[m] = U(4, N - m, jj, kk);
    double d;
    if constexpr (FD4)
      d = (+25.0 * P[0] - 48.0 * P[1] + 36.0 * P[2] - 16.0 * P[3] + 3.0 * P[4]) * r4r;
    else
      d = (+3.0 * P[0] - 4.0 * P[1] + P[2]) * rr;
    const double v = T(T_VOUT, N, jj);
    R[4] = -(v * (d + P[0] / G.r_out));
  }
  const uint32_t so = OFF(N, jj, kk) + sysb;
  double vb[5] = {0, 0, 0, 0, 0}, va[5] = {0, 0, 0, 0, 0}, vk[5] = {0, 0, 0, 0, 0};
  if constexpr (Combine<MODE>::keep) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      vb[f] = gld(FP(R_IN, f), so);
      vk[f] = gld(FP(R_KIN, f), so);
    }
  } else if constexpr (CBASE) {
#pragma unroll
    for (int f = CLO; f < CHI; ++f) vb[f] = gld(FP(R_BASE, f), so);
  }
  if constexpr (CAUX) {
#pragma unroll
    for (int f = CLO; f < CHI; ++f) va[f] = gld(FP(R_AUX, f), so);
  }
  if constexpr (MODE == M_RK4_S1 || MODE == M_RK4_S3) {
#pragma unroll
    for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, R[f] * dt);
  } else if constexpr (MODE == M_RK4_S2) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      const double k2 = R[f] * dt;
      gst(FP(R_OUT, f), so, vb[f] + 0.50 * k2);
      gst(FP(R_AUX, f), so, vk[f] + 2.0 * k2);
    }
  } else if constexpr (MODE == M_RK4_S4) {
#pragma unroll
    for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] + (va[f] + 2.0 * vk[f] + R[f] * dt) / 6.0);
  } else if constexpr (MODE == M_RK3_S1) {
#pragma unroll
    for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] + R[f] * dt);
  } else if constexpr (MODE == M_RK3_S2) {
#pragma unroll
    for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, 0.750 * vb[f] + 0.250 * (cur[f] + R[f] * dt));
  } else if constexpr (MODE == M_RK3_S3) {
#pragma unroll
    for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, vb[f] / 3.0 + 2.0 * (cur[f] + R[f] * dt) / 3.0);
  } else if constexpr (MODE == M_PIRK_1B) {
    gst(FP(R_OUT, 4), so, vb[4] + dt * R[4]);
  } else if constexpr (MODE == M_PIRK2_2) {
#pragma unroll
    for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, 0.5 * (vb[f] + cur[f] + dt * R[f]));
  } else if constexpr (MODE == M_PIRK3_2B) {
    gst(FP(R_OUT, 4), so, 0.75 * vb[4] + 0.25 * (va[4] + dt * R[4]));
  } else if constexpr (MODE == M_PIRK3_3) {
    const double third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
#pragma unroll
    for (int f = 0; f < 5; ++f) gst(FP(R_OUT, f), so, third * vb[f] + two_thirds * (cur[f] + dt * R[f]));
  }
}

// one wavefront per workgroup: a column has few items (nj x nk per system),
// and many small workgroups spread them over the CUs
constexpr int SOMM_BLOCK = 64;

__host__ __device__ inline int64_t somm_count(const StageArgs& a) {
  return (int64_t)a.nsys * (a.j_hi - a.j_lo) * (a.k_hi - a.k_lo);
}

template <int MODE, int FD4>
__global__ __launch_bounds__(SOMM_BLOCK) void somm_kernel(const StageArgs a, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * SOMM_BLOCK + threadIdx.x;
  if (e < n) somm_item<MODE, FD4>(a, e);
}

template <int MODE, int FD4>
__global__ __launch_bounds__(SOMM_BLOCK) void somm_kernel_batch(const StageArgs* __restrict__ descs,
                                                         const int* __restrict__ bstart, int ngrids) {
  const int q = find_grid(bstart, ngrids, (int)blockIdx.x);
  const StageArgs& a = descs[q];
  const int64_t e = (int64_t)(blockIdx.x - bstart[q]) * SOMM_BLOCK + threadIdx.x;
  if (e < somm_count(a)) somm_item<MODE, FD4>(a, e);
}

template <int MODE, int FD4>
hipError_t launch_somm_mode(const StageArgs& a, hipStream_t s) {
  const int64_t n = somm_count(a);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((somm_kernel<MODE, FD4>), dim3((unsigned)((n + SOMM_BLOCK - 1) / SOMM_BLOCK)), dim3(SOMM_BLOCK), 0, s, a, n);
  return hipGetLastError();
}

template <int MODE, int FD4>
hipError_t launch_somm_batch_mode(const StageArgs* descs, const int* bstart, int ngrids, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((somm_kernel_batch<MODE, FD4>), dim3((unsigned)nblocks), dim3(SOMM_BLOCK), 0, s, descs, bstart, ngrids);
  return hipGetLastError();
}

// dense [sys][5][nk][nj][ni] <-> pitched fields
// fortran = 1: src is the reference's u(5, i, j, k) (complex pairs when nsys =
// 2: variable, then part, fastest) instead of the dense SoA
__global__ __launch_bounds__(256) void pack_kernel(const Geom G, double* __restrict__ src, double* d0,
                                                   double* d1, double* d2, double* d3, double* d4, int to_dev,
                                                   int fortran, int nsys) {
  const int nbx = (G.ni + 63) / 64;
  const int ii = (blockIdx.x % nbx) * 64 + threadIdx.x;
  const int jj = blockIdx.x / nbx;
  const int kk = blockIdx.y;
  if (ii >= G.ni) return;
  const int c = blockIdx.z % 5;
  const int s = blockIdx.z / 5;
  gdouble* fld = gptr(c == 0 ? d0 : c == 1 ? d1 : c == 2 ? d2 : c == 3 ? d3 : d4);
  gdouble* dns = gptr(src);
  const int64_t p = ((int64_t)kk * G.nj + jj) * G.ni + ii;
  const int64_t dense =
      fortran ? (p * 5 + c) * nsys + s : ((((int64_t)s * 5 + c) * G.nk + kk) * G.nj + jj) * G.ni + ii;
  const int64_t pitched = (int64_t)s * G.sys_stride + (int64_t)kk * G.plane + (int64_t)jj * G.ld + ii;
  if (to_dev)
    fld[pitched] = dns[dense];
  else
    dns[dense] = fld[pitched];
}

template <int MODE>
hipError_t launch_fd4(const StageArgs& a0, hipStream_t s, int nsys) {
  StageArgs a = a0;
  a.tj = TJ4;
  finalize_tiles(a);
  // the last r tile as strips when its columns reach ni-1 within SI4 points
  // (the strip then needs no right halo)
  const int i_last = a.i_lo + (a.nti - 1) * TI;
  if (a.nti > 1 && a.G.ni - i_last <= SI4) {
    a.nti -= 1;
    a.nts = (a.j_hi - a.j_lo + SJ4 - 1) / SJ4;
  }
  dim3 grid((unsigned)((a.nti * a.ntj + a.nts) * a.ntk), (unsigned)nsys, 1);
  dim3 block(TI, TJ4, 1);
  if (a.G.metric_ef)
    hipLaunchKernelGGL((stage4_kernel<MODE, 1>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((stage4_kernel<MODE, 0>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <int MODE, int TJ>
hipError_t launch_mode_tj(const StageArgs& a0, hipStream_t s, int nsys) {
  StageArgs a = a0;
  // the last r tile as strips when its columns reach ni-1 within SI2 points
  // (the strip then needs no right halo)
  const int i_last = a.i_lo + (a.nti - 1) * TI;
  if (strips2(MODE) && a.nti > 1 && a.G.ni - i_last <= SI2) {
    constexpr int TJS = TJ * TI / SI2;
    a.nti -= 1;
    a.nts = (a.j_hi - a.j_lo + TJS - 1) / TJS;
  }
  dim3 grid((unsigned)((a.nti * a.ntj + a.nts) * a.ntk), (unsigned)nsys, 1);
  dim3 block(TI, TJ, 1);
  if (a.G.metric_ef)
    hipLaunchKernelGGL((stage_kernel<MODE, 1, TJ>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((stage_kernel<MODE, 0, TJ>), grid, block, 0, s, a);
  return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(const StageArgs& a, hipStream_t s, int nsys) {
  return a.tj == 8 ? launch_mode_tj<MODE, 8>(a, s, nsys) : launch_mode_tj<MODE, 4>(a, s, nsys);
}

// workgroups of a batched stage launch: one per tile, or nwg (persistent) when fewer
unsigned batch_grid(int total, int nwg) { return (unsigned)(nwg > 0 && nwg < total ? nwg : total); }

template <int MODE, int TJ>
hipError_t launch_batch_tj(const StageArgs* descs, const int* tstart, int ngrids, int total, int nwg, int ef,
                           hipStream_t s) {
  dim3 grid(batch_grid(total, nwg), 1, 1);
  dim3 block(TI, TJ, 1);
  if (ef)
    hipLaunchKernelGGL((stage_kernel_batch<MODE, 1, TJ>), grid, block, 0, s, descs, tstart, ngrids, total);
  else
    hipLaunchKernelGGL((stage_kernel_batch<MODE, 0, TJ>), grid, block, 0, s, descs, tstart, ngrids, total);
  return hipGetLastError();
}

}  // namespace

// The PIRK stages of Pi alone (1B, 3's 2B) also run 4 rows: on the 2048 x
// 128 x 256 grid PIRK2 stage 1B takes 0.807 ms at 4 rows against 0.829 at 8,
// while stages 1A and 2 are faster at 8 (0.977 / 1.594 ms at kc = 4 against
// 1.023 / 1.632 at 4 rows).
int stage_tj(int mode, int tj) {
  return (mode == M_RK4_S4 || mode == M_RK4_S2 || mode == M_PIRK_1B || mode == M_PIRK3_2B) ? 4 : tj;
}

void finalize_tiles(StageArgs& a) {
  if (a.tj != 4 && a.tj != 8) a.tj = 4;
  if (a.kc < 1) a.kc = 1;
  a.nti = (a.i_hi - a.i_lo + TI - 1) / TI;
  a.ntj = (a.j_hi - a.j_lo + a.tj - 1) / a.tj;
  a.nts = 0;
  a.ntk = (a.k_hi - a.k_lo + a.kc - 1) / a.kc;
  // the same number of chunks, balanced: a full box of 257 planes at kc = 32
  // marches as 9 chunks of 29 instead of 8 of 32 and one of a single plane
  a.kc = (a.k_hi - a.k_lo + a.ntk - 1) / a.ntk;
  a.ntk = (a.k_hi - a.k_lo + a.kc - 1) / a.kc;
}

template <int MODE>
hipError_t launch_batch_mode(const StageArgs* descs, const int* tstart, int ngrids, int total, int nwg, int tj, int ef,
                             hipStream_t s) {
  return tj == 8 ? launch_batch_tj<MODE, 8>(descs, tstart, ngrids, total, nwg, ef, s)
                 : launch_batch_tj<MODE, 4>(descs, tstart, ngrids, total, nwg, ef, s);
}

template <int MODE>
hipError_t launch_batch4(const StageArgs* descs, const int* tstart, int ngrids, int total, int nwg, int ef,
                         hipStream_t s) {
  dim3 grid(batch_grid(total, nwg), 1, 1);
  dim3 block(TI, TJ4, 1);
  if (ef)
    hipLaunchKernelGGL((stage4_kernel_batch<MODE, 1>), grid, block, 0, s, descs, tstart, ngrids, total);
  else
    hipLaunchKernelGGL((stage4_kernel_batch<MODE, 0>), grid, block, 0, s, descs, tstart, ngrids, total);
  return hipGetLastError();
}

hipError_t launch_stage_batch(int mode, const StageArgs* descs, const int* tstart, int ngrids, int total, int nwg,
                              int tj, int ef, int fd4, hipStream_t s) {
  if (fd4) {
    switch (mode) {
      case M_RK4_S1: return launch_batch4<M_RK4_S1>(descs, tstart, ngrids, total, nwg, ef, s);
      case M_RK4_S2: return launch_batch4<M_RK4_S2>(descs, tstart, ngrids, total, nwg, ef, s);
      case M_RK4_S3: return launch_batch4<M_RK4_S3>(descs, tstart, ngrids, total, nwg, ef, s);
      case M_RK4_S4: return launch_batch4<M_RK4_S4>(descs, tstart, ngrids, total, nwg, ef, s);
      case M_RK3_S1: return launch_batch4<M_RK3_S1>(descs, tstart, ngrids, total, nwg, ef, s);
      case M_RK3_S2: return launch_batch4<M_RK3_S2>(descs, tstart, ngrids, total, nwg, ef, s);
      case M_RK3_S3: return launch_batch4<M_RK3_S3>(descs, tstart, ngrids, total, nwg, ef, s);
      default: return hipErrorInvalidValue;  // the PIRK stages are refused at the API
    }
  }
  switch (mode) {
    case M_RK4_S1: return launch_batch_mode<M_RK4_S1>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_RK4_S2: return launch_batch_mode<M_RK4_S2>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_RK4_S3: return launch_batch_mode<M_RK4_S3>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_RK4_S4: return launch_batch_mode<M_RK4_S4>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_RK3_S1: return launch_batch_mode<M_RK3_S1>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_RK3_S2: return launch_batch_mode<M_RK3_S2>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_RK3_S3: return launch_batch_mode<M_RK3_S3>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_PIRK_1A: return launch_batch_mode<M_PIRK_1A>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_PIRK_1B: return launch_batch_mode<M_PIRK_1B>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_PIRK2_2: return launch_batch_mode<M_PIRK2_2>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_PIRK3_2A: return launch_batch_mode<M_PIRK3_2A>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_PIRK3_2B: return launch_batch_mode<M_PIRK3_2B>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    case M_PIRK3_3: return launch_batch_mode<M_PIRK3_3>(descs, tstart, ngrids, total, nwg, tj, ef, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage(int mode, const StageArgs& a0, hipStream_t s) {
  StageArgs a = a0;
  const int nsys = a.nsys;
  if (a.G.fd4) {
    switch (mode) {
      case M_RK4_S1: return launch_fd4<M_RK4_S1>(a, s, nsys);
      case M_RK4_S2: return launch_fd4<M_RK4_S2>(a, s, nsys);
      case M_RK4_S3: return launch_fd4<M_RK4_S3>(a, s, nsys);
      case M_RK4_S4: return launch_fd4<M_RK4_S4>(a, s, nsys);
      case M_RK3_S1: return launch_fd4<M_RK3_S1>(a, s, nsys);
      case M_RK3_S2: return launch_fd4<M_RK3_S2>(a, s, nsys);
      case M_RK3_S3: return launch_fd4<M_RK3_S3>(a, s, nsys);
      default: return hipErrorInvalidValue;  // the PIRK stages are refused at the API
    }
  }
  a.tj = stage_tj(mode, a.tj);
  finalize_tiles(a);
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

template <typename I>
hipError_t launch_fill_i(const FillArgs& a, hipStream_t s, int phase, int64_t n) {
  const dim3 grid((unsigned)((n + 255) / 256));
  if (phase == 0)
    hipLaunchKernelGGL((fill_kernel<0, I>), grid, dim3(256), 0, s, a, n);
  else if (phase == 1)
    hipLaunchKernelGGL((fill_kernel<1, I>), grid, dim3(256), 0, s, a, n);
  else
    hipLaunchKernelGGL((fill_kernel<2, I>), grid, dim3(256), 0, s, a, n);
  return hipGetLastError();
}

hipError_t launch_fill(const FillArgs& a, hipStream_t s, int phase, int nsys) {
  const int64_t n = fill_items(a.G, nsys, phase);
  return n < (int64_t{1} << 32) ? launch_fill_i<uint32_t>(a, s, phase, n) : launch_fill_i<int64_t>(a, s, phase, n);
}

template <typename I>
hipError_t launch_fill_batch_i(const FillArgs* descs, const int* bstart, int ngrids, const int64_t* nitems,
                               int nblocks, int phase, hipStream_t s) {
  const dim3 grid((unsigned)nblocks);
  if (phase == 0)
    hipLaunchKernelGGL((fill_kernel_batch<0, I>), grid, dim3(256), 0, s, descs, bstart, ngrids, nitems);
  else if (phase == 1)
    hipLaunchKernelGGL((fill_kernel_batch<1, I>), grid, dim3(256), 0, s, descs, bstart, ngrids, nitems);
  else
    hipLaunchKernelGGL((fill_kernel_batch<2, I>), grid, dim3(256), 0, s, descs, bstart, ngrids, nitems);
  return hipGetLastError();
}

hipError_t launch_fill_batch(const FillArgs* descs, const int* bstart, int ngrids, const int64_t* nitems,
                             int nblocks, int phase, int wide, hipStream_t s) {
  return wide ? launch_fill_batch_i<int64_t>(descs, bstart, ngrids, nitems, nblocks, phase, s)
              : launch_fill_batch_i<uint32_t>(descs, bstart, ngrids, nitems, nblocks, phase, s);
}

int64_t fill_work_items(const Geom& G, int nsys, int phase) { return fill_items(G, nsys, phase); }

hipError_t launch_pack(const Geom& G, double* src, double* const d[5], int nsys, int to_dev, hipStream_t s,
                       int fortran) {
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((G.ni + 63) / 64 * G.nj), (unsigned)G.nk, 5 * nsys), dim3(64),
                     0, s, G, src, d[0], d[1], d[2], d[3], d[4], to_dev, fortran, nsys);
  return hipGetLastError();
}

}  // namespace kg

namespace kg {

bool somm_stage(int mode) { return mode != M_PIRK_1A && mode != M_PIRK3_2A; }

int64_t somm_items(const StageArgs& a) { return somm_count(a); }

int somm_block() { return SOMM_BLOCK; }

#define KG_SOMM_CASES(CALL)                                  \
  switch (mode) {                                            \
    case M_RK4_S1: return CALL(M_RK4_S1);                    \
    case M_RK4_S2: return CALL(M_RK4_S2);                    \
    case M_RK4_S3: return CALL(M_RK4_S3);                    \
    case M_RK4_S4: return CALL(M_RK4_S4);                    \
    case M_RK3_S1: return CALL(M_RK3_S1);                    \
    case M_RK3_S2: return CALL(M_RK3_S2);                    \
    case M_RK3_S3: return CALL(M_RK3_S3);                    \
    default: break;                                          \
  }

hipError_t launch_somm(int mode, const StageArgs& a, hipStream_t s) {
  if (a.G.fd4) {
#define KG_SOMM_FD4(M) launch_somm_mode<M, 1>(a, s)
    KG_SOMM_CASES(KG_SOMM_FD4)
#undef KG_SOMM_FD4
    return hipErrorInvalidValue;
  }
#define KG_SOMM_REF2(M) launch_somm_mode<M, 0>(a, s)
  KG_SOMM_CASES(KG_SOMM_REF2)
#undef KG_SOMM_REF2
  switch (mode) {
    case M_PIRK_1B: return launch_somm_mode<M_PIRK_1B, 0>(a, s);
    case M_PIRK2_2: return launch_somm_mode<M_PIRK2_2, 0>(a, s);
    case M_PIRK3_2B: return launch_somm_mode<M_PIRK3_2B, 0>(a, s);
    case M_PIRK3_3: return launch_somm_mode<M_PIRK3_3, 0>(a, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_somm_batch(int mode, const StageArgs* descs, const int* bstart, int ngrids, int nblocks, int fd4,
                             hipStream_t s) {
  if (fd4) {
#define KG_SOMM_FD4(M) launch_somm_batch_mode<M, 1>(descs, bstart, ngrids, nblocks, s)
    KG_SOMM_CASES(KG_SOMM_FD4)
#undef KG_SOMM_FD4
    return hipErrorInvalidValue;
  }
#define KG_SOMM_REF2(M) launch_somm_batch_mode<M, 0>(descs, bstart, ngrids, nblocks, s)
  KG_SOMM_CASES(KG_SOMM_REF2)
#undef KG_SOMM_REF2
  switch (mode) {
    case M_PIRK_1B: return launch_somm_batch_mode<M_PIRK_1B, 0>(descs, bstart, ngrids, nblocks, s);
    case M_PIRK2_2: return launch_somm_batch_mode<M_PIRK2_2, 0>(descs, bstart, ngrids, nblocks, s);
    case M_PIRK3_2B: return launch_somm_batch_mode<M_PIRK3_2B, 0>(descs, bstart, ngrids, nblocks, s);
    case M_PIRK3_3: return launch_somm_batch_mode<M_PIRK3_3, 0>(descs, bstart, ngrids, nblocks, s);
  }
  return hipErrorInvalidValue;
}
#undef KG_SOMM_CASES

}  // namespace kg
