// Fused Navier-Stokes right-hand side and the RK4 / Euler stage updates.
//
// One thread owns one interior cell (i,j,k) and produces the u-, v- and
// w-momentum RHS there (u at face x_i, v at y_j, w at z_k) straight from the
// 33 padded velocity values around it, instead of the reference's ~220
// whole-array passes (solver/velocity.py:221-383).  The arithmetic follows the
// reference's op order exactly: every torch op rounds on its own, and the
// reference's add_/sub_ with a scalar (metric_staggered.py:411-418) is one
// fused multiply-add, so the fluxes use separate ops and div_visc uses fma().
// The file is built with -ffp-contract=off so the compiler adds no other
// contractions.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "pf_kernels.h"

namespace pf {

template <typename T>
struct Sfs {
  const T* f[6];      // FXX, FYY, FZZ, FXY, FXZ, FYZ at the interior origin, or all NULL
  const T* visc;      // viscosity field VISC_P at the interior origin (field-viscosity kernels)
  T rho;              // density, for the field viscosity's 2*VISC/rho and interp_sc(VISC)/rho
  const T* src[3];    // GX, GY, GZ of a 'source' closure at the interior origin (SRC kernels)
};

// 0.25*(((F(0)+F(-a))+F(-b))+F(-a-b)): interp_sc_xy/xz/yz at an edge.
template <typename T>
__device__ __forceinline__ T edge4(const T* __restrict__ F, int64_t c, int64_t sa, int64_t sb) {
  return (T)0.25 * (((F[c] + F[c - sa]) + F[c - sb]) + F[c - sa - sb]);
}

// RHS at interior cell (i,j,k); c is the padded offset of the cell (from the
// interior origin) in the SFS and viscosity fields.  ADV = false leaves out the advective
// terms (the adjoint RHS applies the viscous operator alone to its field).
// VF = true takes the viscosity from the field sfs.visc (an eddy viscosity,
// domain.py:348-352): 2*VISC/rho at the cell centres of the normal fluxes and
// interp_sc_xy/xz/yz(VISC)/rho at the edges of the cross fluxes
// (velocity.py:248-305), each rounded as the reference's whole-array ops
// round them; VF = false uses the constant coefficients of k.
//
// U, V and W are taps: U(di, dj, dk) is u at (i+di, j+dj, k+dk), read from
// global memory (GlobalTap) or from the LDS planes of the tiled kernel
// (PlaneTap).  The arithmetic is the same for both; the rhs_cell overload
// below takes plain pointers to the padded interior origin.
//
// SRC = true subtracts a 'source' closure's GX, GY, GZ interpolated to the
// faces, 0.5*(G[i] + G[i-1]) along the component's own axis, after the
// advective terms (velocity.py:360-383).
template <typename T, bool SFS, bool ADV, bool VF, bool SRC, class Tap>
__device__ __forceinline__ void rhs_cell_tap(const Tap& U, const Tap& V, const Tap& W, int64_t c,
                                             int64_t sx, int64_t sy, const StepCoef<T>& k,
                                             const Sfs<T>& sfs, T& ru, T& rv, T& rw) {
  const T dxi = k.dxi, dyi = k.dyi, dzi = k.dzi;
  const T half = (T)0.5;
  const T* __restrict__ vis = sfs.visc;
  auto c2 = [&](int64_t o) -> T { return VF ? ((T)2.0 * vis[o]) / sfs.rho : k.c2; };
  auto exy = [&](int64_t o) -> T { return VF ? edge4(vis, o, sx, sy) / sfs.rho : k.exy; };
  auto exz = [&](int64_t o) -> T { return VF ? edge4(vis, o, sx, (int64_t)1) / sfs.rho : k.exz; };
  auto eyz = [&](int64_t o) -> T { return VF ? edge4(vis, o, sy, (int64_t)1) / sfs.rho : k.eyz; };
#define LD(A, di, dj, dk) A(di, dj, dk)
  const T u000 = LD(U, 0, 0, 0), up00 = LD(U, 1, 0, 0), um00 = LD(U, -1, 0, 0);
  const T u0p0 = LD(U, 0, 1, 0), u0m0 = LD(U, 0, -1, 0), u00p = LD(U, 0, 0, 1), u00m = LD(U, 0, 0, -1);
  const T ump0 = LD(U, -1, 1, 0), upm0 = LD(U, 1, -1, 0), um0p = LD(U, -1, 0, 1), up0m = LD(U, 1, 0, -1);
  const T v000 = LD(V, 0, 0, 0), vp00 = LD(V, 1, 0, 0), vm00 = LD(V, -1, 0, 0);
  const T v0p0 = LD(V, 0, 1, 0), v0m0 = LD(V, 0, -1, 0), v00p = LD(V, 0, 0, 1), v00m = LD(V, 0, 0, -1);
  const T vmp0 = LD(V, -1, 1, 0), vpm0 = LD(V, 1, -1, 0), v0mp = LD(V, 0, -1, 1), v0pm = LD(V, 0, 1, -1);
  const T w000 = LD(W, 0, 0, 0), wp00 = LD(W, 1, 0, 0), wm00 = LD(W, -1, 0, 0);
  const T w0p0 = LD(W, 0, 1, 0), w0m0 = LD(W, 0, -1, 0), w00p = LD(W, 0, 0, 1), w00m = LD(W, 0, 0, -1);
  const T wm0p = LD(W, -1, 0, 1), wp0m = LD(W, 1, 0, -1), w0mp = LD(W, 0, -1, 1), w0pm = LD(W, 0, 1, -1);
#undef LD

  // div_vel_over / 3 at this cell and its three lower neighbours
  // (metric_staggered.py:424-433, velocity.py:240-241)
  const T third = (T)3.0;
  const T D_c = ((((up00 - u000) * dxi) + ((v0p0 - v000) * dyi)) + ((w00p - w000) * dzi)) / third;
  const T D_xm = ((((u000 - um00) * dxi) + ((vmp0 - vm00) * dyi)) + ((wm0p - wm00) * dzi)) / third;
  const T D_ym = ((((upm0 - u0m0) * dxi) + ((v000 - v0m0) * dyi)) + ((w0mp - w0m0) * dzi)) / third;
  const T D_zm = ((((up0m - u00m) * dxi) + ((v0pm - v00m) * dyi)) + ((w000 - w00m) * dzi)) / third;

  // ---- u: viscous fluxes (velocity.py:248-269) ----
  T FX_c = (((up00 - u000) * dxi) - D_c) * c2(c);
  T FX_m = (((u000 - um00) * dxi) - D_xm) * c2(c - sx);
  T FY_c = (((u0p0 - u000) * dyi) + ((vp00 - v000) * dxi)) * exy(c);
  T FY_m = (((u000 - u0m0) * dyi) + ((vpm0 - v0m0) * dxi)) * exy(c - sy);
  T FZ_c = (((u00p - u000) * dzi) + ((wp00 - w000) * dxi)) * exz(c);
  T FZ_m = (((u000 - u00m) * dzi) + ((wp0m - w00m) * dxi)) * exz(c - 1);
  if (SFS) {
    FX_c = FX_c + sfs.f[0][c];
    FX_m = FX_m + sfs.f[0][c - sx];
    FY_c = FY_c + edge4(sfs.f[3], c, sx, sy);
    FY_m = FY_m + edge4(sfs.f[3], c - sy, sx, sy);
    FZ_c = FZ_c + edge4(sfs.f[4], c, sx, (int64_t)1);
    FZ_m = FZ_m + edge4(sfs.f[4], c - 1, sx, (int64_t)1);
  }
  T r = fma(FX_c, dxi, (T)0);
  r = fma(FX_m, -dxi, r);
  r = fma(FY_c, dyi, r);
  r = fma(FY_m, -dyi, r);
  r = fma(FZ_c, dzi, r);
  r = fma(FZ_m, -dzi, r);
  // ---- u: advection (velocity.py:322-331, metric_staggered.py:458-526) ----
  if (ADV) {
    const T a = half * (up00 + u000), b = half * (u000 + um00);
    r = r - ((a * a - b * b) * dxi);
    const T uy_p = half * (u0p0 + u000), vx_p = half * (v0p0 + vmp0);
    const T uy_c = half * (u000 + u0m0), vx_c = half * (v000 + vm00);
    r = r - ((uy_p * vx_p - uy_c * vx_c) * dyi);
    const T uz_p = half * (u00p + u000), wx_p = half * (w00p + wm0p);
    const T uz_c = half * (u000 + u00m), wx_c = half * (w000 + wm00);
    r = r - ((uz_p * wx_p - uz_c * wx_c) * dzi);
  }
  if (SRC) r = r - (T)0.5 * (sfs.src[0][c] + sfs.src[0][c - sx]);
  ru = r;

  // ---- v: viscous fluxes (velocity.py:272-293) ----
  FX_c = (((vp00 - v000) * dxi) + ((u0p0 - u000) * dyi)) * exy(c);
  FX_m = (((v000 - vm00) * dxi) + ((ump0 - um00) * dyi)) * exy(c - sx);
  FY_c = (((v0p0 - v000) * dyi) - D_c) * c2(c);
  FY_m = (((v000 - v0m0) * dyi) - D_ym) * c2(c - sy);
  FZ_c = (((v00p - v000) * dzi) + ((w0p0 - w000) * dyi)) * eyz(c);
  FZ_m = (((v000 - v00m) * dzi) + ((w0pm - w00m) * dyi)) * eyz(c - 1);
  if (SFS) {
    FX_c = FX_c + edge4(sfs.f[3], c, sx, sy);
    FX_m = FX_m + edge4(sfs.f[3], c - sx, sx, sy);
    FY_c = FY_c + sfs.f[1][c];
    FY_m = FY_m + sfs.f[1][c - sy];
    FZ_c = FZ_c + edge4(sfs.f[5], c, sy, (int64_t)1);
    FZ_m = FZ_m + edge4(sfs.f[5], c - 1, sy, (int64_t)1);
  }
  r = fma(FX_c, dxi, (T)0);
  r = fma(FX_m, -dxi, r);
  r = fma(FY_c, dyi, r);
  r = fma(FY_m, -dyi, r);
  r = fma(FZ_c, dzi, r);
  r = fma(FZ_m, -dzi, r);
  // ---- v: advection (velocity.py:334-343) ----
  if (ADV) {
    const T vx_p1 = half * (vp00 + v000), uy_p1 = half * (up00 + upm0);
    const T vx_c = half * (v000 + vm00), uy_c = half * (u000 + u0m0);
    r = r - ((vx_p1 * uy_p1 - vx_c * uy_c) * dxi);
    const T a = half * (v0p0 + v000), b = half * (v000 + v0m0);
    r = r - ((a * a - b * b) * dyi);
    const T vz_p = half * (v00p + v000), wy_p = half * (w00p + w0mp);
    const T vz_c = half * (v000 + v00m), wy_c = half * (w000 + w0m0);
    r = r - ((vz_p * wy_p - vz_c * wy_c) * dzi);
  }
  if (SRC) r = r - (T)0.5 * (sfs.src[1][c] + sfs.src[1][c - sy]);
  rv = r;

  // ---- w: viscous fluxes (velocity.py:296-317) ----
  FX_c = (((wp00 - w000) * dxi) + ((u00p - u000) * dzi)) * exz(c);
  FX_m = (((w000 - wm00) * dxi) + ((um0p - um00) * dzi)) * exz(c - sx);
  FY_c = (((w0p0 - w000) * dyi) + ((v00p - v000) * dzi)) * eyz(c);
  FY_m = (((w000 - w0m0) * dyi) + ((v0mp - v0m0) * dzi)) * eyz(c - sy);
  FZ_c = (((w00p - w000) * dzi) - D_c) * c2(c);
  FZ_m = (((w000 - w00m) * dzi) - D_zm) * c2(c - 1);
  if (SFS) {
    FX_c = FX_c + edge4(sfs.f[4], c, sx, (int64_t)1);
    FX_m = FX_m + edge4(sfs.f[4], c - sx, sx, (int64_t)1);
    FY_c = FY_c + edge4(sfs.f[5], c, sy, (int64_t)1);
    FY_m = FY_m + edge4(sfs.f[5], c - sy, sy, (int64_t)1);
    FZ_c = FZ_c + sfs.f[2][c];
    FZ_m = FZ_m + sfs.f[2][c - 1];
  }
  r = fma(FX_c, dxi, (T)0);
  r = fma(FX_m, -dxi, r);
  r = fma(FY_c, dyi, r);
  r = fma(FY_m, -dyi, r);
  r = fma(FZ_c, dzi, r);
  r = fma(FZ_m, -dzi, r);
  // ---- w: advection (velocity.py:346-355) ----
  if (ADV) {
    const T wx_p1 = half * (wp00 + w000), uz_p1 = half * (up00 + up0m);
    const T wx_c = half * (w000 + wm00), uz_c = half * (u000 + u00m);
    r = r - ((wx_p1 * uz_p1 - wx_c * uz_c) * dxi);
    const T wy_p1 = half * (w0p0 + w000), vz_p1 = half * (v0p0 + v0pm);
    const T wy_c = half * (w000 + w0m0), vz_c = half * (v000 + v00m);
    r = r - ((wy_p1 * vz_p1 - wy_c * vz_c) * dyi);
    const T a = half * (w00p + w000), b = half * (w000 + w00m);
    r = r - ((a * a - b * b) * dzi);
  }
  if (SRC) r = r - (T)0.5 * (sfs.src[2][c] + sfs.src[2][c - (int64_t)1]);
  rw = r;
}

// Stencil input read from global memory: p is the field at the cell.
template <typename T>
struct GlobalTap {
  const T* __restrict__ p;
  int64_t sx, sy;
  __device__ __forceinline__ T operator()(int di, int dj, int dk) const {
    return p[di * sx + dj * sy + dk];
  }
};

template <typename T, bool SFS, bool ADV = true, bool VF = false>
__device__ __forceinline__ void rhs_cell(const T* __restrict__ U, const T* __restrict__ V,
                                         const T* __restrict__ W, int64_t c, int64_t sx,
                                         int64_t sy, const StepCoef<T>& k, const Sfs<T>& sfs,
                                         T& ru, T& rv, T& rw) {
  rhs_cell_tap<T, SFS, ADV, VF, false>(GlobalTap<T>{U + c, sx, sy}, GlobalTap<T>{V + c, sx, sy},
                                GlobalTap<T>{W + c, sx, sy}, c, sx, sy, k, sfs, ru, rv, rw);
}

// Stage kinds.
enum { kStage1 = 1, kStage2 = 2, kStage3 = 3, kStage4 = 4, kEuler = 0, kRhsOnly = 5 };

// Stage update of one cell from its RHS r (advancer.py:49-113):
//   Un    : u^n (padded, interior origin)
//   acc   : running k1 + 2k2 + 2k3 (interior)
//   Yout  : next stage state (padded), or the Euler result, or the RHS (interior)
//   Uout  : u* for stage 4 (padded; same arrays as Un, read and written at the own cell only)
// The values a stage reads at its own cell besides the stencil: u^n (all
// kinds but the bare RHS) and the running sum (stages 2-4).
template <typename T, int KIND>
struct CellIn {
  T un[3], ac[3];
  __device__ __forceinline__ void load(const CVec3<T>& Un, const Vec3<T>& acc, uint32_t idx,
                                       int64_t c) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      un[q] = KIND != kRhsOnly ? Un.a[q][c] : (T)0;
      ac[q] = (KIND == kStage2 || KIND == kStage3 || KIND == kStage4) ? acc.a[q][idx] : (T)0;
    }
  }
};

template <typename T, int KIND>
__device__ __forceinline__ void stage_store(const Geom& g, uint32_t idx, int i, int j, int kk,
                                            const T (&r)[3], const CellIn<T, KIND>& in,
                                            const Vec3<T>& acc, const Vec3<T>& Yout,
                                            const Vec3<T>& Uout, const StepCoef<T>& k) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (KIND == kRhsOnly) {
      Yout.a[q][idx] = r[q];
    } else if (KIND == kEuler) {
      store_padded(g, Yout.a[q] - g.poff, i, j, kk, in.un[q] + (r[q] * k.dt));
    } else if (KIND == kStage4) {
      const T s = in.ac[q] + r[q];
      store_padded(g, Uout.a[q] - g.poff, i, j, kk, in.un[q] + (k.dt6 * s));
    } else {
      if (KIND == kStage1)
        acc.a[q][idx] = r[q];
      else
        acc.a[q][idx] = in.ac[q] + ((T)2.0 * r[q]);
      store_padded(g, Yout.a[q] - g.poff, i, j, kk, (k.half_dt * r[q]) + in.un[q]);
    }
  }
}

template <typename T, int KIND>
__device__ __forceinline__ void stage_update(const Geom& g, uint32_t idx, int i, int j, int kk,
                                             int64_t c, const T (&r)[3], const CVec3<T>& Un,
                                             const Vec3<T>& acc, const Vec3<T>& Yout,
                                             const Vec3<T>& Uout, const StepCoef<T>& k) {
  CellIn<T, KIND> in;
  in.load(Un, acc, idx, c);
  stage_store<T, KIND>(g, idx, i, j, kk, r, in, acc, Yout, Uout, k);
}

// One RK stage (or Euler step, or a bare RHS evaluation) of the forward RHS;
// Y is the stage input state (padded, interior origin).
template <typename T, int KIND, bool SFS, bool VF = false>
__device__ __forceinline__ void rk_stage_body(const Geom& g, const CVec3<T>& Y, const CVec3<T>& Un,
                                              const Vec3<T>& acc, const Vec3<T>& Yout,
                                              const Vec3<T>& Uout, const StepCoef<T>& k,
                                              const Sfs<T>& sfs, Blk blk) {
  for (uint32_t idx = blk.id * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += blk.n * blockDim.x) {
    int i, j, kk;
    decode(g, idx, i, j, kk);
    const int64_t c = i * g.psx + j * g.psy + kk;
    T r[3];
    rhs_cell<T, SFS, true, VF>(Y.a[0], Y.a[1], Y.a[2], c, g.psx, g.psy, k, sfs, r[0], r[1], r[2]);
    stage_update<T, KIND>(g, idx, i, j, kk, c, r, Un, acc, Yout, Uout, k);
  }
}

// ---------------------------------------------------------------------------
// Tiled stage.  A block owns a TZ x TY tile of the (y,z) plane and marches
// along x over a chunk of planes.  Each plane of the stage input u, v, w
// (tile plus a one-cell y/z halo with its corners) is read from HBM once and
// kept in an LDS ring of four planes; the cell at x reads its 33 stencil
// values from the ring slots of planes x-1, x and x+1.  A flat one-cell-per-
// thread loop (rk_stage_body, which the batched launches use) re-reads the y
// and x neighbours from L2 or HBM instead (its consecutive blocks land on
// different XCDs): about three times the unique bytes at 512^3.  The loads of plane x+2 are in flight while plane x is computed.
//
// Ring hazard: iteration x writes plane x+1 into slot (x+1) & 3 before its
// barrier.  That slot last held plane x-3, read up to iteration x-2, and
// every thread has finished iteration x-2 before any passes the barrier of
// x-1, so one barrier per plane suffices.
//
// Inputs are padded (two periodic image layers), so every neighbour of an
// interior cell is an in-bounds padded cell; cells of partial tiles past the
// padding read the last padded row or column instead (their values reach only
// cells that store nothing).

// Stencil input read from the LDS ring: m, o, p point at the cell's entry in
// the slots of planes x-1, x and x+1.
template <typename T, int RW>
struct PlaneTap {
  const T* m;
  const T* o;
  const T* p;
  __device__ __forceinline__ T operator()(int di, int dj, int dk) const {
    const T* b = di < 0 ? m : (di > 0 ? p : o);
    return b[dj * RW + dk];
  }
};

template <typename T, int KIND, bool SFS, bool VF, bool SRC, int TZ, int TY>
__global__ void __launch_bounds__(TZ * TY) rk_tile_kernel(Geom g, TileGrid tg, CVec3<T> Y,
                                                          CVec3<T> Un, Vec3<T> acc, Vec3<T> Yout,
                                                          Vec3<T> Uout, StepCoef<T> k,
                                                          Sfs<T> sfs) {
  constexpr int RW = TZ + 2;           // LDS row: tile row plus the two z-halo cells
  constexpr int PL = (TY + 2) * RW;    // one field's plane
  static_assert(TZ * TY >= 2 * TZ + 2 * TY + 4, "halo needs a thread per cell");
  __shared__ T ring[4][3][PL];

  int tz_i, ty_i, cx_i;
  tile_of_block(tg, this_blk(), tz_i, ty_i, cx_i);
  const int z0 = tz_i * TZ, y0 = ty_i * TY, x0 = cx_i * tg.xc;
  const int x1 = min(x0 + tg.xc, g.nx);

  const int t = threadIdx.x;
  const int lz = t % TZ, ly = t / TZ;
  const int z = z0 + lz, y = y0 + ly;
  const bool active = z < g.nz && y < g.ny;
  const int64_t col = (int64_t)min(y, g.ny + 1) * g.psy + min(z, g.nz + 1);
  const int own = (ly + 1) * RW + lz + 1;

  // halo cells: rows y0-1 and y0+TY (2*TZ threads), columns z0-1 and z0+TZ
  // (2*TY threads), then the four corners
  int hslot = -1, hy = 0, hz = 0;
  if (t < 2 * TZ) {
    const bool hi = t >= TZ;
    hy = hi ? y0 + TY : y0 - 1;
    hz = z0 + t % TZ;
    hslot = (hi ? TY + 1 : 0) * RW + t % TZ + 1;
  } else if (t < 2 * TZ + 2 * TY) {
    const int u = t - 2 * TZ;
    const bool hi = u >= TY;
    hy = y0 + u % TY;
    hz = hi ? z0 + TZ : z0 - 1;
    hslot = (u % TY + 1) * RW + (hi ? TZ + 1 : 0);
  } else if (t < 2 * TZ + 2 * TY + 4) {
    const int u = t - 2 * TZ - 2 * TY;
    hy = (u & 1) ? y0 + TY : y0 - 1;
    hz = (u & 2) ? z0 + TZ : z0 - 1;
    hslot = ((u & 1) ? TY + 1 : 0) * RW + ((u & 2) ? TZ + 1 : 0);
  }
  const bool is_halo = hslot >= 0;
  const int64_t hcol = (int64_t)min(hy, g.ny + 1) * g.psy + min(hz, g.nz + 1);

  struct Plane {
    T c[3], h[3];
  };
  auto load_plane = [&](int x) -> Plane {
    Plane P;
    const int64_t xo = (int64_t)x * g.psx;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      P.c[q] = Y.a[q][xo + col];
      P.h[q] = is_halo ? Y.a[q][xo + hcol] : (T)0;
    }
    return P;
  };
  auto put_plane = [&](int x, const Plane& P) {
    T(*s)[PL] = ring[(x + 4) & 3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      s[q][own] = P.c[q];
      if (is_halo) s[q][hslot] = P.h[q];
    }
  };

  put_plane(x0 - 1, load_plane(x0 - 1));
  put_plane(x0, load_plane(x0));
  Plane nxt = load_plane(x0 + 1);
  const int64_t pcol = (int64_t)y * g.psy + z, icol = (int64_t)y * g.isy + z;
  for (int x = x0; x < x1; ++x) {
    put_plane(x + 1, nxt);
    if (x + 1 < x1) nxt = load_plane(x + 2);
    __syncthreads();
    if (active) {
      const int64_t c = (int64_t)x * g.psx + pcol;
      const T* sm = &ring[(x + 3) & 3][0][own];
      const T* so = &ring[x & 3][0][own];
      const T* sp = &ring[(x + 1) & 3][0][own];
      const PlaneTap<T, RW> tu{sm, so, sp}, tv{sm + PL, so + PL, sp + PL},
          tw{sm + 2 * PL, so + 2 * PL, sp + 2 * PL};
      T r[3];
      rhs_cell_tap<T, SFS, true, VF, SRC>(tu, tv, tw, c, g.psx, g.psy, k, sfs, r[0], r[1], r[2]);
      const uint32_t idx = (uint32_t)((int64_t)x * g.isx + icol);
      stage_update<T, KIND>(g, idx, x, y, z, c, r, Un, acc, Yout, Uout, k);
    }
  }
}

// Tiles of the stage kernel: 32 x 8 cells, chunks of at least 8 planes,
// up to 1024 blocks.
inline TileGrid stage_tile_grid(const Geom& g) { return tile_grid_with(g, 32, 1, 1, 8); }

// ---------------------------------------------------------------------------
// Adjoint RHS of the reference's adjoint step (solver/adjoint.py:94-298):
// the viscous operator of the forward RHS applied to the adjoint field, then
// the 15 discrete-exact convective transposes (metric_staggered.py:529-800)
// with the forward state's interpolants, in the reference's order.  Each
// transpose is two subtractions  r -= ((a[m - e_d + s] - a[m + s]) * vi[m + s + vsh]) * coef
// for s = 0 and s = s2.

// Interpolants of the forward state (metric_staggered.py:149-188), at the
// padded offset o:  kind 0 = cell centre along the own axis,
// 1 = the lower edge value 0.5*(f[o] + f[o - e]) along axis stride e.
template <typename T>
__device__ __forceinline__ T centre(const T* __restrict__ F, int64_t o, int64_t e) {
  return (T)0.5 * (F[o + e] + F[o]);
}
template <typename T>
__device__ __forceinline__ T lower(const T* __restrict__ F, int64_t o, int64_t e) {
  return (T)0.5 * (F[o] + F[o - e]);
}

// One adj_conv_* method at cell c: a = adjoint component, F = forward
// component whose interpolant vi is taken along stride ei (CENTRE selects
// the cell-centre interpolant), d = stride of the difference, s2 = shift
// of the second subtraction, vsh = shift of vi.
template <typename T, bool CENTRE>
__device__ __forceinline__ T adj_pair(T r, const T* __restrict__ a, const T* __restrict__ F,
                                      int64_t c, int64_t ei, int64_t d, int64_t s2, int64_t vsh,
                                      T coef) {
  {
    const int64_t o = c + vsh;
    const T vi = CENTRE ? centre(F, o, ei) : lower(F, o, ei);
    r = r - (((a[c - d] - a[c]) * vi) * coef);
  }
  {
    const int64_t o = c + s2 + vsh;
    const T vi = CENTRE ? centre(F, o, ei) : lower(F, o, ei);
    r = r - (((a[c - d + s2] - a[c + s2]) * vi) * coef);
  }
  return r;
}

template <typename T>
struct AdjCoef {
  T dxi, dyi, dzi;      // grad_x/y/z
  T hdx, hdy, hdz;      // 0.5*grad_x etc. as the reference forms them (a Python float)
};

template <typename T>
AdjCoef<T> make_adj_coef(const pf_box* h) {
  AdjCoef<T> a;
  a.dxi = (T)(1.0 / h->dx);
  a.dyi = (T)(1.0 / h->dy);
  a.dzi = (T)(1.0 / h->dz);
  a.hdx = (T)(0.5 * (1.0 / h->dx));
  a.hdy = (T)(0.5 * (1.0 / h->dy));
  a.hdz = (T)(0.5 * (1.0 / h->dz));
  return a;
}

// Adjoint RHS at interior cell c.  A* = adjoint fields, U* = forward fields
// (both padded, interior origin); D = optional closure term (padded).
// VF: the viscous part takes the viscosity field of vs (the reference's
// adjoint applies its forward viscous operator, VISC included, to the
// adjoint field).
template <typename T, bool DELTA, bool VF = false>
__device__ __forceinline__ void adj_rhs_cell(const T* __restrict__ Au, const T* __restrict__ Av,
                                             const T* __restrict__ Aw, const T* __restrict__ Uu,
                                             const T* __restrict__ Uv, const T* __restrict__ Uw,
                                             int64_t c, int64_t sx, int64_t sy,
                                             const StepCoef<T>& k, const AdjCoef<T>& a,
                                             const CVec3<T>& D, const Sfs<T>& vs, T (&r)[3]) {
  rhs_cell<T, false, false, VF>(Au, Av, Aw, c, sx, sy, k, vs, r[0], r[1], r[2]);
  const int64_t z = 1;
  T ru = r[0], rv = r[1], rw = r[2];
  // u (adjoint.py:172-192)
  ru = adj_pair<T, true>(ru, Au, Uu, c, sx, sx, sx, -sx, a.dxi);    // 1a xx, Ubar
  ru = adj_pair<T, false>(ru, Au, Uv, c, sx, sy, sy, 0, a.hdy);     // 1b y,  v^x
  ru = adj_pair<T, false>(ru, Au, Uw, c, sx, z, z, 0, a.hdz);       // 1c z,  w^x
  ru = adj_pair<T, false>(ru, Av, Uv, c, sx, sx, sy, 0, a.hdx);     // 1d,    v^x
  ru = adj_pair<T, false>(ru, Aw, Uw, c, sx, sx, z, 0, a.hdx);      // 1e,    w^x
  // v (adjoint.py:194-214)
  rv = adj_pair<T, false>(rv, Av, Uu, c, sy, sx, sx, 0, a.hdx);     // 2a x,  u^y
  rv = adj_pair<T, true>(rv, Av, Uv, c, sy, sy, sy, -sy, a.dyi);    // 2b yy, Vbar
  rv = adj_pair<T, false>(rv, Av, Uw, c, sy, z, z, 0, a.hdz);       // 2c z,  w^y
  rv = adj_pair<T, false>(rv, Au, Uu, c, sy, sy, sx, 0, a.hdy);     // 2d,    u^y
  rv = adj_pair<T, false>(rv, Aw, Uw, c, sy, sy, z, 0, a.hdy);      // 2e,    w^y
  // w (adjoint.py:216-236)
  rw = adj_pair<T, false>(rw, Aw, Uu, c, z, sx, sx, 0, a.hdx);      // 3a x,  u^z
  rw = adj_pair<T, false>(rw, Aw, Uv, c, z, sy, sy, 0, a.hdy);      // 3b y,  v^z
  rw = adj_pair<T, true>(rw, Aw, Uw, c, z, z, z, -z, a.dzi);        // 3c zz, Wbar
  rw = adj_pair<T, false>(rw, Au, Uu, c, z, z, sx, 0, a.hdz);       // 3d,    u^z
  rw = adj_pair<T, false>(rw, Av, Uv, c, z, z, sy, 0, a.hdz);       // 3e,    v^z
  if (DELTA) {
    ru = ru + D.a[0][c];
    rv = rv + D.a[1][c];
    rw = rw + D.a[2][c];
  }
  r[0] = ru;
  r[1] = rv;
  r[2] = rw;
}

template <typename T, int KIND, bool DELTA, bool VF>
__global__ void __launch_bounds__(kBlock) adj_stage_kernel(Geom g, CVec3<T> Y, CVec3<T> Uf,
                                                           CVec3<T> D, CVec3<T> Un, Vec3<T> acc,
                                                           Vec3<T> Yout, Vec3<T> Uout,
                                                           StepCoef<T> k, AdjCoef<T> a,
                                                           Sfs<T> vs) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, kk;
    decode(g, idx, i, j, kk);
    const int64_t c = i * g.psx + j * g.psy + kk;
    T r[3];
    adj_rhs_cell<T, DELTA, VF>(Y.a[0], Y.a[1], Y.a[2], Uf.a[0], Uf.a[1], Uf.a[2], c, g.psx,
                               g.psy, k, a, D, vs, r);
    stage_update<T, KIND>(g, idx, i, j, kk, c, r, Un, acc, Yout, Uout, k);
  }
}

template <typename T, int KIND>
int launch_stage(pf_box* h, hipStream_t s, const void* const Y[3], const void* const Un[3],
                 void* const Yout[3], void* const Uout[3], const StepCoef<T>& k,
                 const void* const* sfs6, const void* visc = nullptr, double rho = 1.0,
                 const void* const* src3 = nullptr) {
  const Geom& g = h->g;
  CVec3<T> y, un;
  Vec3<T> acc, yo, uo;
  for (int q = 0; q < 3; ++q) {
    y.a[q] = (const T*)Y[q] + g.poff;
    un.a[q] = Un && Un[q] ? (const T*)Un[q] + g.poff : nullptr;
    acc.a[q] = (T*)h->acc[q];
    if (KIND == kRhsOnly)
      yo.a[q] = (T*)Yout[q];
    else
      yo.a[q] = Yout && Yout[q] ? (T*)Yout[q] + g.poff : nullptr;
    uo.a[q] = Uout && Uout[q] ? (T*)Uout[q] + g.poff : nullptr;
  }
  Sfs<T> sf;
  for (int m = 0; m < 6; ++m) sf.f[m] = sfs6 ? (const T*)sfs6[m] + g.poff : nullptr;
  sf.visc = visc ? (const T*)visc + g.poff : nullptr;
  sf.rho = (T)rho;
  for (int q = 0; q < 3; ++q) sf.src[q] = src3 ? (const T*)src3[q] + g.poff : nullptr;
  const TileGrid tg = stage_tile_grid(g);
#define PF_STAGE(SFS, VF, SRC)                                                          \
  hipLaunchKernelGGL((rk_tile_kernel<T, KIND, SFS, VF, SRC, 32, 8>), dim3(tg.nblk), dim3(256), \
                     0, s, g, tg, y, un, acc, yo, uo, k, sf)
  // a 'source' closure is the only closure of its run (constant viscosity, no stresses)
  if (src3) {
    PF_STAGE(false, false, true);
  } else if (visc) {
    if (sfs6) PF_STAGE(true, true, false); else PF_STAGE(false, true, false);
  } else {
    if (sfs6) PF_STAGE(true, false, false); else PF_STAGE(false, false, false);
  }
#undef PF_STAGE
  PF_LAUNCH_CHECK("rk_tile_kernel");
  return 0;
}

template <typename T>
int rk4_impl(pf_box* h, void* const U[3], void* const Ut[3], double dt, double mu, double rho,
             const void* const* sfs6, hipStream_t s, const void* visc = nullptr) {
  const StepCoef<T> k = make_step_coef<T>(h, mu, rho, dt);
  // Stage states alternate between the caller's ut/vt/wt and the box's
  // second buffer so that the last stage state lands in ut/vt/wt
  // (advancer.py:93-96 leaves Y4 in state_uTmp).
  void* const A[3] = {Ut[0], Ut[1], Ut[2]};
  void* const B[3] = {h->ypong[0], h->ypong[1], h->ypong[2]};
  const void* const Uc[3] = {U[0], U[1], U[2]};
  const void* const Ac[3] = {A[0], A[1], A[2]};
  const void* const Bc[3] = {B[0], B[1], B[2]};
  int rc;
  if ((rc = launch_stage<T, kStage1>(h, s, Uc, Uc, A, nullptr, k, sfs6, visc, rho))) return rc;
  if ((rc = launch_stage<T, kStage2>(h, s, Ac, Uc, B, nullptr, k, sfs6, visc, rho))) return rc;
  if ((rc = launch_stage<T, kStage3>(h, s, Bc, Uc, A, nullptr, k, sfs6, visc, rho))) return rc;
  if ((rc = launch_stage<T, kStage4>(h, s, Ac, Uc, nullptr, U, k, sfs6, visc, rho))) return rc;
  return 0;
}

template <typename T>
int euler_impl(pf_box* h, void* const U[3], double dt, double mu, double rho,
               const void* const* sfs6, hipStream_t s, const void* visc = nullptr) {
  const StepCoef<T> k = make_step_coef<T>(h, mu, rho, dt);
  void* const B[3] = {h->ypong[0], h->ypong[1], h->ypong[2]};
  const void* const Uc[3] = {U[0], U[1], U[2]};
  int rc = launch_stage<T, kEuler>(h, s, Uc, Uc, B, nullptr, k, sfs6, visc, rho);
  if (rc) return rc;
  const size_t np = (size_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4) * sizeof(T);
  for (int q = 0; q < 3; ++q) PF_HIP(hipMemcpyAsync(U[q], B[q], np, hipMemcpyDeviceToDevice, s));
  return 0;
}

// ---------------------------------------------------------------------------
// Inputs of a 'source' closure (velocity.py:365-371): the closure receives
// state_u.var_i, state_v.var_i, state_w.var_i right after interp_u_xm,
// interp_v_ym and interp_w_zm.  Those write every plane but the last along
// their own axis; the last plane still holds what the advective terms of the
// same RHS evaluation wrote there last (velocity.py:322-355):
//   u: k >= 1 interp_uv_z (x-plane nxo-1), k = 0 interp_uw_y;
//   v: k >= 1 interp_uv_z (y-plane nyo-1), k = 0 interp_vw_x;
//   w: j >= 1 interp_uw_y (z-plane nzo-1), j = 0 interp_vw_x;
// and one corner of that plane ((nxo-1,0,0), (0,nyo-1,0), (0,0,nzo-1)) is
// written by none of them, so it keeps the value it had (zero from the
// allocation in state.py:125 unless a caller wrote it).  All values are
// 0.5*(f[a] + f[a - e]) as metric_staggered.py:149-188 forms them.
template <typename T>
__global__ void __launch_bounds__(kBlock) source_inputs_kernel(Geom g, CVec3<T> U, Vec3<T> I) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t n = (int64_t)nxo * nyo * nzo;
  const T half = (T)0.5;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(p % nzo);
    const int j = (int)((p / nzo) % nyo);
    const int i = (int)(p / ((int64_t)nzo * nyo));
    const T* u = U.a[0];
    const T* v = U.a[1];
    const T* w = U.a[2];
    // u: x-centre interpolant, or the last x-plane's edge values
    if (i < nxo - 1) I.a[0][p] = half * (u[p + g.psx] + u[p]);
    else if (k >= 1) I.a[0][p] = half * (u[p] + u[p - 1]);
    else if (j >= 1) I.a[0][p] = half * (u[p] + u[p - g.psy]);
    // v
    if (j < nyo - 1) I.a[1][p] = half * (v[p + g.psy] + v[p]);
    else if (k >= 1) I.a[1][p] = half * (v[p] + v[p - 1]);
    else if (i >= 1) I.a[1][p] = half * (v[p] + v[p - g.psx]);
    // w
    if (k < nzo - 1) I.a[2][p] = half * (w[p + 1] + w[p]);
    else if (j >= 1) I.a[2][p] = half * (w[p] + w[p - g.psy]);
    else if (i >= 1) I.a[2][p] = half * (w[p] + w[p - g.psx]);
  }
}

template <typename T>
int source_inputs_impl(pf_box* h, const void* const U[3], void* const I[3], hipStream_t s) {
  CVec3<T> u;
  Vec3<T> in;
  for (int q = 0; q < 3; ++q) {
    u.a[q] = (const T*)U[q];
    in.a[q] = (T*)I[q];
  }
  const int64_t n = (int64_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4);
  int64_t nb = (n + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(source_inputs_kernel<T>, dim3((unsigned)nb), dim3(kBlock), 0, s, h->g, u, in);
  PF_LAUNCH_CHECK("source_inputs_kernel");
  return 0;
}

// The stage input and output buffers of rk4_impl's sequence (advancer.py:75-113):
// stage 1 U -> A, stage 2 A -> B, stage 3 B -> A, stage 4 A -> U, with A the
// caller's ut/vt/wt and B the box's second buffer.
inline const void* rk4_stage_input(pf_box* h, int stage, void* const U[3], void* const Ut[3],
                                   int q) {
  return stage == 1 ? U[q] : (stage == 3 ? h->ypong[q] : Ut[q]);
}

template <typename T>
int rk4_stage_src_impl(pf_box* h, int stage, void* const U[3], void* const Ut[3], double dt,
                       double mu, double rho, const void* const G[3], hipStream_t s) {
  const StepCoef<T> k = make_step_coef<T>(h, mu, rho, dt);
  void* const B[3] = {h->ypong[0], h->ypong[1], h->ypong[2]};
  const void* Y[3];
  for (int q = 0; q < 3; ++q) Y[q] = rk4_stage_input(h, stage, U, Ut, q);
  const void* const Uc[3] = {U[0], U[1], U[2]};
  switch (stage) {
    case 1: return launch_stage<T, kStage1>(h, s, Y, Uc, Ut, nullptr, k, nullptr, nullptr, rho, G);
    case 2: return launch_stage<T, kStage2>(h, s, Y, Uc, B, nullptr, k, nullptr, nullptr, rho, G);
    case 3: return launch_stage<T, kStage3>(h, s, Y, Uc, Ut, nullptr, k, nullptr, nullptr, rho, G);
    default: return launch_stage<T, kStage4>(h, s, Y, Uc, nullptr, U, k, nullptr, nullptr, rho, G);
  }
}

template <typename T, int KIND>
int launch_adj_stage(pf_box* h, hipStream_t s, const void* const Y[3], const void* const Uf[3],
                     const void* const* D, const void* const Un[3], void* const Yout[3],
                     void* const Uout[3], const StepCoef<T>& k, const void* visc = nullptr,
                     double rho = 1.0) {
  const Geom& g = h->g;
  CVec3<T> y, uf, d, un;
  Vec3<T> acc, yo, uo;
  for (int q = 0; q < 3; ++q) {
    y.a[q] = (const T*)Y[q] + g.poff;
    uf.a[q] = (const T*)Uf[q] + g.poff;
    d.a[q] = D ? (const T*)D[q] + g.poff : nullptr;
    un.a[q] = Un && Un[q] ? (const T*)Un[q] + g.poff : nullptr;
    acc.a[q] = (T*)h->acc[q];
    if (KIND == kRhsOnly)
      yo.a[q] = (T*)Yout[q];
    else
      yo.a[q] = Yout && Yout[q] ? (T*)Yout[q] + g.poff : nullptr;
    uo.a[q] = Uout && Uout[q] ? (T*)Uout[q] + g.poff : nullptr;
  }
  const AdjCoef<T> a = make_adj_coef<T>(h);
  Sfs<T> vs = {{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}};
  vs.visc = visc ? (const T*)visc + g.poff : nullptr;
  vs.rho = (T)rho;
  const int nb = launch_blocks(g);
#define PF_ADJ(DELTA, VF)                                                                       \
  hipLaunchKernelGGL((adj_stage_kernel<T, KIND, DELTA, VF>), dim3(nb), dim3(kBlock), 0, s, g, y, \
                     uf, d, un, acc, yo, uo, k, a, vs)
  if (visc) {
    if (D) PF_ADJ(true, true); else PF_ADJ(false, true);
  } else {
    if (D) PF_ADJ(true, false); else PF_ADJ(false, false);
  }
#undef PF_ADJ
  PF_LAUNCH_CHECK("adj_stage_kernel");
  return 0;
}

// Closure term of one adjoint stage, when the step carries a closure.
struct AdjClosure {
  int kind;              // -1 none, 0 gradient model, 1 MLP, 2 'source' ML closure
  const float* params;
  void* work;
  void* delta[3];
  double* dparams;
};

// stage >= 0: a stage of an adjoint RK4 step whose 'source' ML closure term
// runs from the step's cached forward (ml_source_step_begin ... _end).
static int adj_term(pf_box* h, const AdjClosure& cl, const void* const U[3],
                    const void* const Y[3], hipStream_t s, const void* const** D,
                    int stage = -1) {
  if (cl.kind < 0) {
    *D = nullptr;
    return 0;
  }
  if (stage >= 0) {
    if (int rc = ml_source_step_term(h, Y, cl.delta, stage, s)) return rc;
  } else if (int rc = adj_closure_term(h, cl.kind, U, cl.params, Y, cl.work, cl.delta,
                                       cl.dparams, s)) {
    return rc;
  }
  *D = (const void* const*)cl.delta;
  return 0;
}

// Adjoint RK4 (advancer.py:75-113 on the adjoint RHS): the same stage
// structure as the forward step, the forward state U held fixed.
template <typename T>
int adj_rk4_impl(pf_box* h, const void* const U[3], void* const Ua[3], void* const Ut[3],
                 double dt, double mu, double rho, const AdjClosure& cl, hipStream_t s,
                 const void* visc = nullptr) {
  const StepCoef<T> k = make_step_coef<T>(h, mu, rho, dt);
  void* const A[3] = {Ut[0], Ut[1], Ut[2]};
  void* const B[3] = {h->ypong[0], h->ypong[1], h->ypong[2]};
  const void* const Uc[3] = {Ua[0], Ua[1], Ua[2]};
  const void* const Ac[3] = {A[0], A[1], A[2]};
  const void* const Bc[3] = {B[0], B[1], B[2]};
  const void* const* D;
  int rc;
  // the 'source' ML closure: the forward state is the same in the four
  // stages, so its network forward runs once and the step's parameter
  // gradient is one product of the stage-summed layer adjoints (when the
  // box's cache fits; else each stage evaluates the whole term)
  const bool mlc = cl.kind == 2 && ml_source_step_cached(h);
  if (mlc && (rc = ml_source_step_begin(h, U, cl.params, s))) return rc;
  if ((rc = adj_term(h, cl, U, Uc, s, &D, mlc ? 0 : -1))) return rc;
  if ((rc = launch_adj_stage<T, kStage1>(h, s, Uc, U, D, Uc, A, nullptr, k, visc, rho)))
    return rc;
  if ((rc = adj_term(h, cl, U, Ac, s, &D, mlc ? 1 : -1))) return rc;
  if ((rc = launch_adj_stage<T, kStage2>(h, s, Ac, U, D, Uc, B, nullptr, k, visc, rho)))
    return rc;
  if ((rc = adj_term(h, cl, U, Bc, s, &D, mlc ? 2 : -1))) return rc;
  if ((rc = launch_adj_stage<T, kStage3>(h, s, Bc, U, D, Uc, A, nullptr, k, visc, rho)))
    return rc;
  if ((rc = adj_term(h, cl, U, Ac, s, &D, mlc ? 3 : -1))) return rc;
  if ((rc = launch_adj_stage<T, kStage4>(h, s, Ac, U, D, Uc, nullptr, Ua, k, visc, rho)))
    return rc;
  if (mlc && (rc = ml_source_step_end(h, cl.dparams, s))) return rc;
  return 0;
}

template <typename T>
int adj_euler_impl(pf_box* h, const void* const U[3], void* const Ua[3], double dt, double mu,
                   double rho, const AdjClosure& cl, hipStream_t s, const void* visc = nullptr) {
  const StepCoef<T> k = make_step_coef<T>(h, mu, rho, dt);
  void* const B[3] = {h->ypong[0], h->ypong[1], h->ypong[2]};
  const void* const Uc[3] = {Ua[0], Ua[1], Ua[2]};
  const void* const* D;
  int rc;
  if ((rc = adj_term(h, cl, U, Uc, s, &D))) return rc;
  if ((rc = launch_adj_stage<T, kEuler>(h, s, Uc, U, D, Uc, B, nullptr, k, visc, rho)))
    return rc;
  const size_t np = (size_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4) * sizeof(T);
  for (int q = 0; q < 3; ++q) PF_HIP(hipMemcpyAsync(Ua[q], B[q], np, hipMemcpyDeviceToDevice, s));
  return 0;
}

// ---------------------------------------------------------------------------
// Batched RK4 stages: stage s of every box of the list (rk4_impl's buffer
// routing: Y2 -> ut, Y3 -> box buffer, Y4 -> ut, u* -> u).
template <typename T, int KIND>
__global__ void __launch_bounds__(kBlock) batch_stage_kernel(const BoxDesc<T>* __restrict__ D,
                                                             BatchList L) {
  const int i = find_box(L, (int)blockIdx.x);
  const BoxDesc<T>& d = D[L.box[i]];
  const Blk blk{(int)blockIdx.x - L.begin[i], L.begin[i + 1] - L.begin[i]};
  const int64_t o = d.g.poff;
  CVec3<T> y, un;
  Vec3<T> acc, yo, uo;
  for (int q = 0; q < 3; ++q) {
    T* const src = KIND == kStage1 ? d.u[q] : (KIND == kStage3 ? d.yp[q] : d.ut[q]);
    y.a[q] = src + o;
    un.a[q] = d.u[q] + o;
    acc.a[q] = d.acc[q];
    yo.a[q] = (KIND == kStage2 ? d.yp[q] : d.ut[q]) + o;
    uo.a[q] = d.u[q] + o;
  }
  const Sfs<T> none = {{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}};
  rk_stage_body<T, KIND, false>(d.g, y, un, acc, yo, uo, d.k, none, blk);
}

int batch_stage_launch(int dtype, int stage, const void* desc, const BatchList& L, int nblocks,
                       hipStream_t s) {
  if (nblocks <= 0) return 0;
  auto go = [&](auto zero) -> int {
    using T = decltype(zero);
    const BoxDesc<T>* D = (const BoxDesc<T>*)desc;
    switch (stage) {
      case 1: hipLaunchKernelGGL((batch_stage_kernel<T, kStage1>), dim3(nblocks), dim3(kBlock), 0, s, D, L); break;
      case 2: hipLaunchKernelGGL((batch_stage_kernel<T, kStage2>), dim3(nblocks), dim3(kBlock), 0, s, D, L); break;
      case 3: hipLaunchKernelGGL((batch_stage_kernel<T, kStage3>), dim3(nblocks), dim3(kBlock), 0, s, D, L); break;
      case 4: hipLaunchKernelGGL((batch_stage_kernel<T, kStage4>), dim3(nblocks), dim3(kBlock), 0, s, D, L); break;
      default: return fail(PF_EINVAL, "batch_stage_launch: stage is 1..4");
    }
    PF_LAUNCH_CHECK("batch_stage_kernel");
    return 0;
  };
  return dtype == PF_F64 ? go(0.0) : go(0.0f);
}

}  // namespace pf

using namespace pf;

static int check_ptrs(const void* const* p, int n, const char* what) {
  for (int i = 0; i < n; ++i)
    if (!p[i]) return fail(PF_EINVAL, std::string(what) + ": NULL field pointer");
  return 0;
}

extern "C" {

int pf_rk4_step(pf_box* h, void* u, void* v, void* w, void* ut, void* vt, void* wt, double dt,
                double mu, double rho, const void* const* sfs6, void* stream) {
  if (int rc = check_box(h)) return rc;
  void* U[3] = {u, v, w};
  void* Ut[3] = {ut, vt, wt};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_rk4_step")) return rc;
  if (int rc = check_ptrs((const void* const*)Ut, 3, "pf_rk4_step")) return rc;
  if (sfs6)
    if (int rc = check_ptrs(sfs6, 6, "pf_rk4_step(sfs6)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rk4_step: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? rk4_impl<double>(h, U, Ut, dt, mu, rho, sfs6, s)
                            : rk4_impl<float>(h, U, Ut, dt, mu, rho, sfs6, s);
}

int pf_euler_step(pf_box* h, void* u, void* v, void* w, double dt, double mu, double rho,
                  const void* const* sfs6, void* stream) {
  if (int rc = check_box(h)) return rc;
  void* U[3] = {u, v, w};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_euler_step")) return rc;
  if (sfs6)
    if (int rc = check_ptrs(sfs6, 6, "pf_euler_step(sfs6)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_euler_step: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? euler_impl<double>(h, U, dt, mu, rho, sfs6, s)
                            : euler_impl<float>(h, U, dt, mu, rho, sfs6, s);
}

int pf_rk4_step_visc(pf_box* h, void* u, void* v, void* w, void* ut, void* vt, void* wt,
                     double dt, const void* visc, double rho, const void* const* sfs6,
                     void* stream) {
  if (int rc = check_box(h)) return rc;
  void* U[3] = {u, v, w};
  void* Ut[3] = {ut, vt, wt};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_rk4_step_visc")) return rc;
  if (int rc = check_ptrs((const void* const*)Ut, 3, "pf_rk4_step_visc")) return rc;
  if (!visc) return fail(PF_EINVAL, "pf_rk4_step_visc: NULL viscosity field");
  if (sfs6)
    if (int rc = check_ptrs(sfs6, 6, "pf_rk4_step_visc(sfs6)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rk4_step_visc: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? rk4_impl<double>(h, U, Ut, dt, 0.0, rho, sfs6, s, visc)
                            : rk4_impl<float>(h, U, Ut, dt, 0.0, rho, sfs6, s, visc);
}

int pf_euler_step_visc(pf_box* h, void* u, void* v, void* w, double dt, const void* visc,
                       double rho, const void* const* sfs6, void* stream) {
  if (int rc = check_box(h)) return rc;
  void* U[3] = {u, v, w};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_euler_step_visc")) return rc;
  if (!visc) return fail(PF_EINVAL, "pf_euler_step_visc: NULL viscosity field");
  if (sfs6)
    if (int rc = check_ptrs(sfs6, 6, "pf_euler_step_visc(sfs6)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_euler_step_visc: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? euler_impl<double>(h, U, dt, 0.0, rho, sfs6, s, visc)
                            : euler_impl<float>(h, U, dt, 0.0, rho, sfs6, s, visc);
}

int pf_rhs_ns_visc(pf_box* h, const void* u, const void* v, const void* w, const void* visc,
                   double rho, const void* const* sfs6, void* ru, void* rv, void* rw,
                   void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {u, v, w};
  void* R[3] = {ru, rv, rw};
  if (int rc = check_ptrs(U, 3, "pf_rhs_ns_visc")) return rc;
  if (int rc = check_ptrs((const void* const*)R, 3, "pf_rhs_ns_visc")) return rc;
  if (!visc) return fail(PF_EINVAL, "pf_rhs_ns_visc: NULL viscosity field");
  if (sfs6)
    if (int rc = check_ptrs(sfs6, 6, "pf_rhs_ns_visc(sfs6)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rhs_ns_visc: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return launch_stage<double, kRhsOnly>(h, s, U, nullptr, R, nullptr,
                                          make_step_coef<double>(h, 0.0, rho, 0.0), sfs6, visc,
                                          rho);
  return launch_stage<float, kRhsOnly>(h, s, U, nullptr, R, nullptr,
                                       make_step_coef<float>(h, 0.0, rho, 0.0), sfs6, visc, rho);
}

int pf_rhs_ns(pf_box* h, const void* u, const void* v, const void* w, double mu, double rho,
              const void* const* sfs6, void* ru, void* rv, void* rw, void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {u, v, w};
  void* R[3] = {ru, rv, rw};
  if (int rc = check_ptrs(U, 3, "pf_rhs_ns")) return rc;
  if (int rc = check_ptrs((const void* const*)R, 3, "pf_rhs_ns")) return rc;
  if (sfs6)
    if (int rc = check_ptrs(sfs6, 6, "pf_rhs_ns(sfs6)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rhs_ns: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return launch_stage<double, kRhsOnly>(h, s, U, nullptr, R, nullptr,
                                          make_step_coef<double>(h, mu, rho, 0.0), sfs6);
  return launch_stage<float, kRhsOnly>(h, s, U, nullptr, R, nullptr,
                                       make_step_coef<float>(h, mu, rho, 0.0), sfs6);
}

int pf_source_inputs(pf_box* h, const void* u, const void* v, const void* w, void* ui, void* vi,
                     void* wi, void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {u, v, w};
  void* I[3] = {ui, vi, wi};
  if (int rc = check_ptrs(U, 3, "pf_source_inputs")) return rc;
  if (int rc = check_ptrs((const void* const*)I, 3, "pf_source_inputs")) return rc;
  for (int q = 0; q < 3; ++q)
    if (I[q] == U[0] || I[q] == U[1] || I[q] == U[2])
      return fail(PF_EINVAL, "pf_source_inputs: outputs must not alias the velocity");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? source_inputs_impl<double>(h, U, I, s)
                            : source_inputs_impl<float>(h, U, I, s);
}

int pf_rk4_source_inputs(pf_box* h, int stage, void* u, void* v, void* w, void* ut, void* vt,
                         void* wt, void* ui, void* vi, void* wi, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (stage < 1 || stage > 4) return fail(PF_EINVAL, "pf_rk4_source_inputs: stage must be 1..4");
  void* U[3] = {u, v, w};
  void* Ut[3] = {ut, vt, wt};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_rk4_source_inputs")) return rc;
  if (int rc = check_ptrs((const void* const*)Ut, 3, "pf_rk4_source_inputs")) return rc;
  const void* Y[3];
  for (int q = 0; q < 3; ++q) Y[q] = rk4_stage_input(h, stage, U, Ut, q);
  return pf_source_inputs(h, Y[0], Y[1], Y[2], ui, vi, wi, stream);
}

int pf_rk4_stage_src(pf_box* h, int stage, void* u, void* v, void* w, void* ut, void* vt,
                     void* wt, double dt, double mu, double rho, const void* gx, const void* gy,
                     const void* gz, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (stage < 1 || stage > 4) return fail(PF_EINVAL, "pf_rk4_stage_src: stage must be 1..4");
  void* U[3] = {u, v, w};
  void* Ut[3] = {ut, vt, wt};
  const void* G[3] = {gx, gy, gz};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_rk4_stage_src")) return rc;
  if (int rc = check_ptrs((const void* const*)Ut, 3, "pf_rk4_stage_src")) return rc;
  if (int rc = check_ptrs(G, 3, "pf_rk4_stage_src(G)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rk4_stage_src: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? rk4_stage_src_impl<double>(h, stage, U, Ut, dt, mu, rho, G, s)
                            : rk4_stage_src_impl<float>(h, stage, U, Ut, dt, mu, rho, G, s);
}

int pf_euler_step_src(pf_box* h, void* u, void* v, void* w, double dt, double mu, double rho,
                      const void* gx, const void* gy, const void* gz, void* stream) {
  if (int rc = check_box(h)) return rc;
  void* U[3] = {u, v, w};
  const void* G[3] = {gx, gy, gz};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_euler_step_src")) return rc;
  if (int rc = check_ptrs(G, 3, "pf_euler_step_src(G)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_euler_step_src: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  const void* const Uc[3] = {u, v, w};
  void* const B[3] = {h->ypong[0], h->ypong[1], h->ypong[2]};
  int rc;
  if (h->dtype == PF_F64)
    rc = launch_stage<double, kEuler>(h, s, Uc, Uc, B, nullptr,
                                      make_step_coef<double>(h, mu, rho, dt), nullptr, nullptr,
                                      rho, G);
  else
    rc = launch_stage<float, kEuler>(h, s, Uc, Uc, B, nullptr,
                                     make_step_coef<float>(h, mu, rho, dt), nullptr, nullptr, rho,
                                     G);
  if (rc) return rc;
  const size_t np = (size_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4) * h->esize;
  for (int q = 0; q < 3; ++q) PF_HIP(hipMemcpyAsync(U[q], B[q], np, hipMemcpyDeviceToDevice, s));
  return 0;
}

int pf_rhs_ns_src(pf_box* h, const void* u, const void* v, const void* w, double mu, double rho,
                  const void* gx, const void* gy, const void* gz, void* ru, void* rv, void* rw,
                  void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {u, v, w};
  void* R[3] = {ru, rv, rw};
  const void* G[3] = {gx, gy, gz};
  if (int rc = check_ptrs(U, 3, "pf_rhs_ns_src")) return rc;
  if (int rc = check_ptrs((const void* const*)R, 3, "pf_rhs_ns_src")) return rc;
  if (int rc = check_ptrs(G, 3, "pf_rhs_ns_src(G)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rhs_ns_src: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return launch_stage<double, kRhsOnly>(h, s, U, nullptr, R, nullptr,
                                          make_step_coef<double>(h, mu, rho, 0.0), nullptr,
                                          nullptr, rho, G);
  return launch_stage<float, kRhsOnly>(h, s, U, nullptr, R, nullptr,
                                       make_step_coef<float>(h, mu, rho, 0.0), nullptr, nullptr,
                                       rho, G);
}

static int adj_closure_args(pf_box* h, int kind, const float* params, void* work,
                            void* const* delta3, double* dparams, AdjClosure* cl,
                            const char* what) {
  cl->kind = kind;
  cl->params = params;
  cl->work = work;
  cl->dparams = dparams;
  for (int q = 0; q < 3; ++q) cl->delta[q] = delta3 ? delta3[q] : nullptr;
  if (kind < 0) return 0;
  if (kind > 2) return fail(PF_EINVAL, std::string(what) + ": closure kind is -1, 0, 1 or 2");
  if ((kind < 2 && !work) || !delta3 || !delta3[0] || !delta3[1] || !delta3[2])
    return fail(PF_EINVAL, std::string(what) + ": a closure needs work and delta3");
  if (kind >= 1 && !params)
    return fail(PF_EINVAL, std::string(what) + ": the network closure needs params");
  return 0;
}

static int rhs_adj_entry(pf_box* h, const void* u, const void* v, const void* w, const void* ua,
                         const void* va, const void* wa, double mu, const void* visc,
                         double rho, const void* const* delta3, void* ru, void* rv, void* rw,
                         void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {u, v, w};
  const void* A[3] = {ua, va, wa};
  void* R[3] = {ru, rv, rw};
  if (int rc = check_ptrs(U, 3, "pf_rhs_adj")) return rc;
  if (int rc = check_ptrs(A, 3, "pf_rhs_adj")) return rc;
  if (int rc = check_ptrs((const void* const*)R, 3, "pf_rhs_adj")) return rc;
  if (delta3)
    if (int rc = check_ptrs(delta3, 3, "pf_rhs_adj(delta3)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rhs_adj: rho must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return launch_adj_stage<double, kRhsOnly>(h, s, A, U, delta3, nullptr, R, nullptr,
                                              make_step_coef<double>(h, mu, rho, 0.0), visc,
                                              rho);
  return launch_adj_stage<float, kRhsOnly>(h, s, A, U, delta3, nullptr, R, nullptr,
                                           make_step_coef<float>(h, mu, rho, 0.0), visc, rho);
}

int pf_rhs_adj(pf_box* h, const void* u, const void* v, const void* w, const void* ua,
               const void* va, const void* wa, double mu, double rho, const void* const* delta3,
               void* ru, void* rv, void* rw, void* stream) {
  return rhs_adj_entry(h, u, v, w, ua, va, wa, mu, nullptr, rho, delta3, ru, rv, rw, stream);
}

int pf_rhs_adj_visc(pf_box* h, const void* u, const void* v, const void* w, const void* ua,
                    const void* va, const void* wa, const void* visc, double rho,
                    const void* const* delta3, void* ru, void* rv, void* rw, void* stream) {
  if (!visc) return fail(PF_EINVAL, "pf_rhs_adj_visc: NULL viscosity field");
  return rhs_adj_entry(h, u, v, w, ua, va, wa, 0.0, visc, rho, delta3, ru, rv, rw, stream);
}

static int adj_rk4_entry(pf_box* h, const void* u, const void* v, const void* w, void* ua,
                         void* va, void* wa, void* uta, void* vta, void* wta, double dt,
                         double mu, const void* visc, double rho, int closure,
                         const float* params, void* work, void* const* delta3, double* dparams,
                         void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {u, v, w};
  void* A[3] = {ua, va, wa};
  void* At[3] = {uta, vta, wta};
  if (int rc = check_ptrs(U, 3, "pf_adj_rk4_step")) return rc;
  if (int rc = check_ptrs((const void* const*)A, 3, "pf_adj_rk4_step")) return rc;
  if (int rc = check_ptrs((const void* const*)At, 3, "pf_adj_rk4_step")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_adj_rk4_step: rho must be nonzero");
  AdjClosure cl;
  if (int rc = adj_closure_args(h, closure, params, work, delta3, dparams, &cl, "pf_adj_rk4_step"))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? adj_rk4_impl<double>(h, U, A, At, dt, mu, rho, cl, s, visc)
                            : adj_rk4_impl<float>(h, U, A, At, dt, mu, rho, cl, s, visc);
}

int pf_adj_rk4_step(pf_box* h, const void* u, const void* v, const void* w, void* ua, void* va,
                    void* wa, void* uta, void* vta, void* wta, double dt, double mu, double rho,
                    int closure, const float* params, void* work, void* const* delta3,
                    double* dparams, void* stream) {
  return adj_rk4_entry(h, u, v, w, ua, va, wa, uta, vta, wta, dt, mu, nullptr, rho, closure,
                       params, work, delta3, dparams, stream);
}

int pf_adj_rk4_step_visc(pf_box* h, const void* u, const void* v, const void* w, void* ua,
                         void* va, void* wa, void* uta, void* vta, void* wta, double dt,
                         const void* visc, double rho, int closure, const float* params,
                         void* work, void* const* delta3, double* dparams, void* stream) {
  if (!visc) return fail(PF_EINVAL, "pf_adj_rk4_step_visc: NULL viscosity field");
  return adj_rk4_entry(h, u, v, w, ua, va, wa, uta, vta, wta, dt, 0.0, visc, rho, closure,
                       params, work, delta3, dparams, stream);
}

static int adj_euler_entry(pf_box* h, const void* u, const void* v, const void* w, void* ua,
                           void* va, void* wa, double dt, double mu, const void* visc,
                           double rho, int closure, const float* params, void* work,
                           void* const* delta3, double* dparams, void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {u, v, w};
  void* A[3] = {ua, va, wa};
  if (int rc = check_ptrs(U, 3, "pf_adj_euler_step")) return rc;
  if (int rc = check_ptrs((const void* const*)A, 3, "pf_adj_euler_step")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_adj_euler_step: rho must be nonzero");
  AdjClosure cl;
  if (int rc = adj_closure_args(h, closure, params, work, delta3, dparams, &cl,
                                "pf_adj_euler_step"))
    return rc;
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? adj_euler_impl<double>(h, U, A, dt, mu, rho, cl, s, visc)
                            : adj_euler_impl<float>(h, U, A, dt, mu, rho, cl, s, visc);
}

int pf_adj_euler_step(pf_box* h, const void* u, const void* v, const void* w, void* ua, void* va,
                      void* wa, double dt, double mu, double rho, int closure,
                      const float* params, void* work, void* const* delta3, double* dparams,
                      void* stream) {
  return adj_euler_entry(h, u, v, w, ua, va, wa, dt, mu, nullptr, rho, closure, params, work,
                         delta3, dparams, stream);
}

int pf_adj_euler_step_visc(pf_box* h, const void* u, const void* v, const void* w, void* ua,
                           void* va, void* wa, double dt, const void* visc, double rho,
                           int closure, const float* params, void* work, void* const* delta3,
                           double* dparams, void* stream) {
  if (!visc) return fail(PF_EINVAL, "pf_adj_euler_step_visc: NULL viscosity field");
  return adj_euler_entry(h, u, v, w, ua, va, wa, dt, 0.0, visc, rho, closure, params, work,
                         delta3, dparams, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free RK4 step (pf_rk4_step_noalloc).  Its workspace -- the three
// interior accumulators k1 + 2k2 + 2k3 and the second buffer of the stage
// states (three padded fields) that a box otherwise owns -- lives in one
// file-scope device array, so the call needs neither a box nor hipMalloc.
// Each GPU holds its own copy of the array for the life of the process; its
// device address is looked up once per device (hipGetSymbolAddress) and kept.
// Nothing in it needs clearing: stage 1 writes every accumulator cell before
// stages 2-4 read them, and stage 2 writes every cell of the second buffer
// (interior and periodic images) before stage 3 reads it.
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;   // 64 MiB per GPU
constexpr size_t kScratchAlign = 256;
constexpr int kScratchMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchMaxDevices];   // device address of `scratch`, per device

size_t scratch_round(size_t n) { return (n + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

// Bytes of scratch one step needs: 3 interior + 3 padded fields.
size_t rk4_scratch_need(int nx, int ny, int nz, size_t esize) {
  const size_t ni = (size_t)nx * ny * nz * esize;
  const size_t np = (size_t)(nx + 4) * (ny + 4) * (nz + 4) * esize;
  return 3 * scratch_round(ni) + 3 * scratch_round(np);
}

// Device address of `scratch` on the current device (looked up on the first
// call per device, with that device current).
int scratch_on_current_device(unsigned char** out) {
  int dev = -1;
  PF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchMaxDevices)
    return fail(PF_EINVAL, "pf_rk4_step_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(PF_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int64_t pf_rk4_step_noalloc_max_bytes(void) { return (int64_t)kScratchBytes; }

int pf_rk4_step_noalloc(int nx, int ny, int nz, double dx, double dy, double dz, int dtype,
                        void* u, void* v, void* w, void* ut, void* vt, void* wt, double dt,
                        double mu, double rho, const void* const* sfs6, void* stream) {
  // every argument is checked before the first HIP call
  if (nx < 4 || ny < 4 || nz < 4)
    return fail(PF_EINVAL, "pf_rk4_step_noalloc: every dimension must be >= 4 (two overlap cells)");
  if ((int64_t)nx * ny * nz >= (1ll << 31))
    return fail(PF_EINVAL, "pf_rk4_step_noalloc: nx*ny*nz must be < 2^31");
  if (dtype != PF_F64 && dtype != PF_F32) return fail(PF_EINVAL, "pf_rk4_step_noalloc: bad dtype");
  if (!(dx > 0) || !(dy > 0) || !(dz > 0))
    return fail(PF_EINVAL, "pf_rk4_step_noalloc: spacings must be > 0");
  void* U[3] = {u, v, w};
  void* Ut[3] = {ut, vt, wt};
  if (int rc = check_ptrs((const void* const*)U, 3, "pf_rk4_step_noalloc")) return rc;
  if (int rc = check_ptrs((const void* const*)Ut, 3, "pf_rk4_step_noalloc")) return rc;
  if (sfs6)
    if (int rc = check_ptrs(sfs6, 6, "pf_rk4_step_noalloc(sfs6)")) return rc;
  if (!(rho != 0.0)) return fail(PF_EINVAL, "pf_rk4_step_noalloc: rho must be nonzero");
  const size_t esize = dtype == PF_F64 ? 8 : 4;
  const size_t need = rk4_scratch_need(nx, ny, nz, esize);
  if (need > kScratchBytes)
    return fail(PF_EINVAL, "pf_rk4_step_noalloc: a " + std::to_string(nx) + "x" +
                               std::to_string(ny) + "x" + std::to_string(nz) + " box needs " +
                               std::to_string(need) + " bytes of scratch, more than the " +
                               std::to_string(kScratchBytes) + " it has (use pf_rk4_step)");
  unsigned char* base = nullptr;
  if (int rc = scratch_on_current_device(&base)) return rc;

  // A box view on the stack: geometry, spacing and the scratch as workspace.
  pf_box b;
  memset(&b, 0, sizeof(b));
  b.nx = nx; b.ny = ny; b.nz = nz;
  b.dx = dx; b.dy = dy; b.dz = dz;
  b.dtype = dtype;
  b.esize = esize;
  b.g = make_geom(nx, ny, nz);
  size_t off = 0;
  const size_t ni = (size_t)nx * ny * nz * esize;
  const size_t np = (size_t)(nx + 4) * (ny + 4) * (nz + 4) * esize;
  for (int q = 0; q < 3; ++q) {
    b.acc[q] = base + off;
    off += scratch_round(ni);
  }
  for (int q = 0; q < 3; ++q) {
    b.ypong[q] = base + off;
    off += scratch_round(np);
  }
  hipStream_t s = (hipStream_t)stream;
  return dtype == PF_F64 ? rk4_impl<double>(&b, U, Ut, dt, mu, rho, sfs6, s)
                         : rk4_impl<float>(&b, U, Ut, dt, mu, rho, sfs6, s);
}

}  // extern "C"
