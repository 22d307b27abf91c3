// Hand-written reverse-mode kernels of the time step (vector-Jacobian
// products), for closure training: the Navier-Stokes RHS with respect to the
// velocity and to the closure stresses, and the closures with respect to the
// velocity (and the MLP parameters).  The projection P = I - G L^-1 D is
// self-adjoint on the periodic staggered grid (G^T = -D), so its reverse is
// the forward correction kernels applied to the adjoint field.
//
// All adjoint fields are padded arrays whose overlap cells hold periodic
// images, like the forward states.  Reverse passes never scatter with
// atomics: a first kernel computes, per cell, the partial derivatives with
// respect to that cell's stencil inputs into a workspace, and a second kernel
// gathers them per output cell in a fixed order, so the results are
// deterministic.
#include <hip/hip_runtime.h>

#include "pf_box.h"
#include "pf_features.h"

namespace pf {

// Stencil inputs of the RHS at a cell, in this order (offsets di, dj, dk):
//  u: 000 p00 m00 0p0 0m0 00p 00m mp0 pm0 m0p p0m
//  v: 000 p00 m00 0p0 0m0 00p 00m mp0 pm0 0mp 0pm
//  w: 000 p00 m00 0p0 0m0 00p 00m m0p p0m 0mp 0pm
constexpr int kRhsIn = 33;
__constant__ int c_off[kRhsIn][3] = {
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {-1, 1, 0}, {1, -1, 0}, {-1, 0, 1}, {1, 0, -1},
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {-1, 1, 0}, {1, -1, 0}, {0, -1, 1}, {0, 1, -1},
    {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {-1, 0, 1}, {1, 0, -1}, {0, -1, 1}, {0, 1, -1}};

template <typename T>
struct Coef {
  T dxi, dyi, dzi, c2, exy, exz, eyz;
};

template <typename T>
Coef<T> make_coef(const pf_box* h, double mu, double rho) {
  Coef<T> c;
  c.dxi = (T)(1.0 / h->dx);
  c.dyi = (T)(1.0 / h->dy);
  c.dzi = (T)(1.0 / h->dz);
  const T V = (T)mu, R = (T)rho;
  c.c2 = ((T)2.0 * V) / R;
  const T edge = (T)0.25 * (((V + V) + V) + V);
  c.exy = edge / R;
  c.exz = edge / R;
  c.eyz = edge / R;
  return c;
}

// Local reverse of the velocity dependence of the fused RHS (rhs_cell in
// pf_step.hip; velocity.py:234-355): g[k] = sum_c lam_c d R_c / d x_k for the
// 33 stencil inputs x_k.  The closure stresses enter the RHS additively and
// are handled by sfs_vjp_kernel.
template <typename T>
__device__ __forceinline__ void rhs_cell_vjp(const T* __restrict__ U, const T* __restrict__ V,
                                             const T* __restrict__ W, int64_t c, int64_t sx,
                                             int64_t sy, const Coef<T>& k, T lu, T lv, T lw,
                                             T (&g)[kRhsIn]) {
  const T dxi = k.dxi, dyi = k.dyi, dzi = k.dzi, half = (T)0.5;
#define LD(A, di, dj, dk) A[c + (di) * sx + (dj) * sy + (dk)]
  const T u000 = LD(U, 0, 0, 0), up00 = LD(U, 1, 0, 0), um00 = LD(U, -1, 0, 0);
  const T u0p0 = LD(U, 0, 1, 0), u0m0 = LD(U, 0, -1, 0), u00p = LD(U, 0, 0, 1), u00m = LD(U, 0, 0, -1);
  const T upm0 = LD(U, 1, -1, 0), up0m = LD(U, 1, 0, -1);
  const T v000 = LD(V, 0, 0, 0), vp00 = LD(V, 1, 0, 0), vm00 = LD(V, -1, 0, 0);
  const T v0p0 = LD(V, 0, 1, 0), v0m0 = LD(V, 0, -1, 0), v00p = LD(V, 0, 0, 1), v00m = LD(V, 0, 0, -1);
  const T vmp0 = LD(V, -1, 1, 0), v0pm = LD(V, 0, 1, -1);
  const T w000 = LD(W, 0, 0, 0), wp00 = LD(W, 1, 0, 0), wm00 = LD(W, -1, 0, 0);
  const T w0p0 = LD(W, 0, 1, 0), w0m0 = LD(W, 0, -1, 0), w00p = LD(W, 0, 0, 1), w00m = LD(W, 0, 0, -1);
  const T wm0p = LD(W, -1, 0, 1), w0mp = LD(W, 0, -1, 1);
#undef LD
  enum { U000, UP00, UM00, U0P0, U0M0, U00P, U00M, UMP0, UPM0, UM0P, UP0M,
         V000, VP00, VM00, V0P0, V0M0, V00P, V00M, VMP0, VPM0, V0MP, V0PM,
         W000, WP00, WM00, W0P0, W0M0, W00P, W00M, WM0P, WP0M, W0MP, W0PM };
  for (int q = 0; q < kRhsIn; ++q) g[q] = (T)0;
  T Dc = 0, Dxm = 0, Dym = 0, Dzm = 0;   // adjoints of the four div/3 values
  T t;
  // ---- u: viscous (FX, FY, FZ at the cell and its lower neighbour) ----
  t = lu * dxi * k.c2;  g[UP00] += t * dxi; g[U000] -= t * dxi; Dc -= t;
  t = -lu * dxi * k.c2; g[U000] += t * dxi; g[UM00] -= t * dxi; Dxm -= t;
  t = lu * dyi * k.exy;  g[U0P0] += t * dyi; g[U000] -= t * dyi; g[VP00] += t * dxi; g[V000] -= t * dxi;
  t = -lu * dyi * k.exy; g[U000] += t * dyi; g[U0M0] -= t * dyi; g[VPM0] += t * dxi; g[V0M0] -= t * dxi;
  t = lu * dzi * k.exz;  g[U00P] += t * dzi; g[U000] -= t * dzi; g[WP00] += t * dxi; g[W000] -= t * dxi;
  t = -lu * dzi * k.exz; g[U000] += t * dzi; g[U00M] -= t * dzi; g[WP0M] += t * dxi; g[W00M] -= t * dxi;
  // ---- u: advection (rhs -= A) ----
  {
    const T A = -lu;
    const T a = half * (up00 + u000), b = half * (u000 + um00);
    const T ab = A * dxi * (T)2 * a, bb = -A * dxi * (T)2 * b;
    g[UP00] += half * ab; g[U000] += half * ab; g[U000] += half * bb; g[UM00] += half * bb;
    T s = A * dyi;
    const T uy_p = half * (u0p0 + u000), vx_p = half * (v0p0 + vmp0);
    const T uy_c = half * (u000 + u0m0), vx_c = half * (v000 + vm00);
    T x1 = s * vx_p, x2 = s * uy_p, x3 = -s * vx_c, x4 = -s * uy_c;
    g[U0P0] += half * x1; g[U000] += half * x1; g[V0P0] += half * x2; g[VMP0] += half * x2;
    g[U000] += half * x3; g[U0M0] += half * x3; g[V000] += half * x4; g[VM00] += half * x4;
    s = A * dzi;
    const T uz_p = half * (u00p + u000), wx_p = half * (w00p + wm0p);
    const T uz_c = half * (u000 + u00m), wx_c = half * (w000 + wm00);
    x1 = s * wx_p; x2 = s * uz_p; x3 = -s * wx_c; x4 = -s * uz_c;
    g[U00P] += half * x1; g[U000] += half * x1; g[W00P] += half * x2; g[WM0P] += half * x2;
    g[U000] += half * x3; g[U00M] += half * x3; g[W000] += half * x4; g[WM00] += half * x4;
  }
  // ---- v: viscous ----
  t = lv * dxi * k.exy;  g[VP00] += t * dxi; g[V000] -= t * dxi; g[U0P0] += t * dyi; g[U000] -= t * dyi;
  t = -lv * dxi * k.exy; g[V000] += t * dxi; g[VM00] -= t * dxi; g[UMP0] += t * dyi; g[UM00] -= t * dyi;
  t = lv * dyi * k.c2;   g[V0P0] += t * dyi; g[V000] -= t * dyi; Dc -= t;
  t = -lv * dyi * k.c2;  g[V000] += t * dyi; g[V0M0] -= t * dyi; Dym -= t;
  t = lv * dzi * k.eyz;  g[V00P] += t * dzi; g[V000] -= t * dzi; g[W0P0] += t * dyi; g[W000] -= t * dyi;
  t = -lv * dzi * k.eyz; g[V000] += t * dzi; g[V00M] -= t * dzi; g[W0PM] += t * dyi; g[W00M] -= t * dyi;
  // ---- v: advection ----
  {
    const T A = -lv;
    T s = A * dxi;
    const T vx_p1 = half * (vp00 + v000), uy_p1 = half * (up00 + upm0);
    const T vx_c = half * (v000 + vm00), uy_c = half * (u000 + u0m0);
    T x1 = s * uy_p1, x2 = s * vx_p1, x3 = -s * uy_c, x4 = -s * vx_c;
    g[VP00] += half * x1; g[V000] += half * x1; g[UP00] += half * x2; g[UPM0] += half * x2;
    g[V000] += half * x3; g[VM00] += half * x3; g[U000] += half * x4; g[U0M0] += half * x4;
    const T a = half * (v0p0 + v000), b = half * (v000 + v0m0);
    const T ab = A * dyi * (T)2 * a, bb = -A * dyi * (T)2 * b;
    g[V0P0] += half * ab; g[V000] += half * ab; g[V000] += half * bb; g[V0M0] += half * bb;
    s = A * dzi;
    const T vz_p = half * (v00p + v000), wy_p = half * (w00p + w0mp);
    const T vz_c = half * (v000 + v00m), wy_c = half * (w000 + w0m0);
    x1 = s * wy_p; x2 = s * vz_p; x3 = -s * wy_c; x4 = -s * vz_c;
    g[V00P] += half * x1; g[V000] += half * x1; g[W00P] += half * x2; g[W0MP] += half * x2;
    g[V000] += half * x3; g[V00M] += half * x3; g[W000] += half * x4; g[W0M0] += half * x4;
  }
  // ---- w: viscous ----
  t = lw * dxi * k.exz;  g[WP00] += t * dxi; g[W000] -= t * dxi; g[U00P] += t * dzi; g[U000] -= t * dzi;
  t = -lw * dxi * k.exz; g[W000] += t * dxi; g[WM00] -= t * dxi; g[UM0P] += t * dzi; g[UM00] -= t * dzi;
  t = lw * dyi * k.eyz;  g[W0P0] += t * dyi; g[W000] -= t * dyi; g[V00P] += t * dzi; g[V000] -= t * dzi;
  t = -lw * dyi * k.eyz; g[W000] += t * dyi; g[W0M0] -= t * dyi; g[V0MP] += t * dzi; g[V0M0] -= t * dzi;
  t = lw * dzi * k.c2;   g[W00P] += t * dzi; g[W000] -= t * dzi; Dc -= t;
  t = -lw * dzi * k.c2;  g[W000] += t * dzi; g[W00M] -= t * dzi; Dzm -= t;
  // ---- w: advection ----
  {
    const T A = -lw;
    T s = A * dxi;
    const T wx_p1 = half * (wp00 + w000), uz_p1 = half * (up00 + up0m);
    const T wx_c = half * (w000 + wm00), uz_c = half * (u000 + u00m);
    T x1 = s * uz_p1, x2 = s * wx_p1, x3 = -s * uz_c, x4 = -s * wx_c;
    g[WP00] += half * x1; g[W000] += half * x1; g[UP00] += half * x2; g[UP0M] += half * x2;
    g[W000] += half * x3; g[WM00] += half * x3; g[U000] += half * x4; g[U00M] += half * x4;
    s = A * dyi;
    const T wy_p1 = half * (w0p0 + w000), vz_p1 = half * (v0p0 + v0pm);
    const T wy_c = half * (w000 + w0m0), vz_c = half * (v000 + v00m);
    x1 = s * vz_p1; x2 = s * wy_p1; x3 = -s * vz_c; x4 = -s * wy_c;
    g[W0P0] += half * x1; g[W000] += half * x1; g[V0P0] += half * x2; g[V0PM] += half * x2;
    g[W000] += half * x3; g[W0M0] += half * x3; g[V000] += half * x4; g[V00M] += half * x4;
    const T a = half * (w00p + w000), b = half * (w000 + w00m);
    const T ab = A * dzi * (T)2 * a, bb = -A * dzi * (T)2 * b;
    g[W00P] += half * ab; g[W000] += half * ab; g[W000] += half * bb; g[W00M] += half * bb;
  }
  // ---- the four (div u)/3 values ----
  const T third = (T)1 / (T)3;
  T d = Dc * third;
  g[UP00] += d * dxi; g[U000] -= d * dxi; g[V0P0] += d * dyi; g[V000] -= d * dyi; g[W00P] += d * dzi; g[W000] -= d * dzi;
  d = Dxm * third;
  g[U000] += d * dxi; g[UM00] -= d * dxi; g[VMP0] += d * dyi; g[VM00] -= d * dyi; g[WM0P] += d * dzi; g[WM00] -= d * dzi;
  d = Dym * third;
  g[UPM0] += d * dxi; g[U0M0] -= d * dxi; g[V000] += d * dyi; g[V0M0] -= d * dyi; g[W0MP] += d * dzi; g[W0M0] -= d * dzi;
  d = Dzm * third;
  g[UP0M] += d * dxi; g[U00M] -= d * dxi; g[V0PM] += d * dyi; g[V00M] -= d * dyi; g[W000] += d * dzi; g[W00M] -= d * dzi;
}

// Pass 1: per interior cell, the 33 local input adjoints into work[k][cell].
template <typename T>
__global__ void __launch_bounds__(kBlock) rhs_vjp_local_kernel(Geom g, CVec3<T> Y, CVec3<T> L,
                                                               Coef<T> k, T* __restrict__ work) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, kk;
    decode(g, idx, i, j, kk);
    const int64_t c = i * g.psx + j * g.psy + kk;
    T gl[kRhsIn];
    rhs_cell_vjp(Y.a[0], Y.a[1], Y.a[2], c, g.psx, g.psy, k, L.a[0][c], L.a[1][c], L.a[2][c], gl);
#pragma unroll
    for (int q = 0; q < kRhsIn; ++q) work[(int64_t)q * g.ncell + idx] = gl[q];
  }
}

// Pass 2: out_c(j) += sum over the inputs of component c of work[k][j - off_k]
// (periodic), and kprev_c(j) += h * that sum when kprev is given.
template <typename T>
__global__ void __launch_bounds__(kBlock) rhs_vjp_gather_kernel(Geom g, const T* __restrict__ work,
                                                                Vec3<T> out, Vec3<T> kprev, T h) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, kk;
    decode(g, idx, i, j, kk);
    const int64_t c = pidx(g, i, j, kk);
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      T s = (T)0;
#pragma unroll
      for (int q = 0; q < 11; ++q) {
        const int kq = comp * 11 + q;
        int a = i - c_off[kq][0], b = j - c_off[kq][1], e = kk - c_off[kq][2];
        a = a < 0 ? a + g.nx : (a >= g.nx ? a - g.nx : a);
        b = b < 0 ? b + g.ny : (b >= g.ny ? b - g.ny : b);
        e = e < 0 ? e + g.nz : (e >= g.nz ? e - g.nz : e);
        s = s + work[(int64_t)kq * g.ncell + a * g.isx + b * g.isy + e];
      }
      store_padded(g, out.a[comp], i, j, kk, out.a[comp][c] + s);
      if (kprev.a[comp]) store_padded(g, kprev.a[comp], i, j, kk, kprev.a[comp][c] + h * s);
    }
  }
}

// Adjoint of the closure stresses read by the RHS (velocity.py:262-315):
// for every padded cell q, with L(q) = lambda at the interior cell of q
// (zero outside the interior),
//   FXX(q) += dxi*(Lu(q) - Lu(q+x)),  FYY(q) += dyi*(Lv(q) - Lv(q+y)),
//   FZZ(q) += dzi*(Lw(q) - Lw(q+z)),
//   FXY(q) += 0.25*sum of Exy over {q, q+x, q+y, q+x+y}, with the edge
//   adjoint Exy(c) = dyi*(Lu(c) - Lu(c+y)) + dxi*(Lv(c) - Lv(c+x)), and the
//   same pattern for FXZ (u,w; x,z) and FYZ (v,w; y,z).
template <typename T>
__device__ __forceinline__ T lam_at(const Geom& g, const T* L, int a, int b, int e) {
  // padded coordinates -> interior value or 0
  const int i = a - kNover, j = b - kNover, k = e - kNover;
  if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) return (T)0;
  return L[pidx(g, i, j, k)];
}

template <typename T, bool ACC>
__global__ void __launch_bounds__(kBlock) sfs_vjp_kernel(Geom g, CVec3<T> L, Coef<T> k,
                                                         Vec6<T> tb) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t total = (int64_t)nxo * nyo * nzo;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int e = (int)(q % nzo);
    const int64_t t1 = q / nzo;
    const int b = (int)(t1 % nyo);
    const int a = (int)(t1 / nyo);
    const T* Lu = L.a[0];
    const T* Lv = L.a[1];
    const T* Lw = L.a[2];
    auto lu = [&](int x, int y, int z) { return lam_at(g, Lu, x, y, z); };
    auto lv = [&](int x, int y, int z) { return lam_at(g, Lv, x, y, z); };
    auto lw = [&](int x, int y, int z) { return lam_at(g, Lw, x, y, z); };
    auto exy = [&](int x, int y, int z) {
      return k.dyi * (lu(x, y, z) - lu(x, y + 1, z)) + k.dxi * (lv(x, y, z) - lv(x + 1, y, z));
    };
    auto exz = [&](int x, int y, int z) {
      return k.dzi * (lu(x, y, z) - lu(x, y, z + 1)) + k.dxi * (lw(x, y, z) - lw(x + 1, y, z));
    };
    auto eyz = [&](int x, int y, int z) {
      return k.dzi * (lv(x, y, z) - lv(x, y, z + 1)) + k.dyi * (lw(x, y, z) - lw(x, y + 1, z));
    };
    const T q4 = (T)0.25;
    T t[6];
    t[0] = k.dxi * (lu(a, b, e) - lu(a + 1, b, e));
    t[1] = k.dyi * (lv(a, b, e) - lv(a, b + 1, e));
    t[2] = k.dzi * (lw(a, b, e) - lw(a, b, e + 1));
    t[3] = q4 * (exy(a, b, e) + exy(a + 1, b, e) + exy(a, b + 1, e) + exy(a + 1, b + 1, e));
    t[4] = q4 * (exz(a, b, e) + exz(a + 1, b, e) + exz(a, b, e + 1) + exz(a + 1, b, e + 1));
    t[5] = q4 * (eyz(a, b, e) + eyz(a, b + 1, e) + eyz(a, b, e + 1) + eyz(a, b + 1, e + 1));
#pragma unroll
    for (int m = 0; m < 6; ++m) tb.a[m][q] = ACC ? tb.a[m][q] + t[m] : t[m];
  }
}

// Fold a padded adjoint field onto its interior (transpose of the periodic
// overlap fill): every interior cell adds the values held by its images,
// then the images are refreshed.
template <typename T>
__global__ void __launch_bounds__(kBlock) fold_halo_kernel(Geom g, T* __restrict__ a) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    int ii[2], jj[2], kk[2];
    int ni = 1, nj = 1, nk = 1;
    ii[0] = i; jj[0] = j; kk[0] = k;
    if (i < kNover) ii[ni++] = i + g.nx; else if (i >= g.nx - kNover) ii[ni++] = i - g.nx;
    if (j < kNover) jj[nj++] = j + g.ny; else if (j >= g.ny - kNover) jj[nj++] = j - g.ny;
    if (k < kNover) kk[nk++] = k + g.nz; else if (k >= g.nz - kNover) kk[nk++] = k - g.nz;
    T s = (T)0;
    for (int x = 0; x < ni; ++x)
      for (int y = 0; y < nj; ++y)
        for (int z = 0; z < nk; ++z) s = s + a[pidx(g, ii[x], jj[y], kk[z])];
    for (int x = 0; x < ni; ++x)
      for (int y = 0; y < nj; ++y)
        for (int z = 0; z < nk; ++z) a[pidx(g, ii[x], jj[y], kk[z])] = s;
  }
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_rhs_vjp(pf_box* h, const void* u, const void* v, const void* w, double mu, double rho,
               const void* lu, const void* lv, const void* lw, void* work, void* ou, void* ov,
               void* ow, void* ku, void* kv, void* kw, double hcoef, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !lu || !lv || !lw || !work || !ou || !ov || !ow)
    return fail(PF_EINVAL, "pf_rhs_vjp: NULL pointer");
  if ((ku || kv || kw) && !(ku && kv && kw)) return fail(PF_EINVAL, "pf_rhs_vjp: kprev needs all three");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = h->g;
  const int nb = launch_blocks(g);
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    CVec3<T> Y = {{(const T*)u + g.poff, (const T*)v + g.poff, (const T*)w + g.poff}};
    CVec3<T> L = {{(const T*)lu + g.poff, (const T*)lv + g.poff, (const T*)lw + g.poff}};
    hipLaunchKernelGGL(rhs_vjp_local_kernel<T>, dim3(nb), dim3(kBlock), 0, s, g, Y, L,
                       make_coef<T>(h, mu, rho), (T*)work);
    PF_LAUNCH_CHECK("rhs_vjp_local_kernel");
    Vec3<T> O = {{(T*)ou, (T*)ov, (T*)ow}};
    Vec3<T> K = {{(T*)ku, (T*)kv, (T*)kw}};
    hipLaunchKernelGGL(rhs_vjp_gather_kernel<T>, dim3(nb), dim3(kBlock), 0, s, g,
                       (const T*)work, O, K, (T)hcoef);
    PF_LAUNCH_CHECK("rhs_vjp_gather_kernel");
    return 0;
  };
  return h->dtype == PF_F64 ? run(0.0) : run(0.0f);
}

int pf_sfs_vjp(pf_box* h, const void* lu, const void* lv, const void* lw, void* const* taubar6,
               int accumulate, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!lu || !lv || !lw || !taubar6) return fail(PF_EINVAL, "pf_sfs_vjp: NULL pointer");
  for (int m = 0; m < 6; ++m)
    if (!taubar6[m]) return fail(PF_EINVAL, "pf_sfs_vjp: NULL stress adjoint");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = h->g;
  const int64_t total = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nb = (total + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    CVec3<T> L = {{(const T*)lu, (const T*)lv, (const T*)lw}};
    Vec6<T> tb;
    for (int m = 0; m < 6; ++m) tb.a[m] = (T*)taubar6[m];
    if (accumulate)
      hipLaunchKernelGGL((sfs_vjp_kernel<T, true>), dim3((int)nb), dim3(kBlock), 0, s, g, L,
                         make_coef<T>(h, 0.0, 1.0), tb);
    else
      hipLaunchKernelGGL((sfs_vjp_kernel<T, false>), dim3((int)nb), dim3(kBlock), 0, s, g, L,
                         make_coef<T>(h, 0.0, 1.0), tb);
    PF_LAUNCH_CHECK("sfs_vjp_kernel");
    return 0;
  };
  return h->dtype == PF_F64 ? run(0.0) : run(0.0f);
}

int pf_fold_halo(pf_box* h, void* field, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!field) return fail(PF_EINVAL, "pf_fold_halo: NULL field");
  hipStream_t s = (hipStream_t)stream;
  const int nb = launch_blocks(h->g);
  if (h->dtype == PF_F64)
    hipLaunchKernelGGL(fold_halo_kernel<double>, dim3(nb), dim3(kBlock), 0, s, h->g, (double*)field);
  else
    hipLaunchKernelGGL(fold_halo_kernel<float>, dim3(nb), dim3(kBlock), 0, s, h->g, (float*)field);
  PF_LAUNCH_CHECK("fold_halo_kernel");
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Closure reverse passes.  Features are the nine cell-centred gradients of
// grad_center (pf_closure.hip); their reverse includes the stale-value rule
// at overlap index 0.
namespace pf {

constexpr int kNp = 746;                       // MLP parameters
constexpr int kW1 = 0, kB1 = 180, kW2 = 200, kB2 = 600, kW3 = 620, kB3 = 740;

struct MlpCache {
  float x[9], h1[20], h2[20];
};

__device__ __forceinline__ void mlp_forward_f32(const float* __restrict__ P, const float (&x)[9],
                                                MlpCache& c, float (&o)[6]) {
  for (int f = 0; f < 9; ++f) c.x[f] = x[f];
  for (int r = 0; r < 20; ++r) {
    float z = P[kB1 + r];
    for (int k = 0; k < 9; ++k) z = fmaf(P[kW1 + r * 9 + k], x[k], z);
    c.h1[r] = tanhf(z);
  }
  for (int r = 0; r < 20; ++r) {
    float z = P[kB2 + r];
    for (int k = 0; k < 20; ++k) z = fmaf(P[kW2 + r * 20 + k], c.h1[k], z);
    c.h2[r] = tanhf(z);
  }
  for (int r = 0; r < 6; ++r) {
    float z = P[kB3 + r];
    for (int k = 0; k < 20; ++k) z = fmaf(P[kW3 + r * 20 + k], c.h2[k], z);
    o[r] = z;
  }
}

// FP32 forward of the closure MLP (the precision used when its gradients are
// checked): same features and layout as the MFMA kernel.
template <typename T>
__global__ void __launch_bounds__(kBlock) mlp_f32_kernel(Geom g, const T* __restrict__ U,
                                                         const T* __restrict__ V,
                                                         const T* __restrict__ W, T dxi, T dyi,
                                                         T dzi, const float* __restrict__ P,
                                                         float cst, Vec6<T> out) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t total = (int64_t)nxo * nyo * nzo;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int k0 = (int)(q % nzo);
    const int64_t t1 = q / nzo;
    const int b0 = (int)(t1 % nyo);
    const int a0 = (int)(t1 / nyo);
    const int a = a0 == nxo - 1 ? 3 : a0, b = b0 == nyo - 1 ? 3 : b0, k = k0 == nzo - 1 ? 3 : k0;
    T G[3][3];
    grad_center(U, V, W, g.psx, g.psy, a, b, k, dxi, dyi, dzi, G);
    float x[9];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) x[3 * c + d] = (float)G[c][d];
    MlpCache mc;
    float o[6];
    mlp_forward_f32(P, x, mc, o);
    for (int m = 0; m < 6; ++m) out.a[m][q] = (T)(o[m] * cst);
  }
}

// Per padded cell: feature adjoints Gbar (9 arrays) from the stress adjoints;
// for the MLP also the parameter-gradient partials of this wave.
// KIND 0: gradient model (sfsmodel_gradient.py:77-148); 1: MLP (FP32).
template <typename T, int KIND, bool PGRAD>
__global__ void __launch_bounds__(64) closure_vjp_local_kernel(
    Geom g, const T* __restrict__ U, const T* __restrict__ V, const T* __restrict__ W, T dxi,
    T dyi, T dzi, T cst, Vec6<T> tb, const float* __restrict__ P, T* __restrict__ gbar,
    double* __restrict__ wpart) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t total = (int64_t)nxo * nyo * nzo;
  const int lane = threadIdx.x;
  constexpr int kNs = (kNp + 63) / 64;
  double acc[kNs];
  // LDS row of a cell: x (9), h1, h2, z1bar, z2bar (20 each), obar (6), 1
  constexpr int kSx = 0, kSh1 = 9, kSh2 = 29, kSz1 = 49, kSz2 = 69, kSob = 89, kSone = 95;
  __shared__ float sv[PGRAD ? 64 : 1][96];
  // the two row entries whose product is parameter t * 64 + lane's term
  int ia[kNs], ib[kNs];
#pragma unroll
  for (int t = 0; t < kNs; ++t) {
    acc[t] = 0.0;
    const int p = t * 64 + lane;
    if (p < kB1) { ia[t] = kSz1 + p / 9; ib[t] = kSx + p % 9; }
    else if (p < kW2) { ia[t] = kSz1 + (p - kB1); ib[t] = kSone; }
    else if (p < kB2) { ia[t] = kSz2 + (p - kW2) / 20; ib[t] = kSh1 + (p - kW2) % 20; }
    else if (p < kW3) { ia[t] = kSz2 + (p - kB2); ib[t] = kSone; }
    else if (p < kB3) { ia[t] = kSob + (p - kW3) / 20; ib[t] = kSh2 + (p - kW3) % 20; }
    else { ia[t] = kSob + (p < kNp ? p - kB3 : 0); ib[t] = kSone; }
  }
  for (int64_t base = (int64_t)blockIdx.x * 64; base < total; base += (int64_t)gridDim.x * 64) {
    const int64_t q = base + lane;
    const bool valid = q < total;
    float ob[6] = {0, 0, 0, 0, 0, 0}, z2b[20], z1b[20];
    MlpCache mc;
    for (int r = 0; r < 20; ++r) { z2b[r] = 0.f; z1b[r] = 0.f; mc.h1[r] = 0.f; mc.h2[r] = 0.f; }
    for (int f = 0; f < 9; ++f) mc.x[f] = 0.f;
    if (valid) {
      const int k0 = (int)(q % nzo);
      const int64_t t1 = q / nzo;
      const int b0 = (int)(t1 % nyo);
      const int a0 = (int)(t1 / nyo);
      const int a = a0 == nxo - 1 ? 3 : a0, b = b0 == nyo - 1 ? 3 : b0, k = k0 == nzo - 1 ? 3 : k0;
      T G[3][3];
      grad_center(U, V, W, g.psx, g.psy, a, b, k, dxi, dyi, dzi, G);
      T Gb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      T tau[6];
      for (int m = 0; m < 6; ++m) tau[m] = tb.a[m][q];
      if (KIND == 0) {
        const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
        for (int m = 0; m < 6; ++m) {
          const int p = pr[m][0], r = pr[m][1];
          const T s = tau[m] * cst;
          for (int d = 0; d < 3; ++d) {
            Gb[p][d] += s * G[r][d];
            Gb[r][d] += s * G[p][d];
          }
        }
      } else {
        float x[9];
        for (int c = 0; c < 3; ++c)
          for (int d = 0; d < 3; ++d) x[3 * c + d] = (float)G[c][d];
        float o[6];
        mlp_forward_f32(P, x, mc, o);
        for (int m = 0; m < 6; ++m) ob[m] = (float)tau[m] * (float)cst;
        for (int k = 0; k < 20; ++k) {
          float s = 0.f;
          for (int r = 0; r < 6; ++r) s = fmaf(P[kW3 + r * 20 + k], ob[r], s);
          z2b[k] = s * (1.f - mc.h2[k] * mc.h2[k]);
        }
        for (int k = 0; k < 20; ++k) {
          float s = 0.f;
          for (int r = 0; r < 20; ++r) s = fmaf(P[kW2 + r * 20 + k], z2b[r], s);
          z1b[k] = s * (1.f - mc.h1[k] * mc.h1[k]);
        }
        for (int f = 0; f < 9; ++f) {
          float s = 0.f;
          for (int r = 0; r < 20; ++r) s = fmaf(P[kW1 + r * 9 + f], z1b[r], s);
          Gb[f / 3][f % 3] = (T)s;
        }
      }
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) gbar[(int64_t)(3 * c + d) * total + q] = Gb[c][d];
    }
    if (KIND == 1 && PGRAD) {
      // parameter gradients: sums over the wave's 64 cells of per-cell outer
      // products (float products, double sums).  The per-cell vectors go to
      // LDS, and lane l sums parameters l, l + 64, ... over the cells in a
      // fixed order.
      float* row = sv[lane];
      for (int f = 0; f < 9; ++f) row[kSx + f] = mc.x[f];
      for (int r = 0; r < 20; ++r) {
        row[kSh1 + r] = mc.h1[r];
        row[kSh2 + r] = mc.h2[r];
        row[kSz1 + r] = z1b[r];
        row[kSz2 + r] = z2b[r];
      }
      for (int m = 0; m < 6; ++m) row[kSob + m] = ob[m];
      row[kSone] = 1.f;
      __syncthreads();   // the block is one wave
#pragma unroll
      for (int t = 0; t < kNs; ++t) {
        if (t * 64 + lane < kNp) {
          double sum = 0.0;
#pragma unroll 8
          for (int c = 0; c < 64; ++c) sum += (double)(sv[c][ia[t]] * sv[c][ib[t]]);
          acc[t] += sum;
        }
      }
      __syncthreads();   // before the next cells overwrite the rows
    }
  }
  if (KIND == 1 && PGRAD) {
#pragma unroll
    for (int t = 0; t < kNs; ++t) {
      const int p = t * 64 + lane;
      if (p < kNp) wpart[(int64_t)blockIdx.x * kNp + p] = acc[t];
    }
  }
}

// dtheta[p] += sum over waves of wpart[w][p] (fixed order): one block per
// parameter, its threads striding over the waves, then a block reduction.
__global__ void __launch_bounds__(kBlock) mlp_param_reduce_kernel(const double* __restrict__ wpart,
                                                                  int nw, double* __restrict__ dth) {
  const int p = blockIdx.x;
  double v[1] = {0.0};
  for (int w = threadIdx.x; w < nw; w += kBlock) v[0] += wpart[(int64_t)w * kNp + p];
  const int ops[1] = {0};
  block_reduce<1, kBlock>(v, ops);
  if (threadIdx.x == 0) dth[p] += v[0];
}

// Reverse of the feature stencil (grad_center): fbar_c(p) for the padded
// cell p from Gbar at the cells whose features read p.
template <typename T>
__global__ void __launch_bounds__(kBlock) closure_vjp_gather_kernel(Geom g, const T* __restrict__ gbar,
                                                                    T dxi, T dyi, T dzi,
                                                                    Vec3<T> fbar) {
  const int n[3] = {g.nx + 2 * kNover, g.ny + 2 * kNover, g.nz + 2 * kNover};
  const int64_t total = (int64_t)n[0] * n[1] * n[2];
  const int64_t st[3] = {(int64_t)n[1] * n[2], n[2], 1};
  const T hinv[3] = {dxi, dyi, dzi};
  const T half = (T)0.5;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * blockDim.x) {
    int pos[3];
    pos[2] = (int)(p % n[2]);
    const int64_t t1 = p / n[2];
    pos[1] = (int)(t1 % n[1]);
    pos[0] = (int)(t1 / n[1]);
    // Gbar of feature (c,d) at the cell displaced by `dd` from p, or 0 outside the array
    auto GB = [&](int c, int d, const int (&dd)[3]) -> T {
      int64_t q = p;
      for (int a = 0; a < 3; ++a) {
        const int x = pos[a] + dd[a];
        if (x < 0 || x >= n[a]) return (T)0;
        q += dd[a] * st[a];
      }
      return gbar[(int64_t)(3 * c + d) * total + q];
    };
    for (int c = 0; c < 3; ++c) {
      T s = (T)0;
      // direct difference: G(q) = (f[q+e_c] - f[q]) * h
      {
        int dm[3] = {0, 0, 0};
        dm[c] = -1;
        const int d0[3] = {0, 0, 0};
        s = s + hinv[c] * GB(c, c, dm) - hinv[c] * GB(c, c, d0);
      }
      for (int d = 0; d < 3; ++d) {
        if (d == c) continue;
        // adjoints of the cell values at a source cell q (pos_d(q) decides the stale rule)
        auto cm0 = [&](const int (&dd)[3]) -> T {
          const T gb = GB(c, d, dd) * hinv[d];
          const bool stale = pos[d] + dd[d] == 0;
          return half * gb + (stale ? -gb : -half * gb);
        };
        auto cp1 = [&](const int (&dd)[3]) -> T { return half * GB(c, d, dd) * hinv[d]; };
        auto cm1 = [&](const int (&dd)[3]) -> T {
          const bool stale = pos[d] + dd[d] == 0;
          return stale ? (T)0 : -half * GB(c, d, dd) * hinv[d];
        };
        int q0[3] = {0, 0, 0}, q1[3] = {0, 0, 0};
        q1[c] = -1;
        s = s + half * (cm0(q0) + cm0(q1));                 // f[q], f[q+e_c] of cm0(q)
        int r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0};
        r0[d] = -1;
        r1[d] = -1;
        r1[c] = -1;
        s = s + half * (cp1(r0) + cp1(r1));                 // f[q+e_d], f[q+e_d+e_c]
        int m0[3] = {0, 0, 0}, m1[3] = {0, 0, 0};
        m0[d] = 1;
        m1[d] = 1;
        m1[c] = -1;
        s = s + half * (cm1(m0) + cm1(m1));                 // f[q-e_d], f[q-e_d+e_c]
      }
      fbar.a[c][p] = s;
    }
  }
}

}  // namespace pf

namespace pf {

static int closure_vjp_impl(pf_box* h, int kind, const void* const U[3], const float* params,
                            void* const* taubar6, void* gbar9, void* const F[3], double* dparams,
                            double* wpart, hipStream_t s) {
  const Geom& g = h->g;
  const int nw = pf_closure_vjp_waves(h);
  const int64_t total = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nb = (total + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    Vec6<T> tb;
    for (int m = 0; m < 6; ++m) tb.a[m] = (T*)taubar6[m];
    const T dxi = (T)(1.0 / h->dx), dyi = (T)(1.0 / h->dy), dzi = (T)(1.0 / h->dz);
    const T cst = (T)(h->dx * h->dx / 12.0);
    const T* u = (const T*)U[0];
    const T* v = (const T*)U[1];
    const T* w = (const T*)U[2];
    if (kind == 0)
      hipLaunchKernelGGL((closure_vjp_local_kernel<T, 0, false>), dim3(nw), dim3(64), 0, s, g, u,
                         v, w, dxi, dyi, dzi, cst, tb, params, (T*)gbar9, wpart);
    else if (dparams)
      hipLaunchKernelGGL((closure_vjp_local_kernel<T, 1, true>), dim3(nw), dim3(64), 0, s, g, u,
                         v, w, dxi, dyi, dzi, cst, tb, params, (T*)gbar9, wpart);
    else
      hipLaunchKernelGGL((closure_vjp_local_kernel<T, 1, false>), dim3(nw), dim3(64), 0, s, g, u,
                         v, w, dxi, dyi, dzi, cst, tb, params, (T*)gbar9, wpart);
    PF_LAUNCH_CHECK("closure_vjp_local_kernel");
    if (kind == 1 && dparams) {
      hipLaunchKernelGGL(mlp_param_reduce_kernel, dim3(kNp), dim3(kBlock), 0, s, wpart, nw,
                         dparams);
      PF_LAUNCH_CHECK("mlp_param_reduce_kernel");
    }
    Vec3<T> fb = {{(T*)F[0], (T*)F[1], (T*)F[2]}};
    hipLaunchKernelGGL(closure_vjp_gather_kernel<T>, dim3((int)nb), dim3(kBlock), 0, s, g,
                       (const T*)gbar9, dxi, dyi, dzi, fb);
    PF_LAUNCH_CHECK("closure_vjp_gather_kernel");
    return 0;
  };
  return h->dtype == PF_F64 ? run(0.0) : run(0.0f);
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t adj_work_bytes(const pf_box* h) {
  const size_t np = (size_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4) * h->esize;
  return 15 * align256(np) + align256((size_t)pf_closure_vjp_waves(h) * kNp * sizeof(double));
}

int adj_closure_term(pf_box* h, int kind, const void* const U[3], const float* params,
                     const void* const A[3], void* work, void* const D[3], double* dparams,
                     hipStream_t s) {
  if (kind == 2) return ml_source_adj_term(h, U, params, A, D, dparams, s);
  const size_t np = align256((size_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4) * h->esize);
  char* base = (char*)work;
  void* tb[6];
  for (int m = 0; m < 6; ++m) tb[m] = base + m * np;
  void* gbar = base + 6 * np;   // 9 padded arrays
  double* wpart = (double*)(base + 15 * np);
  if (int rc = pf_sfs_vjp(h, A[0], A[1], A[2], tb, 0, s)) return rc;
  if (int rc = closure_vjp_impl(h, kind, U, params, tb, gbar, D, kind == 1 ? dparams : nullptr,
                                wpart, s))
    return rc;
  for (int q = 0; q < 3; ++q)
    if (int rc = pf_fold_halo(h, D[q], s)) return rc;
  return 0;
}

}  // namespace pf

extern "C" {

int pf_sfs_mlp_f32(pf_box* h, const void* u, const void* v, const void* w, const float* params,
                   void* const* sfs6, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !params || !sfs6) return fail(PF_EINVAL, "pf_sfs_mlp_f32: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  const Geom& g = h->g;
  const int64_t total = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nb = (total + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  auto run = [&](auto zero) -> int {
    using T = decltype(zero);
    Vec6<T> o;
    for (int m = 0; m < 6; ++m) o.a[m] = (T*)sfs6[m];
    hipLaunchKernelGGL(mlp_f32_kernel<T>, dim3((int)nb), dim3(kBlock), 0, s, g, (const T*)u,
                       (const T*)v, (const T*)w, (T)(1.0 / h->dx), (T)(1.0 / h->dy),
                       (T)(1.0 / h->dz), params, (float)(h->dx * h->dx / 12.0), o);
    PF_LAUNCH_CHECK("mlp_f32_kernel");
    return 0;
  };
  return h->dtype == PF_F64 ? run(0.0) : run(0.0f);
}

// Number of waves (and rows of the parameter-partial workspace) the MLP
// reverse pass uses for this box.
int pf_closure_vjp_waves(const pf_box* h) {
  if (!h) return 0;
  const int64_t total = (int64_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4);
  int64_t nw = (total + 63) / 64;
  return (int)(nw < 2048 ? nw : 2048);
}

int pf_closure_vjp(pf_box* h, int kind, const void* u, const void* v, const void* w,
                   const float* params, void* const* taubar6, void* gbar9, void* fu, void* fv,
                   void* fw, double* dparams, double* wpart, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (kind != 0 && kind != 1) return fail(PF_EINVAL, "pf_closure_vjp: kind is 0 (gradient) or 1 (MLP)");
  if (!u || !v || !w || !taubar6 || !gbar9 || !fu || !fv || !fw)
    return fail(PF_EINVAL, "pf_closure_vjp: NULL pointer");
  if (kind == 1 && !params) return fail(PF_EINVAL, "pf_closure_vjp: the MLP needs params");
  if (dparams && !wpart) return fail(PF_EINVAL, "pf_closure_vjp: dparams needs wpart");
  const void* U[3] = {u, v, w};
  void* F[3] = {fu, fv, fw};
  return closure_vjp_impl(h, kind, U, params, taubar6, gbar9, F, kind == 1 ? dparams : nullptr,
                          wpart, (hipStream_t)stream);
}

size_t pf_adj_work_bytes(const pf_box* h) { return h ? adj_work_bytes(h) : 0; }

int pf_adj_closure_term(pf_box* h, int kind, const void* u, const void* v, const void* w,
                        const float* params, const void* ua, const void* va, const void* wa,
                        void* work, void* const* delta3, double* dparams, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (kind < 0 || kind > 2) return fail(PF_EINVAL, "pf_adj_closure_term: kind is 0, 1 or 2");
  if (!u || !v || !w || !ua || !va || !wa || !delta3 || !delta3[0] || !delta3[1] || !delta3[2] ||
      (kind < 2 && !work))
    return fail(PF_EINVAL, "pf_adj_closure_term: NULL pointer");
  if (kind >= 1 && !params) return fail(PF_EINVAL, "pf_adj_closure_term: the network needs params");
  const void* U[3] = {u, v, w};
  const void* A[3] = {ua, va, wa};
  void* D[3] = {delta3[0], delta3[1], delta3[2]};
  return adj_closure_term(h, kind, U, params, A, work, D, dparams, (hipStream_t)stream);
}

}  // extern "C"
