// Closures with the 'tensor' contract: six cell-centred stress fields
// FXX, FYY, FZZ, FXY, FXZ, FYZ from the nine cell-centred velocity gradients
// of metric_uniform.grad_vel_center (metric/metric_staggered.py:215-268).
//
// The gradients are evaluated over the padded index range exactly as the
// reference's slice sequence leaves them, including one reference artefact:
// for an off-diagonal gradient d/dx_d of a component interpolated along
// another axis, interp_uvwi_* writes var_i[1:] only, so at overlap index 0
// of axis d the "lower" interpolant is the stale cell value c[0] and the
// gradient there is (0.5*(c[1]+c[0]) - c[0]) / h_d rather than a periodic
// image.  The interior RHS reads those cells through the edge interpolation
// of FXY/FXZ/FYZ (velocity.py:264-315), so the artefact is reproduced.
// Cells at the last overlap index (n+3), whose reference values depend on
// the arrays' history and are never read, receive the periodic image.
#include <hip/hip_runtime.h>

#include "pf_box.h"

namespace pf {

// Nine gradients at padded cell (a,b,k) [all in 0..n+2], padded base pointers.
// G[c][d]: component c (u,v,w) differentiated along d (x,y,z).
template <typename T>
__device__ __forceinline__ void grad_center(const T* __restrict__ U, const T* __restrict__ V,
                                            const T* __restrict__ W, int64_t sx, int64_t sy,
                                            int a, int b, int k, T dxi, T dyi, T dzi,
                                            T (&G)[3][3]) {
  const int64_t c0 = a * sx + b * sy + k;
  const T half = (T)0.5;
  const T* F[3] = {U, V, W};
  const int64_t st[3] = {sx, sy, (int64_t)1};
  const int pos[3] = {a, b, k};
  const T hinv[3] = {dxi, dyi, dzi};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T* f = F[c];
    const int64_t sc = st[c];   // interpolation axis of component c
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d == c) {
        // direct staggered difference: grad[:-1] = (var[1:] - var[:-1]) * h
        G[c][d] = (f[c0 + sc] - f[c0]) * hinv[d];
      } else {
        const int64_t sd = st[d];
        // c(m) = 0.5*(f[m + e_c] + f[m]) at offsets 0 and +1 (and -1) along d
        const T cm0 = half * (f[c0 + sc] + f[c0]);
        const T cp1 = half * (f[c0 + sd + sc] + f[c0 + sd]);
        T e0;
        if (pos[d] == 0) {
          e0 = cm0;   // stale: interp_uvwi_* leaves var_i[0] untouched
        } else {
          const T cm1 = half * (f[c0 - sd + sc] + f[c0 - sd]);
          e0 = half * (cm0 + cm1);
        }
        const T e1 = half * (cp1 + cm0);
        G[c][d] = (e1 - e0) * hinv[d];
      }
    }
  }
}

// Gradient-model stresses (sfsmodel_gradient.py:77-148), one padded cell per thread.
template <typename T>
__global__ void __launch_bounds__(kBlock) sfs_gradient_kernel(Geom g, const T* __restrict__ U,
                                                              const T* __restrict__ V,
                                                              const T* __restrict__ W, T dxi,
                                                              T dyi, T dzi, T cst,
                                                              Vec6<T> out) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t total = (int64_t)nxo * nyo * nzo;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int k0 = (int)(q % nzo);
    const int64_t t1 = q / nzo;
    const int b0 = (int)(t1 % nyo);
    const int a0 = (int)(t1 / nyo);
    // the last overlap plane takes the periodic image (index n+3 -> 3)
    const int a = a0 == nxo - 1 ? 3 : a0;
    const int b = b0 == nyo - 1 ? 3 : b0;
    const int k = k0 == nzo - 1 ? 3 : k0;
    T G[3][3];
    grad_center(U, V, W, g.psx, g.psy, a, b, k, dxi, dyi, dzi, G);
    const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int p = pr[m][0], r = pr[m][1];
      T s = G[p][0] * G[r][0];
      s = s + (G[p][1] * G[r][1]);
      s = s + (G[p][2] * G[r][2]);
      out.a[m][q] = s * cst;
    }
  }
}

template <typename T>
int sfs_gradient_impl(pf_box* h, hipStream_t s, const T* U, const T* V, const T* W,
                      void* const* sfs6) {
  const Geom& g = h->g;
  Vec6<T> o;
  for (int m = 0; m < 6; ++m) o.a[m] = (T*)sfs6[m];
  const int64_t total = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nb = (total + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  const T cst = (T)(h->dx * h->dx / 12.0);   // sfsmodel_gradient.py:63
  hipLaunchKernelGGL(sfs_gradient_kernel<T>, dim3((int)nb), dim3(kBlock), 0, s, g, U, V, W,
                     (T)(1.0 / h->dx), (T)(1.0 / h->dy), (T)(1.0 / h->dz), cst, o);
  PF_LAUNCH_CHECK("sfs_gradient_kernel");
  return 0;
}

}  // namespace pf

using namespace pf;

extern "C" int pf_sfs_gradient(pf_box* h, const void* u, const void* v, const void* w,
                               void* const* sfs6, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !sfs6) return fail(PF_EINVAL, "pf_sfs_gradient: NULL pointer");
  for (int m = 0; m < 6; ++m)
    if (!sfs6[m]) return fail(PF_EINVAL, "pf_sfs_gradient: NULL stress field");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return sfs_gradient_impl<double>(h, s, (const double*)u, (const double*)v, (const double*)w,
                                     sfs6);
  return sfs_gradient_impl<float>(h, s, (const float*)u, (const float*)v, (const float*)w, sfs6);
}
