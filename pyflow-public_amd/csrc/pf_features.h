// Cell-centred velocity gradients of metric_uniform.grad_vel_center
// (metric/metric_staggered.py:215-268), the nine features of the 'tensor'
// closures, evaluated at a padded cell exactly as the reference's slice
// sequence leaves them.
//
// One reference artefact is reproduced: for an off-diagonal gradient d/dx_d
// of a component interpolated along another axis, interp_uvwi_* writes
// var_i[1:] only, so at overlap index 0 of axis d the "lower" interpolant is
// the stale cell value c[0] and the gradient there is
// (0.5*(c[1]+c[0]) - c[0]) / h_d rather than a periodic image.  The interior
// RHS reads those cells through the edge interpolation of FXY/FXZ/FYZ
// (velocity.py:264-315).
#pragma once

#include "pf_common.h"

namespace pf {

// G[c][d]: component c (u,v,w) differentiated along d (x,y,z) at padded cell
// (a,b,k), every index in 0..n+2; U, V, W are padded base pointers.
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

// Padded cell of a flat padded index, with the last overlap plane (n+3)
// mapped to its periodic image (3): the reference's values there depend on
// array history and are never read by the RHS.
__device__ __forceinline__ void feature_cell(int64_t q, int nxo, int nyo, int nzo, int& a, int& b,
                                             int& k) {
  const int k0 = (int)(q % nzo);
  const int64_t t1 = q / nzo;
  const int b0 = (int)(t1 % nyo);
  const int a0 = (int)(t1 / nyo);
  a = a0 == nxo - 1 ? 3 : a0;
  b = b0 == nyo - 1 ? 3 : b0;
  k = k0 == nzo - 1 ? 3 : k0;
}

}  // namespace pf
