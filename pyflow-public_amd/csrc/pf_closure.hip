// Closures with the 'tensor' contract: six cell-centred stress fields
// FXX, FYY, FZZ, FXY, FXZ, FYZ from the nine cell-centred velocity gradients
// of metric_uniform.grad_vel_center (pf_features.h).  Cells at the last
// overlap index (n+3), whose reference values depend on the arrays' history
// and are never read, receive the periodic image.
#include <hip/hip_runtime.h>

#include "pf_box.h"
#include "pf_features.h"

namespace pf {

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

// Smagorinsky eddy viscosity (sfsmodel_smagorinsky.py:65-88) added to the
// molecular viscosity as Domain.forwardStep does (domain.py:348-352):
// VISC = mu + rho*(Cs*Delta)^2*|S| at every padded cell, |S| from the nine
// gradients of grad_vel_center in the reference's op order.
template <typename T>
__global__ void __launch_bounds__(kBlock) smagorinsky_kernel(Geom g, const T* __restrict__ U,
                                                             const T* __restrict__ V,
                                                             const T* __restrict__ W, T dxi,
                                                             T dyi, T dzi, T cs2, T mu,
                                                             T* __restrict__ visc) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t total = (int64_t)nxo * nyo * nzo;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    int a, b, k;
    feature_cell(q, nxo, nyo, nzo, a, b, k);
    T G[3][3];
    grad_center(U, V, W, g.psx, g.psy, a, b, k, dxi, dyi, dzi, G);
    const T d3 = ((G[0][0] + G[1][1]) + G[2][2]) / (T)3.0;
    const T s11 = G[0][0] - d3, s22 = G[1][1] - d3, s33 = G[2][2] - d3;
    const T half = (T)0.5;
    const T s12 = half * (G[0][1] + G[1][0]);
    const T s13 = half * (G[0][2] + G[2][0]);
    const T s23 = half * (G[1][2] + G[2][1]);
    T n = s11 * s11;
    n = n + (s22 * s22);
    n = n + (s33 * s33);
    T o = s12 * s12;
    o = o + (s13 * s13);
    o = o + (s23 * s23);
    const T smag = sqrt_rn(n + ((T)2.0 * o));
    visc[q] = mu + (cs2 * smag);
  }
}

template <typename T>
int smagorinsky_impl(pf_box* h, hipStream_t s, const T* U, const T* V, const T* W, double rho,
                     double Cs, double Delta, double mu, T* visc) {
  const Geom& g = h->g;
  const int64_t total = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nb = (total + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  const double cd = Cs * Delta;
  const T cs2 = (T)(rho * (cd * cd));   // rho * (Cs*Delta)**2, one Python float
  hipLaunchKernelGGL(smagorinsky_kernel<T>, dim3((int)nb), dim3(kBlock), 0, s, g, U, V, W,
                     (T)(1.0 / h->dx), (T)(1.0 / h->dy), (T)(1.0 / h->dz), cs2, (T)mu, visc);
  PF_LAUNCH_CHECK("smagorinsky_kernel");
  return 0;
}

}  // namespace pf

using namespace pf;

extern "C" int pf_sfs_smagorinsky(pf_box* h, const void* u, const void* v, const void* w,
                                  double rho, double Cs, double Delta, double mu, void* visc,
                                  void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !visc) return fail(PF_EINVAL, "pf_sfs_smagorinsky: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return smagorinsky_impl<double>(h, s, (const double*)u, (const double*)v, (const double*)w,
                                    rho, Cs, Delta, mu, (double*)visc);
  return smagorinsky_impl<float>(h, s, (const float*)u, (const float*)v, (const float*)w, rho,
                                 Cs, Delta, mu, (float*)visc);
}

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

// ---------------------------------------------------------------------------
// Closure MLP ('tensor' contract): the nine gradients above -> Linear(9,20) ->
// tanh -> Linear(20,20) -> tanh -> Linear(20,6), scaled by dx^2/12 like the
// gradient model, written as FXX, FYY, FZZ, FXY, FXZ, FYZ.
//
// GEMMs run on MFMA (v_mfma_f32_16x16x16_bf16): BF16 operands, FP32
// accumulation, biases added in FP32, tanh in FP32.  Cells are the N
// dimension (16 per MFMA), neurons the M dimension, so a layer's accumulator
// tile (lane l holds rows 4*(l/16)+i of column l%16) is already the next
// layer's B operand (lane l holds k = 4*(l/16)+i of column l%16): activations
// never move between lanes.  Weights (A operand: lane l holds row l%16,
// k = 4*(l/16)+i) sit in registers for the whole wave.
namespace pf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 9, kMlpHid = 20, kMlpOut = 6;
// packed FP32 parameters: W1[20][9] b1[20] W2[20][20] b2[20] W3[6][20] b3[6]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpHid * kMlpIn, kOffW2 = kOffB1 + kMlpHid,
              kOffB2 = kOffW2 + kMlpHid * kMlpHid, kOffW3 = kOffB2 + kMlpHid,
              kOffB3 = kOffW3 + kMlpOut * kMlpHid, kMlpParams = kOffB3 + kMlpOut;

// float -> bf16, round to nearest even
__device__ __forceinline__ short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (short)0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// A-operand tile of a row-major weight matrix W[rows][cols]: rows r0.., k0..
__device__ __forceinline__ bf16x4 weight_tile(const float* W, int rows, int cols, int r0, int k0,
                                              int lane) {
  bf16x4 a;
  const int r = r0 + (lane & 15);
  for (int i = 0; i < 4; ++i) {
    const int k = k0 + 4 * (lane >> 4) + i;
    a[i] = (r < rows && k < cols) ? f2bf(W[r * cols + k]) : (short)0;
  }
  return a;
}

// bias in accumulator layout: lane l, element i -> row r0 + 4*(l/16) + i
__device__ __forceinline__ f32x4 bias_tile(const float* b, int rows, int r0, int lane) {
  f32x4 c;
  for (int i = 0; i < 4; ++i) {
    const int r = r0 + 4 * (lane >> 4) + i;
    c[i] = r < rows ? b[r] : 0.f;
  }
  return c;
}

__device__ __forceinline__ bf16x4 tanh_bf16(f32x4 c) {
  bf16x4 h;
  for (int i = 0; i < 4; ++i) h[i] = f2bf(tanhf(c[i]));
  return h;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) mlp_closure_kernel(Geom g, const T* __restrict__ U,
                                                             const T* __restrict__ V,
                                                             const T* __restrict__ W, T dxi,
                                                             T dyi, T dzi,
                                                             const float* __restrict__ prm,
                                                             float cst, Vec6<T> out) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t total = (int64_t)nxo * nyo * nzo;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ float feat[kBlock / 64][64][kMlpIn + 1];

  // per-wave constant operands
  const bf16x4 w1a = weight_tile(prm + kOffW1, kMlpHid, kMlpIn, 0, 0, lane);
  const bf16x4 w1b = weight_tile(prm + kOffW1, kMlpHid, kMlpIn, 16, 0, lane);
  bf16x4 w2[2][2], w3[2];
  for (int mt = 0; mt < 2; ++mt)
    for (int ks = 0; ks < 2; ++ks)
      w2[mt][ks] = weight_tile(prm + kOffW2, kMlpHid, kMlpHid, 16 * mt, 16 * ks, lane);
  for (int ks = 0; ks < 2; ++ks) w3[ks] = weight_tile(prm + kOffW3, kMlpOut, kMlpHid, 0, 16 * ks, lane);
  const f32x4 b1a = bias_tile(prm + kOffB1, kMlpHid, 0, lane);
  const f32x4 b1b = bias_tile(prm + kOffB1, kMlpHid, 16, lane);
  const f32x4 b2a = bias_tile(prm + kOffB2, kMlpHid, 0, lane);
  const f32x4 b2b = bias_tile(prm + kOffB2, kMlpHid, 16, lane);
  const f32x4 b3t = bias_tile(prm + kOffB3, kMlpOut, 0, lane);

  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < total; base += (int64_t)gridDim.x * kBlock) {
    // 1. features of one padded cell per lane
    const int64_t q = base + threadIdx.x;
    if (q < total) {
      const int k0 = (int)(q % nzo);
      const int64_t t1 = q / nzo;
      const int b0 = (int)(t1 % nyo);
      const int a0 = (int)(t1 / nyo);
      const int a = a0 == nxo - 1 ? 3 : a0;
      const int b = b0 == nyo - 1 ? 3 : b0;
      const int k = k0 == nzo - 1 ? 3 : k0;
      T G[3][3];
      grad_center(U, V, W, g.psx, g.psy, a, b, k, dxi, dyi, dzi, G);
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) feat[wave][lane][3 * c + d] = (float)G[c][d];
    } else {
      for (int f = 0; f < kMlpIn; ++f) feat[wave][lane][f] = 0.f;
    }
    __syncthreads();
    // 2. four 16-cell MFMA column tiles per wave
    const int64_t wbase = base + 64 * wave;
    for (int nt = 0; nt < 4; ++nt) {
      const int cell = 16 * nt + (lane & 15);
      bf16x4 x;
      for (int i = 0; i < 4; ++i) {
        const int kf = 4 * (lane >> 4) + i;
        x[i] = kf < kMlpIn ? f2bf(feat[wave][cell][kf]) : (short)0;
      }
      f32x4 c1a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1a, x, b1a, 0, 0, 0);
      f32x4 c1b = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1b, x, b1b, 0, 0, 0);
      const bf16x4 h1a = tanh_bf16(c1a), h1b = tanh_bf16(c1b);
      f32x4 c2a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[0][0], h1a, b2a, 0, 0, 0);
      c2a = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[0][1], h1b, c2a, 0, 0, 0);
      f32x4 c2b = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[1][0], h1a, b2b, 0, 0, 0);
      c2b = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[1][1], h1b, c2b, 0, 0, 0);
      const bf16x4 h2a = tanh_bf16(c2a), h2b = tanh_bf16(c2b);
      f32x4 c3 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3[0], h2a, b3t, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3[1], h2b, c3, 0, 0, 0);
      const int64_t qc = wbase + cell;
      if (qc < total) {
        for (int i = 0; i < 4; ++i) {
          const int o = 4 * (lane >> 4) + i;
          if (o < kMlpOut) out.a[o][qc] = (T)(c3[i] * cst);
        }
      }
    }
    __syncthreads();
  }
}

template <typename T>
int sfs_mlp_impl(pf_box* h, hipStream_t s, const T* U, const T* V, const T* W, const float* prm,
                 void* const* sfs6) {
  const Geom& g = h->g;
  Vec6<T> o;
  for (int m = 0; m < 6; ++m) o.a[m] = (T*)sfs6[m];
  const int64_t total = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nb = (total + kBlock - 1) / kBlock;
  if (nb > 16384) nb = 16384;
  const float cst = (float)(h->dx * h->dx / 12.0);
  hipLaunchKernelGGL(mlp_closure_kernel<T>, dim3((int)nb), dim3(kBlock), 0, s, g, U, V, W,
                     (T)(1.0 / h->dx), (T)(1.0 / h->dy), (T)(1.0 / h->dz), prm, cst, o);
  PF_LAUNCH_CHECK("mlp_closure_kernel");
  return 0;
}

}  // namespace pf

extern "C" int pf_sfs_mlp(pf_box* h, const void* u, const void* v, const void* w,
                          const float* params, void* const* sfs6, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !params || !sfs6) return fail(PF_EINVAL, "pf_sfs_mlp: NULL pointer");
  for (int m = 0; m < 6; ++m)
    if (!sfs6[m]) return fail(PF_EINVAL, "pf_sfs_mlp: NULL stress field");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return sfs_mlp_impl<double>(h, s, (const double*)u, (const double*)v, (const double*)w,
                                params, sfs6);
  return sfs_mlp_impl<float>(h, s, (const float*)u, (const float*)v, (const float*)w, params,
                             sfs6);
}

extern "C" int pf_sfs_mlp_param_count(void) { return pf::kMlpParams; }
