// Definitions shared by the 'source' ML closure's forward (pf_source.hip)
// and its reverse (pf_source_train.hip): sizes, the packed parameter layout,
// padded-array wrap-around and the feature kernel.
#pragma once

#include "pf_kernels.h"

namespace pf {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kSrcFeat = 21;                    // features per cell (:372-378)
constexpr int kSrcShifts = 13;                  // own cell + 12 shifted copies (:413-447)
constexpr int kSrcIn = kSrcFeat * kSrcShifts;   // 273 = num_inputs (:74)
constexpr int kSrcHid = 200;                    // H (:73)
constexpr int kSrcOut = 18;                     // fc4 outputs (:87)
constexpr int kSrcUsed = 9;                     // output channels the closure uses
constexpr int kSrcTiles = 13;                   // 16-neuron tiles covering 200
// packed FP32 parameters, torch.nn.Linear layout (W[out][in], then b[out]),
// in the order fc1, fcG, fcG2, fc2, fc3, fc4
constexpr int64_t kOffFc1W = 0;
constexpr int64_t kOffFc1B = kOffFc1W + (int64_t)kSrcHid * kSrcIn;
constexpr int64_t kOffGW = kOffFc1B + kSrcHid;
constexpr int64_t kOffGB = kOffGW + (int64_t)kSrcHid * kSrcIn;
constexpr int64_t kOffG2W = kOffGB + kSrcHid;
constexpr int64_t kOffG2B = kOffG2W + (int64_t)kSrcHid * kSrcIn;
constexpr int64_t kOffFc2W = kOffG2B + kSrcHid;
constexpr int64_t kOffFc2B = kOffFc2W + (int64_t)kSrcHid * kSrcHid;
constexpr int64_t kOffFc3W = kOffFc2B + kSrcHid;
constexpr int64_t kOffFc3B = kOffFc3W + (int64_t)kSrcHid * kSrcHid;
constexpr int64_t kOffFc4W = kOffFc3B + kSrcHid;
constexpr int64_t kOffFc4B = kOffFc4W + (int64_t)kSrcOut * kSrcHid;
constexpr int64_t kSrcParams = kOffFc4B + kSrcOut;

struct SrcGeom {
  int nxo, nyo, nzo;
  int64_t m;      // padded cells
  float d2;       // (float)(2*dx)
  float dd;       // (float)(dx*dx)
};

struct Norm21 {
  float v[kSrcFeat];
};

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Flat index of the cell at (i,j,k) moved by shift sh of the reference's
// concatenation order: own, x+1, x-1, y+1, y-1, z+1, z-1, x+2, x-2, y+2, y-2,
// z+2, z-2 (:449-462), wrapping over the padded extents.
__device__ __forceinline__ int64_t shifted(const SrcGeom& s, int i, int j, int k, int sh) {
  if (sh > 0) {
    const int d = sh <= 6 ? 1 : 2;
    const int q = (sh - 1) % 6;          // 0 x+, 1 x-, 2 y+, 3 y-, 4 z+, 5 z-
    const int o = (q & 1) ? -d : d;
    if (q < 2) i = wrap(i + o, s.nxo);
    else if (q < 4) j = wrap(j + o, s.nyo);
    else k = wrap(k + o, s.nzo);
  }
  return ((int64_t)i * s.nyo + j) * s.nzo + k;
}

__device__ __forceinline__ void decode_padded(const SrcGeom& s, int64_t p, int& i, int& j, int& k) {
  k = (int)(p % s.nzo);
  const int64_t r = p / s.nzo;
  j = (int)(r % s.nyo);
  i = (int)(r / s.nyo);
}

// 1. features, feature-major: Fn[f * m + cell]
template <typename T>
__global__ void __launch_bounds__(kBlock) src_features_kernel(SrcGeom s, const T* __restrict__ u,
                                                              const T* __restrict__ v,
                                                              const T* __restrict__ w,
                                                              float* __restrict__ Fn, Norm21 nrm) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.m;
       p += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decode_padded(s, p, i, j, k);
    const int64_t sy = s.nzo, sx = (int64_t)s.nyo * s.nzo;
    const int64_t xp = (int64_t)wrap(i + 1, s.nxo) * sx + j * sy + k;
    const int64_t xm = (int64_t)wrap(i - 1, s.nxo) * sx + j * sy + k;
    const int64_t yp = (int64_t)i * sx + wrap(j + 1, s.nyo) * sy + k;
    const int64_t ym = (int64_t)i * sx + wrap(j - 1, s.nyo) * sy + k;
    const int64_t zp = (int64_t)i * sx + j * sy + wrap(k + 1, s.nzo);
    const int64_t zm = (int64_t)i * sx + j * sy + wrap(k - 1, s.nzo);
    const T* f3[3] = {u, v, w};
    float F[kSrcFeat];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T* f = f3[q];
      const float c = (float)f[p];
      const float nb[6] = {(float)f[xp], (float)f[xm], (float)f[yp],
                           (float)f[ym], (float)f[zp], (float)f[zm]};
      F[q] = c;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float fp = nb[2 * a], fm = nb[2 * a + 1];
        F[3 + 3 * q + a] = (fp - fm) / s.d2;                      // :439-463
        F[12 + 3 * q + a] = ((fp - 2.0f * c) + fm) / s.dd;        // :412-433
      }
    }
#pragma unroll
    for (int f = 0; f < kSrcFeat; ++f) Fn[f * s.m + p] = F[f] / nrm.v[f];
  }
}

// Workspace of one evaluation: 21 normalised features and 9 network outputs
// per padded cell, FP32.
inline size_t src_ws_need(const pf_box* h) {
  const size_t m = (size_t)(h->nx + 2 * kNover) * (h->ny + 2 * kNover) * (h->nz + 2 * kNover);
  return (size_t)(kSrcFeat + kSrcUsed) * m * sizeof(float);
}

}  // namespace pf
