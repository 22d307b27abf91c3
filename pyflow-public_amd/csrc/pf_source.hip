// The reference's 'source' ML closure (sfsmodel/sfsmodel_ML.py:76-116 and
// :354-593) on the GPU: residual_stress.update(u_P, v_P, w_P) -> GX, GY, GZ.
//
// Three kernels:
//  1. features (Finite_Difference_ConCat, :408-465): per padded cell the 21
//     FP32 values u, v, w, their first differences (f[+1] - f[-1])/(2 dx) and
//     second differences ((f[+1] - 2 f) + f[-1])/(dx dx) along x, y, z, all
//     divided by the normalisation table (:252-258, ML_Closure :449-463).  As
//     in the reference, the wrap-around is that of the padded array itself
//     (torch.cat of its first and last planes) and every direction uses dx;
//  2. the network (NeuralNetworkModel2.forward, :88-116) on MFMA in FP32
//     (v_mfma_f32_16x16x4_f32): input = the cell's 21 values and those of its
//     12 neighbours at +-1 and +-2 along each axis (padded wrap again, :413-447),
//     L1 = tanh(fc1 x), G = tanh(fcG x), Gate2 = tanh(fcG2 x),
//     H3 = tanh(fc3 (G * tanh(fc2 L1))), out = fc4 (H3 * Gate2); only output
//     channels 0-8 are used (:466-589);
//  3. divergence: GX = (d/dx p0 + d/dy p1) + d/dz p2 with central differences
//     (p[+1] - p[-1])/(2 dx) over the padded wrap, likewise GY from p3-p5 and GZ
//     from p6-p8 (:466-589), cast to the box dtype.
//
// The network keeps cells on the MFMA N dimension (16 per wave) and neurons
// on M, so a layer's accumulator tile (lane l holds neurons 4*(l/16)+i of cell
// l%16) is, element by element, the next layer's B operand: the K index of
// k-step (t, i) in lane group g is neuron 16 t + 4 g + i, and the weights (A
// operand) are read with the same permutation.  Two waves share 16 cells,
// each computing half of every layer's 13 neuron tiles (7 and 6), so that a
// wave's activations fit two waves per SIMD; before each hidden layer the
// halves meet in LDS.  The first layer's K runs over (shift, feature) with
// the 21 features of a shift padded to 24, so each k-step reads one
// neighbour.  Weights reach the MFMAs through LDS: each block stages one
// chunk of input columns at a time for its eight waves, the next chunk's
// loads in flight while the current one is consumed; a 128-bit LDS read
// feeds four MFMAs.  Measured at 64^3: 2.96 ms per evaluation with one wave
// per 16 cells (498 registers, one wave per SIMD), 2.50 ms with the halves,
// 2.34 ms with the two halves of a cell group on one SIMD.  Packing the
// first layer's 273 columns into 12 dense chunks instead of 13 padded ones
// (two neighbour indices per chunk) spilled 114 VGPRs: 2.64 ms.
// Summation order differs from the reference's CPU GEMM, so results agree to
// FP32 rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include "pf_source.h"

namespace pf {

// A operand of a weight matrix W[rows][ld] for neuron row r and input column
// kc (zero outside the matrix).
__device__ __forceinline__ float wload(const float* __restrict__ W, int ld, int rows, int cols,
                                       int r, int kc) {
  return (r < rows && kc < cols) ? W[(int64_t)r * ld + kc] : 0.f;
}

__device__ __forceinline__ f32x4 mfma(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// tanh(acc + b) over tiles O0..O0+NT-1 in accumulator layout: element i of
// tile t is neuron 16 t + 4 g + i.
template <int O0, int NT>
__device__ __forceinline__ void bias_tanh(f32x4 (&a)[NT], const float* __restrict__ b, int g) {
#pragma unroll
  for (int o = 0; o < NT; ++o)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int n = 16 * (O0 + o) + 4 * g + i;
      a[o][i] = n < kSrcHid ? tanhf(a[o][i] + b[n]) : 0.f;
    }
}

// Weights go through LDS: a block stages a chunk of input columns of a
// weight matrix once for its eight waves, row-major (chunk[neuron][column]),
// so one 128-bit LDS read gives a lane the A operands of four consecutive
// k-steps.  Pitches (in floats) put the 8 lanes of each pass of a 128-bit
// read on distinct banks: 20 for the 16-column chunks of the hidden layers,
// 28 for the 24-column chunks (21 features of a shift) of the input layers.
constexpr int kPitchHid = 20;
constexpr int kPitchIn = 28;
constexpr int kStageCols = 24;                             // 21 features of a shift, padded
constexpr int kRows = 16 * kSrcTiles;                      // 208 neurons
constexpr int kMatIn = kRows * kPitchIn;                   // one input-layer chunk
constexpr int kStageFloats = 2 * kMatIn;                   // >= kRows * kPitchHid
constexpr int kNetThreads = 512;                           // 4 groups of 16 cells x 2 halves
// A wave computes one half of a layer's 13 neuron tiles for its 16 cells:
// tiles 0-6 (half 0) or 7-12 (half 1).
constexpr int kHalf0 = 7;

typedef float f32x2v __attribute__((ext_vector_type(2)));

// One thread's share of a staged chunk: chunk[mat * kMatIn + r * PITCH + j] =
// W[r][c0 + j] for r < nrows and j < nvalid, else 0 (208 neuron rows, NC
// columns, NMAT matrices).  load() fills registers, store() writes them to
// LDS, so the next chunk's loads are in flight while the current one is
// consumed.
template <int PITCH, int NC, int NMAT>
struct Stager {
  static constexpr int kPer = NMAT * NC * kRows;
  static constexpr int kN = (kPer + kNetThreads - 1) / kNetThreads;
  float v[kN];
  __device__ __forceinline__ void load(const float* __restrict__ W0, const float* __restrict__ W1,
                                       int ld, int nrows, int c0, int nvalid) {
#pragma unroll
    for (int e = 0; e < kN; ++e) {
      const int q = threadIdx.x + kNetThreads * e;
      const int mat = q / (NC * kRows);
      const int qq = q - mat * (NC * kRows);
      const int r = qq / NC, j = qq - r * NC;
      const float* W = mat ? W1 : W0;
      v[e] = (q < kPer && r < nrows && j < nvalid) ? W[(int64_t)r * ld + c0 + j] : 0.f;
    }
  }
  __device__ __forceinline__ void store(float* __restrict__ chunk) const {
#pragma unroll
    for (int e = 0; e < kN; ++e) {
      const int q = threadIdx.x + kNetThreads * e;
      if (q < kPer) {
        const int mat = q / (NC * kRows);
        const int qq = q - mat * (NC * kRows);
        const int r = qq / NC, j = qq - r * NC;
        chunk[mat * kMatIn + r * PITCH + j] = v[e];
      }
    }
  }
};

// One hidden layer: out (tiles O0..O0+NT-1) = W x over the permuted K order
// (see the file comment), x the whole previous layer (13 tiles, in
// registers), one 16-column chunk of W per t, double-buffered in LDS (buf0,
// buf1); one barrier per chunk.  A lane reads W[16 o + col][16 t + 4 g .. +
// 3] with one 128-bit read, one tile ahead.
template <int O0, int NT>
__device__ __forceinline__ void hidden_layer(const f32x4 (&x)[kSrcTiles],
                                             const float* __restrict__ W, float* buf0,
                                             float* buf1, int col, int g, f32x4 (&out)[NT]) {
  auto nvalid = [](int t) { return t < kSrcTiles - 1 ? 16 : kSrcHid - 16 * t; };
  Stager<kPitchHid, 16, 1> st;
  st.load(W, W, kSrcHid, kSrcHid, 0, nvalid(0));
  st.store(buf0);
  __syncthreads();
#pragma unroll
  for (int o = 0; o < NT; ++o) out[o] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < kSrcTiles; ++t) {
    if (t + 1 < kSrcTiles) st.load(W, W, kSrcHid, kSrcHid, 16 * (t + 1), nvalid(t + 1));
    const float* base = ((t & 1) ? buf1 : buf0) + (16 * O0 + col) * kPitchHid + 4 * g;
    f32x4 a = *reinterpret_cast<const f32x4*>(base);
#pragma unroll
    for (int o = 0; o < NT; ++o) {
      f32x4 an = a;
      if (o + 1 < NT) an = *reinterpret_cast<const f32x4*>(base + 16 * (o + 1) * kPitchHid);
#pragma unroll
      for (int i = 0; i < 4; ++i) out[o] = mfma(a[i], x[t][i], out[o]);
      a = an;
    }
    if (t + 1 < kSrcTiles) st.store((t & 1) ? buf0 : buf1);
    __syncthreads();
  }
}

// First-layer product of one or two 273-input matrices over the gathered
// input (output tiles O0..O0+NT-1), one chunk of 21 columns (one shift) per
// matrix and step, double-buffered like hidden_layer.  K order within a
// shift: four k-steps i with lane group g on feature 4 g + i (one 128-bit
// read), then two with 16 + 2 g + i (one 64-bit read; features 21-23 are
// zero).
template <int NMAT, int O0, int NT>
__device__ __forceinline__ void input_layer(const SrcGeom& s, const float* __restrict__ Fn,
                                            bool valid, int ci, int cj, int ck,
                                            const float* __restrict__ W0,
                                            const float* __restrict__ W1, float* buf0,
                                            float* buf1, int col, int g, f32x4 (&a0)[NT],
                                            f32x4 (&a1)[NT]) {
  auto gather = [&](int sh, float (&b)[6]) {
    const int64_t nb = valid ? shifted(s, ci, cj, ck, sh) : 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int f = i < 4 ? 4 * g + i : 16 + 2 * g + (i - 4);
      b[i] = (valid && f < kSrcFeat) ? Fn[(int64_t)f * s.m + nb] : 0.f;
    }
  };
  Stager<kPitchIn, kStageCols, NMAT> st;
  st.load(W0, W1, kSrcIn, kSrcHid, 0, kSrcFeat);
  st.store(buf0);
  float b[6], bn[6];
  gather(0, b);
  __syncthreads();
#pragma unroll
  for (int o = 0; o < NT; ++o) {
    a0[o] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (NMAT > 1) a1[o] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  for (int sh = 0; sh < kSrcShifts; ++sh) {
    const bool more = sh + 1 < kSrcShifts;
    if (more) {
      st.load(W0, W1, kSrcIn, kSrcHid, (sh + 1) * kSrcFeat, kSrcFeat);
      gather(sh + 1, bn);
    }
    const float* c0 = ((sh & 1) ? buf1 : buf0) + (16 * O0 + col) * kPitchIn;
    const float* c1 = c0 + (NMAT > 1 ? kMatIn : 0);
    // the operands of tile o + 1 are read while tile o's MFMAs run
    struct Ops {
      f32x4 wa, va;
      f32x2v wb, vb;
    };
    auto read_ops = [&](int o) -> Ops {
      const int row = 16 * o * kPitchIn;
      Ops q;
      q.wa = *reinterpret_cast<const f32x4*>(c0 + row + 4 * g);
      q.wb = *reinterpret_cast<const f32x2v*>(c0 + row + 16 + 2 * g);
      q.va = q.wa;
      q.vb = q.wb;
      if (NMAT > 1) {
        q.va = *reinterpret_cast<const f32x4*>(c1 + row + 4 * g);
        q.vb = *reinterpret_cast<const f32x2v*>(c1 + row + 16 + 2 * g);
      }
      return q;
    };
    Ops nx = read_ops(0);
#pragma unroll
    for (int o = 0; o < NT; ++o) {
      const Ops cu = nx;
      if (o + 1 < NT) nx = read_ops(o + 1);
      const f32x4 wa = cu.wa, va = cu.va;
      const f32x2v wb = cu.wb, vb = cu.vb;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a0[o] = mfma(wa[i], b[i], a0[o]);
        if (NMAT > 1) a1[o] = mfma(va[i], b[i], a1[o]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a0[o] = mfma(wb[i], b[4 + i], a0[o]);
        if (NMAT > 1) a1[o] = mfma(vb[i], b[4 + i], a1[o]);
      }
    }
    if (more) {
      st.store((sh & 1) ? buf0 : buf1);
#pragma unroll
      for (int i = 0; i < 6; ++i) b[i] = bn[i];
    }
    __syncthreads();
  }
}

// The two halves of a layer meet in LDS: each wave of a cell group writes
// its tiles (one 128-bit slot per lane and tile), and after a barrier reads
// its partner's, so that both hold the whole layer as the next layer's B
// operand.
template <int O0, int NT>
__device__ __forceinline__ void exchange(float* __restrict__ xb, int lane,
                                         const f32x4 (&mine)[NT], f32x4 (&all)[kSrcTiles]) {
#pragma unroll
  for (int o = 0; o < NT; ++o) {
    *reinterpret_cast<f32x4*>(xb + ((O0 + o) * 64 + lane) * 4) = mine[o];
    all[O0 + o] = mine[o];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kSrcTiles; ++t)
    if (t < O0 || t >= O0 + NT) all[t] = *reinterpret_cast<const f32x4*>(xb + (t * 64 + lane) * 4);
}

// The network for one half (H) of a wave pair: 16 cells, neuron tiles
// O0..O0+NT-1 of every layer.
template <int H>
__device__ __forceinline__ void net_half(const SrcGeom& s, const float* __restrict__ Fn,
                                         const float* __restrict__ prm, float* __restrict__ P,
                                         float* buf0, float* buf1, float* xb, float* red,
                                         int lane, int64_t cell, bool valid, int ci, int cj,
                                         int ck) {
  constexpr int O0 = H ? kHalf0 : 0;
  constexpr int NT = H ? kSrcTiles - kHalf0 : kHalf0;
  const int col = lane & 15, g = lane >> 4;
  f32x4 all[kSrcTiles];
  f32x4 h1[NT], gt[NT];
  // L1 = tanh(fc1 x), G = tanh(fcG x)  (:91-100)
  input_layer<2, O0, NT>(s, Fn, valid, ci, cj, ck, prm + kOffFc1W, prm + kOffGW, buf0, buf1,
                         col, g, h1, gt);
  bias_tanh<O0, NT>(h1, prm + kOffFc1B, g);
  bias_tanh<O0, NT>(gt, prm + kOffGB, g);
  exchange<O0, NT>(xb, lane, h1, all);
  // H2 = tanh(fc2 L1); H2_G = G * H2  (:94-104)
  f32x4 hg[NT];
  hidden_layer<O0, NT>(all, prm + kOffFc2W, buf0, buf1, col, g, hg);
  bias_tanh<O0, NT>(hg, prm + kOffFc2B, g);
#pragma unroll
  for (int o = 0; o < NT; ++o) hg[o] = gt[o] * hg[o];
  // Gate2 = tanh(fcG2 x)  (:102)
  f32x4 g2[NT];
  input_layer<1, O0, NT>(s, Fn, valid, ci, cj, ck, prm + kOffG2W, prm + kOffG2W, buf0, buf1,
                         col, g, g2, g2);
  bias_tanh<O0, NT>(g2, prm + kOffG2B, g);
  exchange<O0, NT>(xb, lane, hg, all);
  // H3 = tanh(fc3 H2_G); H4 = H3 * Gate2  (:108-112)
  f32x4 h3[NT];
  hidden_layer<O0, NT>(all, prm + kOffFc3W, buf0, buf1, col, g, h3);
  bias_tanh<O0, NT>(h3, prm + kOffFc3B, g);
#pragma unroll
  for (int o = 0; o < NT; ++o) h3[o] = h3[o] * g2[o];
  // out = fc4 H4  (:114): each half sums its K tiles, one 16-neuron tile of
  // channels 0..15; the halves' sums meet in LDS
  f32x4 o4 = f32x4{0.f, 0.f, 0.f, 0.f};
  const float* W4 = prm + kOffFc4W;
#pragma unroll
  for (int o = 0; o < NT; ++o)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      o4 = mfma(wload(W4, kSrcHid, kSrcOut, kSrcHid, col, 16 * (O0 + o) + 4 * g + i), h3[o][i],
                o4);
  if (H) *reinterpret_cast<f32x4*>(red + lane * 4) = o4;
  __syncthreads();
  if (!H && valid) {
    const f32x4 o1 = *reinterpret_cast<const f32x4*>(red + lane * 4);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = 4 * g + i;
      if (c < kSrcUsed) P[(int64_t)c * s.m + cell] = (o4[i] + o1[i]) + prm[kOffFc4B + c];
    }
  }
}

// 2. the network: 64 cells per 512-thread block, 16 per pair of waves, each
// wave of a pair on half of every layer's neurons (one wave per SIMD would
// otherwise hold the whole layer: 13 tiles per activation); P[c * m + cell]
// for channels c = 0..8.
__global__ void __launch_bounds__(kNetThreads) src_network_kernel(SrcGeom s,
                                                                  const float* __restrict__ Fn,
                                                                  const float* __restrict__ prm,
                                                                  float* __restrict__ P) {
  __shared__ __align__(16) float chunks[2 * kStageFloats];
  __shared__ __align__(16) float xchg[4 * kSrcTiles * 64 * 4];   // per cell group
  __shared__ __align__(16) float red[4 * 64 * 4];
  float* buf0 = chunks;
  float* buf1 = chunks + kStageFloats;
  const int lane = threadIdx.x & 63;
  // wave-uniform (kept in a scalar register, so the halves branch, not mask).
  // The waves of a block are dealt to the CU's four SIMDs in turn, so waves w
  // and w + 4 share a SIMD: making them the two halves of one cell group puts
  // 7 + 6 tiles of MFMA work on every SIMD (with half = w & 1 two SIMDs
  // would carry 7 + 7 and the other two 6 + 6).
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int grp = wave & 3, half = wave >> 2;
  const int col = lane & 15;
  const int64_t cell = ((int64_t)blockIdx.x * 4 + grp) * 16 + col;
  const bool valid = cell < s.m;
  int ci = 0, cj = 0, ck = 0;
  if (valid) decode_padded(s, cell, ci, cj, ck);
  float* xb = xchg + grp * kSrcTiles * 64 * 4;
  float* rb = red + grp * 64 * 4;
  if (half)
    net_half<1>(s, Fn, prm, P, buf0, buf1, xb, rb, lane, cell, valid, ci, cj, ck);
  else
    net_half<0>(s, Fn, prm, P, buf0, buf1, xb, rb, lane, cell, valid, ci, cj, ck);
}

// 3. GX, GY, GZ from the network's channels 0-8
template <typename T>
__global__ void __launch_bounds__(kBlock) src_divergence_kernel(SrcGeom s,
                                                                const float* __restrict__ P,
                                                                T* __restrict__ gx,
                                                                T* __restrict__ gy,
                                                                T* __restrict__ gz) {
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
    T* out[3] = {gx, gy, gz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float* p0 = P + (int64_t)(3 * q) * s.m;
      const float* p1 = p0 + s.m;
      const float* p2 = p1 + s.m;
      const float dxp = (p0[xp] - p0[xm]) / s.d2;
      const float dyp = (p1[yp] - p1[ym]) / s.d2;
      const float dzp = (p2[zp] - p2[zm]) / s.d2;
      out[q][p] = (T)((dxp + dyp) + dzp);                  // :586-588
    }
  }
}

template <typename T>
int ml_source_impl(pf_box* h, const float* prm, const Norm21& nrm, const void* const U[3],
                   void* const G[3], hipStream_t st) {
  SrcGeom s;
  s.nxo = h->nx + 2 * kNover;
  s.nyo = h->ny + 2 * kNover;
  s.nzo = h->nz + 2 * kNover;
  s.m = (int64_t)s.nxo * s.nyo * s.nzo;
  s.d2 = (float)(2.0 * h->dx);
  s.dd = (float)(h->dx * h->dx);
  if (h->src_ws_bytes < src_ws_need(h))
    return fail(PF_EINVAL, "pf_sfs_ml_source: workspace not reserved (call "
                           "pf_sfs_ml_source_reserve on the box first)");
  float* Fn = h->src_ws;
  float* P = Fn + (size_t)kSrcFeat * s.m;
  int64_t nb = (s.m + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(src_features_kernel<T>, dim3((unsigned)nb), dim3(kBlock), 0, st, s,
                     (const T*)U[0], (const T*)U[1], (const T*)U[2], Fn, nrm);
  PF_LAUNCH_CHECK("src_features_kernel");
  const int64_t nnet = (s.m + 63) / 64;
  hipLaunchKernelGGL(src_network_kernel, dim3((unsigned)nnet), dim3(kNetThreads), 0, st, s,
                     (const float*)Fn, prm, P);
  PF_LAUNCH_CHECK("src_network_kernel");
  hipLaunchKernelGGL(src_divergence_kernel<T>, dim3((unsigned)nb), dim3(kBlock), 0, st, s,
                     (const float*)P, (T*)G[0], (T*)G[1], (T*)G[2]);
  PF_LAUNCH_CHECK("src_divergence_kernel");
  return 0;
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_sfs_ml_source(pf_box* h, const float* params, const float* norm21, const void* ui,
                     const void* vi, const void* wi, void* gx, void* gy, void* gz,
                     void* stream) {
  if (int rc = check_box(h)) return rc;
  const void* U[3] = {ui, vi, wi};
  void* G[3] = {gx, gy, gz};
  if (!params) return fail(PF_EINVAL, "pf_sfs_ml_source: NULL params");
  if (!norm21) return fail(PF_EINVAL, "pf_sfs_ml_source: NULL normalisation table");
  for (int q = 0; q < 3; ++q) {
    if (!U[q] || !G[q]) return fail(PF_EINVAL, "pf_sfs_ml_source: NULL field");
    for (int r = 0; r < 3; ++r)
      if (G[q] == U[r]) return fail(PF_EINVAL, "pf_sfs_ml_source: outputs alias the inputs");
  }
  Norm21 nrm;
  for (int f = 0; f < kSrcFeat; ++f) {
    nrm.v[f] = norm21[f];
    if (!(nrm.v[f] != 0.f)) return fail(PF_EINVAL, "pf_sfs_ml_source: zero normalisation entry");
  }
  if (!(h->dx > 0.0)) return fail(PF_EINVAL, "pf_sfs_ml_source: dx must be positive");
  hipStream_t s = (hipStream_t)stream;
  return h->dtype == PF_F64 ? ml_source_impl<double>(h, params, nrm, U, G, s)
                            : ml_source_impl<float>(h, params, nrm, U, G, s);
}

int64_t pf_sfs_ml_source_params(void) { return kSrcParams; }

int pf_sfs_ml_source_reserve(pf_box* h) {
  if (int rc = check_box(h)) return rc;
  const size_t need = src_ws_need(h);
  if (h->src_ws_bytes >= need) return 0;
  if (h->src_ws) {
    PF_HIP(hipFree(h->src_ws));
    h->bytes -= (int64_t)h->src_ws_bytes;
    h->src_ws = nullptr;
    h->src_ws_bytes = 0;
  }
  hipError_t e = hipMalloc((void**)&h->src_ws, need);
  if (e != hipSuccess) {
    h->src_ws = nullptr;
    return check_hip(e, "hipMalloc");
  }
  h->src_ws_bytes = need;
  h->bytes += (int64_t)need;
  return 0;
}

}  // extern "C"
