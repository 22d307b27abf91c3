// Training through the reference's 'source' ML closure: the reverse of
// residual_stress.update (sfsmodel/sfsmodel_ML.py:354-593, NeuralNetworkModel2
// of :76-116) and the closure term of the adjoint RHS (solver/adjoint.py:
// 230-286), which the reference obtains with torch autograd:
//
//   g = sum over all padded cells of GX.uA + GY.vA + GZ.wA,
//   G = closure(var_i of the forward state),  uA = var_i of the adjoint state,
//   delta_q = d g / d var_i_q (FP32),   rhs_q -= face interpolant of delta_q,
//
// with the closure's parameter gradient accumulated on the way (the
// reference's g.backward() fills its model's .grad).  Everything below runs
// in FP32 as the reference's closure does; the parameter gradient is summed
// into an FP64 accumulator.
//
// The reverse runs as:
//  1. the closure's input features of every padded cell (pf_source.h);
//  2. per chunk of cells: the 273 network inputs (the cell's features and
//     those of its 12 neighbours), the forward activations of every layer,
//     the output adjoints from the divergence's transpose, the reverse through
//     the gated network, the input adjoints of the chunk's cells and the
//     parameter-gradient sums; every layer is a GEMM over the chunk's cells
//     (gemm_kernel: v_mfma_f32_16x16x4_f32, cells on M) followed by an
//     elementwise kernel (bias, tanh, gates, tanh' = 1 - y*y);
//  3. the feature adjoints of every cell (the 13 shifted copies gathered
//     back, divided by the normalisation table) and the transpose of the
//     finite differences: d g / d var_i.
// Parameter gradients are split-K GEMMs over the chunk's cells into FP32
// partials, summed in a fixed order into FP64: deterministic results.
#include <hip/hip_runtime.h>

#include <cstring>

#include "pf_source.h"

namespace pf {

extern "C" int pf_source_inputs(pf_box* h, const void* u, const void* v, const void* w, void* ui,
                                void* vi, void* wi, void* stream);
extern "C" int pf_sfs_ml_source_reserve(pf_box* h);

namespace {

constexpr int kChunkRows = 65536;   // cells per chunk of the network reverse
constexpr int kSplitK = 2048;       // cells per split of the parameter-gradient GEMMs
constexpr int kXw = kSrcIn;         // network input width (273)
constexpr int kH = kSrcHid;         // hidden width (200)

inline size_t al256(size_t n) { return (n + 255) / 256 * 256; }

// Layout of the box's training workspace.
struct Srct {
  int64_t m;        // padded cells
  int C;            // rows per chunk
  int splits;       // parameter-gradient splits of a full chunk
  size_t off_fwd, off_adj;            // 3 padded box-dtype arrays each
  size_t off_dX, off_dF, off_dvar;    // FP32: m x 273, 21 x m, 3 x m
  size_t off_X, off_act, off_dout, off_adjb, off_ones, off_part;
  size_t bytes;
};

Srct srct_layout(const pf_box* h) {
  Srct L;
  L.m = (int64_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4);
  L.C = L.m < kChunkRows ? (int)L.m : kChunkRows;
  L.splits = (L.C + kSplitK - 1) / kSplitK;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t at = o;
    o += al256(n);
    return at;
  };
  const size_t pad = (size_t)L.m * h->esize;
  L.off_fwd = take(3 * pad);
  L.off_adj = take(3 * pad);
  L.off_dX = take((size_t)L.m * kXw * 4);
  L.off_dF = take((size_t)L.m * kSrcFeat * 4);
  L.off_dvar = take((size_t)L.m * 3 * 4);
  L.off_X = take((size_t)L.C * kXw * 4);
  L.off_act = take((size_t)L.C * kH * 4 * 7);      // H1, G, Gate2, H2, H2G, H3, H4
  L.off_dout = take((size_t)L.C * kSrcOut * 4);
  L.off_adjb = take((size_t)L.C * kH * 4 * 6);     // dH, dL1, dLG, dLG2, dL2, dL4
  L.off_ones = take((size_t)L.C * 4);
  L.off_part = take((size_t)L.splits * kSrcParams * 4);
  L.bytes = o;
  return L;
}

// ---------------------------------------------------------------------------
// C (op) op(A) op(B): A is M x K, B is K x N after their transposes; row
// major storage with leading dimensions.  TA: A stored as [K][M]; TB: B
// stored as [N][K].  Split-K over blockIdx.z (K range kchunk, C offset cz).
// act: 0 store, 1 store tanh(acc + bias), 2 accumulate, 3 store acc + bias.
struct GemmOp {
  const float* A;
  int64_t lda;
  int ta;
  const float* B;
  int64_t ldb;
  int tb;
  float* C;
  int64_t ldc;
  int M, N, K;
  int kchunk;
  int64_t cz;
  const float* bias;
  int act;
};

constexpr int kGM = 64, kGN = 64, kGK = 16;

// 256 threads, a 64 x 64 tile of C; each wave a 32 x 32 quarter as 2 x 2
// MFMA tiles of 16 x 16 (lane l: A[l%16][l/16], B[l/16][l%16], C rows
// 4*(l/16)+i, column l%16).  The K slice of each step is staged through LDS.
__global__ void __launch_bounds__(256) gemm_kernel(GemmOp op) {
  __shared__ float As[kGK][kGM + 4];
  __shared__ float Bs[kGK][kGN + 4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * kGM, n0 = blockIdx.x * kGN;
  const int kb = blockIdx.z * op.kchunk;
  const int ke = min(op.K, kb + op.kchunk);
  float* C = op.C + (int64_t)blockIdx.z * op.cz;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += kGK) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t * 4 + q;
      const int m = op.ta ? (e & 63) : (e >> 4);
      const int k = op.ta ? (e >> 6) : (e & 15);
      const int gm = m0 + m, gk = k0 + k;
      float v = 0.f;
      if (gm < op.M && gk < ke)
        v = op.ta ? op.A[(int64_t)gk * op.lda + gm] : op.A[(int64_t)gm * op.lda + gk];
      As[k][m] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t * 4 + q;
      const int n = op.tb ? (e >> 4) : (e & 63);
      const int k = op.tb ? (e & 15) : (e >> 6);
      const int gn = n0 + n, gk = k0 + k;
      float v = 0.f;
      if (gn < op.N && gk < ke)
        v = op.tb ? op.B[(int64_t)gn * op.ldb + gk] : op.B[(int64_t)gk * op.ldb + gn];
      Bs[k][n] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGK; ks += 4) {
      const int kk = ks + (lane >> 4);
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk][wm + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk][wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn + j * 16 + (lane & 15);
        if (row >= op.M || col >= op.N) continue;
        float* c = C + (int64_t)row * op.ldc + col;
        const float v = acc[i][j][r];
        if (op.act == 2) {
          *c = *c + v;
        } else if (op.act == 1) {
          *c = tanhf(v + op.bias[col]);
        } else if (op.act == 3) {
          *c = v + op.bias[col];
        } else {
          *c = v;
        }
      }
}

int gemm(const GemmOp& op, int splits, hipStream_t s) {
  if (op.M <= 0 || op.N <= 0 || op.K <= 0) return 0;
  const dim3 grid((op.N + kGN - 1) / kGN, (op.M + kGM - 1) / kGM, splits);
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, op);
  PF_LAUNCH_CHECK("gemm_kernel");
  return 0;
}

GemmOp op_of(const float* A, int64_t lda, int ta, const float* B, int64_t ldb, int tb, float* C,
             int64_t ldc, int M, int N, int K, int act, const float* bias = nullptr) {
  GemmOp o;
  o.A = A; o.lda = lda; o.ta = ta;
  o.B = B; o.ldb = ldb; o.tb = tb;
  o.C = C; o.ldc = ldc;
  o.M = M; o.N = N; o.K = K;
  o.kchunk = K;
  o.cz = 0;
  o.bias = bias;
  o.act = act;
  return o;
}

// ---------------------------------------------------------------------------
// Elementwise and stencil kernels (one thread per element or per cell).

// Centre interpolants of the adjoint state (metric_staggered.py:149-165):
// out_q[p] = 0.5*(a_q[p + e_q] + a_q[p]) except on the last plane along q,
// which nothing writes (the reference's adjoint var_i keeps its zero there).
template <typename T>
__global__ void __launch_bounds__(kBlock) centre_interp_kernel(SrcGeom s, CVec3<T> A,
                                                               Vec3<T> out) {
  const int64_t sy = s.nzo, sx = (int64_t)s.nyo * s.nzo;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.m;
       p += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decode_padded(s, p, i, j, k);
    const T half = (T)0.5;
    if (i < s.nxo - 1) out.a[0][p] = half * (A.a[0][p + sx] + A.a[0][p]);
    if (j < s.nyo - 1) out.a[1][p] = half * (A.a[1][p + sy] + A.a[1][p]);
    if (k < s.nzo - 1) out.a[2][p] = half * (A.a[2][p + 1] + A.a[2][p]);
  }
}

// The network input of chunk row r (cell c0 + r): the 21 normalised features
// of the cell and of its 12 neighbours, in the reference's order (:449-462).
__global__ void __launch_bounds__(kBlock) gather_x_kernel(SrcGeom s, const float* __restrict__ Fn,
                                                          float* __restrict__ X, int64_t c0,
                                                          int rows) {
  const int64_t n = (int64_t)rows * kSrcShifts;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / kSrcShifts), sh = (int)(e % kSrcShifts);
    int i, j, k;
    decode_padded(s, c0 + r, i, j, k);
    const int64_t q = shifted(s, i, j, k, sh);
    float* x = X + (int64_t)r * kXw + sh * kSrcFeat;
#pragma unroll
    for (int f = 0; f < kSrcFeat; ++f) x[f] = Fn[(int64_t)f * s.m + q];
  }
}

// out = a * b over n elements (the gates: H2G = G*H2, H4 = H3*Gate2).
__global__ void __launch_bounds__(kBlock) mul_kernel(int64_t n, const float* __restrict__ a,
                                                     const float* __restrict__ b,
                                                     float* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = a[e] * b[e];
}

// Output adjoints from the divergence (:466-589): G_q = sum over axes a of
// (P_{3q+a}[c+e_a] - P_{3q+a}[c-e_a])/(2 dx) with the padded wrap, so
// dP_{3q+a}[c] = dG_q[c-e_a]/(2 dx) - dG_q[c+e_a]/(2 dx); dG_q = float(the
// adjoint's centre interpolant), since g multiplies G (cast to the box
// dtype) by it.  Channels 9..17 get no adjoint.
template <typename T>
__global__ void __launch_bounds__(kBlock) dout_kernel(SrcGeom s, CVec3<T> gI,
                                                      float* __restrict__ dout, int64_t c0,
                                                      int rows) {
  const int64_t sy = s.nzo, sx = (int64_t)s.nyo * s.nzo;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = c0 + r;
    int i, j, k;
    decode_padded(s, p, i, j, k);
    const int64_t nb[3][2] = {
        {(int64_t)wrap(i - 1, s.nxo) * sx + j * sy + k, (int64_t)wrap(i + 1, s.nxo) * sx + j * sy + k},
        {(int64_t)i * sx + wrap(j - 1, s.nyo) * sy + k, (int64_t)i * sx + wrap(j + 1, s.nyo) * sy + k},
        {(int64_t)i * sx + j * sy + wrap(k - 1, s.nzo), (int64_t)i * sx + j * sy + wrap(k + 1, s.nzo)}};
    float* o = dout + r * kSrcOut;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float gm = (float)gI.a[q][nb[a][0]], gp = (float)gI.a[q][nb[a][1]];
        o[3 * q + a] = gm / s.d2 - gp / s.d2;
      }
#pragma unroll
    for (int c = kSrcUsed; c < kSrcOut; ++c) o[c] = 0.f;
  }
}

// Reverse through H4 = H3*Gate2, H3 = tanh(L4), Gate2 = tanh(LG2):
// dL4 = (dH4*Gate2)*(1 - H3*H3), dLG2 = (dH4*H3)*(1 - Gate2*Gate2).
__global__ void __launch_bounds__(kBlock) rev_out_kernel(int64_t n, const float* __restrict__ dH,
                                                         const float* __restrict__ H3,
                                                         const float* __restrict__ Gate2,
                                                         float* __restrict__ dL4,
                                                         float* __restrict__ dLG2) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float d = dH[e], h3 = H3[e], g2 = Gate2[e];
    dL4[e] = (d * g2) * (1.f - h3 * h3);
    dLG2[e] = (d * h3) * (1.f - g2 * g2);
  }
}

// Reverse through H2G = G*H2, G = tanh(LG), H2 = tanh(L2).
__global__ void __launch_bounds__(kBlock) rev_gate_kernel(int64_t n, const float* __restrict__ dH,
                                                          const float* __restrict__ G,
                                                          const float* __restrict__ H2,
                                                          float* __restrict__ dLG,
                                                          float* __restrict__ dL2) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float d = dH[e], gv = G[e], h2 = H2[e];
    dLG[e] = (d * h2) * (1.f - gv * gv);
    dL2[e] = (d * gv) * (1.f - h2 * h2);
  }
}

// Reverse through H1 = tanh(L1).
__global__ void __launch_bounds__(kBlock) rev_tanh_kernel(int64_t n, const float* __restrict__ dH,
                                                          const float* __restrict__ H,
                                                          float* __restrict__ dL) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float h = H[e];
    dL[e] = dH[e] * (1.f - h * h);
  }
}

// The shift that undoes shift sh (x+1 <-> x-1, ...).
__device__ __forceinline__ int opposite(int sh) {
  if (sh == 0) return 0;
  const int base = sh <= 6 ? 1 : 7;
  return base + (((sh - base) ^ 1));
}

// Feature adjoints: slot sh of cell c reads the features of c + shift(sh),
// so dF[f][p] = sum over slots of dX[p - shift(sh)][sh*21 + f] / norm[f].
__global__ void __launch_bounds__(kBlock) feat_adj_kernel(SrcGeom s, const float* __restrict__ dX,
                                                          float* __restrict__ dF, Norm21 nrm) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.m;
       p += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decode_padded(s, p, i, j, k);
    float acc[kSrcFeat];
#pragma unroll
    for (int f = 0; f < kSrcFeat; ++f) acc[f] = 0.f;
    for (int sh = 0; sh < kSrcShifts; ++sh) {
      const int64_t c = shifted(s, i, j, k, opposite(sh));
      const float* x = dX + c * kXw + sh * kSrcFeat;
#pragma unroll
      for (int f = 0; f < kSrcFeat; ++f) acc[f] += x[f] / nrm.v[f];
    }
#pragma unroll
    for (int f = 0; f < kSrcFeat; ++f) dF[(int64_t)f * s.m + p] = acc[f];
  }
}

// Transpose of the features (pf_source.h, :408-465): feature q is the value,
// 3+3q+a the first difference (f[+e] - f[-e])/(2 dx), 12+3q+a the second
// difference ((f[+e] - 2 f) + f[-e])/(dx dx) along axis a, padded wrap.
__global__ void __launch_bounds__(kBlock) var_adj_kernel(SrcGeom s, const float* __restrict__ dF,
                                                         float* __restrict__ dvar) {
  const int64_t sy = s.nzo, sx = (int64_t)s.nyo * s.nzo;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < s.m;
       p += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decode_padded(s, p, i, j, k);
    const int64_t nb[3][2] = {
        {(int64_t)wrap(i - 1, s.nxo) * sx + j * sy + k, (int64_t)wrap(i + 1, s.nxo) * sx + j * sy + k},
        {(int64_t)i * sx + wrap(j - 1, s.nyo) * sy + k, (int64_t)i * sx + wrap(j + 1, s.nyo) * sy + k},
        {(int64_t)i * sx + j * sy + wrap(k - 1, s.nzo), (int64_t)i * sx + j * sy + wrap(k + 1, s.nzo)}};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float d = dF[(int64_t)q * s.m + p];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float* d1 = dF + (int64_t)(3 + 3 * q + a) * s.m;
        const float* d2 = dF + (int64_t)(12 + 3 * q + a) * s.m;
        d += d1[nb[a][0]] / s.d2 - d1[nb[a][1]] / s.d2;
        d += d2[nb[a][0]] / s.dd + d2[nb[a][1]] / s.dd - 2.0f * (d2[p] / s.dd);
      }
      dvar[(int64_t)q * s.m + p] = d;
    }
  }
}

// The closure term of the adjoint RHS at interior cells (adjoint.py:278-286:
// interp_sc_x/y/z of the FP32 adjoint, then rhs -= it): D_q = -(0.5*(d[c] +
// d[c - e_q])) in FP32, stored in the box dtype.
template <typename T>
__global__ void __launch_bounds__(kBlock) delta_kernel(Geom g, const float* __restrict__ dvar,
                                                       int64_t m, Vec3<T> D) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    const int64_t back[3] = {g.psx, g.psy, 1};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float* d = dvar + q * m;
      const float v = 0.5f * (d[c] + d[c - back[q]]);
      D.a[q][c] = -(T)v;
    }
  }
}

// dst[p] += sum over z of part[z*stride + p], z in order (FP64).
__global__ void __launch_bounds__(kBlock) add_partials_kernel(double* __restrict__ dst,
                                                              const float* __restrict__ part,
                                                              int64_t n, int splits,
                                                              int64_t stride) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int z = 0; z < splits; ++z) a += (double)part[(int64_t)z * stride + p];
    dst[p] += a;
  }
}

__global__ void __launch_bounds__(kBlock) fill_kernel(int64_t n, float* __restrict__ a, float v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    a[e] = v;
}

unsigned grid_of(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

SrcGeom src_geom(const pf_box* h) {
  SrcGeom s;
  s.nxo = h->nx + 2 * kNover;
  s.nyo = h->ny + 2 * kNover;
  s.nzo = h->nz + 2 * kNover;
  s.m = (int64_t)s.nxo * s.nyo * s.nzo;
  s.d2 = (float)(2.0 * h->dx);
  s.dd = (float)(h->dx * h->dx);
  return s;
}

#define PF_ELEM(kernel, n, ...)                                                              \
  do {                                                                                       \
    hipLaunchKernelGGL(kernel, dim3(grid_of(n)), dim3(kBlock), 0, st, __VA_ARGS__);          \
    PF_LAUNCH_CHECK(#kernel);                                                                \
  } while (0)

// d g / d var_i into dvar (3 x m FP32) for g = sum(G(fwdI) . gI), and the
// parameter gradient added to dparams (FP64, may be NULL).
template <typename T>
int ml_vjp_impl(pf_box* h, const float* prm, const void* const fwdI[3], const void* const gI[3],
                float* dvar, double* dparams, hipStream_t st) {
  const Srct L = srct_layout(h);
  const SrcGeom s = src_geom(h);
  char* base = (char*)h->srct_ws;
  Norm21 nrm;
  for (int f = 0; f < kSrcFeat; ++f) nrm.v[f] = h->srct_norm[f];
  float* Fn = h->src_ws;
  float* dX = (float*)(base + L.off_dX);
  float* dF = (float*)(base + L.off_dF);
  float* X = (float*)(base + L.off_X);
  float* act = (float*)(base + L.off_act);
  float* H1 = act;
  float* Gt = act + (size_t)L.C * kH;
  float* G2 = act + (size_t)2 * L.C * kH;
  float* H2 = act + (size_t)3 * L.C * kH;
  float* H2G = act + (size_t)4 * L.C * kH;
  float* H3 = act + (size_t)5 * L.C * kH;
  float* H4 = act + (size_t)6 * L.C * kH;
  float* dout = (float*)(base + L.off_dout);
  float* ab = (float*)(base + L.off_adjb);
  float* dH = ab;
  float* dL1 = ab + (size_t)L.C * kH;
  float* dLG = ab + (size_t)2 * L.C * kH;
  float* dLG2 = ab + (size_t)3 * L.C * kH;
  float* dL2 = ab + (size_t)4 * L.C * kH;
  float* dL4 = ab + (size_t)5 * L.C * kH;
  float* ones = (float*)(base + L.off_ones);
  float* part = (float*)(base + L.off_part);
  const float* W1 = prm + kOffFc1W;
  const float* b1 = prm + kOffFc1B;
  const float* WG = prm + kOffGW;
  const float* bG = prm + kOffGB;
  const float* WG2 = prm + kOffG2W;
  const float* bG2 = prm + kOffG2B;
  const float* W2 = prm + kOffFc2W;
  const float* b2 = prm + kOffFc2B;
  const float* W3 = prm + kOffFc3W;
  const float* b3 = prm + kOffFc3B;
  const float* W4 = prm + kOffFc4W;

  // 1. features of every padded cell
  hipLaunchKernelGGL(src_features_kernel<T>, dim3(grid_of(s.m)), dim3(kBlock), 0, st, s,
                     (const T*)fwdI[0], (const T*)fwdI[1], (const T*)fwdI[2], Fn, nrm);
  PF_LAUNCH_CHECK("src_features_kernel");
  PF_ELEM(fill_kernel, L.C, (int64_t)L.C, ones, 1.f);
  CVec3<T> g3;
  for (int q = 0; q < 3; ++q) g3.a[q] = (const T*)gI[q];

  // 2. chunks of cells
  for (int64_t c0 = 0; c0 < s.m; c0 += L.C) {
    const int R = (int)((s.m - c0) < L.C ? (s.m - c0) : L.C);
    const int64_t nRH = (int64_t)R * kH;
    PF_ELEM(gather_x_kernel, (int64_t)R * kSrcShifts, s, (const float*)Fn, X, c0, R);
    // forward: L1, LG, LG2 from X; L2 from H1; L4 from G*H2; H4 = H3*Gate2
    if (int rc = gemm(op_of(X, kXw, 0, W1, kXw, 1, H1, kH, R, kH, kXw, 1, b1), 1, st)) return rc;
    if (int rc = gemm(op_of(X, kXw, 0, WG, kXw, 1, Gt, kH, R, kH, kXw, 1, bG), 1, st)) return rc;
    if (int rc = gemm(op_of(X, kXw, 0, WG2, kXw, 1, G2, kH, R, kH, kXw, 1, bG2), 1, st)) return rc;
    if (int rc = gemm(op_of(H1, kH, 0, W2, kH, 1, H2, kH, R, kH, kH, 1, b2), 1, st)) return rc;
    PF_ELEM(mul_kernel, nRH, nRH, (const float*)Gt, (const float*)H2, H2G);
    if (int rc = gemm(op_of(H2G, kH, 0, W3, kH, 1, H3, kH, R, kH, kH, 1, b3), 1, st)) return rc;
    PF_ELEM(mul_kernel, nRH, nRH, (const float*)H3, (const float*)G2, H4);
    // reverse: output adjoints, then back through fc4, fc3, the gates, fc2, fc1/fcG/fcG2
    PF_ELEM(dout_kernel<T>, R, s, g3, dout, c0, R);
    if (int rc = gemm(op_of(dout, kSrcOut, 0, W4, kH, 0, dH, kH, R, kH, kSrcOut, 0), 1, st))
      return rc;
    PF_ELEM(rev_out_kernel, nRH, nRH, (const float*)dH, (const float*)H3, (const float*)G2, dL4,
            dLG2);
    if (int rc = gemm(op_of(dL4, kH, 0, W3, kH, 0, dH, kH, R, kH, kH, 0), 1, st)) return rc;
    PF_ELEM(rev_gate_kernel, nRH, nRH, (const float*)dH, (const float*)Gt, (const float*)H2, dLG,
            dL2);
    if (int rc = gemm(op_of(dL2, kH, 0, W2, kH, 0, dH, kH, R, kH, kH, 0), 1, st)) return rc;
    PF_ELEM(rev_tanh_kernel, nRH, nRH, (const float*)dH, (const float*)H1, dL1);
    float* dXc = dX + c0 * kXw;
    if (int rc = gemm(op_of(dL1, kH, 0, W1, kXw, 0, dXc, kXw, R, kXw, kH, 0), 1, st)) return rc;
    if (int rc = gemm(op_of(dLG, kH, 0, WG, kXw, 0, dXc, kXw, R, kXw, kH, 2), 1, st)) return rc;
    if (int rc = gemm(op_of(dLG2, kH, 0, WG2, kXw, 0, dXc, kXw, R, kXw, kH, 2), 1, st))
      return rc;
    if (!dparams) continue;
    // parameter gradients of this chunk: dW = dL^T act and db = dL^T 1 over
    // its cells, split over kSplitK-cell ranges into FP32 partials laid out
    // like the packed parameters, then summed in order into FP64
    const int splits = (R + kSplitK - 1) / kSplitK;
    struct Pg {
      const float* dL;
      int out;
      const float* a;
      int in;
      int64_t w, b;
    };
    const Pg pg[6] = {{dL1, kH, X, kXw, kOffFc1W, kOffFc1B},
                      {dLG, kH, X, kXw, kOffGW, kOffGB},
                      {dLG2, kH, X, kXw, kOffG2W, kOffG2B},
                      {dL2, kH, H1, kH, kOffFc2W, kOffFc2B},
                      {dL4, kH, H2G, kH, kOffFc3W, kOffFc3B},
                      {dout, kSrcOut, H4, kH, kOffFc4W, kOffFc4B}};
    for (const Pg& q : pg) {
      GemmOp o = op_of(q.dL, q.out, 1, q.a, q.in, 0, part + q.w, q.in, q.out, q.in, R, 0);
      o.kchunk = kSplitK;
      o.cz = kSrcParams;
      if (int rc = gemm(o, splits, st)) return rc;
      GemmOp ob = op_of(ones, 1, 1, q.dL, q.out, 0, part + q.b, q.out, 1, q.out, R, 0);
      ob.kchunk = kSplitK;
      ob.cz = kSrcParams;
      if (int rc = gemm(ob, splits, st)) return rc;
    }
    PF_ELEM(add_partials_kernel, kSrcParams, dparams, (const float*)part, (int64_t)kSrcParams,
            splits, (int64_t)kSrcParams);
  }
  // 3. feature adjoints, then d g / d var_i
  PF_ELEM(feat_adj_kernel, s.m, s, (const float*)dX, dF, nrm);
  PF_ELEM(var_adj_kernel, s.m, s, (const float*)dF, dvar);
  return 0;
}

template <typename T>
int ml_adj_term_impl(pf_box* h, const void* const U[3], const float* prm, const void* const A[3],
                     void* const D[3], double* dparams, hipStream_t st) {
  const Srct L = srct_layout(h);
  const SrcGeom s = src_geom(h);
  char* base = (char*)h->srct_ws;
  const size_t pad = (size_t)L.m * sizeof(T);
  void* fwd[3];
  void* adj[3];
  for (int q = 0; q < 3; ++q) {
    fwd[q] = base + L.off_fwd + q * pad;
    adj[q] = base + L.off_adj + q * pad;
  }
  // the forward state's var_i as the reference's adjoint RHS leaves them
  // before its closure block (adjoint.py:170-237: the same interpolations as
  // the forward RHS, so the same stale last planes)
  if (int rc = pf_source_inputs(h, U[0], U[1], U[2], fwd[0], fwd[1], fwd[2], st)) return rc;
  CVec3<T> a3;
  Vec3<T> o3;
  for (int q = 0; q < 3; ++q) {
    a3.a[q] = (const T*)A[q];
    o3.a[q] = (T*)adj[q];
  }
  hipLaunchKernelGGL(centre_interp_kernel<T>, dim3(grid_of(s.m)), dim3(kBlock), 0, st, s, a3, o3);
  PF_LAUNCH_CHECK("centre_interp_kernel");
  float* dvar = (float*)(base + L.off_dvar);
  const void* const fc[3] = {fwd[0], fwd[1], fwd[2]};
  const void* const ac[3] = {adj[0], adj[1], adj[2]};
  if (int rc = ml_vjp_impl<T>(h, prm, fc, ac, dvar, dparams, st)) return rc;
  Vec3<T> d3;
  for (int q = 0; q < 3; ++q) d3.a[q] = (T*)D[q];
  hipLaunchKernelGGL(delta_kernel<T>, dim3(launch_blocks(h->g)), dim3(kBlock), 0, st, h->g,
                     (const float*)dvar, L.m, d3);
  PF_LAUNCH_CHECK("delta_kernel");
  return 0;
}

int check_reserved(const pf_box* h, const char* what) {
  if (!h->srct_ws || h->srct_bytes < srct_layout(h).bytes || !h->src_ws ||
      h->src_ws_bytes < src_ws_need(h))
    return fail(PF_EINVAL, std::string(what) + ": training workspace not reserved (call "
                                               "pf_sfs_ml_source_train_reserve on the box first)");
  return 0;
}

}  // namespace

int ml_source_adj_term(pf_box* h, const void* const U[3], const float* params,
                       const void* const A[3], void* const D[3], double* dparams, hipStream_t s) {
  if (int rc = check_reserved(h, "adjoint closure term of the 'source' ML closure")) return rc;
  if (!params) return fail(PF_EINVAL, "adjoint closure term: NULL params");
  return h->dtype == PF_F64 ? ml_adj_term_impl<double>(h, U, params, A, D, dparams, s)
                            : ml_adj_term_impl<float>(h, U, params, A, D, dparams, s);
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_sfs_ml_source_train_reserve(pf_box* h, const float* norm21) {
  if (int rc = check_box(h)) return rc;
  if (!norm21) return fail(PF_EINVAL, "pf_sfs_ml_source_train_reserve: NULL normalisation table");
  for (int f = 0; f < kSrcFeat; ++f)
    if (!(norm21[f] != 0.f))
      return fail(PF_EINVAL, "pf_sfs_ml_source_train_reserve: zero normalisation entry");
  if (int rc = pf_sfs_ml_source_reserve(h)) return rc;
  const size_t need = srct_layout(h).bytes;
  if (h->srct_bytes < need) {
    if (h->srct_ws) {
      PF_HIP(hipFree(h->srct_ws));
      h->bytes -= (int64_t)h->srct_bytes;
      h->srct_ws = nullptr;
      h->srct_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->srct_ws, need);
    if (e != hipSuccess) {
      h->srct_ws = nullptr;
      return check_hip(e, "hipMalloc");
    }
    // the closure's input arrays keep zeros where nothing writes them (the
    // last planes of the adjoint's centre interpolants, one corner of the
    // forward ones), as the reference's freshly allocated var_i do
    e = hipMemset(h->srct_ws, 0, need);
    if (e != hipSuccess) return check_hip(e, "hipMemset");
    h->srct_bytes = need;
    h->bytes += (int64_t)need;
  }
  for (int f = 0; f < kSrcFeat; ++f) h->srct_norm[f] = norm21[f];
  return 0;
}

int pf_sfs_ml_source_vjp(pf_box* h, const float* params, const void* ui, const void* vi,
                         const void* wi, const void* gx, const void* gy, const void* gz,
                         float* dui, float* dvi, float* dwi, double* dparams, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (int rc = check_reserved(h, "pf_sfs_ml_source_vjp")) return rc;
  if (!params || !ui || !vi || !wi || !gx || !gy || !gz)
    return fail(PF_EINVAL, "pf_sfs_ml_source_vjp: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  const Srct L = srct_layout(h);
  float* dvar = (float*)((char*)h->srct_ws + L.off_dvar);
  const void* const F[3] = {ui, vi, wi};
  const void* const G[3] = {gx, gy, gz};
  const int rc = h->dtype == PF_F64 ? ml_vjp_impl<double>(h, params, F, G, dvar, dparams, s)
                                    : ml_vjp_impl<float>(h, params, F, G, dvar, dparams, s);
  if (rc) return rc;
  float* out[3] = {dui, dvi, dwi};
  for (int q = 0; q < 3; ++q)
    if (out[q])
      PF_HIP(hipMemcpyAsync(out[q], dvar + (size_t)q * L.m, (size_t)L.m * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
  return 0;
}

}  // extern "C"
