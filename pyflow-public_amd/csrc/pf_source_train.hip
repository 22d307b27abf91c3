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
//  1. the closure's input features of every padded cell (pf_source.h) and a
//     copy of the weights in the products' padded layouts (pack_kernel);
//  2. per chunk of cells: the 273 network inputs (the cell's features and
//     those of its 12 neighbours), the forward activations of every layer,
//     the output adjoints from the divergence's transpose, the reverse through
//     the gated network into the input adjoints of the chunk's cells, and the
//     parameter-gradient sums.  Every layer is one tiled FP32 MFMA product
//     (mm_kernel, v_mfma_f32_16x16x4_f32, cells on M) whose epilogue applies
//     the layer's elementwise part (bias and tanh, the gates, tanh' = 1-y*y);
//     fc1, fcG and fcG2 share their input, so their three products are one
//     (624 = 3 x 208 columns), and so are their three input-adjoint products
//     (K = 624);
//  3. the feature adjoints of every cell (the 13 shifted copies gathered
//     back, divided by the normalisation table) and the transpose of the
//     finite differences: d g / d var_i.
// Parameter gradients are split-K products over 1,024-cell ranges of the
// chunk into FP32 partials laid out like the packed parameters (the bias
// gradient is the product's column against a column of ones), summed in a
// fixed order into FP64: deterministic results.
#include <hip/hip_runtime.h>

#include <cstring>

#include "pf_source.h"

namespace pf {

extern "C" int pf_source_inputs(pf_box* h, const void* u, const void* v, const void* w, void* ui,
                                void* vi, void* wi, void* stream);
extern "C" int pf_sfs_ml_source_reserve(pf_box* h);

namespace {

constexpr int kChunkRows = 65536;   // cells per chunk of the network reverse
constexpr int kSplitK = 1024;       // cells per split of the parameter-gradient products
constexpr int kMaxSplits = kChunkRows / kSplitK;
constexpr int kH = kSrcHid;         // hidden width (200)
// Padded widths: every reduction length is a multiple of the 16-deep K
// stage, so products never mask K except over cells; pad columns hold zeros.
constexpr int kHp = 208;            // hidden width padded (13 x 16)
constexpr int kCat = 3 * kHp;       // fc1 | fcG | fcG2 outputs (624)
constexpr int kXld = 288;           // network input width padded (273 -> 18 x 16)
constexpr int kOutp = 32;           // fc4 outputs padded (18 -> 2 x 16)
constexpr int kTileM = 128;         // rows (cells) per product tile
// packed weights: Wcat [624][288] (fc1, fcG, fcG2 rows at 0, 208, 416),
// W2p, W3p [208][208], W4p [32][208]
constexpr int64_t kPkCat = 0;
constexpr int64_t kPkW2 = kPkCat + (int64_t)kCat * kXld;
constexpr int64_t kPkW3 = kPkW2 + (int64_t)kHp * kHp;
constexpr int64_t kPkW4 = kPkW3 + (int64_t)kHp * kHp;
constexpr int64_t kPkFloats = kPkW4 + (int64_t)kOutp * kHp;

inline size_t al256(size_t n) { return (n + 255) / 256 * 256; }

// Layout of the box's training workspace.  The network's forward
// activations (X, H1 | G | Gate2, H2, H2G, H3, H4) are kept either for one
// chunk of cells (rows = Ca) or, when the box's cache fits its budget
// (kCacheBudget unless pf_sfs_ml_source_cache_budget set one), for
// every cell (rows = Ma), together with the sums of the layer adjoints over
// the stages of an adjoint RK4 step: the forward state is the same in all
// four stages, so the step evaluates the forward once and takes the weight
// gradient of its four stages as one product of the summed adjoints with
// the (stage-independent) layer inputs.
constexpr size_t kCacheBudget = (size_t)48 << 30;

struct Srct {
  int64_t m;        // padded cells
  int C;            // rows per chunk
  int Ca;           // rows allocated per chunk array (C rounded up to a tile)
  int64_t Ma;       // m rounded up to a tile
  bool full;        // forward activations and stage sums for every cell
  int64_t Mr;       // rows of the forward-activation arrays (Ma or Ca)
  size_t off_fwd, off_adj;            // 3 padded box-dtype arrays each
  size_t off_dX, off_dF, off_dvar;    // FP32: m x 288, 21 x m, 3 x m
  size_t off_X, off_cat, off_h;       // forward activations, Mr rows
  size_t off_dout, off_dcat, off_dl;  // a chunk's layer adjoints
  size_t off_sout, off_scat, off_sl;  // stage sums of the layer adjoints (full only)
  size_t off_pack, off_part;
  size_t bytes;
};

// Floats per row of the full-cache arrays: forward activations and sums.
constexpr size_t kCacheRowFloats = kXld + kCat + 4 * kHp + kOutp + kCat + 2 * kHp;

Srct srct_layout(const pf_box* h) {
  Srct L;
  L.m = (int64_t)(h->nx + 4) * (h->ny + 4) * (h->nz + 4);
  L.C = L.m < kChunkRows ? (int)L.m : kChunkRows;
  L.Ca = (L.C + kTileM - 1) / kTileM * kTileM;
  L.Ma = (L.m + kTileM - 1) / kTileM * kTileM;
  const size_t budget = h->srct_budget < 0 ? kCacheBudget : (size_t)h->srct_budget;
  L.full = (size_t)L.Ma * kCacheRowFloats * 4 <= budget;
  L.Mr = L.full ? L.Ma : L.Ca;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t at = o;
    o += al256(n);
    return at;
  };
  const size_t pad = (size_t)L.m * h->esize;
  const size_t ca = (size_t)L.Ca, mr = (size_t)L.Mr;
  L.off_fwd = take(3 * pad);
  L.off_adj = take(3 * pad);
  L.off_dX = take((size_t)L.m * kXld * 4);
  L.off_dF = take((size_t)L.m * kSrcFeat * 4);
  L.off_dvar = take((size_t)L.m * 3 * 4);
  L.off_X = take(mr * kXld * 4);
  L.off_cat = take(mr * kCat * 4);          // H1 | G | Gate2
  L.off_h = take(mr * kHp * 4 * 4);         // H2, H2G, H3, H4
  L.off_dout = take(ca * kOutp * 4);
  L.off_dcat = take(ca * kCat * 4);         // dL1 | dLG | dLG2
  L.off_dl = take(ca * kHp * 4 * 2);        // dL4, dL2
  const size_t ms = L.full ? (size_t)L.Ma : 0;
  L.off_sout = take(ms * kOutp * 4);
  L.off_scat = take(ms * kCat * 4);
  L.off_sl = take(ms * kHp * 4 * 2);
  L.off_pack = take((size_t)kPkFloats * 4);
  L.off_part = take((size_t)kMaxSplits * kSrcParams * 4);
  L.bytes = o;
  return L;
}

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// C = A B over a tile grid, A M x K, B K x N (FP32, row-major, leading
// dimensions in floats, every row 16-byte aligned).  ACOL: A stored as
// [K][M]; else [M][K].  BCOL: B stored as [K][N]; else [N][K].  Split-K over
// `splits` ranges of kchunk (only the cell dimension is split; it is the only
// K that is masked).  Epilogues (EPI), with d = the tile's accumulator at
// (row, col):
//   kEpiStore     C = d (col < N)
//   kEpiTanhCat   C = tanh(d + bias) over 3 groups of 208 columns (bias of
//                 group col/208 at col%208 < 200, else 0)
//   kEpiTanhGate  h = tanh(d + bias[col]); C = h; C2 = h * E1
//   kEpiRevPair   C = (d*E1)*(1 - E2*E2); C2 = (d*E2)*(1 - E1*E1)
//   kEpiRevTanh   C = d*(1 - E1*E1)
//   kEpiPartial   the parameter-gradient partials of split z: row = group
//                 row/rg, output o = row%rg (< nout); col < nin: weight
//                 [o][col], col == nin: the bias (B's column nin is ones)
// SUMS (kEpiRevPair, kEpiRevTanh): C and C2 are also added to S1 and S2.
enum { kEpiStore, kEpiTanhCat, kEpiTanhGate, kEpiRevPair, kEpiRevTanh, kEpiPartial };

struct MmOp {
  const float* A;
  int64_t lda;
  const float* B;
  int64_t ldb;
  int M, N, K;
  int kchunk, splits;
  int ones;                 // B column replaced by ones (-1: none)
  int tiles_m, tiles_n;
  float* C;
  int64_t ldc;
  float* C2;
  int64_t ldc2;
  const float* E1;
  int64_t lde1;
  const float* E2;
  int64_t lde2;
  const float* bias[3];
  float* S1;                // stage sums of C and C2 (SUMS): S += value, or
  int64_t lds1;             // S = value when sacc == 0
  float* S2;
  int64_t lds2;
  int sacc;
  int rg, nout, nin;
  float* part;              // slab z at part + z*kSrcParams
  int64_t woff[3], boff[3];
};

// Up to three products of one kind in one launch (the small
// parameter-gradient products share a launch to fill the GPU).
struct MmBatch {
  MmOp op[3];
  int nops;
  int start[4];             // first block of each product
};

constexpr int kPitch = 20;  // LDS floats per row of a 16-deep K slice

// 256 threads = 4 waves stacked along M; wave w owns rows 32w..32w+31 of the
// 128-row tile (2 MFMA row tiles) and all TN*16 columns.  Lane l supplies
// row (or column) l%16 and k = 4*(l/16) + s to MFMA s of a stage, for both
// operands.  A row-major operand ([M][K] or [N][K]) sits in LDS as
// [row][k] (pitch 20: the 128-bit reads of a pass land on distinct banks),
// so one 128-bit read gives a lane its operands of four MFMAs.  A
// column-major one ([K][M] or [K][N]) sits as it arrives, [k][row] with
// pitch rows + 4 (the four k-rows of a 32-bit read pass start 16 banks
// apart), read 32 bits per MFMA -- except the B operand of 208-column tiles,
// whose 52 such reads per stage cost more than transposing it into [row][k]
// on the way in (a thread's four columns at one k, k fastest across lanes,
// so a wave's 32-bit stores hit distinct banks).  Measured at 64^3: k-major
// 316 vs 358 us for the 144-column weight-gradient product, 353 vs 277 us
// for the 208-column one.  Double-buffered: the next stage's global loads
// are in flight in registers while the current one is consumed; one barrier
// per stage.
template <int TN, bool ACOL, bool BCOL, int EPI, bool SUMS = false>
__global__ void __launch_bounds__(256, 2) mm_kernel(MmBatch bt) {
  constexpr int BM = kTileM, BN = TN * 16;
  constexpr int kAF4 = BM * 16 / 4 / 256;              // float4 of A per thread (2)
  constexpr int kBF4 = (BN * 16 / 4 + 255) / 256;       // float4 of B per thread
  constexpr bool KMB = BCOL && TN <= 9;                 // B k-major in LDS
  constexpr int kPA = BM + 4, kPB = BN + 4;             // k-major pitches
  constexpr int kAFl = ACOL ? 16 * kPA : BM * kPitch;   // LDS floats of an A slice
  constexpr int kBFl = KMB ? 16 * kPB : BN * kPitch;
  constexpr int kBuf = kAFl + kBFl;
  __shared__ __align__(16) float lds[2 * kBuf];

  int oi = 0;
  while (oi + 1 < bt.nops && (int)blockIdx.x >= bt.start[oi + 1]) ++oi;
  const MmOp& op = bt.op[oi];
  const int nb = bt.start[oi + 1] - bt.start[oi];
  int b = (int)blockIdx.x - bt.start[oi];
  // XCD-aware order: blocks are dealt to the 8 XCDs round robin; give each
  // XCD a contiguous run of tiles, so the tiles of one row block (and of one
  // split) share that XCD's L2
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const int per = op.tiles_m * op.tiles_n;
  const int z = b / per;
  const int tt = b - z * per;
  const int tm = tt / op.tiles_n, tn = tt - tm * op.tiles_n;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kb = z * op.kchunk;
  const int ke = min(op.K, kb + op.kchunk);
  const int nst = (ke - kb + 15) >> 4;

  const int t = threadIdx.x, l = t & 63, w = t >> 6;
  const int lr = l & 15, lg = l >> 4;
  // the operands' fields, read once (op is indexed by a runtime product number)
  const float* __restrict__ Ag = op.A;
  const float* __restrict__ Bg = op.B;
  const int64_t lda = op.lda, ldb = op.ldb;
  const int Mv = op.M, Nv = op.N, ones = op.ones;

  f32x4 ra[kAF4], rb[kBF4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int e = 0; e < kAF4; ++e) {
      const int idx = t + 256 * e;
      f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
      if (ACOL) {
        const int k = idx / (BM / 4), mq = idx - k * (BM / 4);
        const int gk = k0 + k, gm = m0 + 4 * mq;
        if (gk < ke && gm < Mv)
          v = *reinterpret_cast<const f32x4*>(Ag + (int64_t)gk * lda + gm);
      } else {
        const int mr = idx >> 2, q = idx & 3;
        v = *reinterpret_cast<const f32x4*>(Ag + (int64_t)(m0 + mr) * lda + k0 + 4 * q);
      }
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < kBF4; ++e) {
      const int idx = t + 256 * e;
      f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
      if (idx < BN * 4) {
        if (BCOL) {
          const int k = KMB ? idx / (BN / 4) : (idx & 15);
          const int nq = KMB ? idx - k * (BN / 4) : (idx >> 4);
          const int gk = k0 + k, gn = n0 + 4 * nq;
          if (gk < ke && gn < Nv) {
            v = *reinterpret_cast<const f32x4*>(Bg + (int64_t)gk * ldb + gn);
            if (ones >= gn && ones < gn + 4) v[ones - gn] = 1.f;
          }
        } else {
          const int nr = idx >> 2, q = idx & 3;
          if (n0 + nr < Nv)
            v = *reinterpret_cast<const f32x4*>(Bg + (int64_t)(n0 + nr) * ldb + k0 + 4 * q);
        }
      }
      rb[e] = v;
    }
  };
  auto lstore = [&](float* As) {
    float* Bs = As + kAFl;
#pragma unroll
    for (int e = 0; e < kAF4; ++e) {
      const int idx = t + 256 * e;
      if (ACOL) {
        const int k = idx / (BM / 4), mq = idx - k * (BM / 4);
        *reinterpret_cast<f32x4*>(As + k * kPA + 4 * mq) = ra[e];
      } else {
        const int mr = idx >> 2, q = idx & 3;
        *reinterpret_cast<f32x4*>(As + mr * kPitch + 4 * q) = ra[e];
      }
    }
#pragma unroll
    for (int e = 0; e < kBF4; ++e) {
      const int idx = t + 256 * e;
      if (idx < BN * 4) {
        if (KMB) {
          const int k = idx / (BN / 4), nq = idx - k * (BN / 4);
          *reinterpret_cast<f32x4*>(Bs + k * kPB + 4 * nq) = rb[e];
        } else if (BCOL) {
          const int k = idx & 15, nq = idx >> 4;
#pragma unroll
          for (int c = 0; c < 4; ++c) Bs[(4 * nq + c) * kPitch + k] = rb[e][c];
        } else {
          const int nr = idx >> 2, q = idx & 3;
          *reinterpret_cast<f32x4*>(Bs + nr * kPitch + 4 * q) = rb[e];
        }
      }
    }
  };

  f32x4 acc[2][TN];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  gload(kb);
  lstore(lds);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const bool more = st + 1 < nst;
    if (more) gload(kb + 16 * (st + 1));
    const float* As = lds + (st & 1) * kBuf;
    const float* Bs = As + kAFl;
    f32x4 a0, a1;
    if (ACOL) {
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        a0[s4] = As[(4 * lg + s4) * kPA + 32 * w + lr];
        a1[s4] = As[(4 * lg + s4) * kPA + 32 * w + 16 + lr];
      }
    } else {
      a0 = *reinterpret_cast<const f32x4*>(As + (32 * w + lr) * kPitch + 4 * lg);
      a1 = *reinterpret_cast<const f32x4*>(As + (32 * w + 16 + lr) * kPitch + 4 * lg);
    }
    f32x4 bv[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      if (KMB) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) bv[j][s4] = Bs[(4 * lg + s4) * kPB + 16 * j + lr];
      } else {
        bv[j] = *reinterpret_cast<const f32x4*>(Bs + (16 * j + lr) * kPitch + 4 * lg);
      }
    }
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc[0][j] = mfma4(a0[s4], bv[j][s4], acc[0][j]);
        acc[1][j] = mfma4(a1[s4], bv[j][s4], acc[1][j]);
      }
    if (more) lstore(lds + ((st + 1) & 1) * kBuf);
    __syncthreads();
  }

  // epilogue: lane l holds rows 4*(l/16)+r of each 16-row tile, column l%16.
  // The operands it reads (bias, E1, E2, S1, S2) are loaded unconditionally,
  // one column tile at a time and one tile ahead: the loads of tile j + 1 are
  // in flight while tile j's values are formed and stored (rows <
  // tiles_m*128 and columns < 208 stay inside the chunk arrays for every
  // product that has them; only the stores are masked).
  constexpr bool HE1 = EPI == kEpiTanhGate || EPI == kEpiRevPair || EPI == kEpiRevTanh;
  constexpr bool HE2 = EPI == kEpiRevPair;
  constexpr bool HS1 = SUMS, HS2 = SUMS && EPI == kEpiRevPair;
  const int rbase = m0 + 32 * w + 4 * lg;
  // kernel-argument fields are read once here (an indexed read of the batch
  // inside the loop would be a memory round trip each)
  const float* bias0 = op.bias[0];
  const float* bias1 = op.bias[1];
  const float* bias2 = op.bias[2];
  const int64_t wo0 = op.woff[0], wo1 = op.woff[1], wo2 = op.woff[2];
  const int64_t bo0 = op.boff[0], bo1 = op.boff[1], bo2 = op.boff[2];
  float* slab = op.part + (int64_t)z * kSrcParams;
  const float* __restrict__ E1p = op.E1;
  const float* __restrict__ E2p = op.E2;
  float* __restrict__ Cp = op.C;
  float* __restrict__ C2p = op.C2;
  float* __restrict__ S1p = op.S1;
  float* __restrict__ S2p = op.S2;
  const int64_t lde1 = op.lde1, lde2 = op.lde2, ldc = op.ldc, ldc2 = op.ldc2;
  const int64_t lds1 = op.lds1, lds2 = op.lds2;
  const int Nn = op.N, sacc = op.sacc;
  if constexpr (HE1 || HS1) {
    // prefetch unit u: UI 16-row tiles of column tile u*UI/2 (one where four
    // operands are read, so that two units' operands fit in registers).
    // Unit u's operands sit in buffer u & 1; the loop takes units in pairs,
    // so every buffer is named statically and no copy waits on a load.
    constexpr int UI = HS2 ? 1 : 2, NU = TN * 2 / UI;
    struct EOps {
      float e1[UI][4], e2[UI][4], s1[UI][4], s2[UI][4], bb;
    };
    EOps ob0, ob1;
    auto eload = [&](int u, EOps& o) {
      const int j = u * UI / 2, i0 = u * UI - 2 * j;
      const int col = n0 + 16 * j + lr;
      o.bb = 0.f;
      if (EPI == kEpiTanhGate) {
        const float bv = bias0[col < kH ? col : kH - 1];
        o.bb = col < kH ? bv : 0.f;
      }
#pragma unroll
      for (int ii = 0; ii < UI; ++ii)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = rbase + 16 * (i0 + ii) + r;
          if (HE1) o.e1[ii][r] = E1p[row * lde1 + col];
          if (HE2) o.e2[ii][r] = E2p[row * lde2 + col];
          if (HS1) o.s1[ii][r] = S1p[row * lds1 + col];
          if (HS2) o.s2[ii][r] = S2p[row * lds2 + col];
        }
    };
    auto unit = [&](int u, const EOps& o) {
      const int j = u * UI / 2, i0 = u * UI - 2 * j;
      const int col = n0 + 16 * j + lr;
      const bool cok = col < Nn;
      // values first, without branches; then the masked stores
      float v1[UI][4], v2[UI][4];
#pragma unroll
      for (int ii = 0; ii < UI; ++ii)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float d = acc[i0 + ii][j][r];
          if (EPI == kEpiTanhGate) {
            v1[ii][r] = tanhf(d + o.bb);
            v2[ii][r] = v1[ii][r] * o.e1[ii][r];
          } else if (EPI == kEpiRevPair) {
            const float a1 = o.e1[ii][r], a2 = o.e2[ii][r];
            v1[ii][r] = (d * a1) * (1.f - a2 * a2);
            v2[ii][r] = (d * a2) * (1.f - a1 * a1);
          } else {  // kEpiRevTanh
            const float a1 = o.e1[ii][r];
            v1[ii][r] = d * (1.f - a1 * a1);
          }
        }
#pragma unroll
      for (int ii = 0; ii < UI; ++ii)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rbase + 16 * (i0 + ii) + r;
          if (row >= Mv || !cok) continue;
          Cp[(int64_t)row * ldc + col] = v1[ii][r];
          if (EPI == kEpiTanhGate || EPI == kEpiRevPair)
            C2p[(int64_t)row * ldc2 + col] = v2[ii][r];
          if (HS1) S1p[(int64_t)row * lds1 + col] = sacc ? o.s1[ii][r] + v1[ii][r] : v1[ii][r];
          if (HS2) S2p[(int64_t)row * lds2 + col] = sacc ? o.s2[ii][r] + v2[ii][r] : v2[ii][r];
        }
    };
    eload(0, ob0);
#pragma unroll
    for (int pu = 0; pu < (NU + 1) / 2; ++pu) {
      const int u = 2 * pu;
      if (u + 1 < NU) eload(u + 1, ob1);
      unit(u, ob0);
      if (u + 1 < NU) {
        if (u + 2 < NU) eload(u + 2, ob0);
        unit(u + 1, ob1);
      }
    }
  } else {
    // kEpiStore, kEpiTanhCat, kEpiPartial: no operands beyond the bias
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + 16 * j + lr;
      const bool cok = col < Nn;
      float bb = 0.f;
      if (EPI == kEpiTanhCat) {
        const int g = col / kHp, o = col - g * kHp;
        const float* bp = g == 0 ? bias0 : (g == 1 ? bias1 : bias2);
        const float bv = bp[o < kH ? o : kH - 1];
        bb = o < kH ? bv : 0.f;
      }
      float v1[2][4];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float d = acc[i][j][r];
          v1[i][r] = EPI == kEpiTanhCat ? tanhf(d + bb) : d;
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rbase + 16 * i + r;
          if (row >= Mv || !cok) continue;
          if (EPI == kEpiPartial) {
            const int g = row / op.rg, o = row - g * op.rg;
            if (o >= op.nout) continue;
            const int64_t wo = g == 0 ? wo0 : (g == 1 ? wo1 : wo2);
            const int64_t bo = g == 0 ? bo0 : (g == 1 ? bo1 : bo2);
            if (col < op.nin)
              slab[wo + (int64_t)o * op.nin + col] = v1[i][r];
            else if (col == op.nin)
              slab[bo + o] = v1[i][r];
          } else {
            Cp[(int64_t)row * ldc + col] = v1[i][r];
          }
        }
    }
  }
}

template <int TN, bool ACOL, bool BCOL, int EPI, bool SUMS = false>
int mm_launch(MmBatch bt, hipStream_t s) {
  int total = 0;
  for (int i = 0; i < bt.nops; ++i) {
    MmOp& o = bt.op[i];
    o.tiles_m = (o.M + kTileM - 1) / kTileM;
    o.tiles_n = (o.N + TN * 16 - 1) / (TN * 16);
    bt.start[i] = total;
    total += o.tiles_m * o.tiles_n * o.splits;
  }
  bt.start[bt.nops] = total;
  if (total == 0) return 0;
  hipLaunchKernelGGL((mm_kernel<TN, ACOL, BCOL, EPI, SUMS>), dim3(total), dim3(256), 0, s, bt);
  PF_LAUNCH_CHECK("mm_kernel");
  return 0;
}

MmOp mm_op(const float* A, int64_t lda, const float* B, int64_t ldb, int M, int N, int K,
           float* C, int64_t ldc) {
  MmOp o;
  std::memset(&o, 0, sizeof(o));
  o.A = A; o.lda = lda;
  o.B = B; o.ldb = ldb;
  o.M = M; o.N = N; o.K = K;
  o.kchunk = K;
  o.splits = 1;
  o.ones = -1;
  o.C = C; o.ldc = ldc;
  return o;
}

MmBatch one(const MmOp& o) {
  MmBatch b;
  std::memset(&b, 0, sizeof(b));
  b.op[0] = o;
  b.nops = 1;
  return b;
}

// The weights in the products' padded layouts (zeros in the pads).
__global__ void __launch_bounds__(kBlock) pack_kernel(const float* __restrict__ prm,
                                                      float* __restrict__ pk) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < kPkFloats;
       e += (int64_t)gridDim.x * blockDim.x) {
    float v = 0.f;
    if (e < kPkW2) {
      const int r = (int)(e / kXld), c = (int)(e % kXld);
      const int g = r / kHp, o = r % kHp;
      const int64_t woff[3] = {kOffFc1W, kOffGW, kOffG2W};
      if (o < kH && c < kSrcIn) v = prm[woff[g] + (int64_t)o * kSrcIn + c];
    } else if (e < kPkW4) {
      const int64_t q = e < kPkW3 ? e - kPkW2 : e - kPkW3;
      const int r = (int)(q / kHp), c = (int)(q % kHp);
      const int64_t off = e < kPkW3 ? kOffFc2W : kOffFc3W;
      if (r < kH && c < kH) v = prm[off + (int64_t)r * kH + c];
    } else {
      const int64_t q = e - kPkW4;
      const int r = (int)(q / kHp), c = (int)(q % kHp);
      if (r < kSrcOut && c < kH) v = prm[kOffFc4W + (int64_t)r * kH + c];
    }
    pk[e] = v;
  }
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
    float* x = X + (int64_t)r * kXld + sh * kSrcFeat;
#pragma unroll
    for (int f = 0; f < kSrcFeat; ++f) x[f] = Fn[(int64_t)f * s.m + q];
  }
}

// Output adjoints from the divergence (:466-589): G_q = sum over axes a of
// (P_{3q+a}[c+e_a] - P_{3q+a}[c-e_a])/(2 dx) with the padded wrap, so
// dP_{3q+a}[c] = dG_q[c-e_a]/(2 dx) - dG_q[c+e_a]/(2 dx); dG_q = float(the
// adjoint's centre interpolant), since g multiplies G (cast to the box
// dtype) by it.  Channels 9..17 get no adjoint.
template <typename T>
__global__ void __launch_bounds__(kBlock) dout_kernel(SrcGeom s, CVec3<T> gI,
                                                      float* __restrict__ dout, int64_t c0,
                                                      int rows, float* __restrict__ dsum,
                                                      int sacc) {
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
    float* o = dout + r * kOutp;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float gm = (float)gI.a[q][nb[a][0]], gp = (float)gI.a[q][nb[a][1]];
        o[3 * q + a] = gm / s.d2 - gp / s.d2;
      }
#pragma unroll
    for (int c = kSrcUsed; c < kSrcOut; ++c) o[c] = 0.f;
    if (dsum) {
      float* q = dsum + r * kOutp;
#pragma unroll
      for (int c = 0; c < kSrcUsed; ++c) q[c] = sacc ? q[c] + o[c] : o[c];
    }
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
      const float* x = dX + c * kXld + sh * kSrcFeat;
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

// Device arrays of the reverse, from the workspace layout.
struct Bufs {
  Srct L;
  SrcGeom s;
  Norm21 nrm;
  float *Fn, *dX, *dF, *dvar;
  float *X, *cat, *H2, *H2G, *H3, *H4;      // forward activations (Mr rows)
  float *dout, *dcat, *dL4, *dL2;           // a chunk's layer adjoints (Ca rows)
  float *sout, *scat, *sL4, *sL2;           // stage sums (Ma rows, full only)
  float *pk, *part;
  void *fwd[3], *adj[3];
};

Bufs bufs_of(pf_box* h) {
  Bufs B;
  B.L = srct_layout(h);
  B.s = src_geom(h);
  for (int f = 0; f < kSrcFeat; ++f) B.nrm.v[f] = h->srct_norm[f];
  char* base = (char*)h->srct_ws;
  const Srct& L = B.L;
  auto F = [&](size_t off) { return (float*)(base + off); };
  B.Fn = h->src_ws;
  B.dX = F(L.off_dX);
  B.dF = F(L.off_dF);
  B.dvar = F(L.off_dvar);
  B.X = F(L.off_X);
  B.cat = F(L.off_cat);
  const size_t mr = (size_t)L.Mr * kHp, ca = (size_t)L.Ca * kHp;
  B.H2 = F(L.off_h);
  B.H2G = B.H2 + mr;
  B.H3 = B.H2 + 2 * mr;
  B.H4 = B.H2 + 3 * mr;
  B.dout = F(L.off_dout);
  B.dcat = F(L.off_dcat);
  B.dL4 = F(L.off_dl);
  B.dL2 = B.dL4 + ca;
  const size_t ma = (size_t)L.Ma * kHp;
  B.sout = L.full ? F(L.off_sout) : nullptr;
  B.scat = L.full ? F(L.off_scat) : nullptr;
  B.sL4 = L.full ? F(L.off_sl) : nullptr;
  B.sL2 = L.full ? B.sL4 + ma : nullptr;
  B.pk = F(L.off_pack);
  B.part = F(L.off_part);
  const size_t pad = (size_t)L.m * h->esize;
  for (int q = 0; q < 3; ++q) {
    B.fwd[q] = base + L.off_fwd + q * pad;
    B.adj[q] = base + L.off_adj + q * pad;
  }
  return B;
}

// Features of the closure's inputs and the weights in the padded layouts.
template <typename T>
int prep(const Bufs& B, const float* prm, const void* const fwdI[3], hipStream_t st) {
  hipLaunchKernelGGL(src_features_kernel<T>, dim3(grid_of(B.s.m)), dim3(kBlock), 0, st, B.s,
                     (const T*)fwdI[0], (const T*)fwdI[1], (const T*)fwdI[2], B.Fn, B.nrm);
  PF_LAUNCH_CHECK("src_features_kernel");
  PF_ELEM(pack_kernel, kPkFloats, prm, B.pk);
  return 0;
}

// Row offset of chunk c0 in the forward-activation arrays.
inline int64_t frow(const Bufs& B, int64_t c0) { return B.L.full ? c0 : 0; }

// Forward of chunk c0 (R cells): inputs and activations of every layer.
int fwd_chunk(const Bufs& B, const float* prm, int64_t c0, int R, hipStream_t st) {
  const int64_t f = frow(B, c0);
  float* X = B.X + f * kXld;
  float* cat = B.cat + f * kCat;
  float* H2 = B.H2 + f * kHp;
  float* H2G = B.H2G + f * kHp;
  float* H3 = B.H3 + f * kHp;
  float* H4 = B.H4 + f * kHp;
  PF_ELEM(gather_x_kernel, (int64_t)R * kSrcShifts, B.s, (const float*)B.Fn, X, c0, R);
  // [H1 G Gate2] = tanh(X [fc1 fcG fcG2]^T + b) (:91-102)
  {
    MmOp o = mm_op(X, kXld, B.pk + kPkCat, kXld, R, kCat, kXld, cat, kCat);
    o.bias[0] = prm + kOffFc1B;
    o.bias[1] = prm + kOffGB;
    o.bias[2] = prm + kOffG2B;
    if (int rc = mm_launch<13, false, false, kEpiTanhCat>(one(o), st)) return rc;
  }
  // H2 = tanh(fc2 H1), H2G = G*H2 (:94-104)
  {
    MmOp o = mm_op(cat, kCat, B.pk + kPkW2, kHp, R, kHp, kHp, H2, kHp);
    o.bias[0] = prm + kOffFc2B;
    o.C2 = H2G; o.ldc2 = kHp;
    o.E1 = cat + kHp; o.lde1 = kCat;
    if (int rc = mm_launch<13, false, false, kEpiTanhGate>(one(o), st)) return rc;
  }
  // H3 = tanh(fc3 H2G), H4 = H3*Gate2 (:108-112)
  {
    MmOp o = mm_op(H2G, kHp, B.pk + kPkW3, kHp, R, kHp, kHp, H3, kHp);
    o.bias[0] = prm + kOffFc3B;
    o.C2 = H4; o.ldc2 = kHp;
    o.E1 = cat + 2 * kHp; o.lde1 = kCat;
    if (int rc = mm_launch<13, false, false, kEpiTanhGate>(one(o), st)) return rc;
  }
  return 0;
}

// Reverse of chunk c0 from the output adjoints of gI into the chunk's layer
// adjoints and the input adjoints dX; with sums, the layer adjoints are also
// added to (sacc) or stored in the stage sums.
template <typename T, bool SUMS>
int rev_chunk(const Bufs& B, const void* const gI[3], int64_t c0, int R, int sacc,
              hipStream_t st) {
  const int64_t f = frow(B, c0);
  const float* cat = B.cat + f * kCat;
  const float* H2 = B.H2 + f * kHp;
  const float* H3 = B.H3 + f * kHp;
  CVec3<T> g3;
  for (int q = 0; q < 3; ++q) g3.a[q] = (const T*)gI[q];
  float* sout = SUMS ? B.sout + c0 * kOutp : nullptr;
  float* scat = SUMS ? B.scat + c0 * kCat : nullptr;
  float* sL4 = SUMS ? B.sL4 + c0 * kHp : nullptr;
  float* sL2 = SUMS ? B.sL2 + c0 * kHp : nullptr;
  PF_ELEM(dout_kernel<T>, R, B.s, g3, B.dout, c0, R, sout, sacc);
  // dH4 = dout fc4 -> dL4, dLG2
  {
    MmOp o = mm_op(B.dout, kOutp, B.pk + kPkW4, kHp, R, kHp, kOutp, B.dL4, kHp);
    o.C2 = B.dcat + 2 * kHp; o.ldc2 = kCat;
    o.E1 = cat + 2 * kHp; o.lde1 = kCat;
    o.E2 = H3; o.lde2 = kHp;
    o.S1 = sL4; o.lds1 = kHp;
    o.S2 = scat + (SUMS ? 2 * kHp : 0); o.lds2 = kCat;
    o.sacc = sacc;
    if (int rc = mm_launch<13, false, true, kEpiRevPair, SUMS>(one(o), st)) return rc;
  }
  // dH2G = dL4 fc3 -> dLG, dL2
  {
    MmOp o = mm_op(B.dL4, kHp, B.pk + kPkW3, kHp, R, kHp, kHp, B.dcat + kHp, kCat);
    o.C2 = B.dL2; o.ldc2 = kHp;
    o.E1 = H2; o.lde1 = kHp;
    o.E2 = cat + kHp; o.lde2 = kCat;
    o.S1 = scat + (SUMS ? kHp : 0); o.lds1 = kCat;
    o.S2 = sL2; o.lds2 = kHp;
    o.sacc = sacc;
    if (int rc = mm_launch<13, false, true, kEpiRevPair, SUMS>(one(o), st)) return rc;
  }
  // dH1 = dL2 fc2 -> dL1
  {
    MmOp o = mm_op(B.dL2, kHp, B.pk + kPkW2, kHp, R, kHp, kHp, B.dcat, kCat);
    o.E1 = cat; o.lde1 = kCat;
    o.S1 = scat; o.lds1 = kCat;
    o.sacc = sacc;
    if (int rc = mm_launch<13, false, true, kEpiRevTanh, SUMS>(one(o), st)) return rc;
  }
  // input adjoints of the chunk: dX = [dL1 dLG dLG2] [fc1; fcG; fcG2]
  {
    MmOp o = mm_op(B.dcat, kCat, B.pk + kPkCat, kXld, R, kSrcIn, kCat, B.dX + c0 * kXld, kXld);
    if (int rc = mm_launch<9, false, true, kEpiStore>(one(o), st)) return rc;
  }
  return 0;
}

// Parameter gradients of chunk c0 added to dparams: [dW | db] = dL^T [act | 1]
// over its cells, split over kSplitK-cell ranges into FP32 partials laid out
// like the packed parameters, then summed in order into FP64.  The layer
// adjoints are the chunk's (dcat, dL2, dL4, dout; rows from 0) or the stage
// sums (rows from c0).
int wgrad_chunk(const Bufs& B, const float* dcat, const float* dL2, const float* dL4,
                const float* dout, int64_t c0, int R, double* dparams, hipStream_t st) {
  const int64_t f = frow(B, c0);
  const float* X = B.X + f * kXld;
  const float* H1 = B.cat + f * kCat;
  const float* H2G = B.H2G + f * kHp;
  const float* H4 = B.H4 + f * kHp;
  const int splits = (R + kSplitK - 1) / kSplitK;
  {
    MmOp o = mm_op(dcat, kCat, X, kXld, kCat, kSrcIn + 1, R, nullptr, 0);
    o.kchunk = kSplitK; o.splits = splits; o.ones = kSrcIn;
    o.part = B.part; o.rg = kHp; o.nout = kH; o.nin = kSrcIn;
    o.woff[0] = kOffFc1W; o.woff[1] = kOffGW; o.woff[2] = kOffG2W;
    o.boff[0] = kOffFc1B; o.boff[1] = kOffGB; o.boff[2] = kOffG2B;
    if (int rc = mm_launch<9, true, true, kEpiPartial>(one(o), st)) return rc;
  }
  {
    MmBatch bt;
    std::memset(&bt, 0, sizeof(bt));
    struct Pg {
      const float* dL;
      int64_t ldl;
      int M, nout;
      const float* act;
      int64_t lda;
      int64_t w, b;
    };
    const Pg pg[3] = {{dL2, kHp, kHp, kH, H1, kCat, kOffFc2W, kOffFc2B},
                      {dL4, kHp, kHp, kH, H2G, kHp, kOffFc3W, kOffFc3B},
                      {dout, kOutp, kOutp, kSrcOut, H4, kHp, kOffFc4W, kOffFc4B}};
    for (int i = 0; i < 3; ++i) {
      MmOp o = mm_op(pg[i].dL, pg[i].ldl, pg[i].act, pg[i].lda, pg[i].M, kH + 1, R, nullptr, 0);
      o.kchunk = kSplitK; o.splits = splits; o.ones = kH;
      o.part = B.part; o.rg = pg[i].M; o.nout = pg[i].nout; o.nin = kH;
      o.woff[0] = pg[i].w; o.boff[0] = pg[i].b;
      bt.op[i] = o;
    }
    bt.nops = 3;
    if (int rc = mm_launch<13, true, true, kEpiPartial>(bt, st)) return rc;
  }
  PF_ELEM(add_partials_kernel, kSrcParams, dparams, (const float*)B.part, (int64_t)kSrcParams,
          splits, (int64_t)kSrcParams);
  return 0;
}

// Input adjoints of every cell -> d g / d var_i (dvar).
int finish_inputs(const Bufs& B, hipStream_t st) {
  PF_ELEM(feat_adj_kernel, B.s.m, B.s, (const float*)B.dX, B.dF, B.nrm);
  PF_ELEM(var_adj_kernel, B.s.m, B.s, (const float*)B.dF, B.dvar);
  return 0;
}

// d g / d var_i into dvar (3 x m FP32) for g = sum(G(fwdI) . gI), and the
// parameter gradient added to dparams (FP64, may be NULL).
template <typename T>
int ml_vjp_impl(pf_box* h, const float* prm, const void* const fwdI[3], const void* const gI[3],
                double* dparams, hipStream_t st) {
  const Bufs B = bufs_of(h);
  if (int rc = prep<T>(B, prm, fwdI, st)) return rc;
  for (int64_t c0 = 0; c0 < B.s.m; c0 += B.L.C) {
    const int R = (int)((B.s.m - c0) < B.L.C ? (B.s.m - c0) : B.L.C);
    if (int rc = fwd_chunk(B, prm, c0, R, st)) return rc;
    if (int rc = rev_chunk<T, false>(B, gI, c0, R, 0, st)) return rc;
    if (dparams)
      if (int rc = wgrad_chunk(B, B.dcat, B.dL2, B.dL4, B.dout, c0, R, dparams, st)) return rc;
  }
  return finish_inputs(B, st);
}

// The forward state's var_i as the reference's adjoint RHS leaves them
// before its closure block (adjoint.py:170-237: the same interpolations as
// the forward RHS, so the same stale last planes).
template <typename T>
int fwd_inputs(pf_box* h, const Bufs& B, const void* const U[3], hipStream_t st) {
  return pf_source_inputs(h, U[0], U[1], U[2], B.fwd[0], B.fwd[1], B.fwd[2], st);
}

// The adjoint state's centre interpolants.
template <typename T>
int adj_inputs(const Bufs& B, const void* const A[3], hipStream_t st) {
  CVec3<T> a3;
  Vec3<T> o3;
  for (int q = 0; q < 3; ++q) {
    a3.a[q] = (const T*)A[q];
    o3.a[q] = (T*)B.adj[q];
  }
  hipLaunchKernelGGL(centre_interp_kernel<T>, dim3(grid_of(B.s.m)), dim3(kBlock), 0, st, B.s,
                     a3, o3);
  PF_LAUNCH_CHECK("centre_interp_kernel");
  return 0;
}

template <typename T>
int delta_out(pf_box* h, const Bufs& B, void* const D[3], hipStream_t st) {
  Vec3<T> d3;
  for (int q = 0; q < 3; ++q) d3.a[q] = (T*)D[q];
  hipLaunchKernelGGL(delta_kernel<T>, dim3(launch_blocks(h->g)), dim3(kBlock), 0, st, h->g,
                     (const float*)B.dvar, B.L.m, d3);
  PF_LAUNCH_CHECK("delta_kernel");
  return 0;
}

template <typename T>
int ml_adj_term_impl(pf_box* h, const void* const U[3], const float* prm, const void* const A[3],
                     void* const D[3], double* dparams, hipStream_t st) {
  const Bufs B = bufs_of(h);
  if (int rc = fwd_inputs<T>(h, B, U, st)) return rc;
  if (int rc = adj_inputs<T>(B, A, st)) return rc;
  const void* const fc[3] = {B.fwd[0], B.fwd[1], B.fwd[2]};
  const void* const ac[3] = {B.adj[0], B.adj[1], B.adj[2]};
  if (int rc = ml_vjp_impl<T>(h, prm, fc, ac, dparams, st)) return rc;
  return delta_out<T>(h, B, D, st);
}

// One adjoint RK4 step with the closure (full cache): the forward once ...
template <typename T>
int ml_step_begin_impl(pf_box* h, const void* const U[3], const float* prm, hipStream_t st) {
  const Bufs B = bufs_of(h);
  if (int rc = fwd_inputs<T>(h, B, U, st)) return rc;
  const void* const fc[3] = {B.fwd[0], B.fwd[1], B.fwd[2]};
  if (int rc = prep<T>(B, prm, fc, st)) return rc;
  for (int64_t c0 = 0; c0 < B.s.m; c0 += B.L.C) {
    const int R = (int)((B.s.m - c0) < B.L.C ? (B.s.m - c0) : B.L.C);
    if (int rc = fwd_chunk(B, prm, c0, R, st)) return rc;
  }
  return 0;
}

// ... the reverse of each stage, its layer adjoints summed over the stages ...
template <typename T>
int ml_step_term_impl(pf_box* h, const void* const A[3], void* const D[3], int stage,
                      hipStream_t st) {
  const Bufs B = bufs_of(h);
  if (int rc = adj_inputs<T>(B, A, st)) return rc;
  const void* const ac[3] = {B.adj[0], B.adj[1], B.adj[2]};
  for (int64_t c0 = 0; c0 < B.s.m; c0 += B.L.C) {
    const int R = (int)((B.s.m - c0) < B.L.C ? (B.s.m - c0) : B.L.C);
    if (int rc = rev_chunk<T, true>(B, ac, c0, R, stage > 0, st)) return rc;
  }
  if (int rc = finish_inputs(B, st)) return rc;
  return delta_out<T>(h, B, D, st);
}

// ... and the step's parameter gradient from the sums.
int ml_step_end_impl(pf_box* h, double* dparams, hipStream_t st) {
  const Bufs B = bufs_of(h);
  for (int64_t c0 = 0; c0 < B.s.m; c0 += B.L.C) {
    const int R = (int)((B.s.m - c0) < B.L.C ? (B.s.m - c0) : B.L.C);
    if (int rc = wgrad_chunk(B, B.scat + c0 * kCat, B.sL2 + c0 * kHp, B.sL4 + c0 * kHp,
                             B.sout + c0 * kOutp, c0, R, dparams, st))
      return rc;
  }
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

bool ml_source_step_cached(const pf_box* h) {
  return h->srct_ws && h->srct_bytes >= srct_layout(h).bytes && srct_layout(h).full;
}

int ml_source_step_begin(pf_box* h, const void* const U[3], const float* params, hipStream_t s) {
  if (int rc = check_reserved(h, "adjoint closure term of the 'source' ML closure")) return rc;
  if (!params) return fail(PF_EINVAL, "adjoint closure term: NULL params");
  return h->dtype == PF_F64 ? ml_step_begin_impl<double>(h, U, params, s)
                            : ml_step_begin_impl<float>(h, U, params, s);
}

int ml_source_step_term(pf_box* h, const void* const A[3], void* const D[3], int stage,
                        hipStream_t s) {
  return h->dtype == PF_F64 ? ml_step_term_impl<double>(h, A, D, stage, s)
                            : ml_step_term_impl<float>(h, A, D, stage, s);
}

int ml_source_step_end(pf_box* h, double* dparams, hipStream_t s) {
  return dparams ? ml_step_end_impl(h, dparams, s) : 0;
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

int pf_sfs_ml_source_cache_budget(pf_box* h, int64_t bytes) {
  if (int rc = check_box(h)) return rc;
  if (h->srct_ws)
    return fail(PF_EINVAL, "pf_sfs_ml_source_cache_budget: call it before "
                           "pf_sfs_ml_source_train_reserve");
  h->srct_budget = bytes < 0 ? -1 : bytes;
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
  const int rc = h->dtype == PF_F64 ? ml_vjp_impl<double>(h, params, F, G, dparams, s)
                                    : ml_vjp_impl<float>(h, params, F, G, dparams, s);
  if (rc) return rc;
  float* out[3] = {dui, dvi, dwi};
  for (int q = 0; q < 3; ++q)
    if (out[q])
      PF_HIP(hipMemcpyAsync(out[q], dvar + (size_t)q * L.m, (size_t)L.m * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
  return 0;
}

}  // extern "C"
