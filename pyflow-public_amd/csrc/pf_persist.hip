// BiCGStab (solver/pressure.py:196-297) as one cooperative kernel whose
// Krylov vectors never leave the chip.
//
// The launch-per-reduction solver (KA, KB, KC in pf_krylov.hip) streams its
// vectors through HBM / the Infinity Cache three times per iteration.  On a
// box whose vectors fit in the register files and LDS of the GPU (128^3 FP64:
// p, v, r, t in VGPRs and x, r_hat in LDS, 100 MB of the ~168 MB on chip),
// each block keeps its cells for the whole solve and the blocks meet three
// times per iteration, once per global reduction.  Only tile faces and
// per-block partial sums cross between blocks:
//   * a block of this kernel covers V consecutive x chunks of one tile, i.e.
//     V blocks ("virtual blocks") of the KA/KB/KC grid: in FP64 16 planes
//     per thread (two z cells each) at one wave per SIMD, so that its state
//     fits the 512 VGPRs a lone wave may hold; in FP32 8 planes per thread at
//     two waves per SIMD;
//   * every block publishes the faces of v (after v = A p) and of r (after
//     r = s - omega t) that its neighbours read; a block computes p and s on
//     its halo from those faces and the p halo it formed itself one
//     iteration before, with the same operations as their owners, so the
//     same bits;
//   * a grid sync is an all-to-all of partial sums: every block writes its
//     virtual blocks' partials into their slots of the sync's slot set
//     (initially an empty marker, a signalling-NaN bit pattern that no
//     arithmetic result can take) and polls every slot until none is empty,
//     so the partials arrive with the sync itself; each block then sums them
//     in the fixed order of pf_krylov.hip's reductions, so every block derives
//     the same scalars;
//   * faces and partials are written with device-scope stores (write-through
//     past the XCD's L2) and read with device-scope loads, and a block waits
//     for its stores to complete before it writes its partials, so a block
//     that sees a partial sees the faces written before it.  No L2 write-back
//     or invalidation is needed (agent-scope release / acquire fences made a
//     sync round 4-7x slower on this GPU).
// Virtual block b keeps the tile, thread-to-cell map, plane order and
// per-thread accumulation order of block b of KA/KB/KC, and all vector
// arithmetic is theirs: the iterates, the residual history and the solution
// are bit-identical to the launch-per-reduction path.
//
// Every poll is bounded: a block that waits too long raises an abort flag
// that all other polls check, and every block leaves; the host reports the
// failure.  The launch is cooperative, so all blocks are co-resident.

#include <hip/hip_runtime.h>

#include "pf_kernels.h"

namespace pf {

namespace {

constexpr unsigned long long kSlotEmpty = 0x7FF4DEADBEEF0123ull;
constexpr int kSets = 4;              // slot sets in rotation (sync j uses set j % 4)
// Replicas of each slot set: a block writes its partials into every replica
// and polls replica (b mod 8) mod kReps, so that each replica's cache lines
// carry a quarter of the polls (a bare all-to-all round at 256 blocks: 3.45
// us with one replica, 2.43 with four; 512 blocks x 2 slots: 4.84 -> 4.29).
// Eight, one per XCD, were no faster in the solver (128^3 FP64 24.85 us per
// iteration with four and with eight; FP32 20.1-20.4 against 20.5-20.8).
constexpr int kReps = 4;
constexpr int kSpin = 1 << 21;        // polls before a block gives up
constexpr int kPT = 256;              // threads per block
constexpr int kPZ = 64, kPY = 8;      // tile extents (z, y): two z cells per thread

// Device-scope accesses through global-address-space pointers: a generic
// (flat) access would also count against the LDS wait counter, so every LDS
// wait of the stencils would wait for the face stores in flight as well.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ T ld_dev(const T* p) {
  return __hip_atomic_load((gptr<T>)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_dev(T* p, T v) {
  __hip_atomic_store((gptr<T>)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Fused multiply-add in the operands' own precision (one rounding).
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }
// Keep a value's computation where it stands: the compiler may not move it
// past this point (without it, the stencils' arithmetic is deferred past
// later barriers and every plane's LDS operands stay live at once).
template <typename T>
__device__ __forceinline__ void pin(T& v) {
  asm volatile("" : "+v"(v));
}

// Wait until every memory operation this wave issued has completed.
__device__ __forceinline__ void wait_mem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

template <typename T>
struct alignas(2 * sizeof(T)) P2 {
  T a, b;
};

template <typename T>
__device__ __forceinline__ P2<T> ld2(const T* p) {
  return *reinterpret_cast<const P2<T>*>(p);
}
template <typename T>
__device__ __forceinline__ P2<T> ld2_dev(const T* p) {
  if constexpr (sizeof(T) == 4) {
    const uint64_t w = ld_dev(reinterpret_cast<const uint64_t*>(p));
    P2<T> r;
    __builtin_memcpy(&r, &w, 8);
    return r;
  } else {
    return P2<T>{ld_dev(p), ld_dev(p + 1)};
  }
}
template <typename T>
__device__ __forceinline__ void st2_dev(T* p, const P2<T>& v) {
  if constexpr (sizeof(T) == 4) {
    uint64_t w;
    __builtin_memcpy(&w, &v, 8);
    st_dev(reinterpret_cast<uint64_t*>(p), w);
  } else {
    st_dev(p, v.a);
    st_dev(p + 1, v.b);
  }
}

// Face layout of one block's region (XC planes, 8 rows, 64 columns):
//   [first plane: 8 x 64][last plane: 8 x 64][row y0: XC x 64][row y0+7: XC x 64]
//   [column z0: XC x 8][column z0+63: XC x 8]
template <int XC>
struct FaceMap {
  static constexpr int xlo = 0, xhi = kPY * kPZ, ylo = 2 * kPY * kPZ, yhi = ylo + XC * kPZ,
                       zlo = yhi + XC * kPZ, zhi = zlo + XC * kPY, size = zhi + XC * kPY;
};

// Virtual block id of a tile (inverse of tile_of_block).
__device__ __forceinline__ int block_of_tile(const TileGrid& tg, int nb, int tz, int ty, int cx) {
  tz = (tz + tg.ntz) % tg.ntz;
  ty = (ty + tg.nty) % tg.nty;
  cx = (cx + tg.ncx) % tg.ncx;
  const int bp = (cx * tg.nty + ty) * tg.ntz + tz;
  if ((nb & 7) == 0) {
    const int per = nb >> 3;
    return (bp % per) * 8 + bp / per;
  }
  return bp;
}

}  // namespace

template <typename T>
struct PersistArgs {
  Geom g;
  TileGrid tg;               // the KA/KB/KC grid: its blocks are the virtual blocks
  KrylovCtl* ctl;
  const double* part_init;   // init kernel's partials: <r, r_hat>, max|r| per block
  int npart_init;
  const T* r0;               // r = r_hat = b (interior layout)
  const T* rh;
  T* x;                      // solution out
  T* faces;                  // [2][blocks][face elems]: v, r
  double* slots;             // [kSets][kReps][virtual blocks][2], kSlotEmpty where not yet written
  unsigned* flag;            // [0] abort
  Lap<T> L;
};

namespace {

// One grid sync: publish the NV partials of each of this block's V virtual
// blocks vid[v] (mine[v*NV + q]) for sync j, wait for every virtual block's,
// and reduce them as reduce_partials<NV, 256> reduces a kernel's partials
// (thread t takes virtual blocks t, t+256, ... in order, then the block
// tree).  As soon as the six neighbours' slots nbv (one virtual block of
// each) are filled, their faces are complete: early() then issues this
// thread's face loads, which travel while the wait for the other blocks
// goes on.  Returns false (block-uniform) when the wait was abandoned.
template <int NV, int V, bool SKIP, typename Early>
__device__ __forceinline__ bool grid_sync(double* __restrict__ slots, unsigned* flag, int nblk,
                                          const int (&vid)[V], int j,
                                          const double (&mine)[V * NV], double (&red)[NV],
                                          const int (&op)[NV], const int (&nbv)[6],
                                          Early&& early) {
  const size_t stride = (size_t)nblk * 2;   // one replica of a slot set
  double* set = slots + ((size_t)(j % kSets) * kReps + (blockIdx.x & 7) % kReps) * stride;
  wait_mem();        // this thread's face stores (and loads) are complete
  __syncthreads();   // ... and every other thread's of the block
  if (threadIdx.x < kReps) {   // thread r writes replica r
    double* ws = slots + ((size_t)(j % kSets) * kReps + threadIdx.x) * stride;
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
      for (int q = 0; q < NV; ++q) st_dev(ws + 2 * vid[v] + q, mine[v * NV + q]);
  }
  constexpr int MB = 4;   // up to 1024 virtual blocks
  double got[MB][NV];
  bool ok = false;
  // first the neighbours (their faces), then everyone (the partials)
  bool nb = false;
  for (int it = 0; it < kSpin && !nb; ++it) {
    nb = true;
#pragma unroll
    for (int q = 0; q < 6; ++q)
      nb = nb && ld_dev(reinterpret_cast<const unsigned long long*>(set + 2 * nbv[q])) !=
                     kSlotEmpty;
    if ((it & 63) == 63 && ld_dev(flag) != 0u) break;
  }
  early();   // (after a timeout too: in-bounds loads whose values go unused)
  if constexpr (SKIP) {
    // a slot once seen filled is not read again.  The polls of all blocks
    // fall on the few cache lines of the slot set; leaving out the loads of
    // filled slots pays where there are many (1024 virtual blocks: 128^3
    // FP32, 22.4 -> 22.1 us per iteration) and costs a little where there
    // are fewer (512: 28.5 -> 28.7 us)
    unsigned pend = 0;
#pragma unroll
    for (int u = 0; u < MB; ++u)
      if (threadIdx.x + kPT * u < nblk) pend |= ((1u << NV) - 1u) << (u * NV);
    for (int it = 0; it < kSpin && nb; ++it) {
      // all of this round's loads first, then their tests
      unsigned long long w[MB][NV];
#pragma unroll
      for (int u = 0; u < MB; ++u) {
        const int bb = threadIdx.x + kPT * u;
#pragma unroll
        for (int q = 0; q < NV; ++q)
          w[u][q] = (pend >> (u * NV + q)) & 1u
                        ? ld_dev(reinterpret_cast<const unsigned long long*>(set + 2 * bb + q))
                        : kSlotEmpty;
      }
#pragma unroll
      for (int u = 0; u < MB; ++u)
#pragma unroll
        for (int q = 0; q < NV; ++q) {
          const unsigned bit = 1u << (u * NV + q);
          if ((pend & bit) && w[u][q] != kSlotEmpty) {
            __builtin_memcpy(&got[u][q], &w[u][q], 8);
            pend &= ~bit;
          }
        }
      ok = pend == 0;
      if (ok) break;
      if ((it & 63) == 63 && ld_dev(flag) != 0u) break;
    }
  } else {
    for (int it = 0; it < kSpin && nb; ++it) {
      ok = true;
#pragma unroll
      for (int u = 0; u < MB; ++u) {
        const int bb = threadIdx.x + kPT * u;
        if (bb < nblk) {
#pragma unroll
          for (int q = 0; q < NV; ++q) {
            const unsigned long long w =
                ld_dev(reinterpret_cast<const unsigned long long*>(set + 2 * bb + q));
            __builtin_memcpy(&got[u][q], &w, 8);
            ok = ok && (w != kSlotEmpty);
          }
        }
      }
      if (ok) break;
      if ((it & 63) == 63 && ld_dev(flag) != 0u) break;
    }
  }
  if (__syncthreads_or(!ok)) {
    if (threadIdx.x == 0) st_dev(flag, 1u);
    return false;
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) red[q] = 0.0;
#pragma unroll
  for (int u = 0; u < MB; ++u) {
    const int bb = threadIdx.x + kPT * u;
    if (bb < nblk) {
#pragma unroll
      for (int q = 0; q < NV; ++q) red[q] = op[q] ? nmax(red[q], got[u][q]) : red[q] + got[u][q];
    }
  }
  block_reduce<NV, kPT>(red, op);
  // every block has now published sync j, so all have read sync j-1: its
  // slots can be emptied for reuse at sync j + kSets - 1
  if (threadIdx.x < kReps && j > 0) {
    double* prev = slots + ((size_t)((j - 1) % kSets) * kReps + threadIdx.x) * stride;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      st_dev(reinterpret_cast<unsigned long long*>(prev + 2 * vid[v]), kSlotEmpty);
      st_dev(reinterpret_cast<unsigned long long*>(prev + 2 * vid[v] + 1), kSlotEmpty);
    }
  }
  return true;
}

}  // namespace

// XV planes per virtual block (the KA/KB/KC chunk), V virtual blocks per
// block along x, XC = V * XV planes per thread; WPE waves per SIMD.
template <typename T, int XV, int V, int WPE>
__global__ void __launch_bounds__(kPT, WPE) bicg_persist_kernel(PersistArgs<T> A) {
  constexpr int XC = XV * V;
  using Pr = P2<T>;
  using FM = FaceMap<XC>;
  constexpr int NY = XC / 4;              // y-halo pairs per thread
  constexpr int NZH = 2 * XC * kPY;       // z-halo cells (threads [0, NZH) own one)
  const Geom& g = A.g;
  const TileGrid& tg = A.tg;
  const int nblk = tg.nblk;               // virtual blocks
  const int nphys = gridDim.x;
  const int b = blockIdx.x;
  // this block: tile (tzi, tyi), x chunks cg*V .. cg*V + V-1
  const int tzi = b % tg.ntz, tyi = (b / tg.ntz) % tg.nty, cg = b / (tg.ntz * tg.nty);
  const int z0 = tzi * kPZ, y0 = tyi * kPY, x0 = cg * XC;
  int vid[V];
#pragma unroll
  for (int v = 0; v < V; ++v) vid[v] = block_of_tile(tg, nblk, tzi, tyi, cg * V + v);
  // one virtual block of each neighbour block: its partial slot filled means
  // the neighbour's faces are complete
  int nbv[6];
  nbv[0] = block_of_tile(tg, nblk, tzi, tyi, (cg - 1) * V);
  nbv[1] = block_of_tile(tg, nblk, tzi, tyi, (cg + 1) * V);
  nbv[2] = block_of_tile(tg, nblk, tzi, tyi - 1, cg * V);
  nbv[3] = block_of_tile(tg, nblk, tzi, tyi + 1, cg * V);
  nbv[4] = block_of_tile(tg, nblk, tzi - 1, tyi, cg * V);
  nbv[5] = block_of_tile(tg, nblk, tzi + 1, tyi, cg * V);
  auto phys_of = [&](int tz, int ty, int c) -> int {
    tz = (tz + tg.ntz) % tg.ntz;
    ty = (ty + tg.nty) % tg.nty;
    const int ncg = tg.ncx / V;
    c = (c + ncg) % ncg;
    return (c * tg.nty + ty) * tg.ntz + tz;
  };
  int t = threadIdx.x;
  int lz = t % (kPZ / 2), ly = t / (kPZ / 2);
  const size_t FW = FM::size;
  const size_t FB = (size_t)nphys * FW;   // one field's faces
  const bool hasz = t < NZH;

  // Halo values at this thread's halo cells: the two x-halo pairs of its own
  // column (planes x0-1 and x0+XC), NY y-halo pairs (pair q = t + 256 u:
  // plane q / 64, row y0-1 or y0+8 by (q / 32) & 1, columns 2 (q & 31) + 0/1)
  // and, for t < NZH, one z-halo cell (plane t / 16, column z0-1 or z0+64 by
  // (t / 8) & 1, row t & 7).
  struct Halo {
    Pr x[2];
    Pr y[NY];
    T z;
  };
  auto fetch = [&](const T* base, Halo& H) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int nb = phys_of(tzi, tyi, cg + (s ? 1 : -1));
      H.x[s] = ld2_dev(base + (size_t)nb * FW + (s ? FM::xlo : FM::xhi) + ly * kPZ + 2 * lz);
    }
#pragma unroll
    for (int u = 0; u < NY; ++u) {
      const int q = t + kPT * u;
      const int side = (q / 32) & 1;
      const int nb = phys_of(tzi, tyi + (side ? 1 : -1), cg);
      H.y[u] = ld2_dev(base + (size_t)nb * FW + (side ? FM::ylo : FM::yhi) + (q / 64) * kPZ +
                       2 * (q & 31));
    }
    H.z = (T)0;
    if (hasz) {
      const int side = (t / 8) & 1;
      const int nb = phys_of(tzi + (side ? 1 : -1), tyi, cg);
      H.z = ld_dev(base + (size_t)nb * FW + (side ? FM::zlo : FM::zhi) + (t / 16) * kPY + (t & 7));
    }
  };
  // the halo values of r_0 (before any r face exists)
  auto fetch_r0 = [&](Halo& H) {
    auto gidx = [&](int x, int y, int z) -> int64_t {
      x = (x + g.nx) % g.nx;
      y = (y + g.ny) % g.ny;
      z = (z + g.nz) % g.nz;
      return (int64_t)x * g.isx + (int64_t)y * g.isy + z;
    };
    H.x[0] = ld2(A.r0 + gidx(x0 - 1, y0 + ly, z0 + 2 * lz));
    H.x[1] = ld2(A.r0 + gidx(x0 + XC, y0 + ly, z0 + 2 * lz));
#pragma unroll
    for (int u = 0; u < NY; ++u) {
      const int q = t + kPT * u;
      H.y[u] = ld2(A.r0 + gidx(x0 + q / 64, ((q / 32) & 1) ? y0 + kPY : y0 - 1,
                               z0 + 2 * (q & 31)));
    }
    H.z = hasz ? A.r0[gidx(x0 + t / 16, y0 + (t & 7), ((t / 8) & 1) ? z0 + kPZ : z0 - 1)]
               : (T)0;
  };
  __shared__ T s_rh[XC][kPY][kPZ];
  __shared__ T s_x[XC][kPY][kPZ];
  // LDS planes of the stencil's field with their halo: staged GP at a time
  // (FP64: 4, which is what fits beside r_hat and x; FP32: all 16)
  constexpr int GP = sizeof(T) == 8 ? 4 : XC;
  __shared__ T plane[GP][kPY + 2][kPZ + 2];

  // 7-point operator of the field F(x) (own cells of plane x) with halo H,
  // operands as compute_plane in pf_krylov.hip; epi(x, f, af) receives the
  // field and the operator at plane x.  A group of GP planes is written to
  // LDS, then read, with one barrier between (and one before the next
  // group's writes); hold() pins the caller's sums after each group.
  auto stencil = [&](auto&& F, const Halo& H, auto&& epi, auto&& hold) {
    // recompute the lane masks below here rather than keep them all live
    asm volatile("" : "+v"(t), "+v"(lz), "+v"(ly));
    const int zl = 2 * lz + 1;
    // the y halo of plane x is wave (x % 4)'s pair x / 4: a uniform branch
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int yrow = (t & 32) ? kPY + 1 : 0, ycol = 2 * (t & 31) + 1;
#pragma unroll
    for (int g0 = 0; g0 < XC; g0 += GP) {
      if (g0 > 0) __syncthreads();
#pragma unroll
      for (int x = g0; x < g0 + GP; ++x) {
        T(*pl)[kPZ + 2] = plane[x - g0];
        const Pr fc = F(x);
        pl[ly + 1][zl] = fc.a;
        pl[ly + 1][zl + 1] = fc.b;
        if ((x & 3) == wv) {
          pl[yrow][ycol] = H.y[x >> 2].a;
          pl[yrow][ycol + 1] = H.y[x >> 2].b;
        }
        if (hasz && t / 16 == x) pl[(t & 7) + 1][((t / 8) & 1) ? kPZ + 1 : 0] = H.z;
      }
      __syncthreads();
#pragma unroll
      for (int x = g0; x < g0 + GP; ++x) {
        T(*pl)[kPZ + 2] = plane[x - g0];
        const Pr fm = x == 0 ? H.x[0] : F(x - 1);
        const Pr fc = F(x);
        const Pr fq = x == XC - 1 ? H.x[1] : F(x + 1);
        const T af0 = lap7(A.L, fm.a, fc.a, fq.a, pl[ly][zl], pl[ly + 2][zl],
                           pl[ly + 1][zl - 1], fc.b);
        const T af1 = lap7(A.L, fm.b, fc.b, fq.b, pl[ly][zl + 1], pl[ly + 2][zl + 1], fc.a,
                           pl[ly + 1][zl + 2]);
        epi(x, fc, Pr{af0, af1});
      }
      // the group's results are complete here: its LDS operands are dead
      // before the next group's are read
      hold();
    }
  };

  // faces of a field this block publishes.  The x faces (every thread's
  // column) go straight out; the y faces (two rows of threads) and z faces
  // (two lanes per wave) are gathered in the idle LDS plane buffer first, so
  // that all threads store them with full-wave, contiguous stores (the face
  // region from row y0 on is [y0 row][y0+7 row][z0 column][z0+63 column]).
  auto publish = [&](T* base, auto&& F) {
    T* reg = base + (size_t)b * FW;
    st2_dev(reg + FM::xlo + ly * kPZ + 2 * lz, F(0));
    st2_dev(reg + FM::xhi + ly * kPZ + 2 * lz, F(XC - 1));
    constexpr int NYZ = FM::size - FM::ylo;
    static_assert(NYZ <= GP * (kPY + 2) * (kPZ + 2), "face staging exceeds the plane buffer");
    T* stg = &plane[0][0][0];
    __syncthreads();   // the plane buffer is free
    if (ly == 0 || ly == kPY - 1) {
      T* d = stg + (ly == 0 ? 0 : FM::yhi - FM::ylo) + 2 * lz;
#pragma unroll
      for (int x = 0; x < XC; ++x) {
        const Pr f = F(x);
        d[x * kPZ] = f.a;
        d[x * kPZ + 1] = f.b;
      }
    }
    if (lz == 0) {
#pragma unroll
      for (int x = 0; x < XC; ++x) stg[FM::zlo - FM::ylo + x * kPY + ly] = F(x).a;
    } else if (lz == kPZ / 2 - 1) {
#pragma unroll
      for (int x = 0; x < XC; ++x) stg[FM::zhi - FM::ylo + x * kPY + ly] = F(x).b;
    }
    __syncthreads();
    if constexpr (sizeof(T) == 4) {
      static_assert(NYZ % 2 == 0, "");
      for (int i = t; i < NYZ / 2; i += kPT)
        st_dev(reinterpret_cast<uint64_t*>(reg + FM::ylo) + i,
               reinterpret_cast<const uint64_t*>(stg)[i]);
    } else {
      for (int i = t; i < NYZ; i += kPT) st_dev(reg + FM::ylo + i, stg[i]);
    }
    __syncthreads();   // the plane buffer may be written again
  };

  // own cells: p, v, r, t in registers; r_hat and x in LDS
  Pr P[XC], V_[XC], R[XC], Tt[XC];
  const Pr zero2 = {(T)0, (T)0};
  {
    const int64_t col = (int64_t)(y0 + ly) * g.isy + z0 + 2 * lz;
#pragma unroll
    for (int x = 0; x < XC; ++x) {
      const int64_t q = (int64_t)(x0 + x) * g.isx + col;
      R[x] = ld2(A.r0 + q);
      const Pr rr = ld2(A.rh + q);
      s_rh[x][ly][2 * lz] = rr.a;
      s_rh[x][ly][2 * lz + 1] = rr.b;
      s_x[x][ly][2 * lz] = (T)0;
      s_x[x][ly][2 * lz + 1] = (T)0;
      P[x] = zero2;
      V_[x] = zero2;
      Tt[x] = zero2;
    }
  }

  const KrylovCtl c0 = *A.ctl;   // written by ctl_init_kernel before the launch
  const bool lead = b == 0 && t == 0;
  double rho1, rho2 = c0.rho2, alpha = c0.alpha, omega = c0.omega, res0, maxr = 0.0;
  int iter = 0, j = 0;
  int status = PF_SOLVE_CONVERGED;
  // Halos.  A block keeps the p halo it formed in KA (the neighbours' p
  // faces, bit for bit: the same operations on the same values) and the v
  // halo fetched during the sync after KA for the next KA, and the r halo
  // fetched during the sync after KC for KB's s halo, instead of fetching any
  // of them again: each sync fetches one face field (v after KA, r after KC)
  // and no p face is published.  hPo, hVo: p and v of the previous
  // iteration's KA on the halo; hRa: r on the halo; hVb: v on the halo for KB.
  // (Fetching p, v and r after KC and v and r after KA, with p and v faces
  // double-buffered: 128^3 FP64 26.5 us per iteration, FP32 21.9; keeping
  // FP64's halos in an L2-resident scratch slice per thread instead of
  // registers: 27.8.)
  Halo hPo, hVo, hRa, hVb;
  {
    double red[2];
    const int ops[2] = {0, 1};
    reduce_partials<2, kPT>(A.part_init, A.npart_init, red, ops);
    rho1 = red[0];
    res0 = red[1];
    if (lead) A.ctl->res0 = res0;
  }
  // v faces are read in the sync after KA only, which every block has
  // finished before any block passes the sync after KB, so before the next
  // KA writes them; r faces are read in the sync after KC only, finished
  // before the next KC writes them
  T* const fv = A.faces;
  T* const fr = A.faces + FB;

  for (;;) {
    // keep per-thread index arithmetic inside the loop instead of holding
    // every address it leads to in registers across iterations
    asm volatile("" : "+v"(t), "+v"(lz), "+v"(ly));
    // ---- KA: convergence of the previous iteration, beta, p, v = A p, <v, r_hat>
    if (iter > 0) {
      const double max_res = maxr / res0;                  // pressure.py:279-280
      if (lead) {
        A.ctl->max_res = max_res;
        if (iter <= c0.hist_cap) *(gptr<double>)(c0.hist + iter - 1) = max_res;
      }
      const bool cap = iter >= c0.maxit;                   // pressure.py:283-287
      const bool conv = max_res <= c0.tol;
      if (cap || conv) {
        status = conv ? PF_SOLVE_CONVERGED : PF_SOLVE_MAXIT;
        break;
      }
    }
    if (rho2 * omega == 0.0) {
      status = PF_SOLVE_BREAKDOWN_RHO;
      break;
    }
    const double beta = alpha * rho1 / (rho2 * omega);     // pressure.py:233
    if (lead) A.ctl->rho1 = rho1;
    {
      const T c1 = (T)(-omega), c2 = (T)beta;
      auto pupd = [&](const Pr& pp, const Pr& vv, const Pr& rr) -> Pr {
        return Pr{(fmx(vv.a, c1, pp.a) * c2) + rr.a, (fmx(vv.b, c1, pp.b) * c2) + rr.b};
      };
      // p on the halo from the neighbours' previous p and v and current r
      // (the own cells' update sits between the face loads and their use)
      Halo H;
      if (iter == 0) {
#pragma unroll
        for (int x = 0; x < XC; ++x) P[x] = pupd(P[x], V_[x], R[x]);
        Halo R0;   // p = v = 0
        fetch_r0(R0);
#pragma unroll
        for (int s = 0; s < 2; ++s) H.x[s] = pupd(zero2, zero2, R0.x[s]);
#pragma unroll
        for (int u = 0; u < NY; ++u) H.y[u] = pupd(zero2, zero2, R0.y[u]);
        H.z = (fmx((T)0, c1, (T)0) * c2) + R0.z;
        hRa = R0;
      } else {
#pragma unroll
        for (int x = 0; x < XC; ++x) P[x] = pupd(P[x], V_[x], R[x]);
#pragma unroll
        for (int s = 0; s < 2; ++s) H.x[s] = pupd(hPo.x[s], hVo.x[s], hRa.x[s]);
#pragma unroll
        for (int u = 0; u < NY; ++u) H.y[u] = pupd(hPo.y[u], hVo.y[u], hRa.y[u]);
        H.z = (fmx(hVo.z, c1, hPo.z) * c2) + hRa.z;
      }
      hPo = H;
      double acc[V];
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] = 0.0;
      stencil([&](int x) -> Pr { return P[x]; }, H, [&](int x, const Pr& f, const Pr& af) {
        V_[x] = af;
        acc[x / XV] += (double)af.a * (double)s_rh[x][ly][2 * lz];
        acc[x / XV] += (double)af.b * (double)s_rh[x][ly][2 * lz + 1];
      }, [&]() {
#pragma unroll
        for (int v = 0; v < V; ++v) pin(acc[v]);
      });
      publish(fv, [&](int x) -> Pr { return V_[x]; });
      double gam[1];
      const int ops[1] = {0};
      {
        int opv[V];
#pragma unroll
        for (int v = 0; v < V; ++v) opv[v] = 0;
        block_reduce<V, kPT>(acc, opv);
      }
      if (!grid_sync<1, V, sizeof(T) == 4>(A.slots, A.flag, nblk, vid, j++, acc, gam, ops, nbv, [&]() {
            fetch(fv, hVb);
          }))
        return;
      if (gam[0] == 0.0) {                                    // pressure.py:247-248
        status = PF_SOLVE_BREAKDOWN_A;
        break;
      }
      alpha = rho1 / gam[0];                                  // pressure.py:249
    }
    if (lead) A.ctl->alpha = alpha;

    // ---- KB: s = r - alpha v, t = A s, <s,t>, <t,t>
    const T na = (T)(-alpha), ta = (T)alpha;
    auto sval = [&](int x) -> Pr {
      return Pr{fmx(V_[x].a, na, R[x].a), fmx(V_[x].b, na, R[x].b)};
    };
    {
      Halo S;
      const Halo& hR = hRa;
#pragma unroll
      for (int s = 0; s < 2; ++s)
        S.x[s] = Pr{fmx(hVb.x[s].a, na, hR.x[s].a), fmx(hVb.x[s].b, na, hR.x[s].b)};
#pragma unroll
      for (int u = 0; u < NY; ++u)
        S.y[u] = Pr{fmx(hVb.y[u].a, na, hR.y[u].a), fmx(hVb.y[u].b, na, hR.y[u].b)};
      S.z = fmx(hVb.z, na, hR.z);
      hVo = hVb;
      double ab[2 * V];
#pragma unroll
      for (int q = 0; q < 2 * V; ++q) ab[q] = 0.0;
      stencil(sval, S, [&](int x, const Pr& f, const Pr& af) {
        Tt[x] = af;
        const int v = x / XV;
        ab[2 * v] += (double)f.a * (double)af.a;
        ab[2 * v] += (double)f.b * (double)af.b;
        ab[2 * v + 1] += (double)af.a * (double)af.a;
        ab[2 * v + 1] += (double)af.b * (double)af.b;
      }, [&]() {
#pragma unroll
        for (int q = 0; q < 2 * V; ++q) pin(ab[q]);
      });
      double st[2];
      const int ops[2] = {0, 0};
      {
        int opv[2 * V];
#pragma unroll
        for (int q = 0; q < 2 * V; ++q) opv[q] = 0;
        block_reduce<2 * V, kPT>(ab, opv);
      }
      if (!grid_sync<2, V, sizeof(T) == 4>(A.slots, A.flag, nblk, vid, j++, ab, st, ops, nbv, []() {}))
        return;
      if (st[1] == 0) {                                     // pressure.py:251, 266-267
#pragma unroll
        for (int x = 0; x < XC; ++x) {
          s_x[x][ly][2 * lz] = fmx(P[x].a, ta, s_x[x][ly][2 * lz]);
          s_x[x][ly][2 * lz + 1] = fmx(P[x].b, ta, s_x[x][ly][2 * lz + 1]);
        }
        status = PF_SOLVE_BREAKDOWN_B;
        break;
      }
      omega = st[0] / st[1];                                // pressure.py:268
    }
    if (lead) {
      A.ctl->omega = omega;
      A.ctl->rho2 = rho1;                                   // pressure.py:276
      A.ctl->iter = iter + 1;
    }
    rho2 = rho1;
    ++iter;

    // ---- KC: x += alpha p; x += omega s; r = s - omega t; <r, r_hat>, max|r|
    {
      const T to = (T)omega, tmo = (T)(-omega);
      double cc[2 * V];
#pragma unroll
      for (int q = 0; q < 2 * V; ++q) cc[q] = 0.0;
#pragma unroll
      for (int x = 0; x < XC; ++x) {
        const Pr sv = sval(x);
        T xa = fmx(P[x].a, ta, s_x[x][ly][2 * lz]);
        T xb = fmx(P[x].b, ta, s_x[x][ly][2 * lz + 1]);
        xa = fmx(sv.a, to, xa);
        xb = fmx(sv.b, to, xb);
        s_x[x][ly][2 * lz] = xa;
        s_x[x][ly][2 * lz + 1] = xb;
        const T ra = fmx(Tt[x].a, tmo, sv.a), rb = fmx(Tt[x].b, tmo, sv.b);
        R[x] = Pr{ra, rb};
        const int v = x / XV;
        cc[2 * v] += (double)ra * (double)s_rh[x][ly][2 * lz];
        cc[2 * v] += (double)rb * (double)s_rh[x][ly][2 * lz + 1];
        cc[2 * v + 1] = nmax(cc[2 * v + 1], fabs((double)ra));
        cc[2 * v + 1] = nmax(cc[2 * v + 1], fabs((double)rb));
      }
      publish(fr, [&](int x) -> Pr { return R[x]; });
      double rc[2];
      const int opc[2] = {0, 1};
      {
        int opv[2 * V];
#pragma unroll
        for (int q = 0; q < 2 * V; ++q) opv[q] = q & 1;
        block_reduce<2 * V, kPT>(cc, opv);
      }
      if (!grid_sync<2, V, sizeof(T) == 4>(A.slots, A.flag, nblk, vid, j++, cc, rc, opc, nbv, [&]() {
            fetch(fr, hRa);
          }))
        return;
      rho1 = rc[0];
      maxr = rc[1];
    }
  }
  // solution out; the control block as the last KA leaves it
  {
    const int64_t col = (int64_t)(y0 + ly) * g.isy + z0 + 2 * lz;
#pragma unroll
    for (int x = 0; x < XC; ++x) {
      const int64_t q = (int64_t)(x0 + x) * g.isx + col;
      *reinterpret_cast<Pr*>(A.x + q) = Pr{s_x[x][ly][2 * lz], s_x[x][ly][2 * lz + 1]};
    }
  }
  if (lead) {
    A.ctl->status = status;
    A.ctl->done = 1;
  }
}

__global__ void persist_reset_kernel(double* slots, int n, unsigned* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double e;
    const unsigned long long w = kSlotEmpty;
    __builtin_memcpy(&e, &w, 8);
    slots[i] = e;
  }
  if (i == 0) flag[0] = 0u;
}

// ---------------------------------------------------------------------------
// Host side

namespace {

// Planes per thread of the kernel a box takes: V virtual blocks (chunks of
// the tile grid) of xv planes each.
struct PersistShape {
  int xv, v, nphys;
  size_t face_elems;
};

bool persist_shape(const pf_box* h, const TileGrid& tg, PersistShape& sh) {
  if (tg.tz != kPZ || tg.ty != kPY) return false;
  if (!(tg.xc == 4 || tg.xc == 8 || (tg.xc == 16 && h->esize == 8))) return false;
  if (h->g.nx % tg.xc || h->g.ny % kPY || h->g.nz % kPZ) return false;
  if (tg.nblk > 1024) return false;
  sh.xv = tg.xc;
  // FP64: one wave per SIMD, 16 planes per thread (the state needs the whole
  // register file); FP32 on 4-plane chunks: two waves per SIMD, 8 planes
  // per thread, so that one block's waits are the other's compute (128^3
  // FP32: 27.1 -> 23.3 us per iteration)
  sh.v = (h->esize == 4 && tg.xc == 4) ? 2 : 16 / tg.xc;
  // grids of at most 256 tiles (128^3 FP64 on 16-plane chunks: 256; 64^3:
  // 128): one chunk per block, so that the solve spreads over as many CUs
  // as the grid has tiles
  if (tg.nblk <= 256) sh.v = 1;
  if (tg.ncx % sh.v) return false;
  // the kernels built: one chunk per block of 16 planes (FP64) or 4; two
  // 8-plane chunks per block; two 4-plane chunks per block at two waves per
  // SIMD (FP32)
  const bool built = sh.v == 1 ? (sh.xv == 4 || (sh.xv == 16 && h->esize == 8))
                               : sh.v == 2 && (sh.xv == 8 || (sh.xv == 4 && h->esize == 4));
  if (!built) return false;
  sh.nphys = tg.nblk / sh.v;
  const int xc = sh.v * sh.xv;
  sh.face_elems = xc == 16 ? FaceMap<16>::size : xc == 8 ? FaceMap<8>::size : FaceMap<4>::size;
  return true;
}

// Persistent-solve buffers of a box: [faces: v, r][partial-sum slots][abort flag]
size_t faces_bytes(const PersistShape& sh, size_t esize) {
  return 2 * (size_t)sh.nphys * sh.face_elems * esize;
}

size_t shape_bytes(const PersistShape& sh, size_t esize, int nblk) {
  return faces_bytes(sh, esize) + sizeof(double) * kSets * kReps * 2 * nblk + 256;
}

}  // namespace

// Whether box h's tile grid has a shape the persistent kernel handles, and
// the bytes of its face, slot and flag buffers (0 if not).
size_t persist_bytes(const pf_box* h) {
  const TileGrid tg = make_tile_grid(h->g, (int)h->esize, h->tile_rows);
  PersistShape sh;
  if (!persist_shape(h, tg, sh)) return 0;
  return shape_bytes(sh, h->esize, tg.nblk);
}

// Blocks of the persistent solve of box h (0 if its grid has no persistent
// kernel, its buffers are missing or the box has it switched off).
int persist_blocks(const pf_box* h) {
  if (!h->pers || !h->persist_on) return 0;
  const TileGrid tg = make_tile_grid(h->g, (int)h->esize, h->tile_rows);
  PersistShape sh;
  return persist_shape(h, tg, sh) ? sh.nphys : 0;
}

namespace {

unsigned* flag_ptr(pf_box* h, const TileGrid& tg, const PersistShape& sh) {
  const char* slots = (const char*)h->pers + faces_bytes(sh, h->esize);
  return (unsigned*)(slots + sizeof(double) * kSets * kReps * 2 * tg.nblk);
}

template <typename T, int XV, int V, int WPE>
int persist_launch(pf_box* h, hipStream_t s, const TileGrid& tg, const PersistShape& sh,
                   const Lap<T>& L, const T* r0, const T* rh, T* x, int npart_init, bool* used,
                   hipEvent_t* ev) {
  static int max_blocks[64];   // co-resident blocks per device (0: not yet known)
  const int dev = h->device;
  if (dev < 0 || dev >= 64) return 0;
  if (max_blocks[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bicg_persist_kernel<T, XV, V, WPE>,
                                                     kPT, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      (void)hipGetLastError();
      max_blocks[dev] = -1;
    } else {
      max_blocks[dev] = per_cu * cus;
    }
  }
  if (sh.nphys > max_blocks[dev]) return 0;
  PersistArgs<T> a;
  a.g = h->g;
  a.tg = tg;
  a.ctl = h->ctl;
  a.part_init = part_buf(h, 2);
  a.npart_init = npart_init;
  a.r0 = r0;
  a.rh = rh;
  a.x = x;
  a.faces = (T*)h->pers;
  a.slots = (double*)((char*)h->pers + faces_bytes(sh, sizeof(T)));
  a.flag = flag_ptr(h, tg, sh);
  a.L = L;
  const int ns = kSets * kReps * 2 * tg.nblk;
  hipLaunchKernelGGL(persist_reset_kernel, dim3((ns + 255) / 256), dim3(256), 0, s, a.slots, ns,
                     a.flag);
  PF_LAUNCH_CHECK("persist_reset_kernel");
  void* args[] = {&a};
  if (ev) PF_HIP(hipEventRecord(ev[0], s));
  hipError_t e = hipLaunchCooperativeKernel((const void*)bicg_persist_kernel<T, XV, V, WPE>,
                                            dim3(sh.nphys), dim3(kPT), args, 0, s);
  if (ev && e == hipSuccess) PF_HIP(hipEventRecord(ev[1], s));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;   // not launched: the caller runs the launch-per-reduction path
  }
  *used = true;
  return 0;
}

}  // namespace

template <typename T>
int persist_solve(pf_box* h, hipStream_t s, const TileGrid& tg, const Lap<T>& L, const T* r0,
                  const T* rh, T* x, int npart_init, bool* used, hipEvent_t* ev) {
  *used = false;
  if (!h->pers || !h->persist_on) return 0;
  PersistShape sh;
  if (!persist_shape(h, tg, sh)) return 0;
  if (sh.v == 1) {
    // one chunk per block: 16 planes (FP64 on 16-plane chunks) or 4
    if constexpr (sizeof(T) == 8)
      if (sh.xv == 16)
        return persist_launch<T, 16, 1, 1>(h, s, tg, sh, L, r0, rh, x, npart_init, used, ev);
    if (sh.xv == 4)
      return persist_launch<T, 4, 1, 1>(h, s, tg, sh, L, r0, rh, x, npart_init, used, ev);
    return 0;
  }
  if (sh.xv == 8)
    return persist_launch<T, 8, 2, 1>(h, s, tg, sh, L, r0, rh, x, npart_init, used, ev);
  if constexpr (sizeof(T) == 4)
    if (sh.v == 2)
      return persist_launch<T, 4, 2, 2>(h, s, tg, sh, L, r0, rh, x, npart_init, used, ev);
  return 0;   // other shapes: the launch-per-reduction kernels
}

template int persist_solve<double>(pf_box*, hipStream_t, const TileGrid&, const Lap<double>&,
                                   const double*, const double*, double*, int, bool*,
                                   hipEvent_t*);
template int persist_solve<float>(pf_box*, hipStream_t, const TileGrid&, const Lap<float>&,
                                  const float*, const float*, float*, int, bool*, hipEvent_t*);

// Abort flag of the last persistent solve (synchronises the stream).
int persist_aborted(pf_box* h, hipStream_t s, bool* aborted) {
  *aborted = false;
  if (!h->pers) return 0;
  const TileGrid tg = make_tile_grid(h->g, (int)h->esize, h->tile_rows);
  PersistShape sh;
  if (!persist_shape(h, tg, sh)) return 0;
  unsigned* host = reinterpret_cast<unsigned*>(h->stats_host);
  PF_HIP(hipMemcpyAsync(host, flag_ptr(h, tg, sh), sizeof(unsigned), hipMemcpyDeviceToHost, s));
  if (int rc = wait_stream(s, h->wait_ev)) return rc;
  *aborted = host[0] != 0;
  return 0;
}

}  // namespace pf
