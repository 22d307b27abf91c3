// Shared device/host helpers for the PyFlow box kernels (gfx950, wave64).
//
// Layouts (all row-major, z fastest, as the reference's torch tensors):
//   padded   : (nx+4) x (ny+4) x (nz+4), two overlap cells per side
//              (library/parallel.py:201, data/state.py:123); the interior
//              origin sits at (2,2,2).
//   interior : nx x ny x nz, contiguous (the reference's rhs_*, source_P,
//              and this library's Krylov vectors).
// Kernels read padded fields through their overlap cells exactly like the
// reference's shifted slices do, and read interior fields with periodic
// index wrap.  Every kernel that writes a padded field also writes the
// periodic images of each interior value into the overlap cells, which is
// what state_P.update_border (data/state.py:148-150) leaves behind.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pf {

constexpr int kNover = 2;
constexpr int kBlock = 256;          // threads per block for streaming kernels
constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Unsigned division by a runtime constant (Granlund-Montgomery, round-up
// multiplier).  Exact for 0 <= n < 2^31 and 1 <= d < 2^31.
struct FastDiv {
  uint32_t d, m, s;
  FastDiv() : d(1), m(0), s(0) {}
  explicit FastDiv(uint32_t div) : d(div) {
    s = 0;
    while ((1ull << s) < div) ++s;
    uint64_t two_l = 1ull << s;
    m = (uint32_t)(((two_l - div) << 32) / div + 1);
  }
  __device__ __forceinline__ uint32_t div(uint32_t n) const {
    uint32_t t = __umulhi(n, m);
    return (t + n) >> s;
  }
};

// Small pointer bundles passed to kernels by value.
template <typename T>
struct Vec3 {
  T* a[3];
};
template <typename T>
struct CVec3 {
  const T* a[3];
};
template <typename T>
struct Vec6 {
  T* a[6];
};

// Geometry of one box as the kernels see it.
struct Geom {
  int nx, ny, nz;
  int64_t ncell;           // nx*ny*nz
  int64_t psx, psy;        // padded strides
  int64_t poff;            // padded offset of interior (0,0,0)
  int64_t isx, isy;        // interior strides
  FastDiv fd_yz, fd_z;     // divide by ny*nz and by nz
};

inline Geom make_geom(int nx, int ny, int nz) {
  Geom g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.ncell = (int64_t)nx * ny * nz;
  g.psy = nz + 2 * kNover;
  g.psx = (int64_t)(ny + 2 * kNover) * g.psy;
  g.poff = kNover * g.psx + kNover * g.psy + kNover;
  g.isy = nz;
  g.isx = (int64_t)ny * nz;
  g.fd_yz = FastDiv((uint32_t)(ny * nz));
  g.fd_z = FastDiv((uint32_t)nz);
  return g;
}

__device__ __forceinline__ void decode(const Geom& g, uint32_t idx, int& i, int& j, int& k) {
  uint32_t q = g.fd_yz.div(idx);
  uint32_t r = idx - q * (uint32_t)(g.ny * g.nz);
  uint32_t q2 = g.fd_z.div(r);
  i = (int)q;
  j = (int)q2;
  k = (int)(r - q2 * (uint32_t)g.nz);
}

__device__ __forceinline__ int64_t pidx(const Geom& g, int i, int j, int k) {
  return g.poff + i * g.psx + j * g.psy + k;
}

// Write v at interior (i,j,k) of a padded field and at every periodic image
// of that cell inside the overlap (up to 7 extra stores at a corner).
template <typename T>
__device__ __forceinline__ void store_padded(const Geom& g, T* a, int i, int j, int k, T v) {
  int ii[2], jj[2], kk[2];
  int ni = 1, nj = 1, nk = 1;
  ii[0] = i; jj[0] = j; kk[0] = k;
  if (i < kNover) ii[ni++] = i + g.nx; else if (i >= g.nx - kNover) ii[ni++] = i - g.nx;
  if (j < kNover) jj[nj++] = j + g.ny; else if (j >= g.ny - kNover) jj[nj++] = j - g.ny;
  if (k < kNover) kk[nk++] = k + g.nz; else if (k >= g.nz - kNover) kk[nk++] = k - g.nz;
  for (int a0 = 0; a0 < ni; ++a0)
    for (int b0 = 0; b0 < nj; ++b0)
      for (int c0 = 0; c0 < nk; ++c0)
        a[pidx(g, ii[a0], jj[b0], kk[c0])] = v;
}

// Periodic neighbour offsets in the interior layout.
struct Nbr {
  int64_t xm, xp, ym, yp, zm, zp;   // offsets added to the flat index
};

__device__ __forceinline__ Nbr nbr(const Geom& g, int i, int j, int k) {
  Nbr n;
  n.xm = (i == 0) ? (int64_t)(g.nx - 1) * g.isx : -g.isx;
  n.xp = (i == g.nx - 1) ? -(int64_t)(g.nx - 1) * g.isx : g.isx;
  n.ym = (j == 0) ? (int64_t)(g.ny - 1) * g.isy : -g.isy;
  n.yp = (j == g.ny - 1) ? -(int64_t)(g.ny - 1) * g.isy : g.isy;
  n.zm = (k == 0) ? (int64_t)(g.nz - 1) : -1;
  n.zp = (k == g.nz - 1) ? -(int64_t)(g.nz - 1) : 1;
  return n;
}

// ---------------------------------------------------------------------------
// Deterministic reductions.  A block reduces its threads' values with a fixed
// tree (wave butterfly, then LDS across waves); kernels write one partial per
// block, and whoever consumes them reduces the partials again with the same
// fixed assignment and tree.  Results therefore depend only on the data and
// the (fixed) block count, never on scheduling.

// Sum over the wave as the xor butterfly v += shfl_xor(v, off), off = 32,
// 16, ..., 1, with the same pairs in the same order and so the same bits,
// without LDS: the 32- and 16-lane steps swap halves of the wave and of each
// 32-lane half between two copies (v_permlane32_swap / v_permlane16_swap,
// one copy then holds a lane's own value and the other its partner's, and
// addition commutes); the 8- and 4-lane steps rotate each 16-lane row, whose
// source lane holds the same value as the xor partner once the wider steps
// are done; the 2- and 1-lane steps are quad permutations.
__device__ __forceinline__ double wave_sum(double v) {
  auto swap_add = [](double x, bool wide) -> double {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    const auto L = wide ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                        : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto H = wide ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                        : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double((int)H[0], (int)L[0]) + __hiloint2double((int)H[1], (int)L[1]);
  };
  auto dpp_add = [](double x, auto ctrl) -> double {
    constexpr int c = decltype(ctrl)::value;
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), c, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), c, 0xf, 0xf, false);
    return x + __hiloint2double(hi, lo);
  };
  v = swap_add(v, true);
  v = swap_add(v, false);
  v = dpp_add(v, std::integral_constant<int, 0x128>{});   // row_ror:8
  v = dpp_add(v, std::integral_constant<int, 0x124>{});   // row_ror:4
  v = dpp_add(v, std::integral_constant<int, 0x4E>{});    // quad_perm [2,3,0,1]
  v = dpp_add(v, std::integral_constant<int, 0xB1>{});    // quad_perm [1,0,3,2]
  return v;
}
// NaN-propagating max (torch.max propagates NaN; fmax would drop it).
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }

// Square root rounded to nearest, as IEEE 754 (and the CPU torch.sqrt the
// reference uses) defines it.  The device square root may be off by an ulp;
// one remainder step corrects it: r = x - y*y exactly by an fma, then
// y + r/(2y), also by an fma (Markstein's correction, IBM J. Res. Dev. 34(1),
// 1990).  x = 0 returns 0.  Floats go through double: rounding a correctly
// rounded double square root to float gives the correctly rounded float one
// (53 >= 2*24 + 2).
__device__ __forceinline__ double sqrt_rn(double x) {
  if (!(x > 0.0)) return sqrt(x);
  const double y = sqrt(x);
  const double r = fma(-y, y, x);
  return fma(r, 0.5 / y, y);
}
__device__ __forceinline__ float sqrt_rn(float x) { return (float)sqrt_rn((double)x); }

// Max over the wave as the xor butterfly v = nmax(v, shfl_xor(v, off)), off
// = 32, ..., 1, in the same way as wave_sum (the lane's own value stays the
// first operand; a row rotation's source lane holds the xor partner's value
// whenever nmax of two values does not depend on their order, i.e. except
// for two NaNs of different payloads or two zeros of opposite signs - the
// maxima taken here are of absolute values).
__device__ __forceinline__ double wave_max(double v) {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  auto swap_max = [&](double x, bool wide) -> double {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    const auto L = wide ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                        : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto H = wide ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                        : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double d0 = __hiloint2double((int)H[0], (int)L[0]);
    const double d1 = __hiloint2double((int)H[1], (int)L[1]);
    // d0 holds the own value in the lower half of each swapped pair of
    // halves, d1 in the upper half
    const bool upper = (lane & (wide ? 32 : 16)) != 0;
    return upper ? nmax(d1, d0) : nmax(d0, d1);
  };
  auto dpp_max = [](double x, auto ctrl) -> double {
    constexpr int c = decltype(ctrl)::value;
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), c, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), c, 0xf, 0xf, false);
    return nmax(x, __hiloint2double(hi, lo));
  };
  v = swap_max(v, true);
  v = swap_max(v, false);
  v = dpp_max(v, std::integral_constant<int, 0x128>{});   // row_ror:8
  v = dpp_max(v, std::integral_constant<int, 0x124>{});   // row_ror:4
  v = dpp_max(v, std::integral_constant<int, 0x4E>{});    // quad_perm [2,3,0,1]
  v = dpp_max(v, std::integral_constant<int, 0xB1>{});    // quad_perm [1,0,3,2]
  return v;
}

// Reduce NV values per thread over the block (op per slot: 0 = sum, 1 = max).
// Result valid in every thread.  Uses NV*(kBlock/kWave) doubles of LDS.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const int (&op)[NV]) {
  __shared__ double sh[NV][NT / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = op[q] ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q][wid] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double a = sh[q][0];
#pragma unroll
    for (int w = 1; w < NT / kWave; ++w) a = op[q] ? nmax(a, sh[q][w]) : a + sh[q][w];
    v[q] = a;
  }
  __syncthreads();
}

// Reduce NV partial slots written by nblk blocks (layout part[b*NV + q]).
template <int NV, int NT = kBlock>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ part, int nblk,
                                                double (&v)[NV], const int (&op)[NV]) {
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += NT) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double x = part[(int64_t)b * NV + q];
      v[q] = op[q] ? nmax(v[q], x) : v[q] + x;
    }
  }
  block_reduce<NV, NT>(v, op);
}

template <int NV>
__device__ __forceinline__ void write_partials(double* __restrict__ part, const double (&v)[NV],
                                               int bid) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) part[(int64_t)bid * NV + q] = v[q];
  }
}

template <int NV>
__device__ __forceinline__ void write_partials(double* __restrict__ part, const double (&v)[NV]) {
  write_partials<NV>(part, v, (int)blockIdx.x);
}

// The block of a launch that works on one box: its index among the box's
// blocks and their number.  A single-box launch uses (blockIdx.x, gridDim.x);
// a batched launch over several boxes gives each box the same decomposition
// it has alone, so per-box results do not depend on the batch.
struct Blk {
  int id, n;
};

__device__ __forceinline__ Blk this_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x}; }

}  // namespace pf
