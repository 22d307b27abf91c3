// Pressure Poisson path: divergence source, BiCGStab, Jacobi, projection
// and the per-step monitor values.
//
// BiCGStab (solver/pressure.py:196-297) runs as three kernels per iteration,
// one per global reduction of the recurrence:
//   KA: p = (p - omega*v)*beta + r, v = A p,        partial <v, r_hat>
//   KB: s = r - alpha*v,            t = A s,        partials <s,t>, <t,t>
//   KC: s, t again; x += alpha*p; x += omega*s; r' = s - omega*t,
//                                                   partials <r',r_hat>, max|r'|
// KB stores nothing: KC recomputes s and t = A s from r and v (the same
// operations, so the same bits), which costs a stencil but saves writing and
// re-reading s and t (15 instead of 17 vector passes per iteration).  r is
// double-buffered because KC reads r around each cell while writing r'.
// The scalars never leave the device: every block of a kernel reduces the
// previous kernel's per-block partials itself (same data, same order, same
// bits in every block), and block 0 records them in the box's KrylovCtl.
// KA also checks the previous iteration's convergence, so the host only
// polls the control block once per batch of iterations.  Once a solve is
// done every later kernel of the batch returns at once.
//
// The stencil of A is evaluated on the fly from p (resp. s) recomputed at the
// six neighbours, so p and s are never re-read after being written in the
// same kernel.  All vector arithmetic follows the reference's op order
// (add_/sub_ with a scalar are single FMAs in torch, everything else rounds
// per op); only the dot-product summation order differs from torch's CPU sum.
#include <hip/hip_runtime.h>

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "pf_kernels.h"

namespace pf {

enum { kPartA = 0, kPartB = 1, kPartC = 2, kPartS = 3 };

// ---------------------------------------------------------------------------
// Control block
__global__ void ctl_init_kernel(KrylovCtl* ctl, double tol, int maxit, int npart, double* hist,
                                int hist_cap) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl->rho1 = 1.0; ctl->rho2 = 1.0; ctl->alpha = 0.0; ctl->omega = 1.0;
    ctl->res0 = 0.0; ctl->max_res = 0.0; ctl->tol = tol;
    ctl->iter = 0; ctl->maxit = maxit; ctl->done = 0; ctl->status = PF_SOLVE_CONVERGED;
    ctl->halt = 0;
    ctl->npart = npart;
    ctl->hist = hist; ctl->hist_cap = hist ? hist_cap : 0;
  }
}

__device__ __forceinline__ bool lead(const Blk& b) { return b.id == 0 && threadIdx.x == 0; }

// Block-uniform snapshot of the control block: thread 0 reads it once and
// broadcasts through LDS, so that a write by block 0 of the same kernel can
// never split a block between an early return and a __syncthreads().
__device__ __forceinline__ KrylovCtl snapshot(const KrylovCtl* ctl) {
  __shared__ KrylovCtl sc;
  if (threadIdx.x == 0) sc = *ctl;
  __syncthreads();
  KrylovCtl c = sc;
  __syncthreads();
  return c;
}

// ---------------------------------------------------------------------------
// Source and Krylov start: b = (src*(rho/dt))*(-vol) (pressure.py:205-208),
// r = r_hat = b, x = p = v = 0 (pressure.py:202, 217-218).  With FROM_DIV the
// source is first computed as div(u) (metric_staggered.py:438-453).
template <typename T, bool FROM_DIV>
__device__ __forceinline__ void bicg_init_kernel_body(
    Geom g, T* __restrict__ src, const T* __restrict__ U, const T* __restrict__ V,
    const T* __restrict__ W, T dxi, T dyi, T dzi, T sc1, T sc2, T* __restrict__ r,
    T* __restrict__ rh, T* __restrict__ x, T* __restrict__ p0, T* __restrict__ v0,
    double* __restrict__ partC, Blk blk) {
  double acc[2] = {0.0, 0.0};
  const int ops[2] = {0, 1};
  for (uint32_t idx = blk.id * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += blk.n * blockDim.x) {
    T b;
    if (FROM_DIV) {
      int i, j, k;
      decode(g, idx, i, j, k);
      const int64_t c = i * g.psx + j * g.psy + k;
      b = (U[c + g.psx] - U[c]) * dxi;
      b = b + ((V[c + g.psy] - V[c]) * dyi);
      b = b + ((W[c + 1] - W[c]) * dzi);
    } else {
      b = src[idx];
    }
    b = b * sc1;
    b = b * sc2;
    src[idx] = b;
    r[idx] = b;
    rh[idx] = b;
    x[idx] = (T)0;
    p0[idx] = (T)0;
    v0[idx] = (T)0;
    acc[0] += (double)b * (double)b;
    acc[1] = nmax(acc[1], fabs((double)b));
  }
  block_reduce<2>(acc, ops);
  write_partials<2>(partC, acc, blk.id);
}

template <typename T, bool FROM_DIV>
__global__ void __launch_bounds__(kBlock) bicg_init_kernel(
    Geom g, T* __restrict__ src, const T* __restrict__ U, const T* __restrict__ V,
    const T* __restrict__ W, T dxi, T dyi, T dzi, T sc1, T sc2, T* __restrict__ r,
    T* __restrict__ rh, T* __restrict__ x, T* __restrict__ p0, T* __restrict__ v0,
    double* __restrict__ partC) {
  bicg_init_kernel_body<T, FROM_DIV>(g, src, U, V, W, dxi, dyi, dzi, sc1, sc2, r, rh, x, p0, v0, partC, this_blk());
}

// ---------------------------------------------------------------------------
// Tiled KA / KB.  Both apply the 7-point operator to a field that is itself
// recomputed from three (KA) or two (KB) stored vectors.  A block owns a
// TZ x TY tile of the (y,z) plane and marches along x over a chunk of planes:
// each thread keeps its column's field at x-1, x, x+1 in registers, the
// current plane's tile plus a one-cell y/z halo goes through LDS, and the
// loads of plane x+1 (and of the next halo) are issued before the barrier of
// plane x.  Cells of partial tiles wrap periodically, so the LDS plane always
// holds valid neighbours; only their outputs are masked.
enum { kModeA = 0, kModeB = 1, kModeC = 2 };

// Field recomputed at a cell: KA p = (p_old - omega v_old)*beta + r,
// KB and KC s = r - alpha v.
template <typename T, int MODE>
struct FieldOp {
  const T* __restrict__ r;
  const T* __restrict__ a;   // KA: p_old; KB: unused
  const T* __restrict__ v;   // KA: v_old; KB: v
  T c1, c2;                  // KA: (-omega, beta); KB, KC: (-alpha, -)
  struct Raw {
    T v, a, r;
  };
  // the stored values at a cell (loaded early, combined later)
  __device__ __forceinline__ Raw load(int64_t q) const {
    Raw w;
    w.v = v[q];
    w.a = MODE == kModeA ? a[q] : (T)0;
    w.r = r[q];
    return w;
  }
  __device__ __forceinline__ T eval(const Raw& w) const {
    if (MODE == kModeA) return (fma(w.v, c1, w.a) * c2) + w.r;
    return fma(w.v, c1, w.r);
  }
  __device__ __forceinline__ T at(int64_t q) const { return eval(load(q)); }
};

// Scalars of a tiled kernel: the FieldOp coefficients, and for KC alpha,
// omega, -omega and the <t,t> == 0 breakdown flag.
template <typename T>
struct TileScal {
  T c1 = (T)0, c2 = (T)0;
  T ta = (T)0, to = (T)0, tmo = (T)0;
  bool brk = false;
};

// The scalars of KA / KB / KC from the control block and the previous
// kernel's partials; false when the block must return (solve done,
// converged, capped or broken down).
template <typename T, int MODE, int NT>
__device__ __forceinline__ bool tile_scalars(KrylovCtl* __restrict__ ctl,
                                             const double* __restrict__ part_in, int nblk_in,
                                             const Blk& blk, TileScal<T>& sc) {
  const KrylovCtl c0 = snapshot(ctl);
  if (c0.done) return false;
  T& c1 = sc.c1;
  T& c2 = sc.c2;
  if (MODE == kModeC) {
    double red[2];
    const int ops[2] = {0, 0};
    reduce_partials<2, NT>(part_in, nblk_in, red, ops);
    const double buf1 = red[0], buf2 = red[1];
    sc.ta = (T)c0.alpha;
    c1 = (T)(-c0.alpha);
    if (buf2 == 0) {                                  // pressure.py:251, 266-267
      sc.brk = true;
      if (lead(blk)) {
        ctl->status = PF_SOLVE_BREAKDOWN_B;
        ctl->halt = 1;
      }
      return true;
    }
    const double omega = buf1 / buf2;                 // pressure.py:268
    if (lead(blk)) {
      ctl->omega = omega;
      ctl->rho2 = c0.rho1;                            // pressure.py:276
      ctl->iter = c0.iter + 1;
      ctl->npart = blk.n;
    }
    sc.to = (T)omega;
    sc.tmo = (T)(-omega);
    return true;
  }
  if (MODE == kModeA) {
    if (c0.halt) {
      if (lead(blk)) ctl->done = 1;
      return false;
    }
    double red[2];
    const int ops[2] = {0, 1};
    reduce_partials<2, NT>(part_in, c0.npart, red, ops);   // init's or KC's partials
    const double rho1 = red[0];
    if (c0.iter == 0) {
      if (lead(blk)) ctl->res0 = red[1];
    } else {
      const double max_res = red[1] / c0.res0;       // pressure.py:279-280
      if (lead(blk)) {
        ctl->max_res = max_res;
        if (c0.iter <= c0.hist_cap) c0.hist[c0.iter - 1] = max_res;
      }
      const bool cap = c0.iter >= c0.maxit;           // pressure.py:283-287
      const bool conv = max_res <= c0.tol;
      if (cap || conv) {
        if (lead(blk)) {
          ctl->status = conv ? PF_SOLVE_CONVERGED : PF_SOLVE_MAXIT;
          ctl->done = 1;
        }
        return false;
      }
    }
    if (c0.rho2 * c0.omega == 0.0) {
      // beta would be x/0: the reference's iterates turn NaN from here on
      if (lead(blk)) {
        ctl->status = PF_SOLVE_BREAKDOWN_RHO;
        ctl->done = 1;
      }
      return false;
    }
    const double beta = c0.alpha * rho1 / (c0.rho2 * c0.omega);   // pressure.py:233
    if (lead(blk)) ctl->rho1 = rho1;
    c1 = (T)(-c0.omega);
    c2 = (T)beta;
  } else {
    double red[1];
    const int ops[1] = {0};
    reduce_partials<1, NT>(part_in, nblk_in, red, ops);
    const double gamma = red[0];
    if (gamma == 0.0) {                                // pressure.py:247-248
      if (lead(blk)) {
        ctl->status = PF_SOLVE_BREAKDOWN_A;
        ctl->done = 1;
      }
      return false;
    }
    const double alpha = c0.rho1 / gamma;              // pressure.py:249
    if (lead(blk)) ctl->alpha = alpha;
    c1 = (T)(-alpha);
  }
  return true;
}

// Two cells of a z row, loaded and stored with one vector access.
template <typename T>
struct alignas(2 * sizeof(T)) Pair {
  T a, b;
};

// Output store of the tile kernels.  NT: non-temporal (NTS of the bodies),
// used by the rotating schedule of large grids and by batched launches,
// whose Krylov sets do not fit the Infinity Cache
// (512^3 FP32: KA 674 -> 659, KB 230 -> 217, KC 691 -> 674 us); on the
// short-chunk grids (128^3 FP64, whose set fits) it was slower (KA 20.6 ->
// 22.5 us), the next kernel's reads no longer finding the data there.
// Otherwise, in FP32, the stores write through the XCD's L2 at device
// scope (relaxed agent-scope stores: global_store ... sc1), so that no dirty
// lines of the output are left in the eight L2s for the end of the kernel
// to write back before the next kernel, on any XCD, may read them (128^3
// FP32: KA 13.6 -> 12.6 us, KC 11.9 -> 11.1 us, the step 191 -> 199
// Mcell-steps/s).  FP64 keeps plain 16-byte stores: as two 8-byte
// write-through stores per pair its 128^3 kernels ran half again slower,
// and as one 16-byte write-through buffer store no faster.
#ifndef PF_KRYLOV_WT
#define PF_KRYLOV_WT 1
#endif

template <bool NT, typename T>
__device__ __forceinline__ void store2(T* p, const Pair<T>& v) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 w = {v.a, v.b};
    __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
  } else if constexpr (PF_KRYLOV_WT && sizeof(T) == 4) {
    uint64_t w;
    __builtin_memcpy(&w, &v, 8);
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), w, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  } else {
    *reinterpret_cast<Pair<T>*>(p) = v;
  }
}

template <bool NT, typename T>
__device__ __forceinline__ void store1(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else if constexpr (PF_KRYLOV_WT && sizeof(T) == 4)
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    *p = v;
}

// One cell per thread; PIPE selects the load schedule as in bicg_tile2_body.
template <typename T, int MODE, int TZ, int TY, bool PIPE, bool NTS = !PIPE>
__device__ __forceinline__ void bicg_tile_kernel_body(
    Geom g, TileGrid tg, int nblk_in, KrylovCtl* __restrict__ ctl,
    const double* __restrict__ part_in, double* __restrict__ part_out,
    const T* __restrict__ r, const T* __restrict__ rh, const T* __restrict__ a_old,
    const T* __restrict__ v_in, T* __restrict__ f_out, T* __restrict__ af_out, Lap<T> L, Blk blk) {
  constexpr int NT = TZ * TY;
  // coefficients are set once the scalars are known (below)
  FieldOp<T, MODE> F{r, a_old, v_in, (T)0, (T)0};

  int tz_i, ty_i, cx_i;
  tile_of_block(tg, blk, tz_i, ty_i, cx_i);
  const int z0 = tz_i * TZ, y0 = ty_i * TY, x0 = cx_i * tg.xc;
  const int x1 = min(x0 + tg.xc, g.nx);

  const int t = threadIdx.x;
  const int lz = t % TZ, ly = t / TZ;
  const int zr = z0 + lz, yr = y0 + ly;
  const bool active = zr < g.nz && yr < g.ny;
  const int z = zr % g.nz, y = yr % g.ny;
  const int64_t col = (int64_t)y * g.isy + z;

  // halo assignment: 2*TZ y-halo cells, then 2*TY z-halo cells
  int hz = -1, hy = -1, hslot_y = 0, hslot_z = 0;
  bool is_halo = false;
  if (t < 2 * TZ) {
    is_halo = true;
    hslot_z = t % TZ + 1;
    hslot_y = (t < TZ) ? 0 : TY + 1;
    hz = (z0 + t % TZ) % g.nz;
    hy = (t < TZ) ? (y0 - 1 + g.ny) % g.ny : (y0 + TY) % g.ny;
  } else if (t < 2 * TZ + 2 * TY) {
    is_halo = true;
    const int u = t - 2 * TZ;
    hslot_y = u % TY + 1;
    hslot_z = (u < TY) ? 0 : TZ + 1;
    hy = (y0 + u % TY) % g.ny;
    hz = (u < TY) ? (z0 - 1 + g.nz) % g.nz : (z0 + TZ) % g.nz;
  }
  // with PIPE, threads without a halo cell load their own cell and discard it
  const int64_t hcol = is_halo ? (int64_t)hy * g.isy + hz : (PIPE ? col : 0);

  // Two LDS planes used alternately: plane x is written to buffer x & 1, so
  // one barrier per plane suffices (a thread writing buffer (x+1) & 1 has
  // passed the barrier of plane x, after every read of plane x-1).
  __shared__ T plane[2][TY + 2][TZ + 2];
  using Raw = typename FieldOp<T, MODE>::Raw;
  auto xoff = [&](int x) -> int64_t {
    x = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
    return (int64_t)x * g.isx;
  };
  // The stored vectors of the first planes do not depend on the solver
  // scalars: their loads are issued before the control block and the
  // previous kernel's partials are read and reduced, and overlap them.
  const Raw rm0 = F.load(xoff(x0 - 1) + col);
  const Raw rc0 = F.load(xoff(x0) + col);
  Raw rh0 = {};
  if (PIPE || is_halo) rh0 = F.load(xoff(x0) + hcol);
  // PIPE: register set c/h = column and halo of plane x+1, rh/x/p = the own
  // cell at plane x.  !PIPE: plane x+1 is loaded ahead (x+2 as well for KB)
  // and the loads rotate into place, the own cell is read in its iteration.
  struct Ahead {
    Raw c, h;
    T rh, x, p;
  };
  auto load_own = [&](Ahead& a, int64_t q) {
    if (MODE != kModeB && (PIPE || active)) a.rh = rh[q];
    if (MODE == kModeC && (PIPE || active)) {
      a.x = f_out[q];
      a.p = a_old[q];
    }
  };
  Ahead A = {}, B = {};
  constexpr bool kDeep = !PIPE && MODE == kModeB;
  Raw rn = {}, hrn = {}, rn2 = {}, hrn2 = {};
  if (PIPE) {
    A.c = F.load(xoff(x0 + 1) + col);
    A.h = F.load(xoff(min(x0 + 1, x1 - 1)) + hcol);
    load_own(A, (int64_t)x0 * g.isx + col);
  } else {
    rn = F.load(xoff(x0 + 1) + col);
    if (is_halo && x0 + 1 < x1) hrn = F.load(xoff(x0 + 1) + hcol);
    if (kDeep && x0 + 2 <= x1) rn2 = F.load(xoff(x0 + 2) + col);
    if (kDeep && is_halo && x0 + 2 < x1) hrn2 = F.load(xoff(x0 + 2) + hcol);
  }

  TileScal<T> sc;
  if (!tile_scalars<T, MODE, NT>(ctl, part_in, nblk_in, blk, sc)) return;
  F.c1 = sc.c1;
  F.c2 = sc.c2;
  if (MODE == kModeC && sc.brk) {       // <t,t> == 0: only x += alpha*p
    if (active)
      for (int x = x0; x < x1; ++x) {
        const int64_t q = (int64_t)x * g.isx + col;
        f_out[q] = fma(a_old[q], sc.ta, f_out[q]);
      }
    return;
  }
  T fm = F.eval(rm0);
  T fc = F.eval(rc0);
  T hv = (PIPE || is_halo) ? F.eval(rh0) : (T)0;
  double acc0 = 0.0, acc1 = 0.0;
  // plane x from fm, fc (planes x-1, x), fp (x+1) and the halo hv; then the
  // register window moves on by one plane
  auto compute_plane = [&](int x, T fp, T hn, const Ahead& cur) {
    const int64_t xo = (int64_t)x * g.isx;
    const T rhc = cur.rh, xc = cur.x, pc = cur.p;
    T(*pl)[TZ + 2] = plane[x & 1];
    pl[ly + 1][lz + 1] = fc;
    if (is_halo) pl[hslot_y][hslot_z] = hv;
    __syncthreads();
    if (active) {
      const T fym = pl[ly][lz + 1], fyp = pl[ly + 2][lz + 1];
      const T fzm = pl[ly + 1][lz], fzp = pl[ly + 1][lz + 2];
      const T af = lap7(L, fm, fc, fp, fym, fyp, fzm, fzp);
      if (MODE == kModeA) {
        store1<NTS>(f_out + xo + col, fc);
        store1<NTS>(af_out + xo + col, af);
        acc0 += (double)af * (double)rhc;
      } else if (MODE == kModeB) {
        acc0 += (double)fc * (double)af;
        acc1 += (double)af * (double)af;
      } else {
        // KC: fc = s, af = t = A s
        T xv = fma(pc, sc.ta, xc);                  // pressure.py:251
        xv = fma(fc, sc.to, xv);                    // pressure.py:270
        store1<NTS>(f_out + xo + col, xv);
        const T rnew = fma(af, sc.tmo, fc);         // pressure.py:274-275
        store1<NTS>(af_out + xo + col, rnew);
        acc0 += (double)rnew * (double)rhc;
        acc1 = nmax(acc1, fabs((double)rnew));
      }
    }
    fm = fc;
    fc = fp;
    hv = hn;
  };
  if constexpr (PIPE) {
    auto plane_step = [&](int x, const Ahead& cur, Ahead& nxt) {
      nxt.c = F.load(xoff(min(x + 2, x1)) + col);
      nxt.h = F.load(xoff(min(x + 2, x1 - 1)) + hcol);
      load_own(nxt, (int64_t)min(x + 1, x1 - 1) * g.isx + col);
      compute_plane(x, F.eval(cur.c), F.eval(cur.h), cur);
    };
    int x = x0;
    for (; x + 1 < x1; x += 2) {
      plane_step(x, A, B);
      plane_step(x + 1, B, A);
    }
    if (x < x1) plane_step(x, A, B);
  } else {
    for (int x = x0; x < x1; ++x) {
      const int64_t xo = (int64_t)x * g.isx;
      const int xl = kDeep ? x + 3 : x + 2;
      Raw rnn = {}, hrnn = {};
      if (xl <= x1) rnn = F.load(xoff(xl) + col);
      if (is_halo && xl < x1) hrnn = F.load(xoff(xl) + hcol);
      Ahead cur = {};
      load_own(cur, xo + col);
      const T fp = F.eval(rn);
      const T hn = (is_halo && x + 1 < x1) ? F.eval(hrn) : (T)0;
      compute_plane(x, fp, hn, cur);
      if (kDeep) {
        rn = rn2;
        hrn = hrn2;
        rn2 = rnn;
        hrn2 = hrnn;
      } else {
        rn = rnn;
        hrn = hrnn;
      }
    }
  }
  if (MODE == kModeA) {
    double accv[1] = {acc0};
    const int ops1[1] = {0};
    block_reduce<1, NT>(accv, ops1);
    write_partials<1>(part_out, accv, blk.id);
  } else {
    // KB: <s,t>, <t,t>; KC: <r,r_hat>, max|r|
    double accv[2] = {acc0, acc1};
    const int ops2[2] = {0, MODE == kModeC ? 1 : 0};
    block_reduce<2, NT>(accv, ops2);
    write_partials<2>(part_out, accv, blk.id);
  }
}

// KA / KB with two z-neighbour cells per thread: a 256-thread block owns a
// TZ x TY tile (TZ/2 threads per row, TY = 512/TZ rows), so the y halo is a
// quarter of the tile instead of a half, and every load and store of the
// interior columns is a 2-element vector access.  Used when TZ divides nz
// (no periodic wrap inside a row).  Per-cell arithmetic is the scalar
// kernel's; only the order of the per-thread partial sums differs.
//
// Two load schedules, chosen per grid at launch (same arithmetic, same bits):
// PIPE (short x chunks): every load runs one plane ahead of its use, in two
// register sets that swap roles in a loop unrolled by two.  A register whose
// load is in flight cannot be copied without waiting for that load, so the
// rotating schedule below leaves each iteration waiting on its own loads;
// the swap avoids that.  Every load is unconditional (past the chunk's end
// the loads repeat the last plane they need, threads without a halo cell
// read a discarded pair), because the compiler waits for every load in
// flight where a load or its use sits under a branch.
// !PIPE (long chunks, 512^3): iteration x issues the column loads of plane
// x+2 (x+3 for KB) at its top and rotates them into place at its end, and
// reads the own-cell vectors in their iteration (one plane ahead for FP32
// KC), so it waits on its own loads.  On long chunks, with four waves per
// SIMD, this schedule still reaches the bandwidth of a plain stream, and
// PIPE was slower there (512^3 FP32: KA 683 -> 735 us, KC 690 -> 828 us).
template <typename T, int MODE, int TZ, int TY, bool PIPE, bool NTS = !PIPE>
__device__ __forceinline__ void bicg_tile2_body(
    Geom g, TileGrid tg, int nblk_in, KrylovCtl* __restrict__ ctl,
    const double* __restrict__ part_in, double* __restrict__ part_out,
    const T* __restrict__ r, const T* __restrict__ rh, const T* __restrict__ a_old,
    const T* __restrict__ v_in, T* __restrict__ f_out, T* __restrict__ af_out, Lap<T> L, Blk blk) {
  constexpr int NZT = TZ / 2;
  constexpr int NT = NZT * TY;
  // coefficients are set once the scalars are known (below)
  FieldOp<T, MODE> F{r, a_old, v_in, (T)0, (T)0};
  using Raw = typename FieldOp<T, MODE>::Raw;
  struct Raw2 {
    Raw a, b;
  };
  auto ld2 = [&](const T* p, int64_t q) -> Pair<T> {
    return *reinterpret_cast<const Pair<T>*>(p + q);
  };
  auto load2 = [&](int64_t q) -> Raw2 {
    const Pair<T> v = ld2(F.v, q), r2 = ld2(F.r, q);
    Pair<T> a = {(T)0, (T)0};
    if (MODE == kModeA) a = ld2(F.a, q);
    Raw2 w;
    w.a.v = v.a; w.a.a = a.a; w.a.r = r2.a;
    w.b.v = v.b; w.b.a = a.b; w.b.r = r2.b;
    return w;
  };
  auto eval2 = [&](const Raw2& w) -> Pair<T> { return Pair<T>{F.eval(w.a), F.eval(w.b)}; };

  int tz_i, ty_i, cx_i;
  tile_of_block(tg, blk, tz_i, ty_i, cx_i);
  const int z0 = tz_i * TZ, y0 = ty_i * TY, x0 = cx_i * tg.xc;
  const int x1 = min(x0 + tg.xc, g.nx);

  const int t = threadIdx.x;
  const int lz = t % NZT, ly = t / NZT;
  const int yr = y0 + ly;
  const bool active = yr < g.ny;
  const int64_t col = (int64_t)(yr % g.ny) * g.isy + (z0 + 2 * lz);

  // halo: threads [0, 2*NZT) load the y-halo rows in pairs, the next 2*TY
  // threads one z-halo cell each: with PIPE as one element of the aligned
  // pair that holds it (nz is a multiple of TZ, so the pair lies inside the
  // row), and the other threads read the tile's first pair and discard it.
  int hslot_y = 0, hslot_z = 0;
  bool hpair = false, hone = false, hhi = false;
  int64_t hcol = PIPE ? (int64_t)y0 * g.isy + z0 : 0;
  if (t < 2 * NZT) {
    hpair = true;
    hslot_y = (t < NZT) ? 0 : TY + 1;
    hslot_z = 2 * (t % NZT) + 1;
    const int hy = (t < NZT) ? (y0 - 1 + g.ny) % g.ny : (y0 + TY) % g.ny;
    hcol = (int64_t)hy * g.isy + z0 + 2 * (t % NZT);
  } else if (t < 2 * NZT + 2 * TY) {
    hone = true;
    const int u = t - 2 * NZT;
    hslot_y = u % TY + 1;
    hslot_z = (u < TY) ? 0 : TZ + 1;
    const int hy = (y0 + u % TY) % g.ny;
    int hz = (u < TY) ? (z0 - 1 + g.nz) % g.nz : (z0 + TZ) % g.nz;
    if (PIPE && u < TY) {   // z0 - 1 is the second element of the pair at z0 - 2
      hhi = true;
      hz = (z0 - 2 + g.nz) % g.nz;
    }
    hcol = (int64_t)hy * g.isy + hz;
  }
  auto loadh = [&](int64_t q) -> Raw2 {
    if (PIPE || hpair) return load2(q);
    Raw2 w = {};
    w.a = F.load(q);
    return w;
  };
  auto evalh = [&](const Raw2& w) -> Pair<T> {
    const T e0 = F.eval(hhi ? w.b : w.a), e1 = F.eval(w.b);
    return Pair<T>{e0, hpair ? e1 : (T)0};
  };
  const bool is_halo = hpair || hone;

  __shared__ T plane[2][TY + 2][TZ + 2];
  auto xoff = [&](int x) -> int64_t {
    x = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
    return (int64_t)x * g.isx;
  };
  // the vectors KA and KC read at the cell itself: r_hat (KA, KC), x and p (KC)
  struct Own {
    Pair<T> rh, x, p;
  };
  auto load_own = [&](int64_t q) -> Own {
    Own o = {{(T)0, (T)0}, {(T)0, (T)0}, {(T)0, (T)0}};
    if (MODE != kModeB && (PIPE || active)) o.rh = ld2(rh, q);
    if (MODE == kModeC && (PIPE || active)) {
      o.x = ld2(f_out, q);
      o.p = ld2(a_old, q);
    }
    return o;
  };
  // The stored vectors of the first planes do not depend on the solver
  // scalars: their loads are issued before the control block and the
  // previous kernel's partials are read and reduced, and overlap them.
  const Raw2 rm0 = load2(xoff(x0 - 1) + col);
  const Raw2 rc0 = load2(xoff(x0) + col);
  Raw2 rh0 = {};
  if (PIPE || is_halo) rh0 = loadh(xoff(x0) + hcol);
  // PIPE: register set c/h = column and halo of plane x+1, o = own cells of
  // plane x.  !PIPE: KB loads planes x+1 and x+2 ahead (268 -> 232 us at
  // 512^3 FP32), KA and KC x+1 (with x+2 as well, KC went 690 -> 849 us at
  // 512^3 FP32); FP32 KC loads its own-cell vectors one plane ahead (730 ->
  // 690 us at 512^3; ahead of time KA went 675 -> 743 us there).
  struct Ahead {
    Raw2 c, h;
    Own o;
  };
  Ahead A = {}, B = {};
  constexpr bool kDeep = !PIPE && MODE == kModeB;
  constexpr bool kAhead = !PIPE && MODE == kModeC && sizeof(T) == 4;
  Raw2 rn = {}, hrn = {}, rn2 = {}, hrn2 = {};
  Own own = {};
  if (PIPE) {
    A.c = load2(xoff(x0 + 1) + col);
    A.h = loadh(xoff(min(x0 + 1, x1 - 1)) + hcol);
    A.o = load_own((int64_t)x0 * g.isx + col);
  } else {
    rn = load2(xoff(x0 + 1) + col);
    if (is_halo && x0 + 1 < x1) hrn = loadh(xoff(x0 + 1) + hcol);
    if (kDeep && x0 + 2 <= x1) rn2 = load2(xoff(x0 + 2) + col);
    if (kDeep && is_halo && x0 + 2 < x1) hrn2 = loadh(xoff(x0 + 2) + hcol);
    if (kAhead) own = load_own((int64_t)x0 * g.isx + col);
  }

  TileScal<T> sc;
  if (!tile_scalars<T, MODE, NT>(ctl, part_in, nblk_in, blk, sc)) return;
  F.c1 = sc.c1;
  F.c2 = sc.c2;
  if (MODE == kModeC && sc.brk) {       // <t,t> == 0: only x += alpha*p
    if (active)
      for (int x = x0; x < x1; ++x) {
        const int64_t q = (int64_t)x * g.isx + col;
        const Pair<T> pv = ld2(a_old, q), xv = ld2(f_out, q);
        *reinterpret_cast<Pair<T>*>(f_out + q) =
            Pair<T>{fma(pv.a, sc.ta, xv.a), fma(pv.b, sc.ta, xv.b)};
      }
    return;
  }

  Pair<T> fm = eval2(rm0);
  Pair<T> fc = eval2(rc0);
  Pair<T> hv = {(T)0, (T)0};
  if (PIPE || is_halo) hv = evalh(rh0);
  double acc0 = 0.0, acc1 = 0.0;
  const int zl = 2 * lz + 1;   // LDS column of the first cell
  // plane x from fm, fc (planes x-1, x), fp (x+1) and the halo hv; then the
  // register window moves on by one plane
  auto compute_plane = [&](int x, const Pair<T>& fp, const Pair<T>& hn, const Own& cur) {
    const int64_t xo = (int64_t)x * g.isx;
    const Pair<T> rhc = cur.rh, xc = cur.x, pc = cur.p;
    T(*pl)[TZ + 2] = plane[x & 1];
    pl[ly + 1][zl] = fc.a;
    pl[ly + 1][zl + 1] = fc.b;
    if (is_halo) {
      pl[hslot_y][hslot_z] = hv.a;
      if (hpair) pl[hslot_y][hslot_z + 1] = hv.b;
    }
    __syncthreads();
    if (active) {
      const T af0 = lap7(L, fm.a, fc.a, fp.a, pl[ly][zl], pl[ly + 2][zl], pl[ly + 1][zl - 1],
                         fc.b);
      const T af1 = lap7(L, fm.b, fc.b, fp.b, pl[ly][zl + 1], pl[ly + 2][zl + 1], fc.a,
                         pl[ly + 1][zl + 2]);
      if (MODE == kModeA) {
        store2<NTS>(f_out + xo + col, fc);
        store2<NTS>(af_out + xo + col, Pair<T>{af0, af1});
        acc0 += (double)af0 * (double)rhc.a;
        acc0 += (double)af1 * (double)rhc.b;
      } else if (MODE == kModeB) {
        acc0 += (double)fc.a * (double)af0;
        acc0 += (double)fc.b * (double)af1;
        acc1 += (double)af0 * (double)af0;
        acc1 += (double)af1 * (double)af1;
      } else {
        // KC: fc = s, af = t = A s (pressure.py:251, 270, 274-275)
        T xa = fma(pc.a, sc.ta, xc.a);
        T xb = fma(pc.b, sc.ta, xc.b);
        xa = fma(fc.a, sc.to, xa);
        xb = fma(fc.b, sc.to, xb);
        store2<NTS>(f_out + xo + col, Pair<T>{xa, xb});
        const T ra = fma(af0, sc.tmo, fc.a), rb = fma(af1, sc.tmo, fc.b);
        store2<NTS>(af_out + xo + col, Pair<T>{ra, rb});
        acc0 += (double)ra * (double)rhc.a;
        acc0 += (double)rb * (double)rhc.b;
        acc1 = nmax(acc1, fabs((double)ra));
        acc1 = nmax(acc1, fabs((double)rb));
      }
    }
    fm = fc;
    fc = fp;
    hv = hn;
  };
  if constexpr (PIPE) {
    auto plane_step = [&](int x, const Ahead& cur, Ahead& nxt) {
      nxt.c = load2(xoff(min(x + 2, x1)) + col);
      nxt.h = loadh(xoff(min(x + 2, x1 - 1)) + hcol);
      nxt.o = load_own((int64_t)min(x + 1, x1 - 1) * g.isx + col);
      compute_plane(x, eval2(cur.c), evalh(cur.h), cur.o);
    };
    int x = x0;
    for (; x + 1 < x1; x += 2) {
      plane_step(x, A, B);
      plane_step(x + 1, B, A);
    }
    if (x < x1) plane_step(x, A, B);
  } else {
    for (int x = x0; x < x1; ++x) {
      const int64_t xo = (int64_t)x * g.isx;
      // the loads of plane x + 1 + (kDeep ? 2 : 1)
      const int xl = kDeep ? x + 3 : x + 2;
      Raw2 rn3 = {}, hrn3 = {};
      if (xl <= x1) rn3 = load2(xoff(xl) + col);
      if (is_halo && xl < x1) hrn3 = loadh(xoff(xl) + hcol);
      const Own cur = kAhead ? own : load_own(xo + col);
      if (kAhead && x + 1 < x1) own = load_own(xo + g.isx + col);
      const Pair<T> fp = eval2(rn);
      Pair<T> hn = {(T)0, (T)0};
      if (is_halo && x + 1 < x1) hn = evalh(hrn);
      compute_plane(x, fp, hn, cur);
      if (kDeep) {
        rn = rn2;
        hrn = hrn2;
        rn2 = rn3;
        hrn2 = hrn3;
      } else {
        rn = rn3;
        hrn = hrn3;
      }
    }
  }
  if (MODE == kModeA) {
    double accv[1] = {acc0};
    const int ops1[1] = {0};
    block_reduce<1, NT>(accv, ops1);
    write_partials<1>(part_out, accv, blk.id);
  } else {
    // KB: <s,t>, <t,t>; KC: <r,r_hat>, max|r|
    double accv[2] = {acc0, acc1};
    const int ops2[2] = {0, MODE == kModeC ? 1 : 0};
    block_reduce<2, NT>(accv, ops2);
    write_partials<2>(part_out, accv, blk.id);
  }
}

template <typename T, int MODE, int TZ, int TY, bool PIPE>
__global__ void __launch_bounds__(TZ / 2 * TY) bicg_tile2_kernel(
    Geom g, TileGrid tg, int nblk_in, KrylovCtl* __restrict__ ctl,
    const double* __restrict__ part_in, double* __restrict__ part_out,
    const T* __restrict__ r, const T* __restrict__ rh, const T* __restrict__ a_old,
    const T* __restrict__ v_in, T* __restrict__ f_out, T* __restrict__ af_out, Lap<T> L) {
  bicg_tile2_body<T, MODE, TZ, TY, PIPE>(g, tg, nblk_in, ctl, part_in, part_out, r, rh, a_old, v_in,
                                   f_out, af_out, L, this_blk());
}

template <typename T, int MODE, int TZ, int TY, bool PIPE>
__global__ void __launch_bounds__(TZ * TY) bicg_tile_kernel(
    Geom g, TileGrid tg, int nblk_in, KrylovCtl* __restrict__ ctl,
    const double* __restrict__ part_in, double* __restrict__ part_out,
    const T* __restrict__ r, const T* __restrict__ rh, const T* __restrict__ a_old,
    const T* __restrict__ v_in, T* __restrict__ f_out, T* __restrict__ af_out, Lap<T> L) {
  bicg_tile_kernel_body<T, MODE, TZ, TY, PIPE>(g, tg, nblk_in, ctl, part_in, part_out, r, rh, a_old, v_in, f_out, af_out, L, this_blk());
}

template <typename T, int MODE>
void launch_tile(const TileGrid& tg, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const Geom& g,
                 int nblk_in, KrylovCtl* ctl, const double* pin, double* pout, const T* r,
                 const T* rh, const T* a_old, const T* v_in, T* f_out, T* af_out,
                 const Lap<T>& L) {
#define PF_TILE_P(TZ, TY, P)                                                                 \
  hipExtLaunchKernelGGL((bicg_tile_kernel<T, MODE, TZ, TY, P>), dim3(tg.nblk), dim3(TZ * TY), 0, \
                        s, e0, e1, 0, g, tg, nblk_in, ctl, pin, pout, r, rh, a_old, v_in, f_out, \
                        af_out, L)
#define PF_TILE2_P(TZ, TY, P)                                                                \
  hipExtLaunchKernelGGL((bicg_tile2_kernel<T, MODE, TZ, TY, P>), dim3(tg.nblk),                \
                        dim3(TZ / 2 * TY), 0, s, e0, e1, 0, g, tg, nblk_in, ctl, pin, pout, r,   \
                        rh, a_old, v_in, f_out, af_out, L)
#define PF_TILE(TZ, TY)                                                                      \
  do {                                                                                       \
    if (pipe) PF_TILE_P(TZ, TY, true); else PF_TILE_P(TZ, TY, false);                        \
  } while (0)
#define PF_TILE2(TZ, TY)                                                                     \
  do {                                                                                       \
    if (pipe) PF_TILE2_P(TZ, TY, true); else PF_TILE2_P(TZ, TY, false);                      \
  } while (0)
  // the one-plane-ahead register schedule on short x chunks (measured: 128^3
  // FP64 step 132.4 -> 136.3 Mcell-steps/s, 128^3 FP32 184.5 -> 190.5, 64^3
  // FP64 76.9 -> 80.0); the rotating one on long chunks (512^3: 256 planes),
  // where it is the faster of the two
  const bool pipe = tg.xc <= kPipeChunk;
  // tz * ty == 2 * kBlock: two cells per thread
  const int key = tg.tz * 1000 + tg.ty;
  switch (key) {
    case 64004: PF_TILE(64, 4); break;
    case 32008: PF_TILE(32, 8); break;
    case 16016: PF_TILE(16, 16); break;
    case 64008: PF_TILE2(64, 8); break;
    case 32016: PF_TILE2(32, 16); break;
    case 16032: PF_TILE2(16, 32); break;
    default: PF_TILE(8, 32); break;
  }
#undef PF_TILE2
#undef PF_TILE
#undef PF_TILE2_P
#undef PF_TILE_P
}

// DP (padded) := x with periodic overlap (pressure.py:293), then
// p += DP over the whole padded array (pressure.py:294).
template <typename T, bool ADD_P>
__device__ __forceinline__ void scatter_padded_kernel_body(Geom g, const T* __restrict__ x,
                                                                 T* __restrict__ DP,
                                                                 T* __restrict__ p, Blk blk) {
  const int nxo = g.nx + 2 * kNover, nyo = g.ny + 2 * kNover, nzo = g.nz + 2 * kNover;
  const int64_t total = (int64_t)nxo * nyo * nzo;
  for (int64_t a = blk.id * (int64_t)blockDim.x + threadIdx.x; a < total;
       a += (int64_t)blk.n * blockDim.x) {
    const int ko = (int)(a % nzo);
    const int64_t t1 = a / nzo;
    const int jo = (int)(t1 % nyo);
    const int io = (int)(t1 / nyo);
    int i = io - kNover, j = jo - kNover, k = ko - kNover;
    i = i < 0 ? i + g.nx : (i >= g.nx ? i - g.nx : i);
    j = j < 0 ? j + g.ny : (j >= g.ny ? j - g.ny : j);
    k = k < 0 ? k + g.nz : (k >= g.nz ? k - g.nz : k);
    const T d = x[i * g.isx + j * g.isy + k];
    DP[a] = d;
    if (ADD_P) p[a] = p[a] + d;
  }
}

template <typename T, bool ADD_P>
__global__ void __launch_bounds__(kBlock) scatter_padded_kernel(Geom g, const T* __restrict__ x,
                                                                 T* __restrict__ DP,
                                                                 T* __restrict__ p) {
  scatter_padded_kernel_body<T, ADD_P>(g, x, DP, p, this_blk());
}

// u -= (dt/rho) * grad_P(DP) on the interior (metric_staggered.py:390-401,
// state.py:198-211); DP is the solver's interior-layout solution.
template <typename T>
__device__ __forceinline__ void project_kernel_body(Geom g, const T* __restrict__ x, T* U,
                                                         T* V, T* W, T dxi, T dyi, T dzi,
                                                         T neg_c, Blk blk) {
  for (uint32_t idx = blk.id * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += blk.n * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const Nbr n = nbr(g, i, j, k);
    const T d = x[idx];
    const int64_t c = pidx(g, i, j, k);
    const T gx = (d - x[idx + n.xm]) * dxi;
    const T gy = (d - x[idx + n.ym]) * dyi;
    const T gz = (d - x[idx + n.zm]) * dzi;
    store_padded(g, U, i, j, k, fma(gx, neg_c, U[c]));
    store_padded(g, V, i, j, k, fma(gy, neg_c, V[c]));
    store_padded(g, W, i, j, k, fma(gz, neg_c, W[c]));
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) project_kernel(Geom g, const T* __restrict__ x, T* U,
                                                         T* V, T* W, T dxi, T dyi, T dzi,
                                                         T neg_c) {
  project_kernel_body<T>(g, x, U, V, W, dxi, dyi, dzi, neg_c, this_blk());
}

// Same projection with DP given as a padded field (Jacobi's pOld).
template <typename T>
__global__ void __launch_bounds__(kBlock) project_padded_kernel(Geom g, const T* __restrict__ D,
                                                                T* U, T* V, T* W, T dxi, T dyi,
                                                                T dzi, T neg_c) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    const T d = D[c];
    const T gx = (d - D[c - g.psx]) * dxi;
    const T gy = (d - D[c - g.psy]) * dyi;
    const T gz = (d - D[c - 1]) * dzi;
    store_padded(g, U, i, j, k, fma(gx, neg_c, U[c]));
    store_padded(g, V, i, j, k, fma(gy, neg_c, V[c]));
    store_padded(g, W, i, j, k, fma(gz, neg_c, W[c]));
  }
}

// Monitor values: max|u|,|v|,|w|, sum(u^2+v^2+w^2), max|div| (+ div into divg).
template <typename T>
__device__ __forceinline__ void stats_kernel_body(Geom g, const T* __restrict__ U,
                                                       const T* __restrict__ V,
                                                       const T* __restrict__ W, T dxi, T dyi,
                                                       T dzi, T* __restrict__ divg,
                                                       double* __restrict__ part, Blk blk) {
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int ops[5] = {1, 1, 1, 0, 1};
  for (uint32_t idx = blk.id * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += blk.n * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = i * g.psx + j * g.psy + k;
    const T u = U[c], v = V[c], w = W[c];
    acc[0] = nmax(acc[0], fabs((double)u));
    acc[1] = nmax(acc[1], fabs((double)v));
    acc[2] = nmax(acc[2], fabs((double)w));
    acc[3] += (double)(((u * u) + (v * v)) + (w * w));   // PyFlow.py:388-390
    T d = (U[c + g.psx] - u) * dxi;
    d = d + ((V[c + g.psy] - v) * dyi);
    d = d + ((W[c + 1] - w) * dzi);
    if (divg) divg[idx] = d;
    acc[4] = nmax(acc[4], fabs((double)d));
  }
  block_reduce<5>(acc, ops);
  write_partials<5>(part, acc, blk.id);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) stats_kernel(Geom g, const T* __restrict__ U,
                                                       const T* __restrict__ V,
                                                       const T* __restrict__ W, T dxi, T dyi,
                                                       T dzi, T* __restrict__ divg,
                                                       double* __restrict__ part) {
  stats_kernel_body<T>(g, U, V, W, dxi, dyi, dzi, divg, part, this_blk());
}

struct Ops8 {
  int v[8];
};

template <int NV>
__global__ void __launch_bounds__(kBlock) finalize_kernel(const double* __restrict__ part,
                                                          int nblk, Ops8 o,
                                                          double* __restrict__ out) {
  int ops[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) ops[q] = o.v[q];
  double v[NV];
  reduce_partials<NV>(part, nblk, v, ops);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; ++q) out[q] = v[q];
}

// B (interior) = 7-point operator applied to padded A (pressure.py:63-84).
template <typename T>
__global__ void __launch_bounds__(kBlock) apply_lap_kernel(Geom g, const T* __restrict__ A,
                                                           T* __restrict__ B, Lap<T> L) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    B[idx] = lap7(L, A[c - g.psx], A[c], A[c + g.psx], A[c - g.psy], A[c + g.psy], A[c - 1], A[c + 1]);
  }
}

// ---------------------------------------------------------------------------
// Jacobi (pressure.py:115-161)
// init: p = 0 (and the ping-pong partner), src *= rho/dt, max|src|, src *= dx^2
template <typename T>
__global__ void __launch_bounds__(kBlock) jacobi_init_kernel(Geom g, T* __restrict__ src,
                                                             const T* __restrict__ U,
                                                             const T* __restrict__ V,
                                                             const T* __restrict__ W,
                                                             int from_div, T dxi, T dyi, T dzi,
                                                             T sc1, T dx2,
                                                             double* __restrict__ part) {
  double acc[1] = {0.0};
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    T b;
    if (from_div) {
      int i, j, k;
      decode(g, idx, i, j, k);
      const int64_t c = i * g.psx + j * g.psy + k;
      b = (U[c + g.psx] - U[c]) * dxi;
      b = b + ((V[c + g.psy] - V[c]) * dyi);
      b = b + ((W[c + 1] - W[c]) * dzi);
    } else {
      b = src[idx];
    }
    b = b * sc1;
    acc[0] = nmax(acc[0], fabs((double)b));
    src[idx] = b * dx2;
  }
  const int ops[1] = {1};
  block_reduce<1>(acc, ops);
  write_partials<1>(part, acc);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) zero_padded_kernel(int64_t n, T* __restrict__ a) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    a[q] = (T)0;
}

// One sweep: out = ((((((-o(i+1) - o(i-1)) - o(j+1)) - o(j-1)) - o(k+1)) - o(k-1)) + src)*DInv
template <typename T>
__global__ void __launch_bounds__(kBlock) jacobi_sweep_kernel(Geom g, const T* __restrict__ O,
                                                              const T* __restrict__ src,
                                                              T* __restrict__ out, T dinv) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    T a = -O[c + g.psx];
    a = a - O[c - g.psx];
    a = a - O[c + g.psy];
    a = a - O[c - g.psy];
    a = a - O[c + 1];
    a = a - O[c - 1];
    a = a + src[idx];
    store_padded(g, out, i, j, k, a * dinv);
  }
}

// Final residual: src /= dx^2; res = -(A p) + src; max|res|.
template <typename T>
__global__ void __launch_bounds__(kBlock) jacobi_resid_kernel(Geom g, const T* __restrict__ P,
                                                              T* __restrict__ src, T dx2, Lap<T> L,
                                                              double* __restrict__ part) {
  double acc[1] = {0.0};
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t c = pidx(g, i, j, k);
    const T s = src[idx] / dx2;
    src[idx] = s;
    T b = lap7(L, P[c - g.psx], P[c], P[c + g.psx], P[c - g.psy], P[c + g.psy], P[c - 1], P[c + 1]);
    b = b * (T)(-1.0);
    b = b + s;
    acc[0] = nmax(acc[0], fabs((double)b));
  }
  const int ops[1] = {1};
  block_reduce<1>(acc, ops);
  write_partials<1>(part, acc);
}

// ---------------------------------------------------------------------------
// Batched launches: every box of the list runs the single-box body with its
// own block decomposition (see BoxDesc).

#define PF_BATCH_BOX                                                \
  const int bi_ = find_box(L, (int)blockIdx.x);                     \
  const BoxDesc<T>& d = D[L.box[bi_]];                              \
  const Blk blk{(int)blockIdx.x - L.begin[bi_], L.begin[bi_ + 1] - L.begin[bi_]}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_init_kernel(const BoxDesc<T>* __restrict__ D,
                                                            BatchList L) {
  PF_BATCH_BOX;
  const int64_t o = d.g.poff;
  bicg_init_kernel_body<T, true>(d.g, d.src, d.u[0] + o, d.u[1] + o, d.u[2] + o, d.dxi, d.dyi,
                                 d.dzi, d.sc1, d.sc2, d.rk[0], d.rh, d.x, d.pk[0], d.vk[0], d.pC,
                                 blk);
}

// Iteration j has parity j & 1.  KA of iteration j (launched with parity
// j & 1): p[1-parity] from p[parity], v[parity] and r[parity]; KB and KC of
// iteration j (launched with parity j & 1): s, t from r[parity] and
// v[1-parity]; KC also reads p[1-parity] and writes r[1-parity].
template <typename T, int MODE, int TZ, int TY, bool PAIR, bool NTS>
__global__ void __launch_bounds__(PAIR ? TZ / 2 * TY : TZ * TY) batch_tile_kernel(
    const BoxDesc<T>* __restrict__ D, BatchList L, int parity) {
  PF_BATCH_BOX;
  const int old = parity, nw = 1 - parity, cur = 1 - parity;
  const T* a_old = MODE == kModeA ? d.pk[old] : (MODE == kModeC ? d.pk[cur] : (const T*)nullptr);
  const T* v_in = MODE == kModeA ? d.vk[old] : d.vk[cur];
  T* f_out = MODE == kModeA ? d.pk[nw] : (MODE == kModeC ? d.x : (T*)nullptr);
  T* af_out = MODE == kModeA ? d.vk[nw] : (MODE == kModeC ? d.rk[1 - parity] : (T*)nullptr);
  const double* pin = MODE == kModeA ? d.pC : (MODE == kModeB ? d.pA : d.pB);
  double* pout = MODE == kModeA ? d.pA : (MODE == kModeB ? d.pB : d.pC);
  const T* r = d.rk[parity];
  // served boxes are small (short x chunks): the one-plane-ahead schedule.
  // NTS: non-temporal output stores, for launches whose boxes together
  // outgrow the Infinity Cache (all resident boxes at once); launches over a
  // cache-residency group keep their outputs cacheable for the group's next
  // kernel.
  if (PAIR)
    bicg_tile2_body<T, MODE, TZ, TY, true, NTS>(d.g, d.tg, d.tg.nblk, d.ctl, pin, pout, r, d.rh,
                                                a_old, v_in, f_out, af_out, d.L, blk);
  else
    bicg_tile_kernel_body<T, MODE, TZ, TY, true, NTS>(d.g, d.tg, d.tg.nblk, d.ctl, pin, pout, r,
                                                      d.rh, a_old, v_in, f_out, af_out, d.L, blk);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_scatter_kernel(const BoxDesc<T>* __restrict__ D,
                                                               BatchList L) {
  PF_BATCH_BOX;
  scatter_padded_kernel_body<T, true>(d.g, d.x, d.DP, d.p, blk);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_project_kernel(const BoxDesc<T>* __restrict__ D,
                                                               BatchList L) {
  PF_BATCH_BOX;
  project_kernel_body<T>(d.g, d.x, d.u[0], d.u[1], d.u[2], d.dxi, d.dyi, d.dzi, d.neg_c, blk);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) batch_stats_kernel(const BoxDesc<T>* __restrict__ D,
                                                             BatchList L) {
  PF_BATCH_BOX;
  const int64_t o = d.g.poff;
  stats_kernel_body<T>(d.g, d.u[0] + o, d.u[1] + o, d.u[2] + o, d.dxi, d.dyi, d.dzi, (T*)nullptr,
                       d.pS, blk);
}
#undef PF_BATCH_BOX

// Per-box control over the boxes ids[0..n) (all boxes 0..n when ids is
// NULL): kBatchCtlInit resets their control blocks for a new solve (one
// thread per box); kBatchCtlFlags writes each box's done flag to flags[i]
// (one thread per box); kBatchFinalize writes the monitor values and the
// solver's outcome into d.out (one block per box).
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_ctl_kernel(const BoxDesc<T>* __restrict__ D,
                                                           const int* __restrict__ ids, int n,
                                                           int what, int* __restrict__ flags) {
  if (what == kBatchFinalize) {
    const int b = ids ? ids[blockIdx.x] : (int)blockIdx.x;
    const BoxDesc<T>& d = D[b];
    const int ops[5] = {1, 1, 1, 0, 1};
    double v[5];
    reduce_partials<5>(d.pS, d.nblk, v, ops);
    if (threadIdx.x == 0) {
      for (int q = 0; q < 5; ++q) d.out[q] = v[q];
      d.out[5] = d.ctl->max_res;
      d.out[6] = (double)d.ctl->iter;
      d.out[7] = (double)d.ctl->status;
    }
    return;
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = ids ? ids[i] : i;
  KrylovCtl* c = D[b].ctl;
  if (what == kBatchCtlInit) {
    c->rho1 = 1.0; c->rho2 = 1.0; c->alpha = 0.0; c->omega = 1.0;
    c->res0 = 0.0; c->max_res = 0.0; c->tol = D[b].tol;
    c->iter = 0; c->maxit = D[b].maxit; c->done = 0; c->status = PF_SOLVE_CONVERGED;
    c->halt = 0;
    c->npart = D[b].nblk;
    c->hist = nullptr; c->hist_cap = 0;
  } else {
    flags[i] = c->done;
  }
}

int batch_krylov_launch(int dtype, int what, int parity, const void* desc, const BatchList& L,
                        int nblocks, int tz, int ty, hipStream_t s, bool nts) {
  if (nblocks <= 0) return 0;
  auto go = [&](auto zero) -> int {
    using T = decltype(zero);
    const BoxDesc<T>* D = (const BoxDesc<T>*)desc;
    const dim3 gb(nblocks), tb(kBlock);
    switch (what) {
      case kBatchInit: hipLaunchKernelGGL(batch_init_kernel<T>, gb, tb, 0, s, D, L); break;
      case kBatchScatter: hipLaunchKernelGGL(batch_scatter_kernel<T>, gb, tb, 0, s, D, L); break;
      case kBatchProject: hipLaunchKernelGGL(batch_project_kernel<T>, gb, tb, 0, s, D, L); break;
      case kBatchStats: hipLaunchKernelGGL(batch_stats_kernel<T>, gb, tb, 0, s, D, L); break;
      case kBatchKA:
      case kBatchKB:
      case kBatchKC: {
#define PF_BTILE_N(TZ, TY, PAIR, NT)                                                           \
  if (what == kBatchKA)                                                                          \
    hipLaunchKernelGGL((batch_tile_kernel<T, kModeA, TZ, TY, PAIR, NT>), gb,                     \
                       dim3(PAIR ? TZ / 2 * TY : TZ * TY), 0, s, D, L, parity);                  \
  else if (what == kBatchKB)                                                                     \
    hipLaunchKernelGGL((batch_tile_kernel<T, kModeB, TZ, TY, PAIR, NT>), gb,                     \
                       dim3(PAIR ? TZ / 2 * TY : TZ * TY), 0, s, D, L, parity);                  \
  else                                                                                           \
    hipLaunchKernelGGL((batch_tile_kernel<T, kModeC, TZ, TY, PAIR, NT>), gb,                     \
                       dim3(PAIR ? TZ / 2 * TY : TZ * TY), 0, s, D, L, parity)
#define PF_BTILE(TZ, TY, PAIR)                                                                 \
  do {                                                                                           \
    if (nts) { PF_BTILE_N(TZ, TY, PAIR, true); } else { PF_BTILE_N(TZ, TY, PAIR, false); }       \
  } while (0)
        switch (tz * 1000 + ty) {
          case 64004: PF_BTILE(64, 4, false); break;
          case 32008: PF_BTILE(32, 8, false); break;
          case 16016: PF_BTILE(16, 16, false); break;
          case 8032: PF_BTILE(8, 32, false); break;
          case 64008: PF_BTILE(64, 8, true); break;
          case 32016: PF_BTILE(32, 16, true); break;
          case 16032: PF_BTILE(16, 32, true); break;
          default: return fail(PF_EINVAL, "batch_krylov_launch: unsupported tile shape");
        }
#undef PF_BTILE
#undef PF_BTILE_N
        break;
      }
      default: return fail(PF_EINVAL, "batch_krylov_launch: unknown kernel");
    }
    PF_LAUNCH_CHECK("batch Krylov kernel");
    return 0;
  };
  return dtype == PF_F64 ? go(0.0) : go(0.0f);
}

int batch_ctl_launch(int dtype, int what, const void* desc, const int* ids, int n, int* flags,
                     hipStream_t s) {
  if (n <= 0) return 0;
  auto go = [&](auto zero) -> int {
    using T = decltype(zero);
    const BoxDesc<T>* D = (const BoxDesc<T>*)desc;
    const int nb = what == kBatchFinalize ? n : (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(batch_ctl_kernel<T>, dim3(nb), dim3(kBlock), 0, s, D, ids, n, what, flags);
    PF_LAUNCH_CHECK("batch_ctl_kernel");
    return 0;
  };
  return dtype == PF_F64 ? go(0.0) : go(0.0f);
}

// ---------------------------------------------------------------------------
// Host drivers
template <typename T>
int finalize(pf_box* h, hipStream_t s, const double* part, int nv, const int* ops, double* out) {
  Ops8 o;
  for (int q = 0; q < 8; ++q) o.v[q] = q < nv ? ops[q] : 0;
  switch (nv) {
    case 1: hipLaunchKernelGGL(finalize_kernel<1>, dim3(1), dim3(kBlock), 0, s, part, h->nblk, o, h->stats_dev); break;
    case 5: hipLaunchKernelGGL(finalize_kernel<5>, dim3(1), dim3(kBlock), 0, s, part, h->nblk, o, h->stats_dev); break;
    default: return fail(PF_EINVAL, "finalize: unsupported width");
  }
  PF_LAUNCH_CHECK("finalize_kernel");
  PF_HIP(hipMemcpyAsync(h->stats_host, h->stats_dev, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  if (int rc = wait_stream(s, h->wait_ev)) return rc;
  for (int q = 0; q < nv; ++q) out[q] = h->stats_host[q];
  return 0;
}

template <typename T>
int bicgstab_impl(pf_box* h, hipStream_t s, T* src, const T* U, const T* V, const T* W,
                  double src_scale, double tol, int maxit, double* out_res, int* out_iters,
                  bool vol_scale = true, double* hist = nullptr, int hist_cap = -1) {
  const Geom& g = h->g;
  const int nb = h->nblk;
  const double vol = h->dx * h->dy * h->dz;
  const Lap<T> L = make_lap<T>(h);
  T* x = (T*)h->x;
  T* Rr[2] = {(T*)h->r, (T*)h->s};   // r of even / odd iterations
  T* rh = (T*)h->rh;
  T* P[2] = {(T*)h->p[0], (T*)h->p[1]};
  T* Vv[2] = {(T*)h->v[0], (T*)h->v[1]};
  double* pA = part_buf(h, kPartA);
  double* pB = part_buf(h, kPartB);
  double* pC = part_buf(h, kPartC);

  if (hist_cap < 0) {
    hist = h->hist;
    hist_cap = h->hist_cap;
  }
  hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(1), 0, s, h->ctl, tol, maxit, nb, hist,
                     hist_cap);
  PF_LAUNCH_CHECK("ctl_init_kernel");
  const T dxi = (T)(1.0 / h->dx), dyi = (T)(1.0 / h->dy), dzi = (T)(1.0 / h->dz);
  const T sc1 = (T)src_scale, sc2 = vol_scale ? (T)(-1.0 * vol) : (T)1;
  if (U)
    hipLaunchKernelGGL((bicg_init_kernel<T, true>), dim3(nb), dim3(kBlock), 0, s, g, src,
                       U + g.poff, V + g.poff, W + g.poff, dxi, dyi, dzi, sc1, sc2, Rr[0], rh, x,
                       P[0], Vv[0], pC);
  else
    hipLaunchKernelGGL((bicg_init_kernel<T, false>), dim3(nb), dim3(kBlock), 0, s, g, src,
                       (const T*)nullptr, (const T*)nullptr, (const T*)nullptr, dxi, dyi, dzi,
                       sc1, sc2, Rr[0], rh, x, P[0], Vv[0], pC);
  PF_LAUNCH_CHECK("bicg_init_kernel");

  // KA of iteration 0, then batches of (KB, KC, KA) until the device says done.
  // With profiling on, each launch carries a start/stop event pair recorded
  // by the dispatch itself (hipExtLaunchKernelGGL); pairs are read back after
  // the batch's host sync and only launches that did work (iteration index
  // below the completed count) are accumulated.
  struct Rec {
    int kind, iter, slot;
  };
  Rec recs[3 * 1024 + 1];
  int nrec = 0;
  int slot[3] = {0, 0, 0};
  auto launch = [&](int kind, int iter, auto&& go) {
    // every prof_on-th launch of each kind carries the event pair
    const bool prof = h->prof_on && (h->prof_seq[kind]++ % h->prof_on) == 0 &&
                      slot[kind] < h->prof_cap;
    hipEvent_t* e = prof ? h->prof_ev + ((size_t)kind * h->prof_cap + slot[kind]) * 2 : nullptr;
    go(prof ? e[0] : nullptr, prof ? e[1] : nullptr);
    if (prof) recs[nrec++] = Rec{kind, iter, slot[kind]++};
  };
  int it = 0;  // iterations whose KA has been launched
  const TileGrid tg = make_tile_grid(g, (int)h->esize, h->tile_rows);
  const int nbA = tg.nblk;   // blocks of KA, KB and KC
  // KA of iteration j reads r[j & 1] (its partial count comes from the
  // control block: the init kernel's for j = 0, KC's after)
  auto ka = [&](int j, int old, int nw, hipEvent_t e0, hipEvent_t e1) {
    launch_tile<T, kModeA>(tg, s, e0, e1, g, nbA, h->ctl, pC, pA, Rr[j & 1], rh, P[old], Vv[old],
                           P[nw], Vv[nw], L);
  };
  {
    // the whole solve as one cooperative launch when the box's vectors fit
    // on chip (pf_persist.hip): same iterates, same bits
    bool used = false;
    hipEvent_t* pe = h->prof_on ? h->prof_pev : nullptr;
    if (int rc = persist_solve<T>(h, s, tg, L, Rr[0], rh, x, nb, &used, pe)) return rc;
    h->last_persist = used ? 1 : 0;
    if (used) {
      PF_HIP(hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
      bool aborted = false;
      if (int rc = persist_aborted(h, s, &aborted)) return rc;   // synchronises s
      if (aborted)
        return fail(PF_EHIP, "bicgstab: a block of the persistent solve stopped waiting for the "
                             "others (grid sync timed out)");
      const KrylovCtl c = *h->ctl_host;
      if (pe) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe[0], pe[1]) == hipSuccess) {
          h->prof_pms += ms;
          h->prof_pn += 1;
          h->prof_pit += c.iter;
        }
      }
      h->last_iters = c.iter;
      if (out_res) *out_res = c.max_res;
      if (out_iters) *out_iters = c.iter;
      return c.status;
    }
  }
  h->last_persist = 0;
  launch(0, 0, [&](hipEvent_t e0, hipEvent_t e1) { ka(0, 0, 1, e0, e1); });
  PF_LAUNCH_CHECK("bicg_tile_kernel");
  const int cap = maxit < 0 ? 0 : maxit;
  bool first = true;
  for (;;) {
    int m = h->chunk;
    if (first && h->last_iters > m) m = h->last_iters - 1;
    if (h->prof_on && m > h->prof_cap - 1) m = h->prof_cap - 1;
    if (m > 1024) m = 1024;
    first = false;
    for (int q = 0; q < m; ++q) {
      const int cur = (it + 1) & 1;   // p, v written by iteration `it`
      const int rc = it & 1;          // r read by iteration `it`
      launch(1, it, [&](hipEvent_t e0, hipEvent_t e1) {
        launch_tile<T, kModeB>(tg, s, e0, e1, g, nbA, h->ctl, pA, pB, Rr[rc], rh,
                               (const T*)nullptr, Vv[cur], (T*)nullptr, (T*)nullptr, L);
      });
      launch(2, it, [&](hipEvent_t e0, hipEvent_t e1) {
        launch_tile<T, kModeC>(tg, s, e0, e1, g, nbA, h->ctl, pB, pC, Rr[rc], rh, P[cur],
                               Vv[cur], x, Rr[1 - rc], L);
      });
      ++it;
      const int nxt = (it + 1) & 1;
      launch(0, it, [&](hipEvent_t e0, hipEvent_t e1) { ka(it, cur, nxt, e0, e1); });
    }
    PF_LAUNCH_CHECK("bicgstab batch");
    PF_HIP(hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
    if (int rc = wait_stream(s, h->wait_ev)) return rc;
    const KrylovCtl cb = *h->ctl_host;
    for (int q = 0; q < nrec; ++q) {
      const Rec& rr = recs[q];
      const bool worked = rr.kind == 0 ? (rr.iter < cb.iter || !cb.done) : rr.iter < cb.iter;
      if (!worked) continue;
      float ms = 0.f;
      hipEvent_t* e = h->prof_ev + ((size_t)rr.kind * h->prof_cap + rr.slot) * 2;
      if (hipEventElapsedTime(&ms, e[0], e[1]) == hipSuccess) {
        h->prof_ms[rr.kind] += ms;
        h->prof_n[rr.kind] += 1;
      }
    }
    nrec = 0;
    slot[0] = slot[1] = slot[2] = 0;
    if (cb.done) break;
    if (it > cap + 2)
      return fail(PF_EHIP, "bicgstab: device control block did not finish within the iteration cap");
  }
  const KrylovCtl c = *h->ctl_host;
  h->last_iters = c.iter;
  if (out_res) *out_res = c.max_res;
  if (out_iters) *out_iters = c.iter;
  return c.status;
}

template <typename T>
int scatter_dp(pf_box* h, hipStream_t s, T* DP, T* p) {
  const Geom& g = h->g;
  const int64_t total = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nb = (total + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  if (p)
    hipLaunchKernelGGL((scatter_padded_kernel<T, true>), dim3((int)nb), dim3(kBlock), 0, s, g,
                       (const T*)h->x, DP, p);
  else
    hipLaunchKernelGGL((scatter_padded_kernel<T, false>), dim3((int)nb), dim3(kBlock), 0, s, g,
                       (const T*)h->x, DP, p);
  PF_LAUNCH_CHECK("scatter_padded_kernel");
  return 0;
}

template <typename T>
int jacobi_impl(pf_box* h, hipStream_t s, T* src, const T* U, const T* V, const T* W, T* pOld,
                T* p, double src_scale, int nit, double* out_res) {
  const Geom& g = h->g;
  const int nb = h->nblk;
  const int nbf = launch_blocks(g);
  const T dxi = (T)(1.0 / h->dx), dyi = (T)(1.0 / h->dy), dzi = (T)(1.0 / h->dz);
  const T dx2 = (T)(h->dx * h->dx);
  double* pS = part_buf(h, kPartS);
  const int64_t np = (int64_t)(g.nx + 4) * (g.ny + 4) * (g.nz + 4);
  int64_t nbz = (np + kBlock - 1) / kBlock;
  if (nbz > 65536) nbz = 65536;
  // p = 0 (pressure.py:121); the sweep reading iterate 0 reads zeros.
  T* bufs[2] = {p, pOld};
  // iterate j lives in bufs[(nit - j) & 1] so that p = iterate nit, pOld = iterate nit-1
  T* it0 = bufs[nit & 1];
  hipLaunchKernelGGL(zero_padded_kernel<T>, dim3((int)nbz), dim3(kBlock), 0, s, np, p);
  if (it0 != p) hipLaunchKernelGGL(zero_padded_kernel<T>, dim3((int)nbz), dim3(kBlock), 0, s, np, it0);
  hipLaunchKernelGGL(jacobi_init_kernel<T>, dim3(nb), dim3(kBlock), 0, s, g, src,
                     U ? U + g.poff : nullptr, V ? V + g.poff : nullptr, W ? W + g.poff : nullptr,
                     U ? 1 : 0, dxi, dyi, dzi, (T)src_scale, dx2, pS);
  PF_LAUNCH_CHECK("jacobi_init_kernel");
  double res0;
  const int opm[1] = {1};
  if (int rc = finalize<T>(h, s, pS, 1, opm, &res0)) return rc;
  const T dinv = (T)(-1.0 / 6.0);
  for (int j = 1; j <= nit; ++j) {
    const T* in = bufs[(nit - j + 1) & 1];
    T* out = bufs[(nit - j) & 1];
    hipLaunchKernelGGL(jacobi_sweep_kernel<T>, dim3(nbf), dim3(kBlock), 0, s, g, in, src, out, dinv);
  }
  PF_LAUNCH_CHECK("jacobi_sweep_kernel");
  const Lap<T> L = make_lap<T>(h);
  hipLaunchKernelGGL(jacobi_resid_kernel<T>, dim3(nb), dim3(kBlock), 0, s, g, (const T*)p, src,
                     dx2, L, pS);
  PF_LAUNCH_CHECK("jacobi_resid_kernel");
  double resP;
  if (int rc = finalize<T>(h, s, pS, 1, opm, &resP)) return rc;
  if (out_res) *out_res = resP / res0;
  return 0;
}

template <typename T>
int stats_impl(pf_box* h, hipStream_t s, const T* U, const T* V, const T* W, T* divg,
               double* out5) {
  const Geom& g = h->g;
  double* pS = part_buf(h, kPartS);
  hipLaunchKernelGGL(stats_kernel<T>, dim3(h->nblk), dim3(kBlock), 0, s, g, U + g.poff,
                     V + g.poff, W + g.poff, (T)(1.0 / h->dx), (T)(1.0 / h->dy), (T)(1.0 / h->dz),
                     divg, pS);
  PF_LAUNCH_CHECK("stats_kernel");
  const int ops[5] = {1, 1, 1, 0, 1};
  double v[5];
  if (int rc = finalize<T>(h, s, pS, 5, ops, v)) return rc;
  const double npts = (double)g.ncell;
  out5[0] = v[0]; out5[1] = v[1]; out5[2] = v[2];
  out5[3] = v[3] / npts * 0.5;
  out5[4] = v[4];
  return 0;
}

template <typename T>
int div_impl(pf_box* h, hipStream_t s, const T* U, const T* V, const T* W, T* divg) {
  double out5[5];
  // the stats kernel computes the divergence on the way; its reductions are cheap
  return stats_impl<T>(h, s, U, V, W, divg, out5);
}

// ---------------------------------------------------------------------------
// Mixed-precision iterative refinement (PF_POISSON_BICGSTAB_MPIR).  The
// reference's BiCGStab keeps its shadow residual fixed for the whole solve;
// on long solves <r, r_hat> decays to the rounding floor of the inner product
// and the recurrence breaks down (FP64 at 512^3 after ~900 iterations, FP32
// already near 1e-3).  Refinement restarts it: the iterate x and the
// right-hand side b are kept in FP64, each outer step solves A d = r for the
// FP64 residual r = b - A x (rounded to the box precision) with the box's
// BiCGStab to a loose tolerance, and adds d to x in FP64.  The stop test is
// max|b - A x| / max|b| <= tol on the FP64 residual.
template <typename T, bool FROM_DIV>
__global__ void __launch_bounds__(kBlock) mp_init_kernel(
    Geom g, const T* __restrict__ src, const T* __restrict__ U, const T* __restrict__ V,
    const T* __restrict__ W, double dxi, double dyi, double dzi, double sc1, double sc2,
    double* __restrict__ b, double* __restrict__ x, T* __restrict__ r_out,
    double* __restrict__ part) {
  double m = 0.0;
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    double v;
    if (FROM_DIV) {
      int i, j, k;
      decode(g, idx, i, j, k);
      const int64_t c = i * g.psx + j * g.psy + k;
      v = ((double)U[c + g.psx] - (double)U[c]) * dxi;
      v = v + ((double)V[c + g.psy] - (double)V[c]) * dyi;
      v = v + ((double)W[c + 1] - (double)W[c]) * dzi;
    } else {
      v = (double)src[idx];
    }
    v = v * sc1;
    v = v * sc2;
    b[idx] = v;
    x[idx] = 0.0;
    r_out[idx] = (T)v;
    m = nmax(m, fabs(v));
  }
  double a[1] = {m};
  const int ops[1] = {1};
  block_reduce<1>(a, ops);
  write_partials<1>(part, a);
}

// x += d, then (second kernel) r = b - A x with max|r|: two passes, so that
// every neighbour of the stencil reads the updated x.
template <typename T>
__global__ void __launch_bounds__(kBlock) mp_update_kernel(Geom g, double* __restrict__ x,
                                                           const T* __restrict__ d) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x)
    x[idx] = x[idx] + (double)d[idx];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) mp_residual_kernel(
    Geom g, const double* __restrict__ x, const double* __restrict__ b, Lap<double> L,
    T* __restrict__ r_out, double* __restrict__ part) {
  double m = 0.0;
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(g, idx, i, j, k);
    const int64_t ci = (int64_t)i * g.isx, cj = (int64_t)j * g.isy;
    const int64_t im = (int64_t)(i == 0 ? g.nx - 1 : i - 1) * g.isx;
    const int64_t ip = (int64_t)(i == g.nx - 1 ? 0 : i + 1) * g.isx;
    const int64_t jm = (int64_t)(j == 0 ? g.ny - 1 : j - 1) * g.isy;
    const int64_t jp = (int64_t)(j == g.ny - 1 ? 0 : j + 1) * g.isy;
    const int km = k == 0 ? g.nz - 1 : k - 1, kp = k == g.nz - 1 ? 0 : k + 1;
    const double ax = lap7(L, x[im + cj + k], x[ci + cj + k], x[ip + cj + k], x[ci + jm + k],
                           x[ci + jp + k], x[ci + cj + km], x[ci + cj + kp]);
    const double r = b[idx] - ax;
    r_out[idx] = (T)r;
    m = nmax(m, fabs(r));
  }
  double a[1] = {m};
  const int ops[1] = {1};
  block_reduce<1>(a, ops);
  write_partials<1>(part, a);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) mp_finish_kernel(Geom g, const double* __restrict__ x,
                                                           T* __restrict__ out) {
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < (uint32_t)g.ncell;
       idx += gridDim.x * blockDim.x)
    out[idx] = (T)x[idx];
}

// FP64 iterate, right-hand side and inner residual history of the
// refinement solver, reserved by pf_mpir_configure (a set-up call) so that
// the solve itself allocates nothing.
int mp_reserve(pf_box* h) {
  if (h->mp_hist_cap < h->mp_inner_maxit) {
    if (h->mp_hist) {
      PF_HIP(hipFree(h->mp_hist));
      h->bytes -= (int64_t)sizeof(double) * h->mp_hist_cap;
      h->mp_hist = nullptr;
      h->mp_hist_cap = 0;
    }
    hipError_t e = hipMalloc((void**)&h->mp_hist, sizeof(double) * h->mp_inner_maxit);
    if (e != hipSuccess) {
      h->mp_hist = nullptr;
      return check_hip(e, "hipMalloc");
    }
    h->mp_hist_cap = h->mp_inner_maxit;
    h->bytes += (int64_t)sizeof(double) * h->mp_hist_cap;
  }
  if (h->mp_x) return 0;
  const size_t n = (size_t)h->g.ncell * sizeof(double);
  hipError_t e = hipMalloc((void**)&h->mp_x, n);
  if (e != hipSuccess) {
    h->mp_x = nullptr;
    return check_hip(e, "hipMalloc");
  }
  e = hipMalloc((void**)&h->mp_b, n);
  if (e != hipSuccess) {
    hipFree(h->mp_x);
    h->mp_x = nullptr;
    h->mp_b = nullptr;
    return check_hip(e, "hipMalloc");
  }
  h->bytes += (int64_t)(2 * n);
  return 0;
}

// Leaves the solution, rounded to T, in h->x (as bicgstab_impl does) and the
// last inner right-hand side in src.  *out_iters counts inner iterations.
template <typename T>
int mp_impl(pf_box* h, hipStream_t s, T* src, const T* U, const T* V, const T* W,
            double src_scale, double tol, int maxit, double* out_res, int* out_iters) {
  if (!h->mp_x || !h->mp_b || h->mp_hist_cap < h->mp_inner_maxit)
    return fail(PF_EINVAL, "refinement solve: workspace not reserved (call pf_mpir_configure "
                           "on the box first)");
  const Geom& g = h->g;
  const int nb = h->nblk;
  const double vol = h->dx * h->dy * h->dz;
  double* pS = part_buf(h, kPartS);
  const double dxi = 1.0 / h->dx, dyi = 1.0 / h->dy, dzi = 1.0 / h->dz;
  if (U)
    hipLaunchKernelGGL((mp_init_kernel<T, true>), dim3(nb), dim3(kBlock), 0, s, g,
                       (const T*)nullptr, U + g.poff, V + g.poff, W + g.poff, dxi, dyi, dzi,
                       src_scale, -1.0 * vol, h->mp_b, h->mp_x, src, pS);
  else
    hipLaunchKernelGGL((mp_init_kernel<T, false>), dim3(nb), dim3(kBlock), 0, s, g,
                       (const T*)src, (const T*)nullptr, (const T*)nullptr, (const T*)nullptr,
                       dxi, dyi, dzi, src_scale, -1.0 * vol, h->mp_b, h->mp_x, src, pS);
  PF_LAUNCH_CHECK("mp_init_kernel");
  const int ops[1] = {1};
  double bmax = 0.0;
  if (int rc = finalize<T>(h, s, pS, 1, ops, &bmax)) return rc;
  int status = PF_SOLVE_MAXIT;
  double res = 1.0;
  int total = 0;
  if (!(bmax > 0.0)) {   // b == 0: x = 0 is exact
    status = PF_SOLVE_CONVERGED;
    res = 0.0;
  }
  const Lap<double> L = make_lap<double>(h);
  for (int outer = 0; status != PF_SOLVE_CONVERGED && outer < h->mp_outer_maxit && total < maxit;
       ++outer) {
    const int cap = std::min(h->mp_inner_maxit, maxit - total);
    double rin = 0.0;
    int iin = 0;
    int st = bicgstab_impl<T>(h, s, src, nullptr, nullptr, nullptr, 1.0, h->mp_inner_tol, cap,
                              &rin, &iin, false, h->mp_hist, h->mp_hist_cap);
    if (st < 0) return st;
    total += iin;
    if (iin == 0) {
      status = PF_SOLVE_STALLED;
      break;
    }
    if (st != PF_SOLVE_CONVERGED) {
      // The inner recurrence degraded before reaching its tolerance: re-run
      // it (deterministically, same bits) up to its best iterate.
      std::vector<double> hv(iin);
      PF_HIP(hipMemcpyAsync(hv.data(), h->mp_hist, sizeof(double) * iin, hipMemcpyDeviceToHost, s));
      PF_HIP(hipStreamSynchronize(s));
      int kbest = -1;
      for (int k = 0; k < iin; ++k)
        if (hv[k] < 1.0 && (kbest < 0 || hv[k] < hv[kbest])) kbest = k;
      if (kbest < 0) {
        status = PF_SOLVE_STALLED;
        break;
      }
      if (kbest + 1 < iin && total + kbest + 1 <= maxit) {
        st = bicgstab_impl<T>(h, s, src, nullptr, nullptr, nullptr, 1.0, h->mp_inner_tol,
                              kbest + 1, &rin, &iin, false, h->mp_hist, h->mp_hist_cap);
        if (st < 0) return st;
        total += iin;
      } else if (!(rin < 1.0)) {
        status = PF_SOLVE_STALLED;
        break;
      }
    }
    hipLaunchKernelGGL(mp_update_kernel<T>, dim3(nb), dim3(kBlock), 0, s, g, h->mp_x,
                       (const T*)h->x);
    PF_LAUNCH_CHECK("mp_update_kernel");
    hipLaunchKernelGGL(mp_residual_kernel<T>, dim3(nb), dim3(kBlock), 0, s, g,
                       (const double*)h->mp_x, (const double*)h->mp_b, L, src, pS);
    PF_LAUNCH_CHECK("mp_residual_kernel");
    double rmax = 0.0;
    if (int rc = finalize<T>(h, s, pS, 1, ops, &rmax)) return rc;
    res = rmax / bmax;
    if (res <= tol) status = PF_SOLVE_CONVERGED;
  }
  hipLaunchKernelGGL(mp_finish_kernel<T>, dim3(nb), dim3(kBlock), 0, s, g,
                     (const double*)h->mp_x, (T*)h->x);
  PF_LAUNCH_CHECK("mp_finish_kernel");
  if (out_res) *out_res = res;
  if (out_iters) *out_iters = total;
  return status;
}

template <typename T>
int vel_corr_impl(pf_box* h, hipStream_t s, T* U, T* V, T* W, T* DP, T* p, T* src, int solver,
                  double src_scale, double corr_scale, double tol, int iters, double* out_res,
                  int* out_iters) {
  const Geom& g = h->g;
  const int nbf = launch_blocks(g);
  const T dxi = (T)(1.0 / h->dx), dyi = (T)(1.0 / h->dy), dzi = (T)(1.0 / h->dz);
  const T neg_c = (T)(-corr_scale);
  if (solver == PF_POISSON_BICGSTAB || solver == PF_POISSON_BICGSTAB_MPIR) {
    const int st = solver == PF_POISSON_BICGSTAB
                       ? bicgstab_impl<T>(h, s, src, U, V, W, src_scale, tol, iters, out_res,
                                          out_iters)
                       : mp_impl<T>(h, s, src, U, V, W, src_scale, tol, iters, out_res, out_iters);
    if (st < 0) return st;
    if (int rc = scatter_dp<T>(h, s, DP, p)) return rc;
    hipLaunchKernelGGL(project_kernel<T>, dim3(nbf), dim3(kBlock), 0, s, g, (const T*)h->x, U, V,
                       W, dxi, dyi, dzi, neg_c);
    PF_LAUNCH_CHECK("project_kernel");
    return st;
  }
  int rc = jacobi_impl<T>(h, s, src, U, V, W, DP, p, src_scale, iters, out_res);
  if (rc) return rc;
  if (out_iters) *out_iters = iters;
  hipLaunchKernelGGL(project_padded_kernel<T>, dim3(nbf), dim3(kBlock), 0, s, g, (const T*)DP, U,
                     V, W, dxi, dyi, dzi, neg_c);
  PF_LAUNCH_CHECK("project_padded_kernel");
  return 0;
}

}  // namespace pf

using namespace pf;

extern "C" {

int pf_bicgstab_solve(pf_box* h, void* DP, void* p, void* source, double rho, double dt,
                      double tol, int maxit, double* out_res, int* out_iters, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!DP || !p || !source) return fail(PF_EINVAL, "pf_bicgstab_solve: NULL field pointer");
  if (!(dt != 0.0)) return fail(PF_EINVAL, "pf_bicgstab_solve: dt must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  int st;
  if (h->dtype == PF_F64) {
    st = bicgstab_impl<double>(h, s, (double*)source, nullptr, nullptr, nullptr, rho / dt, tol,
                               maxit, out_res, out_iters);
    if (st >= 0) {
      int rc = scatter_dp<double>(h, s, (double*)DP, (double*)p);
      if (rc) return rc;
    }
  } else {
    st = bicgstab_impl<float>(h, s, (float*)source, nullptr, nullptr, nullptr, rho / dt, tol,
                              maxit, out_res, out_iters);
    if (st >= 0) {
      int rc = scatter_dp<float>(h, s, (float*)DP, (float*)p);
      if (rc) return rc;
    }
  }
  return st;
}

int pf_bicgstab_mpir_solve(pf_box* h, void* DP, void* p, void* source, double rho, double dt,
                           double tol, int maxit, double* out_res, int* out_iters, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!DP || !p || !source) return fail(PF_EINVAL, "pf_bicgstab_mpir_solve: NULL field pointer");
  if (!(dt != 0.0)) return fail(PF_EINVAL, "pf_bicgstab_mpir_solve: dt must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  int st;
  if (h->dtype == PF_F64) {
    st = mp_impl<double>(h, s, (double*)source, nullptr, nullptr, nullptr, rho / dt, tol, maxit,
                         out_res, out_iters);
    if (st >= 0)
      if (int rc = scatter_dp<double>(h, s, (double*)DP, (double*)p)) return rc;
  } else {
    st = mp_impl<float>(h, s, (float*)source, nullptr, nullptr, nullptr, rho / dt, tol, maxit,
                        out_res, out_iters);
    if (st >= 0)
      if (int rc = scatter_dp<float>(h, s, (float*)DP, (float*)p)) return rc;
  }
  return st;
}

int pf_mpir_configure(pf_box* h, double inner_tol, int inner_maxit, int outer_maxit) {
  if (int rc = check_box(h)) return rc;
  if (!(inner_tol > 0.0 && inner_tol < 1.0) || inner_maxit < 1 || outer_maxit < 1)
    return fail(PF_EINVAL, "pf_mpir_configure: need 0 < inner_tol < 1, inner_maxit >= 1, "
                           "outer_maxit >= 1");
  h->mp_inner_tol = inner_tol;
  h->mp_inner_maxit = inner_maxit;
  h->mp_outer_maxit = outer_maxit;
  return mp_reserve(h);
}

int pf_mpir_solution(pf_box* h, double* x, double* b, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!h->mp_x || !h->mp_b)
    return fail(PF_EINVAL, "pf_mpir_solution: no refinement workspace (call pf_mpir_configure "
                           "and solve first)");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)h->g.ncell * sizeof(double);
  if (x) PF_HIP(hipMemcpyAsync(x, h->mp_x, n, hipMemcpyDeviceToDevice, s));
  if (b) PF_HIP(hipMemcpyAsync(b, h->mp_b, n, hipMemcpyDeviceToDevice, s));
  return 0;
}

int pf_jacobi_solve(pf_box* h, void* pOld, void* p, void* source, double rho, double dt, int nit,
                    double* out_res, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!pOld || !p || !source) return fail(PF_EINVAL, "pf_jacobi_solve: NULL field pointer");
  if (nit < 1) return fail(PF_EINVAL, "pf_jacobi_solve: nit must be >= 1");
  if (!(dt != 0.0)) return fail(PF_EINVAL, "pf_jacobi_solve: dt must be nonzero");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return jacobi_impl<double>(h, s, (double*)source, nullptr, nullptr, nullptr, (double*)pOld,
                               (double*)p, rho / dt, nit, out_res);
  return jacobi_impl<float>(h, s, (float*)source, nullptr, nullptr, nullptr, (float*)pOld,
                            (float*)p, rho / dt, nit, out_res);
}

int pf_vel_corr(pf_box* h, void* u, void* v, void* w, void* DP, void* p, void* source, int solver,
                double src_scale, double corr_scale, double tol, int iters, double* out_res,
                int* out_iters, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !DP || !p || !source)
    return fail(PF_EINVAL, "pf_vel_corr: NULL field pointer");
  if (solver != PF_POISSON_BICGSTAB && solver != PF_POISSON_JACOBI &&
      solver != PF_POISSON_BICGSTAB_MPIR)
    return fail(PF_EINVAL, "pf_vel_corr: unknown solver");
  if (solver == PF_POISSON_JACOBI && iters < 1)
    return fail(PF_EINVAL, "pf_vel_corr: Jacobi needs at least one sweep");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return vel_corr_impl<double>(h, s, (double*)u, (double*)v, (double*)w, (double*)DP,
                                 (double*)p, (double*)source, solver, src_scale, corr_scale, tol,
                                 iters, out_res, out_iters);
  return vel_corr_impl<float>(h, s, (float*)u, (float*)v, (float*)w, (float*)DP, (float*)p,
                              (float*)source, solver, src_scale, corr_scale, tol, iters, out_res,
                              out_iters);
}

int pf_profile_enable(pf_box* h, int on) {
  if (int rc = check_box(h)) return rc;
  if (on && !h->prof_ev) {
    const int cap = 1024;
    h->prof_ev = new hipEvent_t[(size_t)3 * cap * 2];
    for (int q = 0; q < 3 * cap * 2; ++q) PF_HIP(hipEventCreate(&h->prof_ev[q]));
    h->prof_cap = cap;
    PF_HIP(hipEventCreate(&h->prof_pev[0]));
    PF_HIP(hipEventCreate(&h->prof_pev[1]));
  }
  h->prof_on = on > 0 ? on : 0;
  for (int k = 0; k < 3; ++k) h->prof_seq[k] = 0;
  for (int k = 0; k < 3; ++k) {
    h->prof_ms[k] = 0.0;
    h->prof_n[k] = 0;
  }
  h->prof_pms = 0.0;
  h->prof_pn = 0;
  h->prof_pit = 0;
  return 0;
}
int pf_profile_read_persistent(pf_box* h, double* total_ms, int64_t* launches,
                               int64_t* iterations) {
  if (int rc = check_box(h)) return rc;
  if (total_ms) *total_ms = h->prof_pms;
  if (launches) *launches = h->prof_pn;
  if (iterations) *iterations = h->prof_pit;
  return 0;
}

int pf_residual_history_enable(pf_box* h, int capacity) {
  if (int rc = check_box(h)) return rc;
  if (capacity < 0) return fail(PF_EINVAL, "pf_residual_history_enable: capacity must be >= 0");
  if (h->hist) {
    PF_HIP(hipFree(h->hist));
    h->bytes -= (int64_t)sizeof(double) * h->hist_cap;
    h->hist = nullptr;
    h->hist_cap = 0;
  }
  if (capacity == 0) return 0;
  hipError_t e = hipMalloc((void**)&h->hist, sizeof(double) * capacity);
  if (e != hipSuccess) {
    h->hist = nullptr;
    return check_hip(e, "hipMalloc");
  }
  h->hist_cap = capacity;
  h->bytes += (int64_t)sizeof(double) * capacity;
  return 0;
}

int pf_residual_history(pf_box* h, double* out, int capacity, int* count, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!out || !count || capacity < 0) return fail(PF_EINVAL, "pf_residual_history: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  PF_HIP(hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(KrylovCtl), hipMemcpyDeviceToHost, s));
  PF_HIP(hipStreamSynchronize(s));
  int n = h->ctl_host->iter;
  if (n > h->hist_cap) n = h->hist_cap;
  if (n > capacity) n = capacity;
  if (n > 0) {
    PF_HIP(hipMemcpyAsync(out, h->hist, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
  }
  *count = n;
  return 0;
}

int pf_profile_read(pf_box* h, int kernel, double* total_ms, int64_t* launches) {
  if (int rc = check_box(h)) return rc;
  if (kernel < 0 || kernel > 2) return fail(PF_EINVAL, "pf_profile_read: kernel is 0 (KA), 1 (KB) or 2 (KC)");
  if (total_ms) *total_ms = h->prof_ms[kernel];
  if (launches) *launches = h->prof_n[kernel];
  return 0;
}

int pf_apply_laplacian(pf_box* h, const void* A, void* B, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!A || !B) return fail(PF_EINVAL, "pf_apply_laplacian: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  const int nb = launch_blocks(h->g);
  if (h->dtype == PF_F64)
    hipLaunchKernelGGL(apply_lap_kernel<double>, dim3(nb), dim3(kBlock), 0, s, h->g,
                       (const double*)A, (double*)B, make_lap<double>(h));
  else
    hipLaunchKernelGGL(apply_lap_kernel<float>, dim3(nb), dim3(kBlock), 0, s, h->g,
                       (const float*)A, (float*)B, make_lap<float>(h));
  PF_LAUNCH_CHECK("apply_lap_kernel");
  return 0;
}

int pf_stats(pf_box* h, const void* u, const void* v, const void* w, void* divg, double* out5,
             void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !out5) return fail(PF_EINVAL, "pf_stats: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return stats_impl<double>(h, s, (const double*)u, (const double*)v, (const double*)w,
                              (double*)divg, out5);
  return stats_impl<float>(h, s, (const float*)u, (const float*)v, (const float*)w, (float*)divg,
                           out5);
}

int pf_div_vel(pf_box* h, const void* u, const void* v, const void* w, void* divg, void* stream) {
  if (int rc = check_box(h)) return rc;
  if (!u || !v || !w || !divg) return fail(PF_EINVAL, "pf_div_vel: NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  if (h->dtype == PF_F64)
    return div_impl<double>(h, s, (const double*)u, (const double*)v, (const double*)w,
                            (double*)divg);
  return div_impl<float>(h, s, (const float*)u, (const float*)v, (const float*)w, (float*)divg);
}

}  // extern "C"
