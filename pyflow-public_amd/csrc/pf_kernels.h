// Per-box constants and decompositions shared by the single-box entry points
// and the batched (multi-box) launches, and the descriptor a batched launch
// reads for each of its boxes.
#pragma once

#include "pf_box.h"

namespace pf {

template <typename T>
struct StepCoef {
  T dxi, dyi, dzi;
  T c2;               // 2*VISC/rho          (velocity.py:250)
  T exy, exz, eyz;    // interp_sc_**(VISC)/rho  (velocity.py:252-260)
  T half_dt;          // 0.5*simDt           (advancer.py:81)
  T dt6;              // simDt/6.0           (advancer.py:100)
  T dt;               // simDt               (advancer.py:55)
};

// Scalars rounded exactly as the reference's tensor ops round them: VISC_P is
// ones*mu in the field dtype (domain.py:130-131); edge viscosities are
// 0.5*0.5*(((V+V)+V)+V) (metric_staggered.py:127-144) divided by rho.
template <typename T>
StepCoef<T> make_step_coef(const pf_box* h, double mu, double rho, double dt) {
  StepCoef<T> c;
  c.dxi = (T)(1.0 / h->dx);
  c.dyi = (T)(1.0 / h->dy);
  c.dzi = (T)(1.0 / h->dz);
  const T V = (T)mu;
  const T R = (T)rho;
  c.c2 = ((T)2.0 * V) / R;
  const T edge = (T)0.25 * (((V + V) + V) + V);
  c.exy = edge / R;
  c.exz = edge / R;
  c.eyz = edge / R;
  c.half_dt = (T)(0.5 * dt);
  c.dt6 = (T)(dt / 6.0);
  c.dt = (T)dt;
  return c;
}

template <typename T>
struct Lap {
  T xm, x0, xp, ym, y0, yp, zm, z0, zp;
};

// Laplace coefficients as metric_uniform builds them (metric_staggered.py:85-102):
// 0 + h^-2 (or -2.0*h^-1*h^-1) added in the field dtype, then scaled by -vol.
template <typename T>
Lap<T> make_lap(const pf_box* h) {
  const double vol = h->dx * h->dy * h->dz;
  const double inv[3] = {1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz};
  T c[3][3];
  for (int a = 0; a < 3; ++a) {
    const double hi = inv[a];
    const double vals[3] = {hi * hi, -2.0 * hi * hi, hi * hi};
    for (int b = 0; b < 3; ++b) c[a][b] = (T)vals[b] * (T)(-1.0 * vol);
  }
  Lap<T> L;
  L.xm = c[0][0]; L.x0 = c[0][1]; L.xp = c[0][2];
  L.ym = c[1][0]; L.y0 = c[1][1]; L.yp = c[1][2];
  L.zm = c[2][0]; L.z0 = c[2][1]; L.zp = c[2][2];
  return L;
}

// B = sum in pressure.py:63-84 order: x-, x0, x+, y-, y0, y+, z-, z0, z+.
template <typename T>
__device__ __forceinline__ T lap7(const Lap<T>& L, T xm, T c, T xp, T ym, T yp, T zm, T zp) {
  T b = (T)0 + xm * L.xm;
  b = b + c * L.x0;
  b = b + xp * L.xp;
  b = b + ym * L.ym;
  b = b + c * L.y0;
  b = b + yp * L.yp;
  b = b + zm * L.zm;
  b = b + c * L.z0;
  b = b + zp * L.zp;
  return b;
}

// Krylov tiles whose x chunks have at most this many planes take the
// one-plane-ahead register schedule (PIPE in pf_krylov.hip), longer chunks
// the rotating one.
constexpr int kPipeChunk = 16;

// Tiled KA / KB decomposition (see pf_krylov.hip): a block owns a TZ x TY
// tile of the (y,z) plane and marches along x over a chunk of planes.
struct TileGrid {
  int tz, ty;             // tile extents: tz*ty == kBlock, or 2*kBlock with two cells per thread
  int ntz, nty, ncx;      // tiles along z, y and chunks along x
  int xc;                 // planes per chunk
  int nblk;               // ntz*nty*ncx
};

// Tile extent along z: the widest of 64, 32, 16 that fits nz and leaves at
// most 1/8 of the tiles' cells as periodic padding (a 96-cell axis takes 32,
// not 64 with a half-empty second tile), else 8.
inline int tile_extent_z(int nz) {
  for (int tz = 64; tz >= 16; tz >>= 1) {
    if (tz > nz) continue;
    const int pad = (nz + tz - 1) / tz * tz - nz;
    if (8 * pad <= nz) return tz;
  }
  return 8;
}

// Tile grid for `pair` cells per thread and x chunks of at least `minp`
// planes (a chunk also loads the planes on either side of it), up to
// `target` blocks in all.
inline TileGrid tile_grid_with(const Geom& g, int tz, int pair, int rows, int minp,
                               int target = 1024) {
  TileGrid t;
  t.tz = tz;
  t.ty = pair * rows * kBlock / t.tz;
  t.ntz = (g.nz + t.tz - 1) / t.tz;
  t.nty = (g.ny + t.ty - 1) / t.ty;
  const int tiles = t.ntz * t.nty;
  int cx = target / tiles;
  const int maxc = g.nx >= minp ? g.nx / minp : 1;
  if (cx > maxc) cx = maxc;
  if (cx < 1) cx = 1;
  t.xc = (g.nx + cx - 1) / cx;
  t.ncx = (g.nx + t.xc - 1) / t.xc;
  t.nblk = tiles * t.ncx;
  return t;
}

// Preferred: two z-neighbour cells per thread (when the tiles cover z
// exactly) and chunks of at least 64 bytes per column (8 planes in FP64, 4
// in FP32) - measured best at 128^3 for the launch-per-reduction kernels.
// Grids that would get fewer than 512 blocks that way (below two per CU)
// take 4-plane chunks, then one cell per thread - except 64 x 8 tiles on
// 4-plane chunks that cover the box exactly (64^3: 128 blocks), which the
// persistent solve holds on chip.  FP64 boxes whose 64 x 8 tiles on
// 16-plane chunks cover them exactly in 128 to 256 tiles (128^3: 256) take
// those: the persistent solve then holds one chunk per block on at least
// half the CUs, with half the partial-sum slots of 8-plane chunks per grid
// sync (128^3 FP64: 28.5 -> 27.1 us per iteration; its launch kernels, the
// fallback, 139.8 -> 135.5 Mcell-steps/s).
inline TileGrid make_tile_grid(const Geom& g, int esize, int rows = 1) {
  const int tz = tile_extent_z(g.nz);
  const bool can_pair = rows == 1 && tz >= 16 && g.nz % tz == 0;
  const int minp = esize >= 8 ? 8 : 4;
  if (can_pair) {
    TileGrid t;
    if (esize >= 8 && tz == 64 && g.nx % 16 == 0 && g.ny % 8 == 0) {
      t = tile_grid_with(g, tz, 2, rows, 16);
      if (t.xc == 16 && t.nblk >= 128 && t.nblk <= 256) return t;
    }
    t = tile_grid_with(g, tz, 2, rows, minp);
    if (t.nblk >= 512) return t;
    if (minp > 4) {
      t = tile_grid_with(g, tz, 2, rows, 4);
      if (t.nblk >= 512) return t;
    }
    if (t.tz == 64 && t.ty == 8 && t.xc == 4 && g.nx % 4 == 0 && g.ny % 8 == 0) return t;
  }
  return tile_grid_with(g, tz, 1, rows, 4);
}

// Tile of a block.  The hardware deals consecutive block ids round robin
// over the 8 XCDs; the ids are renumbered so that each XCD gets a contiguous
// run of tiles, z fastest, then y.  A run then covers whole z rows of tiles,
// so the z-halo cells (single-cell loads whose cache lines belong to the
// z-neighbour tile) and all but the run's edge y halos are lines that a
// neighbour tile on the same XCD brings into that XCD's L2.
__device__ __forceinline__ void tile_of_block(const TileGrid& tg, const Blk& blk, int& tz_i,
                                              int& ty_i, int& cx_i) {
  const int nb = blk.n;
  int b = blk.id;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  tz_i = b % tg.ntz;
  const int rest = b / tg.ntz;
  ty_i = rest % tg.nty;
  cx_i = rest / tg.nty;
}

// ---------------------------------------------------------------------------
// Batched launches.  One launch covers several boxes; each box keeps the
// block decomposition it has in a single-box launch (same blocks, same
// partial sums), so its results are bit-identical alone and in a batch.
template <typename T>
struct BoxDesc {
  Geom g;
  TileGrid tg;
  int nblk;                 // blocks of the reducing kernels
  T* u[3];                  // state (padded)
  T* ut[3];                 // RK stage temporaries (padded): caller's ...Tmp
  T* yp[3];                 // and the box's second buffer
  T* acc[3];                // running k1 + 2k2 + 2k3 (interior)
  T* DP;                    // pressure correction (padded)
  T* p;                     // pressure (padded), += DP
  T* src;                   // Poisson source (interior)
  T *x, *rk[2], *rh, *pk[2], *vk[2];       // Krylov vectors (interior)
  KrylovCtl* ctl;
  double *pA, *pB, *pC, *pS;                 // partial-sum buffers
  double* out;              // 8 doubles: maxU, maxV, maxW, sum(u^2..), maxDiv, res, iters, status
  StepCoef<T> k;
  Lap<T> L;
  T sc1, sc2;               // source scaling rho/dt and -vol
  T dxi, dyi, dzi, neg_c;   // projection: -(dt/rho)
  double tol;
  int maxit;
};

// Boxes of one batched launch and the first block of each (begin[n] = total).
struct BatchList {
  const int* box;
  const int* begin;
  int n;
};

// The list entry owning block `bid` (largest i with begin[i] <= bid).
__device__ __forceinline__ int find_box(const BatchList& L, int bid) {
  int lo = 0, hi = L.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.begin[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Batched launchers, one per kernel family (defined with the kernels).
// Each returns 0 or a negative PF_E* code.  `parity` = iteration index & 1.
int batch_stage_launch(int dtype, int stage, const void* desc, const BatchList& L, int nblocks,
                       hipStream_t s);
int batch_krylov_launch(int dtype, int what, int parity, const void* desc, const BatchList& L,
                        int nblocks, int tz, int ty, hipStream_t s, bool nts = true);
int batch_ctl_launch(int dtype, int what, const void* desc, const int* ids, int n, int* flags,
                     hipStream_t s);

enum { kBatchInit = 0, kBatchKA = 1, kBatchKB = 2, kBatchKC = 3, kBatchScatter = 4,
       kBatchProject = 5, kBatchStats = 6 };
enum { kBatchCtlInit = 0, kBatchCtlFlags = 1, kBatchFinalize = 2 };

// Persistent BiCGStab (pf_persist.hip).  persist_bytes: size of the buffers
// box h needs for it (0: its tile grid is not a shape the kernel handles).
// persist_blocks: blocks of box h's persistent solve (0: none).
// persist_solve: after the init kernel, run the whole solve as one
// cooperative launch; *used = false when it was not launched (the caller
// then runs the launch-per-reduction path); ev, if not NULL, is an event
// pair recorded around the launch.  persist_aborted: whether the
// last persistent solve gave up waiting for a block.
size_t persist_bytes(const pf_box* h);
int persist_blocks(const pf_box* h);
template <typename T>
int persist_solve(pf_box* h, hipStream_t s, const TileGrid& tg, const Lap<T>& L, const T* r0,
                  const T* rh, T* x, int npart_init, bool* used, hipEvent_t* ev);
int persist_aborted(pf_box* h, hipStream_t s, bool* aborted);

}  // namespace pf
