// Internal definition of the box handle and error plumbing shared by the
// translation units of libpyflow_hip.so.  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "pf_common.h"
#include "../../include/pyflow_hip.h"

namespace pf {

// Device-resident control block of one BiCGStab solve.  Block 0 of the
// Krylov kernels is the only writer of each field; every block reads the
// fields written by earlier kernels (kernel boundaries order the accesses).
struct KrylovCtl {
  double rho1, rho2, alpha, omega;
  double res0;      // max|r0|
  double max_res;   // max|r_k| / max|r0| of the last completed iteration
  double tol;
  int iter;         // completed iterations
  int maxit;
  int done;        // set only by KA/KB, whose early returns are then uniform
  int status;       // PF_SOLVE_* code
  int halt;         // set by KC on a <t,t> breakdown; the next KA turns it into done
  int npart;       // per-block partials in the <r,r_hat>, max|r| buffer (init or KC grid)
  int hist_cap;     // entries of hist (0: no history kept)
  double* hist;     // hist[k-1] = max|r_k| / max|r0| for k = 1 .. min(iter, hist_cap)
};

constexpr int kMaxPartSlots = 8;     // doubles per block per reduction
constexpr int kNumPartBufs = 4;      // A, B, C, stats

}  // namespace pf

struct pf_box {
  int nx, ny, nz;
  double dx, dy, dz;
  int dtype;          // PF_F64 / PF_F32
  size_t esize;
  int device;
  pf::Geom g;
  int nblk;           // blocks of the reducing kernels (fixed per box)
  // Krylov vectors, interior layout
  void* x;
  void* r;
  void* rh;
  void* p[2];
  void* v[2];
  void* s;
  void* t;
  // RK workspace
  void* acc[3];       // interior
  void* ypong[3];     // padded
  // reductions and control
  double* part;       // kNumPartBufs * part_stride
  int64_t part_stride;  // doubles per partial buffer
  pf::KrylovCtl* ctl;
  pf::KrylovCtl* ctl_host;   // pinned
  double* stats_dev;         // 8 doubles
  double* stats_host;        // pinned, 8 doubles
  int last_iters;
  double* hist;              // per-iteration relative residuals of the last solve (optional)
  int hist_cap;              // entries of hist
  // mixed-precision refinement (allocated on first use): FP64 iterate and
  // right-hand side, interior layout
  double* mp_x;
  double* mp_b;
  double* mp_hist;           // residual history of the current inner solve
  int mp_hist_cap;
  double mp_inner_tol;       // relative tolerance of each inner BiCGStab solve
  int mp_inner_maxit;        // iteration cap of each inner solve
  int mp_outer_maxit;        // refinement steps
  // 'source' ML closure workspace (allocated on first use): 21 normalised
  // features and 9 network outputs per padded cell, FP32
  float* src_ws;
  size_t src_ws_bytes;
  // its reverse (training through the closure), reserved by
  // pf_sfs_ml_source_train_reserve: activations and adjoints of a chunk of
  // cells (or the activations of every cell and stage sums of the layer
  // adjoints, when they fit), the input adjoints of every cell,
  // parameter-gradient partials, the closure's input arrays; and the
  // normalisation table the reverse divides by
  void* srct_ws;
  size_t srct_bytes;
  float srct_norm[21];
  int64_t srct_budget;       // bytes the all-cell cache may take (-1: default)
  int chunk;                 // iterations per launch batch between host checks
  int tile_rows;             // block rows of the tiled Krylov kernels: 256*tile_rows threads
  // persistent BiCGStab (pf_persist.hip): face, slot and flag buffers, when
  // the box's tile grid has a shape that kernel handles (else NULL)
  void* pers;
  int persist_on;            // 1: solves take the persistent kernel when it fits
  int last_persist;          // the last BiCGStab solve ran as the persistent kernel
  int64_t bytes;             // device bytes owned by the handle
  // optional per-launch timing of the Krylov kernels (measurement only)
  int prof_on;               // 0 = off, else time every prof_on-th launch
  int64_t prof_seq[3];       // launches seen per kernel kind while profiling
  int prof_cap;              // event pairs per kernel class
  hipEvent_t* prof_ev;       // [3][prof_cap][2]
  double prof_ms[3];         // accumulated milliseconds of KA, KB, KC
  int64_t prof_n[3];         // timed launches that did work
  hipEvent_t prof_pev[2];    // start/stop of the persistent solve's launch
  double prof_pms;           // accumulated milliseconds of persistent solves
  int64_t prof_pn;           // timed persistent solves
  int64_t prof_pit;          // BiCGStab iterations inside them
  hipEvent_t wait_ev;        // completion marker the per-step host waits spin on
};

namespace pf {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);
int check_box(const pf_box* h);
// Wait until stream s has done its work: spin on event ev (recorded here)
// for up to 50 ms, then block.  The runtime's own waits spin only briefly
// and then sleep until an interrupt, whose wake-up took milliseconds on
// some hosts (128^3 FP32 steps: 6.6 -> 8.6-10.1 ms).
// ev == NULL: hipStreamSynchronize.
int wait_stream(hipStream_t s, hipEvent_t ev);
int launch_blocks(const Geom& g);   // grid size for flat one-cell-per-thread kernels

// Closure term of the adjoint RHS (pf_adjoint.hip): for the adjoint state A
// at the forward state U, D = (d tau/du)^T (d R/d tau)^T A, folded onto the
// interior and refreshed in the overlap; dparams (may be NULL) accumulates
// the MLP parameter gradient.  work holds adj_work_bytes(h) bytes.
int adj_closure_term(pf_box* h, int kind, const void* const U[3], const float* params,
                     const void* const A[3], void* work, void* const D[3], double* dparams,
                     hipStream_t s);
// The same term for the reference's 'source' ML closure (pf_source_train.hip,
// adjoint.py:230-286): D = -(face interpolant of d g/d var_i) with
// g = sum over padded cells of G(var_i(U)) . (centre interpolant of A).
int ml_source_adj_term(pf_box* h, const void* const U[3], const float* params,
                       const void* const A[3], void* const D[3], double* dparams, hipStream_t s);
// The same over the four stages of an adjoint RK4 step, when the box's
// training workspace holds the forward activations of every cell
// (ml_source_step_cached): _begin evaluates the network on U once, _term
// gives stage `stage`'s D and adds its layer adjoints to the stage sums,
// _end adds the step's parameter gradient (from the sums) to dparams.
bool ml_source_step_cached(const pf_box* h);
int ml_source_step_begin(pf_box* h, const void* const U[3], const float* params, hipStream_t s);
int ml_source_step_term(pf_box* h, const void* const A[3], void* const D[3], int stage,
                        hipStream_t s);
int ml_source_step_end(pf_box* h, double* dparams, hipStream_t s);
size_t adj_work_bytes(const pf_box* h);

inline double* part_buf(pf_box* h, int which) {
  return h->part + (size_t)which * h->part_stride;
}

}  // namespace pf

#define PF_HIP(call)                                             \
  do {                                                           \
    hipError_t e_ = (call);                                      \
    if (e_ != hipSuccess) return pf::check_hip(e_, #call);       \
  } while (0)

#define PF_LAUNCH_CHECK(what)                                    \
  do {                                                           \
    hipError_t e_ = hipGetLastError();                           \
    if (e_ != hipSuccess) return pf::check_hip(e_, what);        \
  } while (0)
