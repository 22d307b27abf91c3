/*
 * libpyflow_hip.so — C-ABI of the MI355X (gfx950) kernels behind PyFlow's
 * single-box time-step path.
 *
 * The reference (MacRTL/PyFlow-public) is pure Python + PyTorch; its "boundary"
 * for this path is the set of Python classes listed beside each entry point
 * below (paths relative to the reference's src/).  The Python package
 * pyflow_amd binds these entry points with ctypes and keeps the reference's
 * class names, signatures and in-place semantics.
 *
 * Conventions
 *   - Every function returns 0 on success, a negative PF_E* code on error
 *     (pf_last_error() then describes it), or, for solvers, a PF_SOLVE_*
 *     status >= 0.
 *   - Field arguments are raw device pointers.  "padded" fields are
 *     (nx+4)*(ny+4)*(nz+4) contiguous arrays, z fastest, two overlap cells per
 *     side, byte-compatible with state_P.var (data/state.py:123).  "interior"
 *     fields are nx*ny*nz contiguous arrays (the reference's rhs_u,
 *     source_P).  Element type is the box's dtype.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 *     and is stream-ordered.  Functions that return host scalars (solvers,
 *     statistics) synchronise on the stream before returning, like the
 *     reference's .cpu() reductions.
 *   - Every kernel that writes a padded field also fills its overlap cells
 *     with periodic images, as state_P.update_border does.
 *   - Device memory is allocated only by pf_box_create and by the set-up
 *     calls that reserve optional workspace in a box (pf_mpir_configure,
 *     pf_sfs_ml_source_reserve, pf_sfs_ml_source_train_reserve,
 *     pf_residual_history_enable; pf_profile_enable
 *     creates events).  Every other entry point allocates nothing and fails
 *     with PF_EINVAL when the workspace it needs was not reserved, so a box's
 *     pf_box_device_bytes is final once its set-up calls are done.  A box may
 *     be used by one host thread and one stream at a time.
 */
#ifndef PYFLOW_HIP_H
#define PYFLOW_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_ABI_VERSION 1

/* dtypes */
#define PF_F64 0
#define PF_F32 1

/* error codes */
#define PF_EINVAL  (-1)
#define PF_EHIP    (-2)
#define PF_ENOMEM  (-3)

/* solver status codes (returned by the Poisson entry points) */
#define PF_SOLVE_CONVERGED   0   /* max|r|/max|r0| <= tol                    */
#define PF_SOLVE_MAXIT       1   /* iteration cap reached                    */
#define PF_SOLVE_BREAKDOWN_A 2   /* <v, r_hat> == 0  (pressure.py:247-248)   */
#define PF_SOLVE_BREAKDOWN_B 3   /* <t, t> == 0      (pressure.py:266-267)   */
/* rho2*omega == 0: beta (pressure.py:233) would divide by zero and every later
 * iterate would be NaN; the solve stops with the last finite iterate.  This
 * happens once <r, r_hat> has decayed to the rounding floor of the inner
 * product and lands on exactly 0 (long solves, e.g. 512^3 beyond ~900
 * iterations).                                                              */
#define PF_SOLVE_BREAKDOWN_RHO 4
#define PF_SOLVE_STALLED       5   /* refinement: an inner solve made no progress */

/* field layouts */
#define PF_LAYOUT_PADDED   0
#define PF_LAYOUT_INTERIOR 1

/* pf_state_op operations */
#define PF_EW_ZAXPY     0   /* out = (a*X) + Y                  state.py:157-165 */
#define PF_EW_COPY      1   /* out = X                          state.py:168-195 */
#define PF_EW_SUBSCALED 2   /* out = out - a*X (one FMA)        state.py:198-211 */

/* pf_metric_pair kinds */
#define PF_PAIR_INTERP 0    /* dst = coef*(a + b) */
#define PF_PAIR_DIFF   1    /* dst = (a - b)*coef */

/* velocity-correction solver selectors */
#define PF_POISSON_BICGSTAB 0
#define PF_POISSON_JACOBI   1
#define PF_POISSON_BICGSTAB_MPIR 2   /* see pf_bicgstab_mpir_solve */

typedef struct pf_box pf_box;

/* ABI version and last error (thread-local text). */
int pf_abi_version(void);
const char* pf_last_error(void);

/*
 * Create / destroy a box: a triply periodic uniform grid nx*ny*nz with
 * spacings dx, dy, dz and element type dtype, plus all device workspace.
 * Replaces the single-rank parts of parallel.decomp (library/parallel.py:118-284),
 * geometry.uniform (geometry/geometry.py:76-116) and the workspace allocations of
 * solver_bicgstab.__init__ (solver/pressure.py:184-190) and RK4
 * (solver/advancer.py:63-73).
 */
int pf_box_create(pf_box** out, int nx, int ny, int nz,
                  double dx, double dy, double dz, int dtype);
int pf_box_destroy(pf_box* box);
/* Device bytes held by the box (for admission control). */
int64_t pf_box_device_bytes(const pf_box* box);
/* Set the grid spacing after creation: parallel.decomp is built before the
 * geometry in PyFlow.run (core/PyFlow.py:81-104), so the box learns dx, dy,
 * dz from geometry.uniform. */
int pf_box_set_spacing(pf_box* box, double dx, double dy, double dz);
/* Whether the box's BiCGStab solves (pf_bicgstab_solve, pf_vel_corr, the
 * refinement's inner solves) may run as one persistent cooperative kernel that
 * keeps the Krylov vectors on chip (default on).  Same iterates and bits as
 * the launch-per-reduction kernels, which run when it is off, when the box's
 * tile grid is not one the persistent kernel handles, and when the GPU cannot
 * hold all of its blocks at once.
 * Returns 1 if the box has the persistent kernel's buffers, 0 if not, or an
 * error.  No counterpart in the reference (a tuning switch of
 * solver_bicgstab.solve, solver/pressure.py:196-297). */
int pf_box_set_persistent(pf_box* box, int on);
/* 1 if the box's last BiCGStab solve ran as the persistent kernel, 0 if as
 * launches per reduction (or an error).  No counterpart in the reference. */
int pf_solve_last_persistent(const pf_box* box);

/*
 * Periodic overlap fill of a padded field.
 * Replaces state_P.update_border / decomp.communicate_border
 * (data/state.py:148-150, library/parallel.py:301-414).
 */
int pf_fill_halo(pf_box* box, void* field, void* stream);

/*
 * Interior update of a padded field `out` (state_P methods ZAXPY, copy,
 * update, vel_corr; data/state.py:157-211).  X and Y are padded or interior
 * fields as given by their layout flags; with fill_halo != 0 the overlap
 * cells receive periodic images (update_border), otherwise only the interior
 * is written (state_P.copy).
 */
int pf_state_op(pf_box* box, int op, double a, void* out, const void* X, int x_layout,
                const void* Y, int y_layout, int fill_halo, void* stream);

/*
 * Fused momentum RHS + time advance.
 *   pf_rk4_step: the reference's four-stage advancer RK4.step
 *     (solver/advancer.py:75-113) with rhs_NavierStokes.evaluate
 *     (solver/velocity.py:221-383) fused into each stage.  u,v,w (padded) hold
 *     u^n on entry and u* on exit; ut,vt,wt (padded) receive the last stage
 *     state, as state_uTmp does.
 *   pf_euler_step: Euler.step (solver/advancer.py:49-57).
 *   mu, rho: constant molecular viscosity and density (domain.py:125-131).
 *   sfs6: NULL, or 6 padded cell-centred stress fields in the order
 *     FXX, FYY, FZZ, FXY, FXZ, FYZ (the 'tensor' closure contract,
 *     solver/velocity.py:262-315).
 */
int pf_rk4_step(pf_box* box, void* u, void* v, void* w, void* ut, void* vt, void* wt,
                double dt, double mu, double rho, const void* const* sfs6, void* stream);
int pf_euler_step(pf_box* box, void* u, void* v, void* w,
                  double dt, double mu, double rho, const void* const* sfs6, void* stream);

/*
 * pf_rk4_step without a box and without any allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops or harnesses that forbid hipMalloc).  The
 * box is described by its sizes, spacings and dtype; u,v,w,ut,vt,wt (padded)
 * and sfs6 mean what they mean for pf_rk4_step and receive the same bits.
 * The step's workspace (3 interior + 3 padded fields of the dtype, each
 * rounded up to 256 bytes) lives in a fixed per-GPU device array of
 * pf_rk4_step_noalloc_max_bytes() bytes (64 MiB: up to 96^3 in FP64 or
 * 128^3 in FP32) that each GPU the library runs on holds until the process
 * ends; larger boxes are rejected with PF_EINVAL before any GPU work.  All
 * calls on one GPU share that array: calls that can overlap on one GPU
 * (other streams, other threads) race and must be ordered by the caller;
 * calls on one stream are ordered by it.
 */
int pf_rk4_step_noalloc(int nx, int ny, int nz, double dx, double dy, double dz, int dtype,
                        void* u, void* v, void* w, void* ut, void* vt, void* wt, double dt,
                        double mu, double rho, const void* const* sfs6, void* stream);
int64_t pf_rk4_step_noalloc_max_bytes(void);

/*
 * One evaluation of the Navier-Stokes RHS into interior arrays ru, rv, rw.
 * Replaces rhs_NavierStokes.evaluate (solver/velocity.py:221-383).
 */
int pf_rhs_ns(pf_box* box, const void* u, const void* v, const void* w,
              double mu, double rho, const void* const* sfs6,
              void* ru, void* rv, void* rw, void* stream);

/*
 * The same three entry points with a spatially varying viscosity field
 * VISC_P (padded, the field Domain.forwardStep builds as mu + eddy viscosity,
 * domain.py:348-352) in place of the constant mu: the normal viscous fluxes
 * use 2*VISC/rho at their cell centres and the cross fluxes
 * interp_sc_xy/xz/yz(VISC)/rho at their edges (velocity.py:248-305,
 * metric_staggered.py:128-144), as the reference rounds them.
 */
int pf_rk4_step_visc(pf_box* box, void* u, void* v, void* w, void* ut, void* vt, void* wt,
                     double dt, const void* visc, double rho, const void* const* sfs6,
                     void* stream);
int pf_euler_step_visc(pf_box* box, void* u, void* v, void* w, double dt, const void* visc,
                       double rho, const void* const* sfs6, void* stream);
int pf_rhs_ns_visc(pf_box* box, const void* u, const void* v, const void* w, const void* visc,
                   double rho, const void* const* sfs6, void* ru, void* rv, void* rw,
                   void* stream);

/*
 * 'source' closures (sfsmodel modelType 'source', velocity.py:360-383).  The
 * reference evaluates such a closure inside every RHS evaluation: it
 * interpolates u, v, w to cell centres into state.var_i (interp_u_xm,
 * interp_v_ym, interp_w_zm), calls sfsmodel.update(var_i...), and subtracts
 * GX, GY, GZ interpolated to the faces (interp_sc_x/y/z) from the RHS.
 *
 * pf_source_inputs writes into ui, vi, wi (padded) exactly what the
 * reference's var_i arrays hold at that call: the centre interpolants, and on
 * the last plane along each component's axis the edge interpolants the
 * advective terms of the same evaluation left there (one corner cell per
 * array, never written by the reference's RHS, is left unchanged).
 * pf_rhs_ns_src and pf_euler_step_src are pf_rhs_ns / pf_euler_step with the
 * closure term (constant viscosity, no stresses: a source closure is the
 * run's only closure).  An RK4 step with such a closure is four calls of
 * pf_rk4_stage_src (stage 1..4), each preceded by pf_rk4_source_inputs of
 * the same stage (the inputs of that stage's state) and the closure's own
 * evaluation; the four stages together equal pf_rk4_step's sequence
 * (advancer.py:75-113).  gx, gy, gz: padded, box dtype.
 */
int pf_source_inputs(pf_box* box, const void* u, const void* v, const void* w, void* ui, void* vi,
                     void* wi, void* stream);
int pf_rhs_ns_src(pf_box* box, const void* u, const void* v, const void* w, double mu, double rho,
                  const void* gx, const void* gy, const void* gz, void* ru, void* rv, void* rw,
                  void* stream);
int pf_euler_step_src(pf_box* box, void* u, void* v, void* w, double dt, double mu, double rho,
                      const void* gx, const void* gy, const void* gz, void* stream);
int pf_rk4_source_inputs(pf_box* box, int stage, void* u, void* v, void* w, void* ut, void* vt,
                         void* wt, void* ui, void* vi, void* wi, void* stream);
int pf_rk4_stage_src(pf_box* box, int stage, void* u, void* v, void* w, void* ut, void* vt,
                     void* wt, double dt, double mu, double rho, const void* gx, const void* gy,
                     const void* gz, void* stream);

/*
 * The reference's 'source' ML closure, residual_stress.update(u_P, v_P, w_P)
 * of sfsmodel/sfsmodel_ML.py:354-404 with ML_Closure (:468-593),
 * Finite_Difference_ConCat (:408-465) and NeuralNetworkModel2 (:76-116):
 * ui, vi, wi (padded, box dtype; the var_i arrays pf_source_inputs writes) ->
 * gx, gy, gz (padded, box dtype).  Features and network in FP32 as the
 * reference evaluates them; the network on MFMA.
 * params: device FP32, pf_sfs_ml_source_params() values: torch.nn.Linear
 * layout (weight[out][in] then bias[out]) of fc1, fcG, fcG2 (273 -> 200),
 * fc2, fc3 (200 -> 200), fc4 (200 -> 18), in that order.  norm21: host, the
 * 21 FP32 normalisation constants (sfsmodel_ML.py:252-258).  Needs the
 * box's workspace of 30 floats per padded cell, reserved once by
 * pf_sfs_ml_source_reserve (PF_EINVAL otherwise).
 */
int pf_sfs_ml_source(pf_box* box, const float* params, const float* norm21, const void* ui,
                     const void* vi, const void* wi, void* gx, void* gy, void* gz, void* stream);
int64_t pf_sfs_ml_source_params(void);
/* Reserve the closure's workspace in the box (set-up call, allocates; made
 * where residual_stress.__init__ allocates its workspace arrays,
 * sfsmodel_ML.py:145-151).  Idempotent. */
int pf_sfs_ml_source_reserve(pf_box* box);

/*
 * Training through the 'source' ML closure: the reverse of pf_sfs_ml_source,
 * the autograd part of the reference's adjoint closure term
 * (solver/adjoint.py:241-273, sfsmodel_ML.py:354-593 differentiated).
 * For g = sum over all padded cells of GX*gx + GY*gy + GZ*gz with
 * (GX, GY, GZ) = closure(ui, vi, wi): dui, dvi, dwi (FP32 padded, each may be
 * NULL) receive dg/d(ui, vi, wi), and dparams (FP64, pf_sfs_ml_source_params()
 * values in the packed order, may be NULL) accumulates dg/dparams, as the
 * reference's g.backward() accumulates its model's gradients.  In FP32 like
 * the reference's closure; summation order differs from torch's CPU
 * autograd.  Needs the training workspace of pf_sfs_ml_source_train_reserve.
 * pf_adj_closure_term, pf_adj_rk4_step and pf_adj_euler_step take this
 * closure as kind 2: delta = -(face interpolant of dg/d var_i), with var_i the
 * forward state's inputs as its RHS leaves them and (gx, gy, gz) the adjoint
 * state's centre interpolants (adjoint.py:230-286).
 */
int pf_sfs_ml_source_vjp(pf_box* box, const float* params, const void* ui, const void* vi,
                         const void* wi, const void* gx, const void* gy, const void* gz,
                         float* dui, float* dvi, float* dwi, double* dparams, void* stream);
/* Set-up call: reserve the reverse's workspace in the box (about 1.2 GB at
 * 64^3, 3.8 GB at 128^3, plus the all-cell cache below when it fits its
 * budget; allocates) and keep the 21-entry normalisation table
 * (sfsmodel_ML.py:252-258) for the adjoint closure term.  Idempotent. */
int pf_sfs_ml_source_train_reserve(pf_box* box, const float* norm21);
/* Set-up call, before pf_sfs_ml_source_train_reserve: the bytes the
 * closure's all-cell cache may take (default 48 GiB; negative = default).
 * With the cache (11.3 KB per padded cell: 3.5 GB at 64^3, 26 GB at 128^3)
 * the workspace holds the network's forward activations of every cell, and
 * pf_adj_rk4_step with closure 2 evaluates the network on the forward state
 * once per step (the reference's four adjoint stages see the same forward
 * state) and its parameter gradient as one product of the stage-summed
 * layer adjoints; without it every stage evaluates the whole term.  The
 * adjoint fields are the same bits either way; the parameter gradient
 * differs by FP32 rounding (sum then product instead of product then sum).
 * No counterpart in the reference. */
int pf_sfs_ml_source_cache_budget(pf_box* box, int64_t bytes);

/*
 * Smagorinsky eddy viscosity plus the molecular one, written to the padded
 * viscosity field: visc = mu + rho*(Cs*Delta)^2*|S| with
 * |S| = sqrt(S11^2+S22^2+S33^2 + 2(S12^2+S13^2+S23^2)) of the trace-free
 * strain of the nine grad_vel_center gradients.  Replaces
 * stress_constCs.eddyVisc (sfsmodel/sfsmodel_smagorinsky.py:65-88) followed by
 * VISC_P.copy_(muMolec + muEddy) (data/domain.py:348-352).
 */
int pf_sfs_smagorinsky(pf_box* box, const void* u, const void* v, const void* w, double rho,
                       double Cs, double Delta, double mu, void* visc, void* stream);

/*
 * Interior cell-centred divergence of padded u, v, w into interior `divg`.
 * Replaces metric_uniform.div_vel (metric/metric_staggered.py:438-453).
 */
int pf_div_vel(pf_box* box, const void* u, const void* v, const void* w, void* divg, void* stream);

/*
 * BiCGStab pressure solve.  Replaces solver_bicgstab.solve
 * (solver/pressure.py:196-297): DP (padded) := solution of
 * (-vol*L7) DP = -vol*(rho/dt)*source from DP = 0, p (padded) += DP, and
 * `source` (interior) is scaled in place by rho/dt and then by -vol, as the
 * reference does.  Stops when max|r|/max|r0| <= tol or after maxit
 * iterations; breakdowns end the solve silently like the reference.
 * Outputs (host pointers, each may be NULL): *out_res = the reference's
 * returned max_res, *out_iters = iterations executed.
 * Returns a PF_SOLVE_* status (>= 0) or an error (< 0).
 */
int pf_bicgstab_solve(pf_box* box, void* DP, void* p, void* source,
                      double rho, double dt, double tol, int maxit,
                      double* out_res, int* out_iters, void* stream);

/*
 * The same solve by mixed-precision iterative refinement (no counterpart in
 * the reference; the remedy for long solves where the reference's BiCGStab
 * breaks down, see PF_SOLVE_BREAKDOWN_RHO, or where FP32 vectors diverge):
 * the iterate x and the right-hand side b = -vol*(rho/dt)*source are kept in
 * FP64; each refinement step rounds the FP64 residual r = b - A x to the box
 * precision, solves A d = r with the box's BiCGStab (fresh shadow residual)
 * to the relative tolerance of pf_mpir_configure, and adds d to x in FP64.
 * Stops when max|b - A x| / max|b| <= tol (the FP64 residual of x, returned
 * in *out_res), after the refinement-step cap or when `maxit` inner
 * iterations are used up (PF_SOLVE_MAXIT), or when an inner solve makes no
 * progress (PF_SOLVE_STALLED).  An inner solve that stops unconverged is
 * re-run up to its best iterate (the solves are deterministic), so a
 * degrading inner recurrence costs iterations, not accuracy.  DP := x rounded to the box precision, p += DP.  `source` ends
 * holding the last inner right-hand side.  *out_iters counts inner
 * iterations.  Uses the FP64 workspace (16 bytes per cell) that
 * pf_mpir_configure reserves; PF_EINVAL if it was never called on the box.
 * pf_vel_corr with solver = PF_POISSON_BICGSTAB_MPIR runs the same solve
 * (with b from div(u) computed in FP64).
 */
int pf_bicgstab_mpir_solve(pf_box* box, void* DP, void* p, void* source,
                           double rho, double dt, double tol, int maxit,
                           double* out_res, int* out_iters, void* stream);

/*
 * Refinement settings of a box (defaults: inner_tol 1e-1, inner_maxit 500,
 * outer_maxit 100).  Short inner solves keep each BiCGStab run well before
 * the point where its recurrence degrades (measured on a 512^3 FP32 solve to
 * 1e-8: inner tolerance 1e-1 took 1797 inner iterations, 1e-2 took 2309, and
 * inner solves run to 3e-3 or beyond diverged).  Set-up call: reserves the
 * refinement workspace (FP64 iterate and right-hand side, 16 bytes per cell,
 * and an inner residual history of inner_maxit doubles) so that the solves
 * allocate nothing.
 */
int pf_mpir_configure(pf_box* box, double inner_tol, int inner_maxit, int outer_maxit);

/*
 * The last refinement solve's FP64 iterate x and right-hand side
 * b = -vol*(rho/dt)*source (interior layout, nx*ny*nz doubles each; either
 * pointer may be NULL), copied on `stream` into caller device memory.  No
 * counterpart in the reference: for callers that need the solution beyond
 * the box precision, or its residual max|b - A x|/max|b| recomputed on
 * their side.
 */
int pf_mpir_solution(pf_box* box, double* x, double* b, void* stream);

/*
 * B (interior) = (-vol*L7) A for a padded A whose overlap is already filled.
 * Replaces Pressure_Base.applyOperator (solver/pressure.py:54-84) after its
 * communicate_border call; same nine-term summation order.
 */
int pf_apply_laplacian(pf_box* box, const void* A, void* B, void* stream);

/*
 * Jacobi pressure solve.  Replaces solver_jacobi.solve
 * (solver/pressure.py:115-161), including its conventions: p is zeroed and
 * iterated nit times, pOld holds iterate nit-1 on exit (the field the
 * reference's velocity correction then uses), and the returned residual is
 * the reference's diagnostic value.
 */
int pf_jacobi_solve(pf_box* box, void* pOld, void* p, void* source,
                    double rho, double dt, int nit, double* out_res, void* stream);

/*
 * Velocity correction of Domain.velCorr (data/domain.py:373-402), fused:
 * source := div(u*) (metric_uniform.div_vel), source scaled by src_scale
 * (the solver's rho/simDt, pressure.py:125/205), solve for DP with the
 * selected solver (PF_POISSON_BICGSTAB or PF_POISSON_BICGSTAB_MPIR: also
 * -vol scaling and p += DP;
 * PF_POISSON_JACOBI: the reference's Jacobi conventions, DP = pOld), then
 * u -= corr_scale * grad(DP) with corr_scale = the domain's simDt/rho,
 * using the staggered gradient of metric_uniform.grad_P
 * (metric_staggered.py:390-401) and state_P.vel_corr (data/state.py:198-211).
 * u,v,w,DP,p padded; source interior.  `iters` is the BiCGStab cap or the
 * Jacobi sweep count.  Returns a PF_SOLVE_* status or an error.
 */
int pf_vel_corr(pf_box* box, void* u, void* v, void* w, void* DP, void* p, void* source,
                int solver, double src_scale, double corr_scale, double tol, int iters,
                double* out_res, int* out_iters, void* stream);

/*
 * Per-step monitor values of PyFlow.run (core/PyFlow.py:378-395):
 * out5 = {max|U|, max|V|, max|W|, TKE, max|div|} over the interior, with
 * TKE = 0.5*sum(u^2+v^2+w^2)/N.  If divg is not NULL the interior divergence
 * is also written there (the reference writes it into source_P).
 */
int pf_stats(pf_box* box, const void* u, const void* v, const void* w, void* divg,
             double* out5, void* stream);

/*
 * Gradient closure ('tensor' model) of sfsmodel_gradient.residual_stress.update
 * (sfsmodel/sfsmodel_gradient.py:70-148): writes the six padded cell-centred
 * stress fields FXX, FYY, FZZ, FXY, FXZ, FYZ from padded u, v, w using the
 * cell-centred gradients of metric_uniform.grad_vel_center
 * (metric_staggered.py:215-268).
 */
int pf_sfs_gradient(pf_box* box, const void* u, const void* v, const void* w,
                    void* const* sfs6, void* stream);

/*
 * metric_uniform building blocks (metric/metric_staggered.py:112-526) with
 * the reference's slice semantics over whole padded arrays (overlap cells
 * included, untouched cells left as they are):
 *   pf_metric_pair: along `axis`, with upper = 1 the box [0, n-1) gets
 *     OP(src[p+e], src[p]) (interp_u_xm/v_ym/w_zm, grad_vel_visc); with
 *     upper = 0 the box [1, n) gets OP(src[p], src[p-e]) (interp_sc_x/y/z,
 *     interp_vw_x, interp_uw_y, interp_uv_z, interp_uvwi_*, grad_P).
 *     dst == src is allowed (interp_uvwi_*: reads a copy).
 *   pf_metric_edge4: interp_sc_xy (plane 0), _xz (1), _yz (2).
 *   pf_div_visc: rhs (interior) += flux divergence (metric_staggered.py:406-418).
 *   pf_div_vel_over: cell-centre divergence over the padded array (:424-433).
 *   pf_vel_conv: rhs (interior) -= ((A1*B1 - A0*B0)/h)*sign on padded var_i
 *     arrays; diagonal = 1 for vel_conv_xx/yy/zz (cells i, i-1), 0 for
 *     vel_conv_x/y/z (cells i+1, i) (:458-526).
 */
int pf_metric_pair(pf_box* box, int kind, int axis, int upper, void* dst, const void* src,
                   double coef, void* stream);
int pf_metric_edge4(pf_box* box, int plane, void* dst, const void* src, void* stream);
int pf_div_visc(pf_box* box, const void* FX, const void* FY, const void* FZ, void* rhs,
                void* stream);
int pf_div_vel_over(pf_box* box, const void* u, const void* v, const void* w, void* divg_o,
                    void* stream);
int pf_vel_conv(pf_box* box, int axis, int diagonal, const void* A, const void* B, void* rhs,
                double sign, void* stream);

/*
 * Closure MLP with the same 'tensor' contract (no counterpart in the
 * reference, whose ML closure is a 'source' model): the nine gradients of
 * grad_vel_center -> Linear(9,20) -> tanh -> Linear(20,20) -> tanh ->
 * Linear(20,6), scaled by dx^2/12, into the six stress fields.  GEMMs on
 * MFMA with BF16 operands and FP32 accumulation.  `params` is a device FP32
 * vector of pf_sfs_mlp_param_count() values: W1[20][9] b1[20] W2[20][20]
 * b2[20] W3[6][20] b3[6].
 */
int pf_sfs_mlp(pf_box* box, const void* u, const void* v, const void* w, const float* params,
               void* const* sfs6, void* stream);
int pf_sfs_mlp_param_count(void);
/* The same MLP in FP32 arithmetic (the precision its reverse pass
 * differentiates); same arguments as pf_sfs_mlp. */
int pf_sfs_mlp_f32(pf_box* box, const void* u, const void* v, const void* w, const float* params,
                   void* const* sfs6, void* stream);

/*
 * Adjoint step of closure training (replaces Domain.adjointStep,
 * data/domain.py:409-438, with rhs_adjPredictor.evaluate,
 * solver/adjoint.py:94-298, and adj_conv_*, metric/metric_staggered.py:529-800).
 * u, v, w: the forward state (padded) the adjoint is linearised about, held
 * fixed; ua, va, wa: the adjoint state (padded).  The adjoint RHS is the
 * viscous operator of the forward RHS applied to the adjoint field plus the
 * 15 convective transposes, bit-for-bit in the reference's op order, plus an
 * optional closure term delta3 (padded, added last).
 *   pf_rhs_adj: RHS into ru, rv, rw (interior arrays).
 *   pf_adj_rk4_step / pf_adj_euler_step: the reference's RK4 / Euler on the
 *     adjoint RHS (advancer.py:49-113); uta..wta receive the last stage state.
 *     closure = -1 (none), 0 (gradient model), 1 (MLP with `params`) or 2
 *     (the reference's 'source' ML closure with `params`): the closure term
 *     of every stage is recomputed from that stage's adjoint state with
 *     pf_adj_closure_term; then delta3 (three padded scratch arrays) is
 *     required, work (pf_adj_work_bytes bytes) too for kinds 0 and 1, and
 *     dparams (the closure's packed parameter count of doubles, may be NULL)
 *     accumulates the parameter gradient of every stage (the reference
 *     accumulates its closure's parameter gradients over every adjoint RHS
 *     evaluation the same way, adjoint.py:266-270).
 *   pf_adj_closure_term: kinds 0 and 1: delta = (d tau/du)^T (d R/d tau)^T a
 *     for a 'tensor' closure tau at the forward state, folded onto the
 *     interior with periodic overlap values; kind 2: the reference's own
 *     term (adjoint.py:230-286), see pf_sfs_ml_source_vjp.
 *   pf_adj_conv: one adj_conv_* method on padded arrays: for s in {0, e_s2},
 *     rhs -= ((a[m - e_d + s] - a[m + s]) * vi[m + s + vsh]) * coef with
 *     vsh = -e_d when centre != 0 (adj_conv_xx/yy/zz) and 0 otherwise.
 */
int pf_rhs_adj(pf_box* box, const void* u, const void* v, const void* w, const void* ua,
               const void* va, const void* wa, double mu, double rho, const void* const* delta3,
               void* ru, void* rv, void* rw, void* stream);
int pf_adj_rk4_step(pf_box* box, const void* u, const void* v, const void* w, void* ua, void* va,
                    void* wa, void* uta, void* vta, void* wta, double dt, double mu, double rho,
                    int closure, const float* params, void* work, void* const* delta3,
                    double* dparams, void* stream);
int pf_adj_euler_step(pf_box* box, const void* u, const void* v, const void* w, void* ua,
                      void* va, void* wa, double dt, double mu, double rho, int closure,
                      const float* params, void* work, void* const* delta3, double* dparams,
                      void* stream);
/* The three with the viscosity field VISC_P (padded) in place of mu: the
 * reference's adjoint applies its forward viscous operator, the field
 * included, to the adjoint velocity (solver/adjoint.py:100-170). */
int pf_rhs_adj_visc(pf_box* box, const void* u, const void* v, const void* w, const void* ua,
                    const void* va, const void* wa, const void* visc, double rho,
                    const void* const* delta3, void* ru, void* rv, void* rw, void* stream);
int pf_adj_rk4_step_visc(pf_box* box, const void* u, const void* v, const void* w, void* ua,
                         void* va, void* wa, void* uta, void* vta, void* wta, double dt,
                         const void* visc, double rho, int closure, const float* params,
                         void* work, void* const* delta3, double* dparams, void* stream);
int pf_adj_euler_step_visc(pf_box* box, const void* u, const void* v, const void* w, void* ua,
                           void* va, void* wa, double dt, const void* visc, double rho,
                           int closure, const float* params, void* work, void* const* delta3,
                           double* dparams, void* stream);
size_t pf_adj_work_bytes(const pf_box* box);
int pf_adj_closure_term(pf_box* box, int kind, const void* u, const void* v, const void* w,
                        const float* params, const void* ua, const void* va, const void* wa,
                        void* work, void* const* delta3, double* dparams, void* stream);
int pf_adj_conv(pf_box* box, int d_axis, int s2_axis, int centre, const void* a, const void* vi,
                void* rhs, double coef, void* stream);

/*
 * Reverse-mode (vector-Jacobian) kernels of the forward step, for exact
 * gradients through it (no counterpart in the reference, which differentiates
 * only its closure, by torch autograd, adjoint.py:241-286).  Adjoint fields
 * are padded arrays with periodic overlap values; every pass is a per-cell
 * local reverse followed by a fixed-order gather (no atomics), so results
 * are deterministic.
 *   pf_rhs_vjp: for lambda = (lu, lv, lw), out += J_u(R)(u)^T lambda (R the
 *     RHS without closure) and, when ku..kw are given, kprev += hcoef * the
 *     same (the RK stage chain).  work: 33 * nx*ny*nz elements.
 *   pf_sfs_vjp: taubar (6 padded) = (d R/d tau)^T lambda, or += when
 *     accumulate != 0.
 *   pf_fold_halo: transpose of the periodic overlap fill, in place.
 *   pf_closure_vjp: f = (d tau/du)^T taubar for kind 0 (gradient model) or 1
 *     (MLP, FP32), before folding; gbar9: 9 padded arrays of feature
 *     adjoints; dparams (+=, 746 doubles) with wpart
 *     (pf_closure_vjp_waves(box) * 746 doubles) for the MLP parameters.
 */
int pf_rhs_vjp(pf_box* box, const void* u, const void* v, const void* w, double mu, double rho,
               const void* lu, const void* lv, const void* lw, void* work, void* ou, void* ov,
               void* ow, void* ku, void* kv, void* kw, double hcoef, void* stream);
int pf_sfs_vjp(pf_box* box, const void* lu, const void* lv, const void* lw, void* const* taubar6,
               int accumulate, void* stream);
int pf_fold_halo(pf_box* box, void* field, void* stream);
int pf_closure_vjp_waves(const pf_box* box);
int pf_closure_vjp(pf_box* box, int kind, const void* u, const void* v, const void* w,
                   const float* params, void* const* taubar6, void* gbar9, void* fu, void* fv,
                   void* fw, double* dparams, double* wpart, void* stream);

/*
 * Batched time steps of many resident boxes (the serving path; no
 * counterpart in the reference, which steps one box per process,
 * core/PyFlow.py:352-425).  Each box of a batch is idle or inside a time
 * step (one Domain.forwardStep: RK4, BiCGStab velCorr, monitor values;
 * closure-free).  Every kernel launch covers all the boxes it concerns;
 * the Krylov iterations run from a hipGraph of pf_batch_chunk_iterations()
 * iterations over all resident boxes, captured once per composition.  Each
 * box keeps the block decomposition of its single-box launches, so its
 * results are bit-identical to pf_rk4_step + pf_vel_corr + pf_stats on the
 * same box, alone or in any mix.  All boxes of a batch share its dtype; a
 * box may be in one batch at a time (its workspace is used by the batched
 * kernels).
 *   pf_batch_set: the n resident boxes.  fields: n x 8 device pointers (u, v,
 *     w, ut, vt, wt, DP, p; padded), sources: n interior arrays (left holding
 *     the scaled Poisson source), params: n x 5 values (dt, mu, rho, tol,
 *     maxit).  Boxes already resident keep their phase (a solve in progress
 *     continues); new boxes join idle.  Synchronous; re-captures the graph.
 *   pf_batch_advance: starts a step for every box with start[i] != 0 (it must
 *     be idle), then runs Krylov chunks until at least one running box's solve
 *     is done, finishes those boxes' steps and returns with finished[i] = 1
 *     for them and, in out[8i..8i+8) (may be NULL): max|u|, max|v|, max|w|,
 *     sum(u^2+v^2+w^2), max|div u|, the solver's max residual, iterations and
 *     status.  Boxes whose solves take longer keep running across calls, so a
 *     fast box starts its next step while slow ones iterate.  out_chunks:
 *     graph replays in this call.  Ordered after the caller's work on
 *     `stream`; returns when the finishing boxes' steps are complete.
 *   pf_batch_set_group_bytes: inside a chunk the resident boxes run in
 *     groups whose Krylov vectors (nine interior vectors per box) together
 *     take at most `bytes` (default 320 MiB, measured best for the 256 MB
 *     Infinity Cache; a box larger than that is a group of its own; 0 = one
 *     group of all boxes, with non-temporal output stores); each group runs
 *     all chunk iterations before the next, so its Krylov vectors stay in
 *     that cache from one kernel to the next.  Only the order of independent
 *     boxes' launches changes, never a result.  Synchronous; re-captures the
 *     graph.
 *   pf_batch_set_persistent: on > 0, a resident box whose persistent solve
 *     fills the GPU (pf_box_set_persistent; 128^3 FP32 and FP64) solves each
 *     step's Poisson problem as that one cooperative launch instead of on the
 *     chunk kernels (same iterates, same bits), at most `on` such solves per
 *     pf_batch_advance call (default 1): a box over the cap waits for a later
 *     call, first come first served, and the other boxes get a chunk between
 *     persistent solves.  A larger cap raises throughput and the other
 *     boxes' step latency.  on = 0: every box runs on the chunk kernels.
 *     Returns 1 if a resident box now takes the persistent solve, else 0;
 *     PF_EINVAL for on < 0.  Synchronous; re-captures the graph.
 */
typedef struct pf_batch pf_batch;
int pf_batch_create(pf_batch** out, int capacity, int dtype);
int pf_batch_destroy(pf_batch* batch);
int pf_batch_set(pf_batch* batch, int n, pf_box* const* boxes, void* const* fields,
                 void* const* sources, const double* params);
int pf_batch_advance(pf_batch* batch, const int* start, int* finished, double* out,
                     int* out_chunks, void* stream);
int pf_batch_chunk_iterations(void);
int pf_batch_set_group_bytes(pf_batch* batch, int64_t bytes);
int pf_batch_set_persistent(pf_batch* batch, int on);

/*
 * Measurement support (no counterpart in the reference): with period > 0,
 * every period-th launch of each of the three BiCGStab kernels (0 = KA:
 * p update + A p + <v,r_hat>; 1 = KB: s + A s + <s,t>,<t,t>; 2 = KC: x, r
 * updates + <r,r_hat>, max|r|) carries a start/stop hipEvent pair recorded by
 * the dispatch itself; pf_profile_read returns the accumulated milliseconds
 * and the number of timed launches that did work.  period = 0 turns it off.
 * Solves that run as the persistent kernel have no per-kernel launches; see
 * pf_profile_read_persistent.
 * Enabling resets the accumulators.  Creates events on first use.
 */
int pf_profile_enable(pf_box* box, int period);
int pf_profile_read(pf_box* box, int kernel, double* total_ms, int64_t* launches);
/* The same for solves that ran as the persistent kernel (pf_box_set_persistent):
 * while profiling is on, each such launch carries a start/stop event pair;
 * returns their accumulated milliseconds, the number of launches and the
 * BiCGStab iterations they performed. */
int pf_profile_read_persistent(pf_box* box, double* total_ms, int64_t* launches,
                               int64_t* iterations);

/*
 * Residual history of the BiCGStab solves of one box: the value the
 * reference computes at pressure.py:279-280 after every iteration (and prints
 * there when its commented-out line 289 is enabled).  pf_residual_history_enable
 * allocates room for `capacity` iterations on the device (0 frees it); every
 * later pf_bicgstab_solve / pf_vel_corr on the box then stores
 * max|r_k|/max|r0| for k = 1 .. min(iterations, capacity), and
 * pf_residual_history copies the last solve's values into out[0..*count).
 * Synchronizes the stream.  (Refinement solves, PF_POISSON_BICGSTAB_MPIR,
 * keep their inner histories to themselves and record nothing here.)
 */
int pf_residual_history_enable(pf_box* box, int capacity);
int pf_residual_history(pf_box* box, double* out, int capacity, int* count, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYFLOW_HIP_H */
