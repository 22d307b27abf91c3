"""Exact gradients through the time step: Domain.forwardStep (closure
stresses from u^n, the reference's RK4, projection; data/domain.py:342-402)
as a torch.autograd.Function whose backward is hand-written HIP kernels.

The reference differentiates only its closure (by torch autograd inside its
adjoint RHS, solver/adjoint.py:241-286) and integrates a separate adjoint
equation backwards (Domain.adjointStep, see adjoint.py of this package).
This module instead gives the exact reverse of the discrete forward step,
so a loss on u^{n+K} has gradients with respect to u^n and to the closure
MLP's parameters that agree with finite differences of the step itself:

* projection: P = I - G L^-1 D is self-adjoint on the periodic staggered
  grid (G^T = -D, L = D G symmetric), so its reverse is the same fused
  correction kernel applied to the adjoint field;
* RK4 (advancer.py:75-113): stage states Y2..Y4 are recomputed from the
  checkpointed u^n (rhs + stage-update kernels, bit-identical to the
  forward's), then each stage's vector-Jacobian product pf_rhs_vjp
  accumulates into du and into the previous stage's adjoint (x dt/2);
* closure: pf_sfs_vjp accumulates the stress adjoints of all four stages,
  pf_closure_vjp maps them back through the closure (gradient model, or the
  MLP in FP32 with its parameter gradient) and pf_fold_halo folds the
  overlap contributions onto the interior.

Each autograd node saves only its step's input velocity (three interior
arrays): a K-step window holds K checkpoints and recomputes the stage states
in the backward pass.

The MLP forward can run on MFMA in BF16 (mlp_forward='bf16', the closure
kernel of Domain) or in FP32 (mlp_forward='fp32', the arithmetic its reverse
pass differentiates; use it when gradients are compared with finite
differences).
"""

import torch

from . import _native


class StepVJP:
    """Forward step and its reverse on one box, with all device scratch
    allocated once.  closure: None, 'gradient' or 'mlp'."""

    def __init__(self, box, mu, rho, dt, closure=None, tol=1e-12, maxit=3000,
                 mlp_forward='fp32'):
        if closure not in (None, 'gradient', 'mlp'):
            raise ValueError("closure is None, 'gradient' or 'mlp'")
        if mlp_forward not in ('fp32', 'bf16'):
            raise ValueError("mlp_forward is 'fp32' or 'bf16'")
        self.box = box
        self.mu, self.rho, self.dt = float(mu), float(rho), float(dt)
        self.closure = closure
        self.tol, self.maxit = float(tol), int(maxit)
        self.mlp_forward = mlp_forward
        dev, dt_ = box.device, box.dtype
        P = box.padded_shape
        I = box.interior_shape
        z = lambda shape: torch.zeros(shape, dtype=dt_, device=dev)   # noqa: E731
        self.U = [z(P) for _ in range(3)]          # u^n, then the step result
        self.T = [z(P) for _ in range(3)]          # RK stage temporaries
        self.Y = [[z(P) for _ in range(3)] for _ in range(3)]   # Y2, Y3, Y4
        self.k = [z(I) for _ in range(3)]          # stage RHS
        self.lam = [z(P) for _ in range(3)]        # adjoint of the step result
        self.acc = [z(P) for _ in range(3)]        # adjoint of u^n
        self.lk = [[z(P) for _ in range(3)] for _ in range(4)]  # adjoints of k1..k4
        self.DP, self.p = z(P), z(P)
        self.src = z(I)
        self.work = torch.empty(33 * I[0] * I[1] * I[2], dtype=dt_, device=dev)
        self.sfs = [z(P) for _ in range(6)] if closure else None
        self.taubar = [z(P) for _ in range(6)] if closure else None
        self.gbar = torch.zeros((9,) + P, dtype=dt_, device=dev) if closure else None
        self.f = [z(P) for _ in range(3)] if closure else None
        self.wpart = (torch.zeros(box.closure_vjp_waves() * 746, dtype=torch.float64, device=dev)
                      if closure == 'mlp' else None)
        self.iterations = []

    # -- pieces ----------------------------------------------------------------
    def _load(self, dst, fields):
        for d, s in zip(dst, fields):
            d[2:-2, 2:-2, 2:-2].copy_(s)
            self.box.fill_halo(d)

    def _stresses(self, U, params):
        if self.closure == 'gradient':
            self.box.sfs_gradient(*U, self.sfs)
        elif self.closure == 'mlp':
            if self.mlp_forward == 'fp32':
                self.box.sfs_mlp_f32(*U, params, self.sfs)
            else:
                self.box.sfs_mlp(*U, params, self.sfs)
        return self.sfs

    def _project(self, F):
        res, its, st = self.box.vel_corr(*F, self.DP, self.p, self.src,
                                         _native.PF_POISSON_BICGSTAB, self.rho / self.dt,
                                         self.dt / self.rho, self.tol, self.maxit)
        self.iterations.append(its)

    # -- forward -----------------------------------------------------------------
    def forward(self, u, v, w, params=None):
        """u^{n+1} from u^n (interior arrays); returns new interior arrays."""
        U = self.U
        self._load(U, (u, v, w))
        sfs = self._stresses(U, params) if self.closure else None
        self.box.rk4_step(*U, *self.T, self.dt, self.mu, self.rho, sfs)
        self._project(U)
        return tuple(x[2:-2, 2:-2, 2:-2].clone() for x in U)

    # -- reverse -----------------------------------------------------------------
    def vjp(self, u, v, w, params, gu, gv, gw):
        """(du, dv, dw, dparams) = the step's Jacobian^T applied to (gu, gv, gw)."""
        box, h = self.box, 0.5 * self.dt
        U = self.U
        self._load(U, (u, v, w))
        # lambda* = P lambda (self-adjoint projection)
        self._load(self.lam, (gu, gv, gw))
        self._project(self.lam)
        sfs = self._stresses(U, params) if self.closure else None
        # stage states Y2..Y4 exactly as the forward kernels produce them
        prev = U
        for s in range(3):
            box.rhs_ns(*prev, self.mu, self.rho, *self.k, sfs)
            for q in range(3):
                box_op_zaxpy(box, h, self.k[q], U[q], self.Y[s][q])
            prev = self.Y[s]
        # adjoints of k1..k4 and of u^n
        w6, w3 = self.dt / 6.0, self.dt / 3.0
        for q in range(3):
            lam = self.lam[q]
            self.acc[q].copy_(lam)
            self.lk[3][q].copy_(lam).mul_(w6)
            self.lk[2][q].copy_(lam).mul_(w3)
            self.lk[1][q].copy_(lam).mul_(w3)
            self.lk[0][q].copy_(lam).mul_(w6)
        stages = [U] + self.Y      # Y1 .. Y4
        for s in (3, 2, 1, 0):
            kprev = self.lk[s - 1] if s > 0 else None
            box.rhs_vjp(*stages[s], self.mu, self.rho, self.lk[s], self.work, self.acc, kprev,
                        h if s > 0 else 0.0)
            if self.closure:
                box.sfs_vjp(self.lk[s], self.taubar, accumulate=(s != 3))
        dparams = None
        if self.closure:
            kind = 0 if self.closure == 'gradient' else 1
            if kind == 1:
                dparams = torch.zeros(params.numel(), dtype=torch.float64, device=params.device)
            box.closure_vjp(kind, *U, params if kind == 1 else None, self.taubar, self.gbar,
                            self.f, dparams, self.wpart)
            for q in range(3):
                box.fold_halo(self.f[q])
                self.acc[q].add_(self.f[q])
        out = tuple(a[2:-2, 2:-2, 2:-2].clone() for a in self.acc)
        return out + (dparams,)


def box_op_zaxpy(box, a, X, Y, out):
    """out (padded) = a*X + Y (X interior, Y padded) with the overlap filled:
    the reference's state_P.ZAXPY (state.py:157-165) as the RK stages use it."""
    from .state import field_arg
    from . import _native as N
    xp, xl = field_arg(box, X, 'X')
    yp, yl = field_arg(box, Y, 'Y')
    N.check(N.lib().pf_state_op(box.h, 0, float(a), box.padded(out, 'out'), xp, xl, yp, yl, 1,
                                box.stream()), 'pf_state_op')


class _StepFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stepper, params, u, v, w):
        out = stepper.forward(u, v, w, params)
        ctx.stepper = stepper
        ctx.save_for_backward(u, v, w, params)
        return out

    @staticmethod
    def backward(ctx, gu, gv, gw):
        u, v, w, params = ctx.saved_tensors
        st = ctx.stepper
        g = [x if x is not None else torch.zeros_like(u) for x in (gu, gv, gw)]
        du, dv, dw, dp = st.vjp(u, v, w, params, *g)
        dparams = dp.to(params.dtype) if (dp is not None and params is not None) else None
        return None, dparams, du, dv, dw


def step(stepper, u, v, w, params=None):
    """One differentiable time step on interior velocity arrays."""
    return _StepFunction.apply(stepper, params, u, v, w)


def rollout(stepper, u, v, w, nsteps, params=None):
    """nsteps differentiable time steps (each step checkpoints its input)."""
    for _ in range(nsteps):
        u, v, w = step(stepper, u, v, w, params)
    return u, v, w


def main(argv=None):
    """Time one closure-training window: K checkpointed steps forward and the
    reverse sweep back to u^0 and the MLP parameters, from a Passot-Pouquet
    field (k0 = 4, u_rms = 1, dt at CFL 0.3):
    python -m pyflow_amd.autodiff --N 64 --steps 20 --dtype f32 --closure mlp --mlp-forward bf16"""
    import argparse
    import math
    import numpy as np
    from .initial_conditions import passot_pouquet, cfl_dt
    from .sfsmodel_mlp import init_params
    ap = argparse.ArgumentParser()
    ap.add_argument('--N', type=int, default=64)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--dtype', choices=('f32', 'f64'), default='f32')
    ap.add_argument('--closure', choices=('none', 'gradient', 'mlp'), default='mlp')
    ap.add_argument('--nu', type=float, default=0.005)
    ap.add_argument('--tol', type=float, default=None,
                    help='Poisson tolerance (default 1e-6 in FP32, 1e-12 in FP64)')
    ap.add_argument('--reps', type=int, default=2)
    ap.add_argument('--mlp-forward', choices=('fp32', 'bf16'), default='fp32',
                    help='MLP forward arithmetic: FP32, or BF16 on MFMA (FP32 accumulation)')
    a = ap.parse_args(argv)
    dt_ = torch.float32 if a.dtype == 'f32' else torch.float64
    tol = a.tol if a.tol is not None else (1e-6 if a.dtype == 'f32' else 1e-12)
    closure = None if a.closure == 'none' else a.closure
    n, K = a.N, a.steps
    h = 2 * math.pi / n
    dev = torch.device('cuda')
    u0 = passot_pouquet(n, k0=4.0, urms=1.0, seed=3)
    dt = cfl_dt(*u0, h)
    box = _native.Box(n, n, n, h, h, h, dtype=dt_, device=dev)
    st = StepVJP(box, a.nu, 1.0, dt, closure=closure, tol=tol, mlp_forward=a.mlp_forward)
    rng = np.random.Generator(np.random.PCG64(7))
    c = [torch.tensor(rng.standard_normal((n, n, n)), dtype=dt_, device=dev) for _ in range(3)]
    params = None
    if closure == 'mlp':
        params = torch.tensor(init_params(13), device=dev, requires_grad=True)
    fwd, rev = [], []
    for rep in range(a.reps + 1):
        x = [torch.tensor(t, dtype=dt_, device=dev, requires_grad=True) for t in u0]
        if params is not None:
            params.grad = None
        st.iterations = []
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        torch.cuda.synchronize()
        e[0].record()
        out = rollout(st, *x, K, params)
        loss = sum(torch.sum(o * cc) for o, cc in zip(out, c))
        e[1].record()
        loss.backward()
        e[2].record()
        torch.cuda.synchronize()
        if rep > 0:            # the first window warms up
            fwd.append(e[0].elapsed_time(e[1]))
            rev.append(e[1].elapsed_time(e[2]))
    its = st.iterations
    nf = len(its) // 2
    f, r = min(fwd), min(rev)
    print('closure-training window %d^3 %s closure=%s (MLP forward %s): %d steps, forward '
          '%.1f ms (%.2f ms/step), reverse %.1f ms (%.2f ms/step), window %.1f ms; '
          'Poisson iterations per solve: forward %.1f, reverse %.1f'
          % (n, a.dtype, a.closure, a.mlp_forward, K, f, f / K, r, r / K, f + r,
             sum(its[:nf]) / max(nf, 1), sum(its[nf:]) / max(len(its) - nf, 1)), flush=True)
    return 0


if __name__ == '__main__':
    import sys
    sys.exit(main())
