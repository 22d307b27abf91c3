"""solver/adjoint.py:37-298 on the GPU: the RHS of the reference's adjoint
step (closure training).

rhs_adjPredictor.evaluate is one fused HIP kernel (pf_rhs_adj): the viscous
operator of the forward RHS applied to the adjoint field plus the 15
discrete-exact convective transposes (metric_staggered.py:529-800) with the
forward state's interpolants, in the reference's op order.  The adjoint
advancer of this package (advancer.RK4 / Euler with an AdjointRHS) fuses
the RHS into its stage kernels (pf_adj_rk4_step / pf_adj_euler_step).

Closure term (adjoint.py:241-286).  The reference's term is the transpose
of its 'source' ML closure, obtained by torch autograd: g = sum G . u_A over
the interior, with G the network's source evaluated on the forward state's
interpolants and u_A the adjoint state's, differentiated with respect to the
forward velocity (the term) and the network parameters (accumulated).  This
package computes the same two derivatives with hand-written reverse kernels
(closure kind 2, pf_sfs_ml_source_vjp inside pf_adj_closure_term): the
network's backward pass on MFMA, the features' and the interpolation's
transposes, recomputed for every stage's adjoint state like the reference's.
Its parameter gradient accumulates in FP64 into sfsmodel.grad and the
closure's finalize() scales it by -simDt as the reference does.

For the 'tensor' closures of this package (gradient model, MLP; kinds 0 and
1) the term is the exact transpose of the closure's share of the forward
RHS, delta = (d tau/du)^T (d R/d tau)^T u_adj, also from hand-written
reverse kernels.  For the MLP each evaluation accumulates the parameter
gradient d(sum tau . taubar)/d theta into sfsmodel.grad; its finalize()
scales it by +simDt (the stresses enter the RHS with a plus sign, the
'source' closure with a minus sign).
"""

import torch

from .velocity import constant_viscosity


def closure_kind(sfsmodel):
    """-1 (no closure term), 0 (gradient model), 1 (MLP) or 2 ('source' ML
    closure, sfsmodel_ML)."""
    mt = getattr(sfsmodel, 'modelType', None)
    if mt == 'source':
        return 2
    if sfsmodel is None or mt != 'tensor':
        return -1
    return 1 if hasattr(sfsmodel, 'params') else 0


class AdjointScratch:
    """Device work of the closure term: the three padded delta arrays and,
    for the tensor closures, their reverse-pass workspace (the 'source'
    closure's lives in the box, reserved by the closure), allocated once per
    box."""

    def __init__(self, box, decomp, kind=0):
        self.work = None
        if kind < 2:
            self.work = torch.empty(box.adj_work_bytes(), dtype=torch.uint8,
                                    device=decomp.device)
        shape = (decomp.nxo_, decomp.nyo_, decomp.nzo_)
        self.delta = [torch.zeros(shape, dtype=decomp.prec, device=decomp.device) for _ in range(3)]


class rhs_adjPredictor:
    def __init__(self, decomp, metric, rho, VISC, sfsmodel, state_u, state_v, state_w,
                 scratch=None):
        self.decomp = decomp
        self.box = decomp.box
        self.metric = metric
        self.rho = rho
        self.VISC = VISC
        self.sfsmodel = sfsmodel
        self.state_u = state_u
        self.state_v = state_v
        self.state_w = state_w
        self.kind = closure_kind(sfsmodel) if getattr(sfsmodel, 'use', True) else -1
        self.scratch = scratch
        shape = (decomp.nx_, decomp.ny_, decomp.nz_)
        self.rhs_u = torch.zeros(shape, dtype=decomp.prec, device=decomp.device)
        self.rhs_v = torch.zeros(shape, dtype=decomp.prec, device=decomp.device)
        self.rhs_w = torch.zeros(shape, dtype=decomp.prec, device=decomp.device)

    def mu(self):
        """The constant viscosity, or the padded VISC field when it varies
        (the kernels then read it per cell, as the forward ones do)."""
        mu = constant_viscosity(self.VISC)
        return self.VISC if mu is None else mu

    def closure_args(self):
        """(kind, params, work, delta3, dparams) for the fused kernels."""
        if self.kind < 0:
            return -1, None, None, None, None
        if self.scratch is None:
            self.scratch = AdjointScratch(self.box, self.decomp, self.kind)
        s = self.sfsmodel
        params = getattr(s, 'params', None) if self.kind >= 1 else None
        dparams = s.param_grad() if self.kind >= 1 and hasattr(s, 'param_grad') else None
        return self.kind, params, self.scratch.work, self.scratch.delta, dparams

    def evaluate(self, state_u_adj, state_v_adj, state_w_adj):
        """adjoint.py:100-298"""
        kind, params, work, delta, dparams = self.closure_args()
        U = (self.state_u.var, self.state_v.var, self.state_w.var)
        A = (state_u_adj.var, state_v_adj.var, state_w_adj.var)
        if kind >= 0:
            self.box.adj_closure_term(kind, *U, params, *A, work, delta, dparams)
        self.box.rhs_adj(*U, *A, self.mu(), self.rho, self.rhs_u, self.rhs_v, self.rhs_w,
                         delta if kind >= 0 else None)


class AdjointRHS:
    """adjoint.py:40-50: one rhs_adjPredictor per stage (sharing the closure
    scratch)."""

    def __init__(self, inputConfig, decomp, metric, rho, VISC, sfsmodel, state_u, state_v,
                 state_w):
        self.rhs1 = rhs_adjPredictor(decomp, metric, rho, VISC, sfsmodel, state_u, state_v, state_w)
        if inputConfig.advancerName[:-1] == 'RK':
            sc = None
            if self.rhs1.kind >= 0:
                sc = AdjointScratch(decomp.box, decomp, self.rhs1.kind)
                self.rhs1.scratch = sc
            self.rhs2 = rhs_adjPredictor(decomp, metric, rho, VISC, sfsmodel, state_u, state_v,
                                         state_w, sc)
            self.rhs3 = rhs_adjPredictor(decomp, metric, rho, VISC, sfsmodel, state_u, state_v,
                                         state_w, sc)
            self.rhs4 = rhs_adjPredictor(decomp, metric, rho, VISC, sfsmodel, state_u, state_v,
                                         state_w, sc)
