"""solver/advancer.py:41-113 on the GPU.

With this package's rhs_NavierStokes (or, for the adjoint step,
rhs_adjPredictor) the whole step is fused: each RK4 stage is one kernel
that evaluates the RHS and applies the stage update, and the overlap cells
are written by the same kernel.  Any other RHS class (duck
typed with evaluate() and rhs_u/v/w) runs the reference's sequence with the
state kernels.
"""

import torch

from .adjoint import rhs_adjPredictor
from .velocity import rhs_NavierStokes


def _fusable(r):
    return isinstance(r, rhs_NavierStokes)


def _adjoint(r):
    return isinstance(r, rhs_adjPredictor)


class Euler:
    def __init__(self, state_u, state_v, state_w, rhsClass):
        self.state_u = state_u
        self.state_v = state_v
        self.state_w = state_w
        self.rhsClass = rhsClass

    def step(self, simDt):
        """advancer.py:49-57"""
        r = self.rhsClass.rhs1
        if _adjoint(r):
            kind, params, work, delta, dparams = r.closure_args()
            r.box.adj_euler_step(r.state_u.var, r.state_v.var, r.state_w.var, self.state_u.var,
                                 self.state_v.var, self.state_w.var, simDt, r.mu(), r.rho, kind,
                                 params, work, delta, dparams)
            return
        if _fusable(r) and r.source_model() is not None:
            src = r.source_model()
            S = (self.state_u, self.state_v, self.state_w)
            r.box.source_inputs(*(st.var for st in S), *(st.var_i for st in S))
            src.update(*(st.var_i for st in S))
            r.box.euler_step_src(*(st.var for st in S), simDt, r.mu(), r.rho,
                                 (src.GX, src.GY, src.GZ))
            return
        if _fusable(r):
            vf = r.visc_field()
            if vf is not None:
                r.box.euler_step_visc(self.state_u.var, self.state_v.var, self.state_w.var,
                                      simDt, vf, r.rho, r.sfs_fields())
            else:
                r.box.euler_step(self.state_u.var, self.state_v.var, self.state_w.var,
                                 simDt, r.mu(), r.rho, r.sfs_fields())
            return
        r.evaluate(self.state_u, self.state_v, self.state_w)
        for st, k in ((self.state_u, r.rhs_u), (self.state_v, r.rhs_v), (self.state_w, r.rhs_w)):
            st.ZAXPY(simDt, k, st.interior())


class RK4:
    """advancer.py:75-113.  With a 'source' closure each stage's closure is
    evaluated on that stage's state (stage 1 on state_u, stages 2-4 on
    state_uTmp, whose var_i arrays receive the inputs), each stage's RHS
    object calling its own closure as rhs_NavierStokes.evaluate does."""

    def __init__(self, state_u, state_uTmp, state_v, state_vTmp, state_w, state_wTmp, rhsClass):
        self.state_u = state_u
        self.state_v = state_v
        self.state_w = state_w
        self.state_uTmp = state_uTmp
        self.state_vTmp = state_vTmp
        self.state_wTmp = state_wTmp
        self.rhsClass = rhsClass

    def step(self, simDt):
        """advancer.py:75-113: k1 = R(u^n); Y_{s+1} = u^n + dt/2*k_s (s = 1..3,
        stage 4 included); u = u^n + dt/6*(k1 + 2k2 + 2k3 + k4)."""
        rc = self.rhsClass
        if all(_adjoint(getattr(rc, n)) for n in ('rhs1', 'rhs2', 'rhs3', 'rhs4')):
            r = rc.rhs1
            kind, params, work, delta, dparams = r.closure_args()
            r.box.adj_rk4_step(r.state_u.var, r.state_v.var, r.state_w.var, self.state_u.var,
                               self.state_v.var, self.state_w.var, self.state_uTmp.var,
                               self.state_vTmp.var, self.state_wTmp.var, simDt, r.mu(), r.rho,
                               kind, params, work, delta, dparams)
            return
        if all(_fusable(getattr(rc, n)) for n in ('rhs1', 'rhs2', 'rhs3', 'rhs4')):
            r = rc.rhs1
            if r.source_model() is not None:
                self._step_source(simDt)
                return
            vf = r.visc_field()
            if vf is not None:
                r.box.rk4_step_visc(self.state_u.var, self.state_v.var, self.state_w.var,
                                    self.state_uTmp.var, self.state_vTmp.var,
                                    self.state_wTmp.var, simDt, vf, r.rho, r.sfs_fields())
            else:
                r.box.rk4_step(self.state_u.var, self.state_v.var, self.state_w.var,
                               self.state_uTmp.var, self.state_vTmp.var, self.state_wTmp.var,
                               simDt, r.mu(), r.rho, r.sfs_fields())
            return
        U = (self.state_u, self.state_v, self.state_w)
        Tm = (self.state_uTmp, self.state_vTmp, self.state_wTmp)
        if any(_fusable(getattr(rc, n)) and getattr(rc, n).source_model() is not None
               for n in ('rhs1', 'rhs2', 'rhs3', 'rhs4')):
            raise NotImplementedError("a 'source' closure needs this package's RHS for all stages")
        rc.rhs1.evaluate(*U)
        prev = rc.rhs1
        for nxt in (rc.rhs2, rc.rhs3, rc.rhs4):
            for t, u, k in zip(Tm, U, (prev.rhs_u, prev.rhs_v, prev.rhs_w)):
                t.ZAXPY(0.5 * simDt, k, u.interior())
            nxt.evaluate(*Tm)
            prev = nxt
        for i, u in enumerate(U):
            ks = [getattr(getattr(rc, n), 'rhs_' + 'uvw'[i]) for n in ('rhs1', 'rhs2', 'rhs3', 'rhs4')]
            acc = ks[0] + 2.0 * ks[1] + 2.0 * ks[2] + ks[3]
            u.ZAXPY(simDt / 6.0, acc, u.interior())

    def _step_source(self, simDt):
        rc = self.rhsClass
        U = (self.state_u, self.state_v, self.state_w)
        Tm = (self.state_uTmp, self.state_vTmp, self.state_wTmp)
        for stage, name in enumerate(('rhs1', 'rhs2', 'rhs3', 'rhs4'), start=1):
            r = getattr(rc, name)
            src = r.source_model()
            holder = U if stage == 1 else Tm
            I = [st.var_i for st in holder]
            r.box.rk4_source_inputs(stage, *(st.var for st in U), *(st.var for st in Tm), *I)
            src.update(*I)
            r.box.rk4_stage_src(stage, *(st.var for st in U), *(st.var for st in Tm), simDt,
                                r.mu(), r.rho, (src.GX, src.GY, src.GZ))
