"""solver/velocity.py:45-76 and :189-383 on the GPU.

rhs_NavierStokes.evaluate is one fused HIP kernel (pf_rhs_ns, or
pf_rhs_ns_visc for a spatially varying VISC field); the RK4 and Euler
advancers of this package do not call it at all but fuse the same RHS into
their stage kernels (pf_rk4_step / pf_euler_step and their _visc variants).

A 'source' closure (velocity.py:360-383) is evaluated inside every RHS
evaluation, as the reference does: pf_source_inputs writes the closure's
inputs into state.var_i exactly as the reference's interpolations leave them,
the closure's update() fills GX, GY, GZ, and the fused kernel subtracts their
face interpolants (pf_rhs_ns_src; the advancers run pf_rk4_stage_src /
pf_euler_step_src stage by stage with the closure evaluated in between).
"""

import torch


def constant_viscosity(VISC):
    """The molecular viscosity of a constant VISC field: a float, or the
    padded tensor Domain builds as ones*mu (domain.py:130-131), which it tags
    with `pf_const_mu`.  Returns None for a varying (eddy-viscosity) field."""
    if isinstance(VISC, (int, float)):
        return float(VISC)
    return getattr(VISC, 'pf_const_mu', None)


class rhs_NavierStokes:
    def __init__(self, decomp, metric, rho, VISC, sfsmodel):
        self.decomp = decomp
        self.box = decomp.box
        self.metric = metric
        self.rho = rho
        self.VISC = VISC
        self.sfsmodel = sfsmodel
        self._shape = (decomp.nx_, decomp.ny_, decomp.nz_)
        self._rhs = None

    def _arrays(self):
        if self._rhs is None:
            self._rhs = [torch.zeros(self._shape, dtype=self.decomp.prec, device=self.decomp.device)
                         for _ in range(3)]
        return self._rhs

    rhs_u = property(lambda self: self._arrays()[0])
    rhs_v = property(lambda self: self._arrays()[1])
    rhs_w = property(lambda self: self._arrays()[2])

    def mu(self):
        mu = constant_viscosity(self.VISC)
        if mu is None:
            raise NotImplementedError('this operation needs a constant viscosity; VISC is a '
                                      'spatially varying field')
        return mu

    def visc_field(self):
        """The padded VISC field when it varies in space (eddy viscosity), else None."""
        return self.VISC if constant_viscosity(self.VISC) is None else None

    def sfs_fields(self):
        """The six stresses of a 'tensor' closure, or None (velocity.py:262-315)."""
        m = getattr(self.sfsmodel, 'modelType', None)
        if m == 'tensor':
            s = self.sfsmodel
            return (s.FXX, s.FYY, s.FZZ, s.FXY, s.FXZ, s.FYZ)
        return None

    def source_model(self):
        """The closure when it is a 'source' model (velocity.py:360-383), else None."""
        if getattr(self.sfsmodel, 'modelType', None) == 'source':
            return self.sfsmodel
        return None

    def evaluate(self, state_u, state_v, state_w):
        """velocity.py:221-383"""
        ru, rv, rw = self._arrays()
        src = self.source_model()
        if src is not None:
            self.box.source_inputs(state_u.var, state_v.var, state_w.var,
                                   state_u.var_i, state_v.var_i, state_w.var_i)
            src.update(state_u.var_i, state_v.var_i, state_w.var_i)
            self.box.rhs_ns_src(state_u.var, state_v.var, state_w.var, self.mu(), self.rho,
                                (src.GX, src.GY, src.GZ), ru, rv, rw)
            return
        vf = self.visc_field()
        if vf is not None:
            self.box.rhs_ns_visc(state_u.var, state_v.var, state_w.var, vf, self.rho,
                                 ru, rv, rw, self.sfs_fields())
        else:
            self.box.rhs_ns(state_u.var, state_v.var, state_w.var, self.mu(), self.rho,
                            ru, rv, rw, self.sfs_fields())


class ForwardRHS:
    """velocity.py:45-76 (Navier-Stokes mode): one RHS object per stage."""

    def __init__(self, inputConfig, decomp, metric, rho, VISC, sfsmodel):
        mode = getattr(inputConfig, 'equationMode', 'NS')
        if mode != 'NS':
            raise NotImplementedError("equationMode '%s': only 'NS' is part of this build" % mode)
        self.rhs1 = rhs_NavierStokes(decomp, metric, rho, VISC, sfsmodel)
        if inputConfig.advancerName[:-1] == 'RK':
            self.rhs2 = rhs_NavierStokes(decomp, metric, rho, VISC, sfsmodel)
            self.rhs3 = rhs_NavierStokes(decomp, metric, rho, VISC, sfsmodel)
            self.rhs4 = rhs_NavierStokes(decomp, metric, rho, VISC, sfsmodel)
