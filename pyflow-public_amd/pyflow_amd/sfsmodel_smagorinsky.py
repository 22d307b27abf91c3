"""sfsmodel/sfsmodel_smagorinsky.py:39-88: constant-coefficient Smagorinsky
eddy viscosity, evaluated by one HIP kernel (pf_sfs_smagorinsky) from the
nine cell-centred gradients of grad_vel_center.  Domain also builds this
object as the blank closure when SFSmodel is 'none' (domain.py:244-247)."""

import torch


class stress_constCs:
    def __init__(self, inputConfig, geo):
        self.modelType = 'eddyVisc'
        self.Cs = getattr(inputConfig, 'Cs', 0.18)
        expFilterFacIn = getattr(inputConfig, 'expFilterFac', 1.0)
        if geo.type == 'uniform':
            self.Delta = geo.dx * expFilterFacIn
        else:
            raise Exception('sfsmodel_smagorinsky: geometry type not implemented')
        self.box = geo.decomp.box
        self._mut = None

    def viscosity(self, state_u, state_v, state_w, rho, mu, VISC):
        """VISC = mu + eddyVisc(...) in one kernel (what Domain.forwardStep
        forms with VISC_P.copy_(muMolec + muEddy), domain.py:348-352)."""
        self.box.sfs_smagorinsky(state_u.var, state_v.var, state_w.var, rho, self.Cs, self.Delta,
                                 mu, VISC)

    def eddyVisc(self, state_u, state_v, state_w, rho, metric):
        """sfsmodel_smagorinsky.py:65-88: mu_t = rho*(Cs*Delta)^2*|S| (padded)."""
        if self._mut is None or self._mut.shape != state_u.var.shape:
            self._mut = torch.empty_like(state_u.var)
        self.viscosity(state_u, state_v, state_w, rho, 0.0, self._mut)
        return self._mut
