"""sfsmodel/sfsmodel_smagorinsky.py:39-60.  Used here as the blank closure
object Domain builds when SFSmodel is 'none' (domain.py:244-247); the
eddy-viscosity evaluation itself is not part of this build."""


class stress_constCs:
    def __init__(self, inputConfig, geo):
        self.modelType = 'eddyVisc'
        self.Cs = getattr(inputConfig, 'Cs', 0.18)
        expFilterFacIn = getattr(inputConfig, 'expFilterFac', 1.0)
        if geo.type == 'uniform':
            self.Delta = geo.dx * expFilterFacIn
        else:
            raise Exception('sfsmodel_smagorinsky: geometry type not implemented')

    def eddyVisc(self, state_u, state_v, state_w, rho, metric):
        raise NotImplementedError('Smagorinsky eddy viscosity is not part of this build')
