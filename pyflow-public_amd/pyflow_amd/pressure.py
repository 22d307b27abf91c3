"""solver/pressure.py:44-297 on the GPU: the 7-point operator, Jacobi and
BiCGStab run as HIP kernels with device-resident Krylov scalars."""

from . import _native


class Pressure_Base:
    """pressure.py:44-84"""

    def __init__(self, decomp, metric):
        self.decomp = decomp
        self.box = decomp.box
        self.metric = metric

    def applyOperator(self, A, B):
        """B (interior) = (-vol*L7) A after the overlap fill of A (pressure.py:54-84)."""
        self.decomp.communicate_border(A)
        self.box.apply_laplacian(A, B)


class solver_jacobi:
    """pressure.py:92-161"""

    def __init__(self, comms, decomp, metric, geo, rho, simDt, Num_pressure_iterations):
        self.comms = comms
        self.pressure = Pressure_Base(decomp, metric)
        self.box = decomp.box
        self.Num_pressure_iterations = Num_pressure_iterations
        self.dx = geo.dx
        self.rho = rho
        self.simDt = simDt
        self.kind = _native.PF_POISSON_JACOBI

    def solve(self, state_pOld_P, state_p_P, source_P):
        return self.box.jacobi(state_pOld_P.var, state_p_P.var, source_P, self.rho, self.simDt,
                               self.Num_pressure_iterations)


class solver_bicgstab:
    """pressure.py:164-297.  After solve(), `iterations` and `status` hold the
    iteration count and the stop reason (the reference returns only max_res)."""

    def __init__(self, comms, decomp, metric, geo, rho, simDt, min_residual, max_iterations):
        self.comms = comms
        self.pressure = Pressure_Base(decomp, metric)
        self.box = decomp.box
        self.min_residual = min_residual
        self.max_iterations = max_iterations
        self.dx = geo.dx
        self.rho = rho
        self.simDt = simDt
        self.vol = metric.vol
        self.kind = _native.PF_POISSON_BICGSTAB
        self.iterations = 0
        self.status = None

    def solve(self, state_DP_P, state_p_P, source_P):
        res, its, st = self.box.bicgstab(state_DP_P.var, state_p_P.var, source_P, self.rho,
                                         self.simDt, self.min_residual, self.max_iterations)
        self.iterations = its
        self.status = _native.SOLVE_STATUS.get(st, st)
        return res
