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


class solver_bicgstab_mpir(solver_bicgstab):
    """pSolverMode 'bicgstab_mpir' (no counterpart in the reference): the
    same Poisson problem solved by mixed-precision iterative refinement around
    the box's BiCGStab -- FP64 iterate and residual, inner solves in the box
    precision restarted with a fresh shadow residual each refinement step.
    `min_residual` bounds the FP64 residual max|b - A x|/max|b|;
    `max_iterations` caps the inner iterations of one solve.  Optional
    `inner_tol`, `inner_maxit`, `outer_maxit` tune the refinement."""

    def __init__(self, comms, decomp, metric, geo, rho, simDt, min_residual, max_iterations,
                 inner_tol=1e-1, inner_maxit=500, outer_maxit=100):
        super().__init__(comms, decomp, metric, geo, rho, simDt, min_residual, max_iterations)
        self.kind = _native.PF_POISSON_BICGSTAB_MPIR
        self.box.mpir_configure(inner_tol, inner_maxit, outer_maxit)

    def solve(self, state_DP_P, state_p_P, source_P):
        res, its, st = self.box.bicgstab_mpir(state_DP_P.var, state_p_P.var, source_P, self.rho,
                                              self.simDt, self.min_residual, self.max_iterations)
        self.iterations = its
        self.status = _native.SOLVE_STATUS.get(st, st)
        return res
