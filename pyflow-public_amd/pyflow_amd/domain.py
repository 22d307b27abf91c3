"""data/domain.py:56-438 on the GPU.

Domain.forwardStep runs the closure kernel (for 'tensor' closures), the
fused advancer and the fused velocity correction: per RK4 step that is four
stage kernels, the BiCGStab kernels and one projection kernel, with no
host round trip except the solver's batched convergence polls.
"""

import numpy as np
import torch

from . import _native
from . import adjoint
from . import advancer
from . import pressure
from . import sfsmodel_gradient
from . import sfsmodel_smagorinsky
from . import state
from . import velocity


class Domain:
    def __init__(self, inputConfig, comms, decomp, data_IC, geometry, metric, names, startTime):
        nx_, ny_, nz_ = decomp.nx_, decomp.ny_, decomp.nz_
        nxo_, nyo_, nzo_ = decomp.nxo_, decomp.nyo_, decomp.nzo_
        self.decomp = decomp
        self.box = decomp.box
        IC_u_np = data_IC[:, :, :, 0]
        IC_v_np = data_IC[:, :, :, 1]
        IC_w_np = data_IC[:, :, :, 2]
        IC_p_np = data_IC[:, :, :, 3]
        simDt = inputConfig.simDt
        self.equationMode = inputConfig.equationMode
        self.metric = metric
        self.useTargetData = getattr(inputConfig, 'useTargetData', False)
        prec = decomp.prec

        # domain.py:101-122
        self.state_u_P = state.state_P(decomp, IC_u_np)
        self.state_v_P = state.state_P(decomp, IC_v_np)
        self.state_w_P = state.state_P(decomp, IC_w_np)
        self.state_p_P = state.state_P(decomp, IC_p_np)
        self.state_DP_P = state.state_P(decomp, IC_p_np)
        self.state_data_all = (self.state_u_P, self.state_v_P, self.state_w_P, self.state_p_P)
        self.data_all_CPU = state.data_all_CPU(decomp, startTime, simDt, names, self.state_data_all)
        if inputConfig.advancerName[:-1] == 'RK':
            self.state_uTmp_P = state.state_P(decomp, IC_u_np)
            self.state_vTmp_P = state.state_P(decomp, IC_v_np)
            self.state_wTmp_P = state.state_P(decomp, IC_w_np)
        else:
            self.state_uTmp_P = self.state_vTmp_P = self.state_wTmp_P = None

        # domain.py:125-131: VISC_P = ones*mu; the tag tells the fused kernels
        # that the field is the constant mu.
        self.mu = inputConfig.mu
        self.rho = inputConfig.rho
        self.source_P = torch.zeros(nx_, ny_, nz_, dtype=prec, device=decomp.device)
        self.VISC_P = torch.ones(nxo_, nyo_, nzo_, dtype=prec, device=decomp.device)
        self.VISC_P.mul_(self.mu)
        self.VISC_P.pf_const_mu = float(self.mu)

        # domain.py:135-191: adjoint state, temporaries and checkpoints
        adjointTraining = getattr(inputConfig, 'adjointTraining', False)
        self.numCheckpointIt = getattr(inputConfig, 'numCheckpointIt', None)
        self.adjointTraining = bool(adjointTraining)
        if adjointTraining:
            if self.numCheckpointIt is None:
                raise Exception("\nAdjoint training requires specification of numCheckpointIt")
            if inputConfig.equationMode != 'NS':
                raise Exception("\nAdjoint training requires Navier-Stokes solver\n")
            zeros = np.zeros((nx_, ny_, nz_))
            self.state_u_adj_P = state.state_P(decomp, zeros)
            self.state_v_adj_P = state.state_P(decomp, zeros)
            self.state_w_adj_P = state.state_P(decomp, zeros)
            self.adjoint_data_all = (self.state_u_adj_P, self.state_v_adj_P, self.state_w_adj_P)
            self.data_adj_CPU = state.data_all_CPU(decomp, startTime, simDt, names[0:3],
                                                   self.adjoint_data_all)
            if inputConfig.advancerName[:-1] == 'RK':
                self.state_uTmp_adj_P = state.state_P(decomp, zeros)
                self.state_vTmp_adj_P = state.state_P(decomp, zeros)
                self.state_wTmp_adj_P = state.state_P(decomp, zeros)
            else:
                self.state_uTmp_adj_P = self.state_vTmp_adj_P = self.state_wTmp_adj_P = None
            # Checkpoints: stored checkpoint-major so each one is a contiguous
            # padded array; check_*_P are views with the reference's index
            # order [nxo, nyo, nzo, numCheckpointIt+1].
            K = self.numCheckpointIt + 1
            self._check = [torch.zeros(K, nxo_, nyo_, nzo_, dtype=prec, device=decomp.device)
                           for _ in range(3)]
            self.check_u_P, self.check_v_P, self.check_w_P = (c.permute(1, 2, 3, 0)
                                                              for c in self._check)
            # domain.py:189-226: target states for training against data
            # files (read per window by the run loop)
            if not getattr(inputConfig, 'adjointVerification', False):
                self.useTargetData = True
                self.numItTargetComp = self.numCheckpointIt
                self.state_u_T = state.state_P(decomp, zeros, need_gradients=False)
                self.state_v_T = state.state_P(decomp, zeros, need_gradients=False)
                self.state_w_T = state.state_P(decomp, zeros, need_gradients=False)
                self.target_data_all = (self.state_u_T, self.state_v_T, self.state_w_T)
                # all the file's names (the header size), three states read
                self.target_data_all_CPU = state.data_all_CPU(decomp, startTime, simDt, names,
                                                              self.target_data_all)

        # domain.py:231-251
        self.use_SFSmodel = True
        SFSmodelName = getattr(inputConfig, 'SFSmodel', 'none')
        if SFSmodelName == 'Smagorinsky':
            self.sfsmodel = sfsmodel_smagorinsky.stress_constCs(inputConfig, geometry)
        elif SFSmodelName == 'gradient':
            self.sfsmodel = sfsmodel_gradient.residual_stress(decomp, geometry, metric)
        elif SFSmodelName == 'MLP':
            from . import sfsmodel_mlp
            self.sfsmodel = sfsmodel_mlp.residual_stress(inputConfig, decomp, geometry, metric)
        elif SFSmodelName == 'ML':
            from . import sfsmodel_ML
            self.sfsmodel = sfsmodel_ML.residual_stress(inputConfig, decomp, geometry)
        else:
            self.use_SFSmodel = False
            self.sfsmodel = sfsmodel_smagorinsky.stress_constCs(inputConfig, geometry)
        if self.sfsmodel.modelType == 'eddyVisc':
            self.muMolec = self.mu

        # domain.py:259-325
        self.forwardRHS = velocity.ForwardRHS(inputConfig, decomp, metric, self.rho, self.VISC_P,
                                              self.sfsmodel)
        if inputConfig.advancerName == 'Euler':
            self.forwardAdvancer = advancer.Euler(self.state_u_P, self.state_v_P, self.state_w_P,
                                                  self.forwardRHS)
        elif inputConfig.advancerName == 'RK4':
            self.forwardAdvancer = advancer.RK4(self.state_u_P, self.state_uTmp_P, self.state_v_P,
                                                self.state_vTmp_P, self.state_w_P,
                                                self.state_wTmp_P, self.forwardRHS)
        else:
            raise Exception('\n Advancer type ' + inputConfig.advancerName +
                            ' unknown! Options are Euler and RK4.')

        # domain.py:263-288
        if adjointTraining:
            self.adjointRHS = adjoint.AdjointRHS(inputConfig, decomp, metric, self.rho,
                                                 self.VISC_P, self.sfsmodel if self.use_SFSmodel
                                                 else None, self.state_u_P, self.state_v_P,
                                                 self.state_w_P)
            if inputConfig.advancerName == 'Euler':
                self.adjointAdvancer = advancer.Euler(self.state_u_adj_P, self.state_v_adj_P,
                                                      self.state_w_adj_P, self.adjointRHS)
            else:
                self.adjointAdvancer = advancer.RK4(self.state_u_adj_P, self.state_uTmp_adj_P,
                                                    self.state_v_adj_P, self.state_vTmp_adj_P,
                                                    self.state_w_adj_P, self.state_wTmp_adj_P,
                                                    self.adjointRHS)

        self.doPressure = False
        self.max_resP = 0.0
        if inputConfig.equationMode == 'NS':
            self.targetDataVelCorr = getattr(inputConfig, 'targetDataVelCorr', False)
            self.doPressure = True
            mode = inputConfig.pSolverMode
            if mode == 'Jacobi':
                self.poisson = pressure.solver_jacobi(comms, decomp, metric, geometry,
                                                      inputConfig.rho, simDt,
                                                      inputConfig.max_pressure_iterations)
            elif mode == 'bicgstab':
                self.poisson = pressure.solver_bicgstab(comms, decomp, metric, geometry,
                                                        inputConfig.rho, simDt,
                                                        inputConfig.min_pressure_residual,
                                                        inputConfig.max_pressure_iterations)
            elif mode == 'bicgstab_mpir':
                self.poisson = pressure.solver_bicgstab_mpir(
                    comms, decomp, metric, geometry, inputConfig.rho, simDt,
                    inputConfig.min_pressure_residual, inputConfig.max_pressure_iterations,
                    getattr(inputConfig, 'mpir_inner_tol', 1e-1),
                    getattr(inputConfig, 'mpir_inner_maxit', 500),
                    getattr(inputConfig, 'mpir_outer_maxit', 100))
            elif mode == 'RedBlackGS':
                raise Exception('\nRed-black GS not yet implemented\n')
            elif mode == 'none':
                self.doPressure = False
            else:
                raise Exception('unknown pSolverMode ' + str(mode))

    def forwardStep(self, simDt):
        """domain.py:342-367"""
        if self.use_SFSmodel:
            if self.sfsmodel.modelType == 'eddyVisc':
                if hasattr(self.sfsmodel, 'viscosity'):
                    self.sfsmodel.viscosity(self.state_u_P, self.state_v_P, self.state_w_P,
                                            self.rho, self.muMolec, self.VISC_P)
                else:
                    muEddy = self.sfsmodel.eddyVisc(self.state_u_P, self.state_v_P,
                                                    self.state_w_P, self.rho, self.metric)
                    self.VISC_P.copy_(self.muMolec + muEddy)
                self.VISC_P.pf_const_mu = None
            elif self.sfsmodel.modelType == 'tensor':
                self.sfsmodel.update(self.state_u_P, self.state_v_P, self.state_w_P, self.metric)
        self.forwardAdvancer.step(simDt)
        if self.doPressure:
            self.velCorr(simDt)

    def velCorr(self, simDt):
        """domain.py:373-402 as one fused call: divergence, solve, projection."""
        P = self.poisson
        if P.kind != _native.PF_POISSON_JACOBI:
            res, its, st = self.box.vel_corr(self.state_u_P.var, self.state_v_P.var,
                                             self.state_w_P.var, self.state_DP_P.var,
                                             self.state_p_P.var, self.source_P, P.kind,
                                             P.rho / P.simDt, simDt / self.rho, P.min_residual,
                                             P.max_iterations)
            P.iterations = its
            P.status = _native.SOLVE_STATUS.get(st, st)
        else:
            res, its, st = self.box.vel_corr(self.state_u_P.var, self.state_v_P.var,
                                             self.state_w_P.var, self.state_DP_P.var,
                                             self.state_p_P.var, self.source_P, P.kind,
                                             P.rho / P.simDt, simDt / self.rho, 0.0,
                                             P.Num_pressure_iterations)
        self.max_resP = res
        return res

    def checkpoint(self, k):
        """Store the forward velocity as checkpoint k (PyFlow.py:333-336, 364-367)."""
        for c, st in zip(self._check, (self.state_u_P, self.state_v_P, self.state_w_P)):
            c[k].copy_(st.var)

    def restore(self, k):
        """Load checkpoint k into the forward velocity (PyFlow.py:548-550, 601-603)."""
        for c, st in zip(self._check, (self.state_u_P, self.state_v_P, self.state_w_P)):
            st.var.copy_(c[k])

    def adjointStep(self, simDt):
        """domain.py:409-438: adjoint projection (the forward correction with
        +simDt/rho on the adjoint field, one fused call), then the adjoint
        advancer with the forward state held where the caller left it."""
        if self.doPressure:
            P = self.poisson
            A = (self.state_u_adj_P.var, self.state_v_adj_P.var, self.state_w_adj_P.var)
            if P.kind != _native.PF_POISSON_JACOBI:
                res, its, st = self.box.vel_corr(*A, self.state_DP_P.var, self.state_p_P.var,
                                                 self.source_P, P.kind, P.rho / P.simDt,
                                                 +simDt / self.rho, P.min_residual,
                                                 P.max_iterations)
                P.iterations = its
                P.status = _native.SOLVE_STATUS.get(st, st)
            else:
                res, its, st = self.box.vel_corr(*A, self.state_DP_P.var, self.state_p_P.var,
                                                 self.source_P, P.kind, P.rho / P.simDt,
                                                 +simDt / self.rho, 0.0,
                                                 P.Num_pressure_iterations)
            self.max_resP = res
        self.adjointAdvancer.step(simDt)

    def TargetDataVelCorr(self, simDt):
        """domain.py:445-470: project the target velocity (the correction with
        +simDt/rho, as the reference writes it), if targetDataVelCorr is set."""
        if not self.targetDataVelCorr:
            return
        P = self.poisson
        T = (self.state_u_T.var, self.state_v_T.var, self.state_w_T.var)
        if P.kind != _native.PF_POISSON_JACOBI:
            res, its, st = self.box.vel_corr(*T, self.state_DP_P.var, self.state_p_P.var,
                                             self.source_P, P.kind, P.rho / P.simDt,
                                             +simDt / self.rho, P.min_residual, P.max_iterations)
            P.iterations = its
            P.status = _native.SOLVE_STATUS.get(st, st)
        else:
            res, its, st = self.box.vel_corr(*T, self.state_DP_P.var, self.state_p_P.var,
                                             self.source_P, P.kind, P.rho / P.simDt,
                                             +simDt / self.rho, 0.0, P.Num_pressure_iterations)
        self.max_resP = res

    def stats(self):
        """Monitor values of core/PyFlow.py:378-395 in one kernel; the
        divergence is left in source_P as the reference's loop leaves it."""
        mu_, mv, mw, tke, mdiv = self.box.stats(self.state_u_P.var, self.state_v_P.var,
                                                self.state_w_P.var, self.source_P)
        return dict(maxU=mu_, maxV=mv, maxW=mw, TKE=tke, maxDiv=mdiv, max_resP=self.max_resP)
