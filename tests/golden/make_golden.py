"""Write the golden input/output vectors under tests/golden/.

Run once in a container that holds the reference tree (read from
/root/reference/src); the fixtures it writes (reference_vectors.npz) are
committed and are all that the tests read.  It imports only the composite
functions that no closed form or textbook routine pins on its own:

* rhs_NavierStokes.evaluate (solver/velocity.py:221-383), with no closure and
  with the gradient closure (sfsmodel/sfsmodel_gradient.py:70-148);
* metric_uniform.grad_vel_center (metric/metric_staggered.py:215-268);
* RK4.step and Euler.step (solver/advancer.py:49-113);
* solver_bicgstab.solve with its residual history (solver/pressure.py:196-297);
* solver_jacobi.solve (solver/pressure.py:115-161);
* Domain.forwardStep with RK4 + BiCGStab and Euler + Jacobi
  (data/domain.py:342-402);
* rhs_adjPredictor.evaluate (solver/adjoint.py:94-298) and Domain.adjointStep
  (data/domain.py:409-438) with RK4 + BiCGStab and Euler + Jacobi, the
  closure term fed by a stand-in 'source' closure whose stresses are zero
  (the reference's ML closure architecture is not part of the build);
* stress_constCs.eddyVisc (sfsmodel/sfsmodel_smagorinsky.py:65-88), the RHS
  with the resulting viscosity field, and Domain.forwardStep with
  SFSmodel 'Smagorinsky' (data/domain.py:348-352);
* the adjoint RHS and an adjoint RK4 step with that viscosity field;
* the 'source' closure contract (solver/velocity.py:360-383): the RHS, an RK4
  step and an Euler step with a recording stand-in closure
  (source_stub.RecordingSource), and the arrays the reference hands it;
* the reference's 'source' ML closure itself (sfsmodel/sfsmodel_ML.py:354-593,
  NeuralNetworkModel2 with the seeded weights of source_stub.ml_source_params)
  on those arrays: GX, GY, GZ and the network's output channels 0-8;
* rhs_adjPredictor.evaluate with that ML closure (the closure term of
  adjoint.py:230-286 by the reference's own autograd): the adjoint RHS and
  the parameter gradients the evaluation leaves in the model.

The reference's MPI layer is replaced by a single-rank periodic stub with the
semantics of library/parallel.py:301-414 on a periodic Cartesian communicator
(the reference module itself imports mpi4py, which is not installed).

Grids are deliberately non-cubic with different spacings per axis, so that a
transposed axis or a swapped spacing fails the comparison.
"""

import io
import math
import os
import sys
import types
import contextlib

import numpy as np
import torch

REF = '/root/reference/src'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'reference_vectors.npz')


class StubDecomp:
    """Single-rank, periodic stand-in for parallel.decomp (parallel.py:118-414)."""

    def __init__(self, nx, ny, nz, prec):
        self.prec = prec
        self.dtypeNumpy = np.float64 if prec == torch.float64 else np.float32
        self.device = torch.device('cpu')
        self.rank = 0
        self.nover = 2
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nx_, self.ny_, self.nz_ = nx, ny, nz
        self.nxo_, self.nyo_, self.nzo_ = nx + 4, ny + 4, nz + 4
        self.imino_ = self.jmino_ = self.kmino_ = 0
        self.imaxo_, self.jmaxo_, self.kmaxo_ = nx + 3, ny + 3, nz + 3
        self.imin_ = self.jmin_ = self.kmin_ = 2
        self.imax_, self.jmax_, self.kmax_ = nx + 1, ny + 1, nz + 1
        self.imin_loc = self.jmin_loc = self.kmin_loc = 0
        self.imax_loc, self.jmax_loc, self.kmax_loc = nx - 1, ny - 1, nz - 1
        self.npx = self.npy = self.npz = 1

    def communicate_border(self, A):
        no = 2
        n1, n2, n3 = A.shape
        A[n1 - no:n1].copy_(A[no:2 * no].clone())
        A[0:no].copy_(A[n1 - 2 * no:n1 - no].clone())
        A[:, n2 - no:n2].copy_(A[:, no:2 * no].clone())
        A[:, 0:no].copy_(A[:, n2 - 2 * no:n2 - no].clone())
        A[:, :, n3 - no:n3].copy_(A[:, :, no:2 * no].clone())
        A[:, :, 0:no].copy_(A[:, :, n3 - 2 * no:n3 - no].clone())


class ZeroSourceClosure:
    """A 'source' closure whose outputs GX, GY, GZ are 0*input: it keeps
    the autograd graph of rhs_adjPredictor's closure term (adjoint.py:241-286)
    intact and contributes zeros."""
    modelType = 'source'

    def update(self, u_V, v_V, w_V, requires_grad=True):
        self.GX = (0.0 * u_V)[..., None]
        self.GY = (0.0 * v_V)[..., None]
        self.GZ = (0.0 * w_V)[..., None]


class StubComms:
    rank = 0
    size = 1

    def parallel_sum(self, x):
        return x

    def parallel_max(self, x):
        return x

    def parallel_min(self, x):
        return x


def stub_geometry(d, L):
    """geometry.uniform (geometry.py:76-116) without its dataReader import."""
    g = types.SimpleNamespace()
    g.type = 'uniform'
    g.device = d.device
    g.prec = d.prec
    g.dtypeNumpy = d.dtypeNumpy
    g.Nx, g.Ny, g.Nz = d.nx, d.ny, d.nz
    g.Lx, g.Ly, g.Lz = L
    g.dx = g.Lx / float(g.Nx); g.dxi = 1.0 / g.dx
    g.dy = g.Ly / float(g.Ny); g.dyi = 1.0 / g.dy
    g.dz = g.Lz / float(g.Nz); g.dzi = 1.0 / g.dz
    for a in ('nover', 'nxo_', 'nx_', 'nyo_', 'ny_', 'nzo_', 'nz_', 'imino_', 'jmino_',
              'kmino_', 'imaxo_', 'jmaxo_', 'kmaxo_', 'imin_', 'imax_', 'jmin_', 'jmax_',
              'kmin_', 'kmax_'):
        setattr(g, a, getattr(d, a))
    return g


def seeded_velocity(shape, L, seed):
    """A smooth periodic field plus noise; not solenoidal on purpose."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nx, ny, nz = shape
    x = (np.arange(nx) + 0.5) / nx * 2 * math.pi
    y = (np.arange(ny) + 0.5) / ny * 2 * math.pi
    z = (np.arange(nz) + 0.5) / nz * 2 * math.pi
    X, Y, Z = np.meshgrid(x, y, z, indexing='ij')
    out = []
    for c in range(3):
        a = rng.normal(size=6)
        f = (a[0] * np.sin(X + a[1]) * np.cos(2 * Y) + a[2] * np.cos(Z + a[3]) * np.sin(Y)
             + a[4] * np.sin(2 * Z) * np.cos(X + a[5]))
        f = f + 0.3 * rng.normal(size=shape)
        out.append(f.astype(np.float64))
    return out


def main():
    for sub in ('data', 'solver', 'metric', 'sfsmodel'):
        sys.path.insert(0, os.path.join(REF, sub))
    import metric_staggered
    import state
    import velocity
    import advancer
    import pressure
    import sfsmodel_gradient
    import sfsmodel_smagorinsky
    import domain
    import adjoint

    torch.set_num_threads(1)
    prec = torch.float64
    shape = (8, 10, 12)
    L = (2 * math.pi, 1.7 * math.pi, 2.3 * math.pi)
    rho, mu, dt = 1.1, 0.013, 0.004
    out = {'shape': np.array(shape), 'L': np.array(L), 'rho': rho, 'mu': mu, 'dt': dt}

    d = StubDecomp(*shape, prec)
    geo = stub_geometry(d, L)
    metric = metric_staggered.metric_uniform(geo)
    comms = StubComms()

    u0, v0, w0 = seeded_velocity(shape, L, 1234)
    out['u0'], out['v0'], out['w0'] = u0, v0, w0

    def states(a, b, c):
        return state.state_P(d, a.copy()), state.state_P(d, b.copy()), state.state_P(d, c.copy())

    VISC = torch.ones(d.nxo_, d.nyo_, d.nzo_, dtype=prec)
    VISC.mul_(mu)

    # 1. RHS without closure
    su, sv, sw = states(u0, v0, w0)
    blank = sfsmodel_smagorinsky.stress_constCs(types.SimpleNamespace(), geo)
    r = velocity.rhs_NavierStokes(d, metric, rho, VISC, blank)
    r.evaluate(su, sv, sw)
    out['rhs_u'], out['rhs_v'], out['rhs_w'] = (r.rhs_u.numpy().copy(), r.rhs_v.numpy().copy(),
                                                r.rhs_w.numpy().copy())

    # 2. grad_vel_center and the gradient closure
    su, sv, sw = states(u0, v0, w0)
    gm = sfsmodel_gradient.residual_stress(d, geo, metric)
    gm.update(su, sv, sw, metric)
    for name, s in (('u', su), ('v', sv), ('w', sw)):
        for ax in ('x', 'y', 'z'):
            out['gc_%s_%s' % (name, ax)] = getattr(s, 'grad_' + ax).numpy().copy()
    for n in ('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ'):
        out['sfs_' + n] = getattr(gm, n).numpy().copy()

    # 3. RHS with the gradient closure
    su, sv, sw = states(u0, v0, w0)
    r = velocity.rhs_NavierStokes(d, metric, rho, VISC, gm)
    r.evaluate(su, sv, sw)
    out['rhs_sfs_u'], out['rhs_sfs_v'], out['rhs_sfs_w'] = (
        r.rhs_u.numpy().copy(), r.rhs_v.numpy().copy(), r.rhs_w.numpy().copy())

    # 4. RK4 and Euler steps (no projection)
    cfg = types.SimpleNamespace(advancerName='RK4', equationMode='NS', pSolverMode='bicgstab')
    su, sv, sw = states(u0, v0, w0)
    tu, tv, tw = states(u0, v0, w0)
    frhs = velocity.ForwardRHS(cfg, d, metric, rho, VISC, blank)
    adv = advancer.RK4(su, tu, sv, tv, sw, tw, frhs)
    adv.step(dt)
    out['rk4_u'], out['rk4_v'], out['rk4_w'] = (su.var.numpy().copy(), sv.var.numpy().copy(),
                                                sw.var.numpy().copy())
    su, sv, sw = states(u0, v0, w0)
    cfg = types.SimpleNamespace(advancerName='Euler', equationMode='NS', pSolverMode='bicgstab')
    frhs = velocity.ForwardRHS(cfg, d, metric, rho, VISC, blank)
    adv = advancer.Euler(su, sv, sw, frhs)
    adv.step(dt)
    out['euler_u'], out['euler_v'], out['euler_w'] = (su.var.numpy().copy(), sv.var.numpy().copy(),
                                                      sw.var.numpy().copy())

    # 5. BiCGStab with the residual history (printed by the reference's own
    #    commented-out line; recovered here by wrapping torch.max)
    su, sv, sw = states(u0, v0, w0)
    zeros = np.zeros(shape)
    sDP = state.state_P(d, zeros.copy())
    sp = state.state_P(d, zeros.copy())
    src = torch.zeros(shape, dtype=prec)
    metric.div_vel(su, sv, sw, src)
    out['div0'] = src.numpy().copy()
    solver = pressure.solver_bicgstab(comms, d, metric, geo, rho, dt, 1e-11, 400)
    hist = []
    real_max = torch.max

    def spy_max(t, *a, **k):
        m = real_max(t, *a, **k)
        if t.shape == tuple(shape) and not a and not k:
            hist.append(float(m))
        return m
    torch.max = spy_max
    try:
        res = solver.solve(sDP, sp, src)
    finally:
        torch.max = real_max
    out['bicg_res'] = float(res)
    out['bicg_hist'] = np.array(hist)  # hist[0] = max|r0|, then max|r_k| (unscaled)
    out['bicg_DP'] = sDP.var.numpy().copy()
    out['bicg_p'] = sp.var.numpy().copy()
    out['bicg_source_after'] = src.numpy().copy()

    # 6. Jacobi
    sDP = state.state_P(d, zeros.copy())
    sp = state.state_P(d, zeros.copy())
    src = torch.tensor(out['div0'].copy())
    jac = pressure.solver_jacobi(comms, d, metric, geo, rho, dt, 25)
    out['jac_res'] = float(jac.solve(sDP, sp, src))
    out['jac_pOld'] = sDP.var.numpy().copy()
    out['jac_p'] = sp.var.numpy().copy()

    # 7. Domain.forwardStep: RK4 + BiCGStab, then Euler + Jacobi
    for tag, advn, solv, tol, nit in (('dom_rk4_bicg', 'RK4', 'bicgstab', 1e-12, 500),
                                      ('dom_euler_jac', 'Euler', 'Jacobi', 0.0, 30)):
        cfg = types.SimpleNamespace(simDt=dt, equationMode='NS', advancerName=advn,
                                    mu=mu, rho=rho, dtypeTorch=prec, SFSmodel='none',
                                    pSolverMode=solv, min_pressure_residual=tol,
                                    max_pressure_iterations=nit)
        data_IC = np.stack([u0, v0, w0, zeros], axis=-1)
        with contextlib.redirect_stdout(io.StringIO()):
            D = domain.Domain(cfg, comms, d, data_IC, geo, metric, ['U', 'V', 'W', 'P'], 0.0)
            for _ in range(2):
                D.forwardStep(dt)
        out[tag + '_u'] = D.state_u_P.var.numpy().copy()
        out[tag + '_v'] = D.state_v_P.var.numpy().copy()
        out[tag + '_w'] = D.state_w_P.var.numpy().copy()
        out[tag + '_p'] = D.state_p_P.var.numpy().copy()
        out[tag + '_DP'] = D.state_DP_P.var.numpy().copy()
        out[tag + '_res'] = float(D.max_resP)

    # 8. Adjoint RHS and adjoint steps at the forward state u0
    a0, b0, c0 = seeded_velocity(shape, L, 4321)
    out['adj_u0'], out['adj_v0'], out['adj_w0'] = a0, b0, c0
    zsrc = ZeroSourceClosure()
    su, sv, sw = states(u0, v0, w0)
    au, av, aw = states(a0, b0, c0)
    ar = adjoint.rhs_adjPredictor(d, metric, rho, VISC, zsrc, su, sv, sw)
    ar.evaluate(au, av, aw)
    out['adjrhs_u'], out['adjrhs_v'], out['adjrhs_w'] = (
        ar.rhs_u.numpy().copy(), ar.rhs_v.numpy().copy(), ar.rhs_w.numpy().copy())
    for tag, advn, solv, tol, nit in (('adj_rk4_bicg', 'RK4', 'bicgstab', 1e-12, 500),
                                      ('adj_euler_jac', 'Euler', 'Jacobi', 0.0, 30)):
        cfg = types.SimpleNamespace(advancerName=advn)
        su, sv, sw = states(u0, v0, w0)
        au, av, aw = states(a0, b0, c0)
        tu, tv, tw = states(a0, b0, c0)
        arhs = adjoint.AdjointRHS(cfg, d, metric, rho, VISC, zsrc, su, sv, sw)
        if advn == 'RK4':
            adv = advancer.RK4(au, tu, av, tv, aw, tw, arhs)
            poisson = pressure.solver_bicgstab(comms, d, metric, geo, rho, dt, tol, nit)
        else:
            adv = advancer.Euler(au, av, aw, arhs)
            poisson = pressure.solver_jacobi(comms, d, metric, geo, rho, dt, nit)
        ns = types.SimpleNamespace(
            doPressure=True, metric=metric, state_u_adj_P=au, state_v_adj_P=av,
            state_w_adj_P=aw, source_P=torch.zeros(shape, dtype=prec), poisson=poisson,
            state_DP_P=state.state_P(d, zeros.copy()), state_p_P=state.state_P(d, zeros.copy()),
            rho=rho, adjointAdvancer=adv, max_resP=0.0)
        domain.Domain.adjointStep(ns, dt)
        out[tag + '_u'] = au.var.numpy().copy()
        out[tag + '_v'] = av.var.numpy().copy()
        out[tag + '_w'] = aw.var.numpy().copy()
        out[tag + '_DP'] = ns.state_DP_P.var.numpy().copy()
        out[tag + '_p'] = ns.state_p_P.var.numpy().copy()
        out[tag + '_res'] = float(ns.max_resP)

    # 9. Smagorinsky eddy viscosity (sfsmodel_smagorinsky.py:65-88), the RHS
    #    with the resulting VISC field (velocity.py:248-305 with a field), and
    #    two Domain.forwardSteps with SFSmodel 'Smagorinsky' (domain.py:348-352)
    su, sv, sw = states(u0, v0, w0)
    smag = sfsmodel_smagorinsky.stress_constCs(types.SimpleNamespace(Cs=0.2), geo)
    mut = smag.eddyVisc(su, sv, sw, rho, metric)
    out['smag_mut'] = mut.numpy().copy()
    VISCf = mu + mut
    out['smag_visc'] = VISCf.numpy().copy()
    su, sv, sw = states(u0, v0, w0)
    r = velocity.rhs_NavierStokes(d, metric, rho, VISCf, smag)
    r.evaluate(su, sv, sw)
    out['rhs_smag_u'], out['rhs_smag_v'], out['rhs_smag_w'] = (
        r.rhs_u.numpy().copy(), r.rhs_v.numpy().copy(), r.rhs_w.numpy().copy())
    cfg = types.SimpleNamespace(simDt=dt, equationMode='NS', advancerName='RK4', mu=mu, rho=rho,
                                dtypeTorch=prec, SFSmodel='Smagorinsky', Cs=0.2,
                                pSolverMode='bicgstab', min_pressure_residual=1e-12,
                                max_pressure_iterations=500)
    data_IC = np.stack([u0, v0, w0, zeros], axis=-1)
    with contextlib.redirect_stdout(io.StringIO()):
        D = domain.Domain(cfg, comms, d, data_IC, geo, metric, ['U', 'V', 'W', 'P'], 0.0)
        for _ in range(2):
            D.forwardStep(dt)
    for c in ('u', 'v', 'w', 'p'):
        out['dom_smag_' + c] = getattr(D, 'state_%s_P' % c).var.numpy().copy()
    out['dom_smag_visc'] = D.VISC_P.numpy().copy()
    out['dom_smag_res'] = float(D.max_resP)

    # 10. Adjoint RHS and an adjoint RK4 step (BiCGStab) with the Smagorinsky
    #     VISC field of section 9 in the viscous part
    su, sv, sw = states(u0, v0, w0)
    au, av, aw = states(a0, b0, c0)
    ar = adjoint.rhs_adjPredictor(d, metric, rho, VISCf, zsrc, su, sv, sw)
    ar.evaluate(au, av, aw)
    out['adjrhs_smag_u'], out['adjrhs_smag_v'], out['adjrhs_smag_w'] = (
        ar.rhs_u.numpy().copy(), ar.rhs_v.numpy().copy(), ar.rhs_w.numpy().copy())
    su, sv, sw = states(u0, v0, w0)
    au, av, aw = states(a0, b0, c0)
    tu, tv, tw = states(a0, b0, c0)
    arhs = adjoint.AdjointRHS(types.SimpleNamespace(advancerName='RK4'), d, metric, rho, VISCf,
                              zsrc, su, sv, sw)
    adv = advancer.RK4(au, tu, av, tv, aw, tw, arhs)
    poisson = pressure.solver_bicgstab(comms, d, metric, geo, rho, dt, 1e-12, 500)
    ns = types.SimpleNamespace(
        doPressure=True, metric=metric, state_u_adj_P=au, state_v_adj_P=av, state_w_adj_P=aw,
        source_P=torch.zeros(shape, dtype=prec), poisson=poisson,
        state_DP_P=state.state_P(d, zeros.copy()), state_p_P=state.state_P(d, zeros.copy()),
        rho=rho, adjointAdvancer=adv, max_resP=0.0)
    domain.Domain.adjointStep(ns, dt)
    for c, a in (('u', au), ('v', av), ('w', aw)):
        out['adj_rk4_smag_' + c] = a.var.numpy().copy()
    out['adj_rk4_smag_res'] = float(ns.max_resP)

    # 11. 'source' closure contract with the recording stand-in: the RHS, the
    #     inputs it receives (var_i after the RHS's interpolations), an RK4 step
    #     (closure evaluated in each stage) and an Euler step
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from source_stub import RecordingSource
    su, sv, sw = states(u0, v0, w0)
    stub = RecordingSource()
    r = velocity.rhs_NavierStokes(d, metric, rho, VISC, stub)
    r.evaluate(su, sv, sw)
    out['rhs_src_u'], out['rhs_src_v'], out['rhs_src_w'] = (
        r.rhs_u.numpy().copy(), r.rhs_v.numpy().copy(), r.rhs_w.numpy().copy())
    for c, t in zip('uvw', stub.calls[0]):
        out['src_in_' + c] = t.numpy().copy()
    cfg = types.SimpleNamespace(advancerName='RK4', equationMode='NS', pSolverMode='bicgstab')
    su, sv, sw = states(u0, v0, w0)
    tu, tv, tw = states(u0, v0, w0)
    stub = RecordingSource()
    frhs = velocity.ForwardRHS(cfg, d, metric, rho, VISC, stub)
    advancer.RK4(su, tu, sv, tv, sw, tw, frhs).step(dt)
    for c, s_ in zip('uvw', (su, sv, sw)):
        out['rk4_src_' + c] = s_.var.numpy().copy()
    for c, t in zip('uvw', stub.calls[3]):
        out['rk4_src_in4_' + c] = t.numpy().copy()
    cfg = types.SimpleNamespace(advancerName='Euler', equationMode='NS', pSolverMode='bicgstab')
    su, sv, sw = states(u0, v0, w0)
    stub = RecordingSource()
    frhs = velocity.ForwardRHS(cfg, d, metric, rho, VISC, stub)
    advancer.Euler(su, sv, sw, frhs).step(dt)
    for c, s_ in zip('uvw', (su, sv, sw)):
        out['euler_src_' + c] = s_.var.numpy().copy()

    # 12. the reference's 'source' ML closure on the inputs of section 11
    import sfsmodel_ML
    from source_stub import ml_source_params, ml_state_dict
    d.comm = None
    cfg = types.SimpleNamespace(simDt=dt)
    with contextlib.redirect_stdout(io.StringIO()):
        ml = sfsmodel_ML.residual_stress(cfg, d, geo)
    ml.model.load_state_dict(ml_state_dict(ml_source_params()))

    class Recorder(torch.nn.Module):
        def __init__(self, m):
            super().__init__()
            self.m = m
            self.out = None

        def forward(self, x):
            y = self.m(x)
            self.out = y.detach().clone()
            return y

    ml.model = Recorder(ml.model)
    ml.update(*(torch.tensor(out['src_in_' + c]) for c in 'uvw'))
    out['mlsrc_GX'], out['mlsrc_GY'], out['mlsrc_GZ'] = (
        ml.GX.numpy().copy(), ml.GY.numpy().copy(), ml.GZ.numpy().copy())
    out['mlsrc_p'] = ml.model.out[..., :9].numpy().copy()

    # 13. the adjoint RHS with the ML closure: its closure term
    #     (adjoint.py:230-286) and the parameter gradients g.backward() leaves
    with contextlib.redirect_stdout(io.StringIO()):
        ml2 = sfsmodel_ML.residual_stress(cfg, d, geo)
    ml2.model.load_state_dict(ml_state_dict(ml_source_params()))
    ml2.model.zero_grad()
    su, sv, sw = states(u0, v0, w0)
    au, av, aw = states(a0, b0, c0)
    ar = adjoint.rhs_adjPredictor(d, metric, rho, VISC, ml2, su, sv, sw)
    ar.evaluate(au, av, aw)
    out['adjrhs_ml_u'], out['adjrhs_ml_v'], out['adjrhs_ml_w'] = (
        ar.rhs_u.numpy().copy(), ar.rhs_v.numpy().copy(), ar.rhs_w.numpy().copy())
    grads = []
    for name, _, _ in (('fc1', 0, 0), ('fcG', 0, 0), ('fcG2', 0, 0), ('fc2', 0, 0),
                       ('fc3', 0, 0), ('fc4', 0, 0)):
        layer = getattr(ml2.model, name)
        grads.append(layer.weight.grad.numpy().ravel().copy())
        grads.append(layer.bias.grad.numpy().ravel().copy())
    out['adjrhs_ml_dparams'] = np.concatenate(grads)

    np.savez_compressed(OUT, **out)
    print('wrote', OUT, 'keys:', len(out))


if __name__ == '__main__':
    main()
