"""Closure training against filtered-DNS targets (PyFlow.py:471-607 with the
MLP closure): targets from les_targets (64^3 DNS, 4x box filter, 16^3 LES),
the run loop's training windows, and the window gradient the reference's
bookkeeping produces checked against the exact reverse of the same window.

Finding pinned here: the reference accumulates the closure's parameter
gradient once per adjoint RHS evaluation, i.e. four times per RK4 step with
equal weight, instead of with the RK4 weights (1/6, 1/3, 1/3, 1/6).  The
window gradient is therefore ~4x the objective's gradient (same direction);
RMSprop divides the scale out.
"""

import math
import types

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def targets(tmp_path_factory):
    from pyflow_amd import les_targets
    out = str(tmp_path_factory.mktemp('targets'))
    return les_targets.generate(out, N=64, ratio=4, steps=8, every=4, seed=5, nu=0.01,
                                verbose=False)


def les_config(t, numIt, params, LR=0.01, numSumDistr=1, sfs='MLP', training=True):
    key = 'mlpParams' if sfs == 'MLP' else 'mlSourceParams'
    return types.SimpleNamespace(
        configName='restart', configFileStr=t['config'], dataFileStr=t['files'][0],
        dataFileType='restart', nproc_x=1, nproc_y=1, nproc_z=1, mu=0.01, rho=1.0,
        simDt=t['dt_les'], numIt=numIt, SFSmodel=sfs, advancerName='RK4',
        equationMode='NS', pSolverMode='bicgstab', min_pressure_residual=1e-12,
        max_pressure_iterations=3000, dtypeTorch=torch.float64, dtypeNumpy=np.float64,
        adjointTraining=training, numCheckpointIt=4, adjointVerification=False,
        targetFileBaseStr=t['base'], startFileIt=0, LR=LR, numSumDistr=numSumDistr,
        writeMonitor=False, **{key: params})


def test_targets_are_divergence_free(targets):
    from pyflow_amd import dataReader as dr
    from pyflow_amd.les_targets import divergence
    assert len(targets['files']) == 3
    for f in targets['files']:
        names, time, data = dr.readNGArestart(f, headerOnly=False, printOut=False)
        u, v, w = (torch.from_numpy(np.ascontiguousarray(data[..., i])) for i in range(3))
        h = 2 * math.pi / 16
        d = float(divergence(u, v, w, h).abs().max()) * h / float(u.abs().max())
        print('%s: t = %.4e, max|div| h / max|u| = %.2e' % (f, time, d))
        assert d < 1e-10


def test_window_gradient_against_exact_reverse(targets):
    from pyflow_amd import PyFlow, autodiff, dataReader as dr, _native
    from pyflow_amd.sfsmodel_mlp import init_params
    params = init_params(11) * 3.0
    res = PyFlow.simulate(les_config(targets, 4, params, LR=0.0), verbose=False)
    D = res.domain
    g_train = D.sfsmodel.last_window_grad.double().cpu().numpy()
    assert len(res.train_history) == 1
    assert np.isfinite(g_train).all()
    # exact reverse of the same window: J = sum_c mean|u_c - u_T,c| after 4 steps
    _, _, d0 = dr.readNGArestart(targets['files'][0], headerOnly=False, printOut=False)
    _, _, d1 = dr.readNGArestart(targets['files'][1], headerOnly=False, printOut=False)
    h = 2 * math.pi / 16
    box = _native.Box(16, 16, 16, h, h, h, dtype=torch.float64)
    st = autodiff.StepVJP(box, 0.01, 1.0, targets['dt_les'], closure='mlp', tol=1e-12,
                          mlp_forward='bf16')
    u0 = [torch.from_numpy(np.ascontiguousarray(d0[..., i])).cuda() for i in range(3)]
    uT = [torch.from_numpy(np.ascontiguousarray(d1[..., i])).cuda() for i in range(3)]
    p = torch.tensor(params, device='cuda', requires_grad=True)
    u1 = autodiff.rollout(st, *u0, 4, p)
    J = sum((a - b).abs().mean() for a, b in zip(u1, uT))
    J.backward()
    g_exact = p.grad.double().cpu().numpy()
    err0 = res.train_history[0]['model_error']
    cos = float(np.dot(g_train, g_exact) / (np.linalg.norm(g_train) * np.linalg.norm(g_exact)))
    ratio = float(np.dot(g_train, g_exact) / np.dot(g_exact, g_exact))
    print('window objective: run loop %.6e, exact reverse %.6e; gradient cosine %.6f, '
          'ratio %.4f' % (err0, float(J), cos, ratio))
    assert abs(err0 - float(J)) <= 1e-9 * abs(float(J))
    assert cos > 0.99
    assert 3.6 < ratio < 4.4


def test_training_windows_update_the_closure(targets):
    from pyflow_amd import PyFlow
    from pyflow_amd.sfsmodel_mlp import init_params
    params = init_params(11) * 3.0
    res = PyFlow.simulate(les_config(targets, 8, params, LR=1e-3), verbose=False)
    D = res.domain
    th = res.train_history
    print('training windows:', th)
    assert len(th) == 2 and D.sfsmodel.epoch == 2
    assert all(np.isfinite(r['model_error']) and r['model_error'] > 0 for r in th)
    dp = D.sfsmodel.params.cpu().numpy() - params
    assert np.isfinite(dp).all() and np.abs(dp).max() > 0
    # RMSprop's first steps move every parameter with a nonzero gradient by
    # about lr/sqrt(1 - alpha) = 10 lr
    assert np.abs(dp).max() < 25e-3


def _window_objective(targets, flat):
    """sum_c mean|u_c - u_T,c| after one 4-step window of the forward run
    with the 'source' ML closure's parameters `flat` (PyFlow.py:518-531)."""
    from pyflow_amd import PyFlow, dataReader as dr
    res = PyFlow.simulate(les_config(targets, 4, flat, sfs='ML', training=False), verbose=False)
    _, _, d1 = dr.readNGArestart(targets['files'][1], headerOnly=False, printOut=False)
    D = res.domain
    J = 0.0
    for i, st in enumerate((D.state_u_P, D.state_v_P, D.state_w_P)):
        t = torch.from_numpy(np.ascontiguousarray(d1[..., i])).cuda()
        J += float((st.interior() - t).abs().mean())
    return J


def test_ml_closure_training_window(targets):
    """The run loop's training window with the reference's 'source' ML
    closure (SFSmodel 'ML'): the objective it reports equals the forward-only
    run's; its window gradient (sum of the eight adjoint-stage gradients times
    -simDt, sfsmodel_ML.py:287-289) points along the objective's descent
    direction, several times larger than the central difference of the
    objective along it: four from the reference's equal-weight accumulation
    over the RK4 stages (as with the MLP), and about two more because the
    reference's g = sum G . u_A runs over the padded arrays (20^3 cells for
    this 16^3 box, adjoint.py:261-263), whose ghost layers repeat the
    periodic field.  Measured 8.66.  The finalize step moves the parameters
    by the RMSprop step of lr 0.01."""
    from pyflow_amd import PyFlow
    from pyflow_amd.sfsmodel_ML import init_params
    flat = init_params(3)
    res = PyFlow.simulate(les_config(targets, 4, flat, sfs='ML'), verbose=False)
    D = res.domain
    assert len(res.train_history) == 1 and D.sfsmodel.epoch == 1
    g = D.sfsmodel.last_window_grad.double().cpu().numpy()
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    err0 = res.train_history[0]['model_error']
    J0 = _window_objective(targets, flat)
    assert abs(err0 - J0) <= 1e-9 * J0
    v = g / np.linalg.norm(g)
    eps = 0.05
    Jp = _window_objective(targets, (flat + eps * v).astype(np.float32))
    Jm = _window_objective(targets, (flat - eps * v).astype(np.float32))
    dJ = (Jp - Jm) / (2 * eps)
    ratio = float(np.linalg.norm(g) / dJ)
    print('window objective %.6e; |g| %.4e; central difference along g %.4e (J+ %.8e, '
          'J- %.8e); ratio %.3f' % (J0, np.linalg.norm(g), dJ, Jp, Jm, ratio))
    assert 5.0 < ratio < 14.0
    dp = D.sfsmodel.params.cpu().numpy() - flat
    assert np.isfinite(dp).all()
    # the first RMSprop step (square_avg 0.01 g^2, numSumDistr 1)
    g32 = g.astype(np.float32)
    want = -0.01 * g32 / (np.sqrt(0.01 * g32 * g32) + 1e-8)
    assert np.allclose(dp, want, rtol=1e-4, atol=1e-7)
