"""GPU parity of training through the reference's 'source' ML closure
(solver/adjoint.py:230-286 with sfsmodel/sfsmodel_ML.py): the closure's
reverse pass (pf_sfs_ml_source_vjp) and the adjoint RHS and step with
closure kind 2.

* the reverse pass on a ragged box against torch autograd through the
  oracle's restatement of the closure (FP32 on both sides; the GPU sums the
  network's products on MFMA in another order): input gradients and
  parameter gradients to rel L2 1e-4;
* rhs_adjPredictor.evaluate with the closure against the reference's own
  adjoint RHS and parameter gradients (golden vectors): the closure's share
  of the RHS and the gradient to rel L2 1e-4;
* a two-step window of Domain.forwardStep / adjointStep (RK4 + BiCGStab at
  1e-12) against the oracle's window: adjoint velocity to rel L2 1e-9 and the
  accumulated parameter gradient (eight adjoint RHS evaluations) to 1e-4;
* the adjoint RK4 step that evaluates the network once per step (all-cell
  cache, stage-summed weight gradient) against the one that evaluates the
  whole term in each stage, on a box of two chunks: adjoint fields bit for
  bit, parameter gradient to FP32 rounding (1e-4; measured 2.0e-5: FP32
  sums over four stages, then one product, instead of four products).

The bars are FP32 rounding of a network whose 273 -> 200 -> 200 -> 18
products are summed in another order; the measured values are printed.
"""

import math
import types

import numpy as np
import pytest
import torch

from helpers import golden, gpu_env, cpu, rel_l2, ml_source_params
from oracle import adjoint_oracle as A
from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu

TOL_TERM = 1e-4
TOL_PARAMS = 1e-4


@pytest.fixture(scope='module')
def G():
    return golden()


def ml_closure(d, g, flat, dt=1e-3, **kw):
    from pyflow_amd import sfsmodel_ML
    cfg = types.SimpleNamespace(simDt=dt, mlSourceParams=flat, adjointTraining=True, **kw)
    return sfsmodel_ML.residual_stress(cfg, d, g)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_reverse_pass_against_autograd_ragged(dtype):
    """d(sum G . Gbar)/d(var_i) and /d(theta) for random var_i and Gbar on a
    13 x 9 x 11 box with three lengths, against autograd through the oracle."""
    shape = (13, 9, 11)
    L = (2 * math.pi, 1.4 * math.pi, 1.9 * math.pi)
    d, g, m = gpu_env(shape, L, dtype)
    og = O.Grid(*shape, Lx=L[0], Ly=L[1], Lz=L[2], dtype=dtype)
    flat, sd = ml_source_params(11)
    rng = np.random.Generator(np.random.PCG64(21))
    pshape = tuple(n + 4 for n in shape)
    ins = [torch.tensor(rng.standard_normal(pshape), dtype=dtype) for _ in range(3)]
    gbar = [torch.tensor(rng.standard_normal(pshape), dtype=dtype) for _ in range(3)]
    ml = ml_closure(d, g, flat)
    dI = [torch.zeros(pshape, dtype=torch.float32, device='cuda') for _ in range(3)]
    dp = torch.zeros(ml.params.numel(), dtype=torch.float64, device='cuda')
    d.box.sfs_ml_source_vjp(ml.params, *(a.to('cuda') for a in ins),
                            [a.to('cuda') for a in gbar], dI, dp)
    torch.set_num_threads(1)
    ref = O.MLSource(og, sd, trainable=True)
    V = [a.clone().requires_grad_(True) for a in ins]
    ref.update(*V, requires_grad=True)
    gsum = sum(torch.sum(Gq[..., 0] * b) for Gq, b in zip((ref.GX, ref.GY, ref.GZ), gbar))
    gsum.backward()
    torch.cuda.synchronize()
    for c, got, v in zip('uvw', dI, V):
        e = rel_l2(cpu(got), v.grad.to(torch.float32).numpy())
        print('%s d/d%s_i rel L2 %.2e' % (dtype, c, e))
        assert e < TOL_TERM, (c, e)
    e = rel_l2(cpu(dp), ref.packed_grad().numpy())
    print('%s d/dtheta rel L2 %.2e' % (dtype, e))
    assert e < TOL_PARAMS


def test_reverse_pass_accumulates_and_needs_reserve():
    """dparams is added to, not overwritten; a box without the reverse
    workspace is refused with the name of the reserve call."""
    from pyflow_amd import _native
    from pyflow_amd.sfsmodel_ML import NORM
    n = 10
    h = 2 * math.pi / n
    flat, _ = ml_source_params()
    params = torch.tensor(flat, dtype=torch.float32, device='cuda')
    gen = torch.Generator(device='cuda').manual_seed(3)
    U = [torch.randn((n + 4,) * 3, dtype=torch.float64, device='cuda', generator=gen)
         for _ in range(3)]
    Gb = [torch.randn((n + 4,) * 3, dtype=torch.float64, device='cuda', generator=gen)
          for _ in range(3)]
    box = _native.Box(n, n, n, h, h, h, dtype=torch.float64)
    with pytest.raises(RuntimeError, match='pf_sfs_ml_source_train_reserve'):
        box.sfs_ml_source_vjp(params, *U, Gb)
    box.sfs_ml_source_train_reserve(NORM)
    once = torch.zeros(params.numel(), dtype=torch.float64, device='cuda')
    box.sfs_ml_source_vjp(params, *U, Gb, dparams=once)
    twice = torch.zeros_like(once)
    box.sfs_ml_source_vjp(params, *U, Gb, dparams=twice)
    box.sfs_ml_source_vjp(params, *U, Gb, dparams=twice)
    torch.cuda.synchronize()
    assert float(once.abs().max()) > 0.0
    assert torch.equal(twice, 2.0 * once)


def env(G, dtype=torch.float64):
    shape = tuple(int(x) for x in G['shape'])
    L = tuple(float(x) for x in G['L'])
    d, g, m = gpu_env(shape, L, dtype)
    return shape, L, d, g, m


def visc(d, mu, dtype=torch.float64):
    V = torch.ones(d.nxo_, d.nyo_, d.nzo_, dtype=dtype, device='cuda') * mu
    V.pf_const_mu = mu
    return V


def test_adjoint_rhs_with_ml_closure_against_reference(G):
    """rhs_adjPredictor.evaluate (adjoint.py:100-298) with the 'source' ML
    closure on the golden box: the closure's share of the adjoint RHS and the
    parameter gradients g.backward() leaves (the reference's own values)."""
    from pyflow_amd import adjoint, state
    shape, L, d, g, m = env(G)
    flat, _ = ml_source_params()
    ml = ml_closure(d, g, flat, float(G['dt']))
    su, sv, sw = (state.state_P(d, G[k].copy()) for k in ('u0', 'v0', 'w0'))
    au, av, aw = (state.state_P(d, G[k].copy()) for k in ('adj_u0', 'adj_v0', 'adj_w0'))
    r = adjoint.rhs_adjPredictor(d, m, float(G['rho']), visc(d, float(G['mu'])), ml, su, sv, sw)
    assert r.kind == 2
    r.evaluate(au, av, aw)
    torch.cuda.synchronize()
    for c in 'uvw':
        got = cpu(getattr(r, 'rhs_' + c)) - G['adjrhs_' + c]
        ref = G['adjrhs_ml_' + c] - G['adjrhs_' + c]
        e = rel_l2(got, ref)
        print('closure term %s: rel L2 %.2e (max|term| %.2e)' % (c, e, np.abs(ref).max()))
        assert e < TOL_TERM, (c, e)
    e = rel_l2(cpu(ml.grad), G['adjrhs_ml_dparams'])
    print('parameter gradient rel L2 %.2e' % e)
    assert e < TOL_PARAMS


@pytest.mark.parametrize('adv', ['RK4', 'Euler'])
def test_adjoint_window_with_ml_closure_against_oracle(G, adv):
    """Two forward steps with checkpoints, then two adjoint steps each on its
    restored forward state (PyFlow.py:333-367, 548-550), BiCGStab at 1e-12,
    the 'source' ML closure in both directions; against the oracle's window
    with the closure's weights as autograd leaves.  RK4 runs the adjoint
    step's closure work from one network forward per step (all-cell cache);
    Euler evaluates the whole term in its single stage."""
    from pyflow_amd import domain
    shape, L, d, g, m = env(G)
    flat, sd = ml_source_params()
    dt, mu, rho = float(G['dt']), float(G['mu']), float(G['rho'])
    cfg = types.SimpleNamespace(simDt=dt, equationMode='NS', advancerName=adv, mu=mu, rho=rho,
                                dtypeTorch=torch.float64, SFSmodel='ML', mlSourceParams=flat,
                                pSolverMode='bicgstab', min_pressure_residual=1e-12,
                                max_pressure_iterations=500, adjointTraining=True,
                                numCheckpointIt=2)
    zeros = np.zeros(shape)
    D = domain.Domain(cfg, None, d, np.stack([G['u0'], G['v0'], G['w0'], zeros], axis=-1),
                      g, m, ['U', 'V', 'W', 'P'], 0.0)
    torch.set_num_threads(1)
    og = O.Grid(*shape, Lx=L[0], Ly=L[1], Lz=L[2])
    B = O.Box(og, G['u0'], G['v0'], G['w0'], mu=mu, rho=rho, dt=dt, advancer=adv, tol=1e-12,
              maxit=500, sfs=O.MLSource(og, sd))
    ck = []
    D.checkpoint(0)
    ck.append([f.var.clone() for f in (B.u, B.v, B.w)])
    for k in range(2):
        D.forwardStep(dt)
        D.checkpoint(k + 1)
        B.forward_step()
        ck.append([f.var.clone() for f in (B.u, B.v, B.w)])
    train = O.MLSource(og, sd, trainable=True)
    AB = A.AdjointBox(B, train)
    for st, f, key in ((D.state_u_adj_P, AB.au, 'adj_u0'), (D.state_v_adj_P, AB.av, 'adj_v0'),
                       (D.state_w_adj_P, AB.aw, 'adj_w0')):
        st.interior().copy_(torch.tensor(G[key], device='cuda'))
        st.update_border()
        og.interior(f.var).copy_(torch.as_tensor(G[key]))
        f.update_border()
    for k in (2, 1):
        D.restore(k - 1)
        D.adjointStep(dt)
        for f, c in zip((B.u, B.v, B.w), ck[k - 1]):
            f.var.copy_(c)
        AB.adjoint_step()
    torch.cuda.synchronize()
    for c, st, f in (('u', D.state_u_adj_P, AB.au), ('v', D.state_v_adj_P, AB.av),
                     ('w', D.state_w_adj_P, AB.aw)):
        e = rel_l2(cpu(st.var), f.var.numpy())
        print('%s adjoint %s after the window: rel L2 %.2e' % (adv, c, e))
        assert e < 1e-9, (c, e)
    e = rel_l2(cpu(D.sfsmodel.grad), train.packed_grad().numpy())
    print('%s window parameter gradient rel L2 %.2e' % (adv, e))
    assert e < TOL_PARAMS


def test_cached_adjoint_step_matches_per_stage_terms():
    """pf_adj_rk4_step with closure 2 on a 38^3 box (42^3 = 74,088 padded
    cells: two chunks of the reverse) with and without the all-cell cache."""
    from pyflow_amd import _native
    from pyflow_amd.sfsmodel_ML import NORM
    n = 38
    h = 2 * math.pi / n
    flat, _ = ml_source_params(5)
    params = torch.tensor(flat, dtype=torch.float32, device='cuda')
    gen = torch.Generator(device='cuda').manual_seed(17)
    shape = (n + 4,) * 3
    U = [torch.randn(shape, dtype=torch.float64, device='cuda', generator=gen) for _ in range(3)]
    A0 = [1e-2 * torch.randn(shape, dtype=torch.float64, device='cuda', generator=gen)
          for _ in range(3)]
    out = []
    for budget in (-1, 0):
        box = _native.Box(n, n, n, h, h, h, dtype=torch.float64)
        box.sfs_ml_source_cache_budget(budget)
        box.sfs_ml_source_train_reserve(NORM)
        A = [a.clone() for a in A0]
        At = [torch.zeros_like(a) for a in A0]
        delta = [torch.zeros_like(a) for a in A0]
        grad = torch.zeros(params.numel(), dtype=torch.float64, device='cuda')
        for _ in range(2):
            box.adj_rk4_step(*U, *A, *At, 2e-3, 0.01, 1.0, 2, params, None, delta, grad)
        torch.cuda.synchronize()
        out.append(([cpu(a) for a in A], cpu(grad), box.device_bytes))
    (Ac, gc, bc), (Ap, gp, bp) = out
    print('device bytes: cached %.2f GB, per-stage %.2f GB' % (bc / 1e9, bp / 1e9))
    assert bc > bp
    for a, b in zip(Ac, Ap):
        assert np.array_equal(a, b)
    e = rel_l2(gc, gp)
    print('parameter gradient, cached vs per-stage: rel L2 %.2e' % e)
    assert 0 < np.abs(gc).max() and e < TOL_PARAMS
