"""Exact gradients through the time step (pyflow_amd.autodiff): the
hand-written reverse kernels behind torch.autograd.Function, checked

* against torch autograd through the oracle's differentiable restatement
  of the same step (oracle.adjoint_oracle.functional_step: the reference's
  op sequence, RK4 and an exact FFT projection) - FP64, rel L2 1e-8 for no
  closure and the gradient model; 1e-4 for the MLP (FP32 MLP arithmetic on
  the GPU, FP64 in the oracle), including the MLP parameter gradient;
* against central finite differences (eps = 1e-4) of the GPU forward step
  along a random direction: rel 1e-6;
* on the LES window (64^3, MLP closure, 20 RK4 steps, checkpointed): the
  FP64 gradient against central differences, and the FP32 gradient (FP32
  storage, FP64 Krylov reductions) against the FP64 one, with the error
  stated;
* the same window with the MLP forward on MFMA in BF16 (FP32 accumulation),
  the arithmetic the closure runs with in Domain: its reverse differentiates
  the closure's FP32 arithmetic at the BF16 forward's states, so the BF16
  window gradient is stated against the FP32-forward gradient and against
  central differences of the BF16 forward;
* a 32^3 window of 20 steps with the MLP against torch autograd through the
  oracle's restatement (FP64).
"""

import math

import numpy as np
import pytest
import torch

from helpers import gpu_env, cpu, rel_l2
from oracle import adjoint_oracle as A
from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu


def _fields(shape, seed, scale=0.5):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [scale * rng.standard_normal(shape) for _ in range(3)]


def _stepper(shape, closure, dtype=torch.float64, mu=0.05, rho=1.0, dt=0.01, tol=1e-12,
             mlp_forward='fp32'):
    from pyflow_amd.autodiff import StepVJP
    d, g, m = gpu_env(shape, (2 * math.pi,) * 3, dtype)
    return StepVJP(d.box, mu, rho, dt, closure=closure, tol=tol, mlp_forward=mlp_forward), d


def _gpu_grad(st, u0, c, K, params=None, dtype=torch.float64):
    from pyflow_amd.autodiff import rollout
    x = [torch.tensor(t, dtype=dtype, device='cuda', requires_grad=True) for t in u0]
    p = None
    if params is not None:
        p = torch.tensor(params, device='cuda', requires_grad=True)
    out = rollout(st, *x, K, p)
    L = sum(torch.sum(o * torch.tensor(cc, dtype=dtype, device='cuda')) for o, cc in zip(out, c))
    L.backward()
    torch.cuda.synchronize()
    return [cpu(t.grad) for t in x], (cpu(p.grad) if p is not None else None), L.item()


def _gpu_loss(st, u0, c, K, params=None, dtype=torch.float64):
    x = [torch.tensor(t, dtype=dtype, device='cuda') for t in u0]
    p = torch.tensor(params, device='cuda') if params is not None else None
    for _ in range(K):
        x = st.forward(*x, p)
    return sum(torch.sum(o.double() * torch.tensor(cc, device='cuda')) for o, cc in zip(x, c)).item()


def _oracle_grad(shape, u0, c, K, closure, params=None, mu=0.05, rho=1.0, dt=0.01):
    g = O.Grid(*shape)
    x = [torch.tensor(t, requires_grad=True) for t in u0]
    theta = None
    if closure == 'gradient':
        clo = A.gradient_closure(g)
    elif closure == 'mlp':
        from pyflow_amd.sfsmodel_mlp import unpack
        theta = [torch.tensor(t.astype(np.float64), requires_grad=True) for t in unpack(params)]
        clo = A.mlp_closure(g, theta)
    else:
        clo = None
    y = x
    for _ in range(K):
        y = A.functional_step(g, *y, mu, rho, dt, closure=clo)
    L = sum(torch.sum(o * torch.tensor(cc)) for o, cc in zip(y, c))
    L.backward()
    dth = np.concatenate([t.grad.numpy().ravel() for t in theta]) if theta else None
    return [t.grad.numpy() for t in x], dth


@pytest.mark.parametrize('closure', [None, 'gradient', 'mlp'])
def test_step_gradient_against_oracle_autograd(closure):
    from pyflow_amd.sfsmodel_mlp import init_params
    shape, K = (12, 12, 12), 3
    u0, c = _fields(shape, 1), _fields(shape, 2, 1.0)
    params = init_params(9) * 3.0 if closure == 'mlp' else None
    st, d = _stepper(shape, closure)
    gg, gp, _ = _gpu_grad(st, u0, c, K, params)
    og, op = _oracle_grad(shape, u0, c, K, closure, params)
    bar = 1e-4 if closure == 'mlp' else 1e-8
    for a, b in zip(gg, og):
        assert rel_l2(a, b) < bar
    if closure == 'mlp':
        assert rel_l2(gp, op) < bar


@pytest.mark.parametrize('closure', [None, 'gradient', 'mlp'])
def test_step_gradient_against_finite_differences(closure):
    from pyflow_amd.sfsmodel_mlp import init_params
    shape, K = (12, 12, 12), 3
    u0, c, dirn = _fields(shape, 3), _fields(shape, 4, 1.0), _fields(shape, 5, 1.0)
    params = init_params(11) * 3.0 if closure == 'mlp' else None
    st, d = _stepper(shape, closure)
    gg, gp, _ = _gpu_grad(st, u0, c, K, params)
    e = 1e-4
    Lp = _gpu_loss(st, [a + e * b for a, b in zip(u0, dirn)], c, K, params)
    Lm = _gpu_loss(st, [a - e * b for a, b in zip(u0, dirn)], c, K, params)
    fd = (Lp - Lm) / (2 * e)
    ad = sum(float(np.sum(a * b)) for a, b in zip(gg, dirn))
    bar = 1e-4 if closure == 'mlp' else 1e-6   # FP32 MLP arithmetic in the loss
    assert abs(fd - ad) <= bar * abs(fd), (fd, ad)
    if closure == 'mlp':
        # one parameter direction
        rng = np.random.Generator(np.random.PCG64(6))
        pd = rng.standard_normal(params.size).astype(np.float32)
        ep = 1e-2
        Lp = _gpu_loss(st, u0, c, K, params + ep * pd)
        Lm = _gpu_loss(st, u0, c, K, params - ep * pd)
        fdp = (Lp - Lm) / (2 * ep)
        adp = float(np.dot(gp.astype(np.float64), pd.astype(np.float64)))
        assert abs(fdp - adp) <= 1e-2 * abs(fdp), (fdp, adp)


def test_les_window_gradient():
    """64^3, MLP closure, 20 checkpointed RK4 steps from a Passot-Pouquet
    field (k0 = 4, u_rms = 1, nu = 0.005, dt at CFL 0.3)."""
    from pyflow_amd.initial_conditions import passot_pouquet, cfl_dt
    from pyflow_amd.sfsmodel_mlp import init_params
    N, K = 64, 20
    shape = (N, N, N)
    u0 = list(passot_pouquet(N, k0=4.0, urms=1.0, seed=3))
    dt = cfl_dt(*u0, 2 * math.pi / N)
    c, dirn = _fields(shape, 7, 1.0), _fields(shape, 8, 1.0)
    params = init_params(13)
    st64, _ = _stepper(shape, 'mlp', torch.float64, mu=0.005, dt=dt, tol=1e-12)
    g64, p64, _ = _gpu_grad(st64, u0, c, K, params)
    e = 1e-4
    Lp = _gpu_loss(st64, [a + e * b for a, b in zip(u0, dirn)], c, K, params)
    Lm = _gpu_loss(st64, [a - e * b for a, b in zip(u0, dirn)], c, K, params)
    fd = (Lp - Lm) / (2 * e)
    ad = sum(float(np.sum(a * b)) for a, b in zip(g64, dirn))
    assert abs(fd - ad) <= 1e-5 * abs(fd), (fd, ad)
    st32, _ = _stepper(shape, 'mlp', torch.float32, mu=0.005, dt=dt, tol=1e-6)
    g32, p32, _ = _gpu_grad(st32, u0, c, K, params, dtype=torch.float32)
    err = max(rel_l2(a, b) for a, b in zip(g32, g64))
    perr = rel_l2(p32, p64)
    print('LES window: FP32 gradient rel L2 vs FP64 %.3e, MLP parameters %.3e' % (err, perr))
    assert err < 1e-3 and perr < 1e-2


def test_les_window_gradient_bf16_forward():
    """64^3, MLP closure with its forward in BF16 on MFMA, 20 checkpointed RK4
    steps from a Passot-Pouquet field (k0 = 4, u_rms = 1, nu = 0.005, dt at
    CFL 0.3).  The closure's share of the RHS is small (stresses scaled by
    h^2/12), so BF16 rounding moves the velocity gradient little and the
    parameter gradient more; both are stated and bounded, and the BF16
    gradient is checked against central differences (eps = 1e-4) of the BF16
    forward itself in FP64 storage."""
    from pyflow_amd.initial_conditions import passot_pouquet, cfl_dt
    from pyflow_amd.sfsmodel_mlp import init_params
    N, K = 64, 20
    shape = (N, N, N)
    u0 = list(passot_pouquet(N, k0=4.0, urms=1.0, seed=3))
    dt = cfl_dt(*u0, 2 * math.pi / N)
    c, dirn = _fields(shape, 7, 1.0), _fields(shape, 8, 1.0)
    params = init_params(13)
    st32, _ = _stepper(shape, 'mlp', torch.float32, mu=0.005, dt=dt, tol=1e-6)
    g32, p32, _ = _gpu_grad(st32, u0, c, K, params, dtype=torch.float32)
    stb, _ = _stepper(shape, 'mlp', torch.float32, mu=0.005, dt=dt, tol=1e-6, mlp_forward='bf16')
    gb, pb, _ = _gpu_grad(stb, u0, c, K, params, dtype=torch.float32)
    err = max(rel_l2(a, b) for a, b in zip(gb, g32))
    perr = rel_l2(pb, p32)
    st64, _ = _stepper(shape, 'mlp', torch.float64, mu=0.005, dt=dt, tol=1e-12, mlp_forward='bf16')
    g64, _, _ = _gpu_grad(st64, u0, c, K, params)
    e = 1e-4
    Lp = _gpu_loss(st64, [a + e * b for a, b in zip(u0, dirn)], c, K, params)
    Lm = _gpu_loss(st64, [a - e * b for a, b in zip(u0, dirn)], c, K, params)
    fd = (Lp - Lm) / (2 * e)
    ad = sum(float(np.sum(a * b)) for a, b in zip(g64, dirn))
    print('LES window, BF16 MLP forward: velocity gradient rel L2 vs the FP32 forward %.3e, '
          'MLP parameters %.3e; FP64-storage gradient vs central differences %.3e'
          % (err, perr, abs(fd - ad) / abs(fd)))
    assert err < 1e-2 and perr < 1e-1
    assert abs(fd - ad) <= 1e-3 * abs(fd), (fd, ad)


def test_window_gradient_against_oracle_autograd_32():
    """A 20-step window at 32^3 with the MLP closure (FP32 MLP arithmetic on
    the GPU, FP64 in the oracle): rel L2 1e-4 on the velocity and the
    parameter gradients."""
    from pyflow_amd.sfsmodel_mlp import init_params
    shape, K = (32, 32, 32), 20
    u0, c = _fields(shape, 21), _fields(shape, 22, 1.0)
    params = init_params(17) * 3.0
    st, d = _stepper(shape, 'mlp', mu=0.02, dt=0.004)
    gg, gp, _ = _gpu_grad(st, u0, c, K, params)
    og, op = _oracle_grad(shape, u0, c, K, 'mlp', params, mu=0.02, dt=0.004)
    ev = max(rel_l2(a, b) for a, b in zip(gg, og))
    ep = rel_l2(gp, op)
    print('32^3, 20 steps, MLP: velocity gradient rel L2 vs oracle autograd %.3e, '
          'parameters %.3e' % (ev, ep))
    assert ev < 1e-4 and ep < 1e-4
