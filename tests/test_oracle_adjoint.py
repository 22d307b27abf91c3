"""The CPU restatement of the adjoint step, pinned three ways:

* against the reference's own outputs (golden vectors of
  rhs_adjPredictor.evaluate and Domain.adjointStep written by
  tests/golden/make_golden.py) - bit for bit where no Krylov reduction is
  involved;
* by the transpose identity <J(u) d, a> = <d, A(u) a> between the adjoint
  RHS A and the Jacobian J of the forward RHS, whose restatement is itself
  pinned bit-exactly.  The forward RHS is quadratic in u, so the central
  difference (R(u + e d) - R(u - e d)) / 2e is J(u) d up to rounding for any e;
* for the tensor-closure term (computed by autograd), by the same identity
  against central differences of the closure's share of the forward RHS
  (exact for the quadratic gradient model; O(e^2) for the MLP).

Also the adjoint-verification protocol of the reference
(core/run_PyFlow_adjoint_verification.py:125-207): the adjoint value at the
perturbed cell against forward differences of the objective.
"""

import math

import numpy as np
import pytest
import torch

from helpers import golden
from oracle import adjoint_oracle as A
from oracle import pyflow_oracle as O


@pytest.fixture(scope='module')
def G():
    torch.set_num_threads(1)
    return golden()


def grid(G):
    shape = tuple(int(x) for x in G['shape'])
    L = tuple(float(x) for x in G['L'])
    return O.Grid(*shape, Lx=L[0], Ly=L[1], Lz=L[2])


def visc(g, mu):
    V = torch.ones(g.nxo, g.nyo, g.nzo, dtype=torch.float64)
    V.mul_(mu)
    return V


def test_adjoint_rhs_golden(G):
    g = grid(G)
    fu, fv, fw = (O.Field(g, G[k]) for k in ('u0', 'v0', 'w0'))
    r = A.AdjointRHS(g, float(G['rho']), visc(g, float(G['mu'])), fu, fv, fw)
    r.evaluate(*(O.Field(g, G[k]) for k in ('adj_u0', 'adj_v0', 'adj_w0')))
    for c in 'uvw':
        assert np.array_equal(getattr(r, 'rhs_' + c).numpy(), G['adjrhs_' + c])


@pytest.mark.parametrize('tag,adv,solver,tol,nit', [
    ('adj_euler_jac', 'Euler', 'Jacobi', 0.0, 30),
    ('adj_rk4_bicg', 'RK4', 'bicgstab', 1e-12, 500)])
def test_adjoint_step_golden(G, tag, adv, solver, tol, nit):
    g = grid(G)
    zeros = np.zeros(tuple(int(x) for x in G['shape']))
    b = O.Box(g, G['u0'], G['v0'], G['w0'], mu=float(G['mu']), rho=float(G['rho']),
              dt=float(G['dt']), advancer=adv, solver=solver, tol=tol, maxit=nit)
    ab = A.AdjointBox(b)
    for f, k in ((ab.au, 'adj_u0'), (ab.av, 'adj_v0'), (ab.aw, 'adj_w0')):
        g.interior(f.var).copy_(torch.as_tensor(G[k]))
        f.update_border()
    ab.adjoint_step()
    got = {'u': ab.au.var, 'v': ab.av.var, 'w': ab.aw.var, 'DP': b.DP.var, 'p': b.p.var}
    for c, t in got.items():
        ref = G['%s_%s' % (tag, c)]
        if solver == 'Jacobi':
            assert np.array_equal(t.numpy(), ref), c
        else:
            # the oracle's BiCGStab issues the reference's reductions: same
            # bits here too, since both run torch CPU with one thread
            assert np.array_equal(t.numpy(), ref), c
    assert b.max_resP == float(G[tag + '_res'])
    del zeros


def _rand(shape, seed):
    gen = torch.Generator().manual_seed(seed)
    return [torch.randn(shape, dtype=torch.float64, generator=gen) for _ in range(3)]


def _rhs(g, rho, V, x, sfs=None):
    F = [O.Field(g, t) for t in x]
    if sfs is not None:
        sfs.update(*F)
    r = O.NavierStokesRHS(g, rho, V, sfs)
    r.evaluate(*F)
    return [r.rhs_u.clone(), r.rhs_v.clone(), r.rhs_w.clone()]


def _dot(a, b):
    return sum(torch.sum(x * y) for x, y in zip(a, b)).item()


def test_adjoint_rhs_is_transpose_of_forward_jacobian():
    g = O.Grid(8, 10, 12, Lx=2 * math.pi, Ly=1.7 * math.pi, Lz=2.3 * math.pi)
    rho, V = 1.1, visc(g, 0.013)
    u, d, ah = _rand((8, 10, 12), 1), _rand((8, 10, 12), 2), _rand((8, 10, 12), 3)
    e = 1e-3
    Rp = _rhs(g, rho, V, [a + e * b for a, b in zip(u, d)])
    Rm = _rhs(g, rho, V, [a - e * b for a, b in zip(u, d)])
    lhs = _dot([(p - m) / (2 * e) for p, m in zip(Rp, Rm)], ah)
    r = A.AdjointRHS(g, rho, V, *(O.Field(g, t) for t in u))
    r.evaluate(*(O.Field(g, t) for t in ah))
    rhs = _dot(d, (r.rhs_u, r.rhs_v, r.rhs_w))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    # a planted mistake (one convective transpose dropped) is caught
    r2 = A.AdjointRHS(g, rho, V, *(O.Field(g, t) for t in u))
    keep = A.adj_conv_1d
    try:
        A.adj_conv_1d = lambda *a, **k: None
        r2.evaluate(*(O.Field(g, t) for t in ah))
    finally:
        A.adj_conv_1d = keep
    assert abs(lhs - _dot(d, (r2.rhs_u, r2.rhs_v, r2.rhs_w))) > 1e-6 * abs(lhs)


@pytest.mark.parametrize('closure', ['gradient', 'mlp'])
def test_closure_term_is_transpose(closure):
    g = O.Grid(8, 10, 12, Lx=2 * math.pi, Ly=1.7 * math.pi, Lz=2.3 * math.pi)
    rho = 1.1
    V0 = visc(g, 0.0)
    u, d, ah = _rand((8, 10, 12), 4), _rand((8, 10, 12), 5), _rand((8, 10, 12), 6)
    if closure == 'gradient':
        make = lambda: O.GradientModel(g)   # noqa: E731
        clo = A.gradient_closure(g)
        e, tol = 1e-3, 1e-11
    else:
        gen = np.random.Generator(np.random.PCG64(7))
        params = [gen.normal(scale=0.4, size=s) for s in ((20, 9), (20,), (20, 20), (20,), (6, 20), (6,))]

        class M:
            modelType = 'tensor'

            def update(self, fu, fv, fw):
                (self.FXX, self.FYY, self.FZZ, self.FXY, self.FXZ,
                 self.FYZ) = O.mlp_closure_forward(g, fu, fv, fw, params)
        make = M
        clo = A.mlp_closure(g, params)
        e, tol = 1e-5, 1e-7
    # with VISC = 0 the forward RHS minus its no-closure value is the closure share
    Rp = _rhs(g, rho, V0, [a + e * b for a, b in zip(u, d)], make())
    Rm = _rhs(g, rho, V0, [a - e * b for a, b in zip(u, d)], make())
    R0p = _rhs(g, rho, V0, [a + e * b for a, b in zip(u, d)])
    R0m = _rhs(g, rho, V0, [a - e * b for a, b in zip(u, d)])
    lhs = _dot([((p - m) - (p0 - m0)) / (2 * e) for p, m, p0, m0 in zip(Rp, Rm, R0p, R0m)], ah)
    dl = A.closure_adjoint_term(g, rho, *(O.Field(g, t) for t in u),
                                *(O.Field(g, t) for t in ah), clo)
    assert abs(lhs - _dot(d, dl)) <= tol * abs(lhs)


def test_adjoint_verification_protocol():
    """run_PyFlow_adjoint_verification.py:125-207 on a 12^3 box, 3 RK4 steps
    with BiCGStab at 1e-13: the adjoint value at the perturbed cell against
    central differences of the objective.  The reference's adjoint is not the
    exact transpose of its RK4 step: agreement is second order in dt (7.5e-5
    relative here, 1.9e-5 at half the step)."""
    n = 12
    g = O.Grid(n, n, n)
    u, v, w = (t * 0.5 for t in _rand((n, n, n), 11))
    kw = dict(mu=0.05, rho=1.0, dt=0.01, advancer='RK4', solver='bicgstab', tol=1e-13, maxit=500)
    obj0, adj = A.adjoint_verification_run(g, u, v, w, 3, 0.0, True, **kw)
    h = 1e-5
    op, _ = A.adjoint_verification_run(g, u, v, w, 3, h, False, **kw)
    om, _ = A.adjoint_verification_run(g, u, v, w, 3, -h, False, **kw)
    fd = (op - om) / (2 * h)
    assert abs(adj - fd) <= 1e-3 * abs(fd)


def test_adjoint_rhs_with_viscosity_field_golden(G):
    """The adjoint RHS with the Smagorinsky VISC field in its viscous part
    (adjoint.py:100-170 with a field)."""
    g = grid(G)
    fu, fv, fw = (O.Field(g, G[k]) for k in ('u0', 'v0', 'w0'))
    V = torch.as_tensor(G['smag_visc'])
    r = A.AdjointRHS(g, float(G['rho']), V, fu, fv, fw)
    r.evaluate(*(O.Field(g, G[k]) for k in ('adj_u0', 'adj_v0', 'adj_w0')))
    for c in 'uvw':
        assert np.array_equal(getattr(r, 'rhs_' + c).numpy(), G['adjrhs_smag_' + c])


def _ml_closure_rhs(G, g, trainable=True):
    from helpers import ml_source_params
    _, sd = ml_source_params()
    fu, fv, fw = (O.Field(g, G[c + '0']) for c in 'uvw')
    au, av, aw = (O.Field(g, G['adj_%s0' % c]) for c in 'uvw')
    ml = O.MLSource(g, sd, trainable=trainable)
    r = A.AdjointRHS(g, float(G['rho']), visc(g, float(G['mu'])), fu, fv, fw, closure=ml)
    r.evaluate(au, av, aw)
    return r, ml


def test_source_ml_closure_adjoint_term_golden(G):
    """rhs_adjPredictor.evaluate with the reference's 'source' ML closure
    (adjoint.py:230-286): the closure's share of the adjoint RHS and the
    parameter gradients its g.backward() leaves, against the reference's own
    (golden vectors).  The oracle issues the same torch CPU ops; autograd may
    accumulate the gradient contributions of a tensor in another order, so
    the bar is FP32 rounding (measured below)."""
    g = grid(G)
    r, ml = _ml_closure_rhs(G, g)
    for c, rr in zip('uvw', (r.rhs_u, r.rhs_v, r.rhs_w)):
        got = rr.numpy() - G['adjrhs_' + c]
        ref = G['adjrhs_ml_' + c] - G['adjrhs_' + c]
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print('closure term %s: rel L2 %.3e (max|term| %.3e)' % (c, e, np.abs(ref).max()))
        assert e < 1e-5
    gp = ml.packed_grad().numpy().astype(np.float64)
    ref = G['adjrhs_ml_dparams'].astype(np.float64)
    e = np.linalg.norm(gp - ref) / np.linalg.norm(ref)
    print('parameter gradient rel L2 %.3e' % e)
    assert e < 1e-5
