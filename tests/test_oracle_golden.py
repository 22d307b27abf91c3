"""The CPU oracle against the reference's own outputs (golden vectors written
by tests/golden/make_golden.py): every function must match bit for bit,
since the oracle issues the same torch CPU ops in the same order."""

import numpy as np
import pytest
import torch

from helpers import golden, recording_source, ml_source_params
from oracle import pyflow_oracle as O


@pytest.fixture(scope='module')
def G():
    torch.set_num_threads(1)
    return golden()


def grid(G):
    shape = tuple(int(x) for x in G['shape'])
    L = tuple(float(x) for x in G['L'])
    return O.Grid(*shape, Lx=L[0], Ly=L[1], Lz=L[2])


def fields(g, G):
    return O.Field(g, G['u0']), O.Field(g, G['v0']), O.Field(g, G['w0'])


def visc(g, G):
    V = torch.ones(g.nxo, g.nyo, g.nzo, dtype=torch.float64)
    V.mul_(float(G['mu']))
    return V


def test_rhs(G):
    g = grid(G)
    r = O.NavierStokesRHS(g, float(G['rho']), visc(g, G))
    r.evaluate(*fields(g, G))
    for c in 'uvw':
        assert np.array_equal(getattr(r, 'rhs_' + c).numpy(), G['rhs_' + c])


def test_gradient_closure_and_rhs(G):
    g = grid(G)
    fu, fv, fw = fields(g, G)
    gm = O.GradientModel(g)
    gm.update(fu, fv, fw)
    for n, f in (('u', fu), ('v', fv), ('w', fw)):
        for a in 'xyz':
            assert np.array_equal(getattr(f, 'grad_' + a).numpy(), G['gc_%s_%s' % (n, a)])
    for n in ('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ'):
        assert np.array_equal(getattr(gm, n).numpy(), G['sfs_' + n])
    r = O.NavierStokesRHS(g, float(G['rho']), visc(g, G), gm)
    r.evaluate(*fields(g, G))
    for c in 'uvw':
        assert np.array_equal(getattr(r, 'rhs_' + c).numpy(), G['rhs_sfs_' + c])


def test_rk4_and_euler(G):
    g = grid(G)
    dt = float(G['dt'])
    fu, fv, fw = fields(g, G)
    tu, tv, tw = fields(g, G)
    O.rk4_step(fu, fv, fw, tu, tv, tw,
               [O.NavierStokesRHS(g, float(G['rho']), visc(g, G)) for _ in range(4)], dt)
    for c, f in zip('uvw', (fu, fv, fw)):
        assert np.array_equal(f.var.numpy(), G['rk4_' + c])
    fu, fv, fw = fields(g, G)
    O.euler_step(fu, fv, fw, O.NavierStokesRHS(g, float(G['rho']), visc(g, G)), dt)
    for c, f in zip('uvw', (fu, fv, fw)):
        assert np.array_equal(f.var.numpy(), G['euler_' + c])


def test_source_closure_contract(G):
    """velocity.py:360-383 with a recording stand-in closure: the arrays it
    receives (var_i, including the last planes the advective interpolations
    leave), the RHS, and RK4 / Euler steps that evaluate it in every stage."""
    g = grid(G)
    rho, dt = float(G['rho']), float(G['dt'])
    stub = recording_source()
    r = O.NavierStokesRHS(g, rho, visc(g, G), stub)
    r.evaluate(*fields(g, G))
    for c, t in zip('uvw', stub.calls[0]):
        assert np.array_equal(t.numpy(), G['src_in_' + c]), c
        assert np.array_equal(getattr(r, 'rhs_' + c).numpy(), G['rhs_src_' + c]), c
    fu, fv, fw = fields(g, G)
    tu, tv, tw = fields(g, G)
    stub = recording_source()
    O.rk4_step(fu, fv, fw, tu, tv, tw, [O.NavierStokesRHS(g, rho, visc(g, G), stub)
                                        for _ in range(4)], dt)
    assert len(stub.calls) == 4
    for c, f, t in zip('uvw', (fu, fv, fw), stub.calls[3]):
        assert np.array_equal(f.var.numpy(), G['rk4_src_' + c]), c
        assert np.array_equal(t.numpy(), G['rk4_src_in4_' + c]), c
    fu, fv, fw = fields(g, G)
    O.euler_step(fu, fv, fw, O.NavierStokesRHS(g, rho, visc(g, G), recording_source()), dt)
    for c, f in zip('uvw', (fu, fv, fw)):
        assert np.array_equal(f.var.numpy(), G['euler_src_' + c]), c


def test_source_inputs_stale_planes(G):
    """The closure's inputs where the centre interpolant does not reach: the
    last plane along each component's own axis holds the edge interpolant
    the advective terms wrote there last, and one corner of it is never
    written (zero from the allocation)."""
    u = np.pad(G['u0'], 2, mode='wrap')
    v = np.pad(G['v0'], 2, mode='wrap')
    w = np.pad(G['w0'], 2, mode='wrap')
    ui, vi, wi = G['src_in_u'], G['src_in_v'], G['src_in_w']
    assert np.array_equal(ui[:-1], 0.5 * (u[1:] + u[:-1]))
    assert np.array_equal(ui[-1, :, 1:], 0.5 * (u[-1, :, 1:] + u[-1, :, :-1]))
    assert np.array_equal(ui[-1, 1:, 0], 0.5 * (u[-1, 1:, 0] + u[-1, :-1, 0]))
    assert ui[-1, 0, 0] == 0.0
    assert np.array_equal(vi[:, :-1], 0.5 * (v[:, 1:] + v[:, :-1]))
    assert np.array_equal(vi[:, -1, 1:], 0.5 * (v[:, -1, 1:] + v[:, -1, :-1]))
    assert np.array_equal(vi[1:, -1, 0], 0.5 * (v[1:, -1, 0] + v[:-1, -1, 0]))
    assert vi[0, -1, 0] == 0.0
    assert np.array_equal(wi[:, :, :-1], 0.5 * (w[:, :, 1:] + w[:, :, :-1]))
    assert np.array_equal(wi[:, 1:, -1], 0.5 * (w[:, 1:, -1] + w[:, :-1, -1]))
    assert np.array_equal(wi[1:, 0, -1], 0.5 * (w[1:, 0, -1] + w[:-1, 0, -1]))
    assert wi[0, 0, -1] == 0.0


def test_ml_source_closure(G):
    """The reference's 'source' ML closure (sfsmodel_ML.py:354-593) on the
    inputs its RHS hands it: network output channels 0-8 and GX, GY, GZ.  The
    oracle issues the same FP32 torch ops (one thread, as the vectors were
    written), so they agree bit for bit."""
    torch.set_num_threads(1)
    g = grid(G)
    _, sd = ml_source_params()
    ml = O.MLSource(g, sd)
    ml.update(*(torch.tensor(G['src_in_' + c]) for c in 'uvw'))
    assert np.array_equal(ml.last_output[..., :9].numpy(), G['mlsrc_p'])
    for n in ('GX', 'GY', 'GZ'):
        assert np.array_equal(getattr(ml, n).numpy(), G['mlsrc_' + n]), n


def test_bicgstab_history(G):
    g = grid(G)
    fu, fv, fw = fields(g, G)
    src = g.interior_zeros()
    O.div_vel(g, fu, fv, fw, src)
    assert np.array_equal(src.numpy(), G['div0'])
    DP = O.Field(g, np.zeros(g.interior_zeros().shape))
    p = O.Field(g, np.zeros(g.interior_zeros().shape))
    bs = O.BiCGStab(g, float(G['rho']), float(G['dt']), 1e-11, 400)
    res = bs.solve(DP, p, src)
    assert res == float(G['bicg_res'])
    hist = G['bicg_hist'][1:] / G['bicg_hist'][0]
    assert np.array_equal(np.array(bs.history), hist)
    assert np.array_equal(DP.var.numpy(), G['bicg_DP'])
    assert np.array_equal(p.var.numpy(), G['bicg_p'])
    assert np.array_equal(src.numpy(), G['bicg_source_after'])


def test_jacobi(G):
    g = grid(G)
    DP = O.Field(g, np.zeros(g.interior_zeros().shape))
    p = O.Field(g, np.zeros(g.interior_zeros().shape))
    src = torch.tensor(G['div0'].copy())
    res = O.Jacobi(g, float(G['rho']), float(G['dt']), 25).solve(DP, p, src)
    assert res == float(G['jac_res'])
    assert np.array_equal(DP.var.numpy(), G['jac_pOld'])
    assert np.array_equal(p.var.numpy(), G['jac_p'])


@pytest.mark.parametrize('tag,adv,solver,tol,nit', [
    ('dom_rk4_bicg', 'RK4', 'bicgstab', 1e-12, 500),
    ('dom_euler_jac', 'Euler', 'Jacobi', 0.0, 30)])
def test_domain_steps(G, tag, adv, solver, tol, nit):
    g = grid(G)
    B = O.Box(g, G['u0'], G['v0'], G['w0'], mu=float(G['mu']), rho=float(G['rho']),
              dt=float(G['dt']), advancer=adv, solver=solver, tol=tol, maxit=nit)
    for _ in range(2):
        B.forward_step()
    for c in ('u', 'v', 'w', 'p', 'DP'):
        assert np.array_equal(getattr(B, c).var.numpy(), G['%s_%s' % (tag, c)]), c
    assert B.max_resP == float(G[tag + '_res'])


def test_smagorinsky_eddy_viscosity_and_rhs(G):
    """stress_constCs.eddyVisc and the RHS with the resulting VISC field."""
    g = grid(G)
    sm = O.Smagorinsky(g, Cs=0.2)
    fu, fv, fw = fields(g, G)
    mut = sm.eddyVisc(fu, fv, fw, float(G['rho']))
    assert np.array_equal(mut.numpy(), G['smag_mut'])
    V = float(G['mu']) + mut
    assert np.array_equal(V.numpy(), G['smag_visc'])
    r = O.NavierStokesRHS(g, float(G['rho']), V, sm)
    r.evaluate(*fields(g, G))
    for c in 'uvw':
        assert np.array_equal(getattr(r, 'rhs_' + c).numpy(), G['rhs_smag_' + c])


def test_smagorinsky_domain_steps(G):
    g = grid(G)
    B = O.Box(g, G['u0'], G['v0'], G['w0'], mu=float(G['mu']), rho=float(G['rho']),
              dt=float(G['dt']), advancer='RK4', solver='bicgstab', tol=1e-12, maxit=500,
              sfs=O.Smagorinsky(g, Cs=0.2))
    for _ in range(2):
        B.forward_step()
    for c in ('u', 'v', 'w', 'p'):
        assert np.array_equal(getattr(B, c).var.numpy(), G['dom_smag_' + c]), c
    assert np.array_equal(B.VISC.numpy(), G['dom_smag_visc'])
    assert B.max_resP == float(G['dom_smag_res'])


def test_smagorinsky_closed_forms():
    """Pins beyond the reference's outputs: a uniform translation has no
    strain (mu_t = 0); a pure shear u = a*y has S_12 = a/2 and
    |S| = sqrt(2*S_12^2) = |a|/sqrt(2), so mu_t = rho*(Cs*h)^2*|a|/sqrt(2) on
    every cell whose stencil sees only the linear field."""
    g = O.Grid(8, 8, 8)
    sm = O.Smagorinsky(g, Cs=0.17)
    one = np.ones((8, 8, 8))
    mut = sm.eddyVisc(O.Field(g, 2.0 * one), O.Field(g, -1.0 * one), O.Field(g, 0.5 * one), 1.3)
    I = (slice(2, -2),) * 3
    assert float(mut[I].abs().max()) == 0.0
    # linear shear, periodic-compatible only in the interior differences: use
    # a small box where the interior cells see only the linear part
    a = 0.7
    y = (np.arange(8) + 0.5) * g.dy
    u = a * y[None, :, None] * one
    f = O.Field(g, u)
    # overlap cells as a linear continuation (not periodic) so every interior
    # stencil sees the same slope
    yo = (np.arange(-2, 10) + 0.5) * g.dy
    f.var.copy_(torch.as_tensor(a * yo[None, :, None] * np.ones((12, 12, 12))))
    z = O.Field(g, 0.0 * one)
    mut = sm.eddyVisc(f, z, O.Field(g, 0.0 * one), 1.3)
    expect = 1.3 * (0.17 * g.dx) ** 2 * a / np.sqrt(2.0)
    inner = (slice(3, -3),) * 3
    assert np.allclose(mut[inner].numpy(), expect, rtol=1e-12, atol=0)


def test_mlp_closure_closed_forms(G):
    """The closure MLP has no reference implementation; its oracle is pinned
    by weights for which the network has a closed form.  (1) Zero weights
    with output biases: every stress is its bias times dx^2/12.  (2) Weights
    that select: hidden unit j of the first layer reads feature perm[j] (a
    9-cycle, so a transposed matrix reads perm^-1) plus a bias, unit j of
    the second reads unit q[j] of the first times a (another 9-cycle), the
    third reads unit m times c and, for m < 3,
    unit m + 6 times c*d, so that every feature reaches a stress with a
    weight of its own: stress m is (c*T(m) + c*d*T(m+6)) * dx^2/12 with
    T(j) = tanh(a*tanh(f[perm[q[j]]] + b1[q[j]])).  The features are the nine
    cell-centred gradients in the order u_x, u_y, u_z, v_x, ..., w_z, from
    the oracle's grad_vel_center (pinned by the gradient model's vectors)."""
    g = grid(G)
    fu, fv, fw = fields(g, G)
    scale = g.dx * g.dx / 12.0
    z = lambda *s: np.zeros(s)   # noqa: E731
    b3 = np.array([0.3, -0.2, 0.1, 0.05, -0.4, 0.25])
    out = O.mlp_closure_forward(g, fu, fv, fw, [z(20, 9), z(20), z(20, 20), z(20), z(6, 20), b3])
    for m in range(6):
        assert torch.allclose(out[m], torch.full_like(out[m], b3[m] * scale), rtol=0, atol=1e-15)

    perm = [(j + 4) % 9 for j in range(9)]
    W1 = z(20, 9)
    for j in range(9):
        W1[j, perm[j]] = 1.0
    b1 = np.concatenate([np.linspace(-0.3, 0.3, 9), z(11)])
    a, c, d = 0.7, 1.9, -0.45
    q = [(j + 2) % 9 for j in range(9)]
    W2 = z(20, 20)
    for j in range(9):
        W2[j, q[j]] = a
    W3 = z(6, 20)
    for m in range(6):
        W3[m, m] = c
        if m < 3:
            W3[m, m + 6] = c * d
    out = O.mlp_closure_forward(g, fu, fv, fw, [W1, b1, W2, z(20), W3, z(6)])
    feats = []
    for f, comp in ((fu, 'u'), (fv, 'v'), (fw, 'w')):
        O.grad_vel_center(g, f, comp)
        feats += [f.grad_x.numpy().copy(), f.grad_y.numpy().copy(), f.grad_z.numpy().copy()]
    T = [np.tanh(a * np.tanh(feats[perm[q[j]]] + b1[q[j]])) for j in range(9)]
    for m in range(6):
        want = c * T[m] + (c * d * T[m + 6] if m < 3 else 0.0)
        np.testing.assert_allclose(out[m].numpy(), want * scale, rtol=1e-13, atol=1e-15)
