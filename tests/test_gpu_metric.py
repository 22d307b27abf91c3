"""metric_uniform methods on the GPU against the oracle's restatement of
metric/metric_staggered.py, one method at a time, on random padded arrays
(so overlap cells and cells the reference leaves untouched are compared
too).  Every method is elementwise arithmetic: results must be bit-exact.
"""

import math

import numpy as np
import pytest
import torch

from helpers import gpu_env, cpu
from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (6, 9, 7)
L = (2 * math.pi, 1.3 * math.pi, 2.6 * math.pi)


@pytest.fixture(scope='module')
def env():
    torch.set_num_threads(1)
    d, g, m = gpu_env(SHAPE, L, torch.float64)
    og = O.Grid(*SHAPE, Lx=L[0], Ly=L[1], Lz=L[2])
    return d, g, m, og


def rnd(shape, seed):
    return np.random.Generator(np.random.PCG64(seed)).standard_normal(shape)


class Pair:
    """A GPU state_P and an oracle Field holding the same random arrays."""

    def __init__(self, d, og, seed):
        from pyflow_amd import state
        P = (SHAPE[0] + 4, SHAPE[1] + 4, SHAPE[2] + 4)
        self.g = state.state_P(d, np.zeros(SHAPE))
        self.o = O.Field(og)
        for k, name in enumerate(('var', 'var_i', 'grad_x', 'grad_y', 'grad_z')):
            a = rnd(P, seed * 10 + k)
            getattr(self.o, name).copy_(torch.tensor(a))
            getattr(self.g, name).copy_(torch.tensor(a, device='cuda'))

    def check(self, *names):
        torch.cuda.synchronize()
        for n in names:
            a = cpu(getattr(self.g, n))
            b = getattr(self.o, n).numpy()
            assert np.array_equal(a, b), (n, np.abs(a - b).max())


@pytest.mark.parametrize('name', ['interp_u_xm', 'interp_v_ym', 'interp_w_zm', 'interp_vw_x',
                                  'interp_uw_y', 'interp_uv_z', 'interp_uvwi_x', 'interp_uvwi_y',
                                  'interp_uvwi_z'])
def test_state_interpolations(env, name):
    d, g, m, og = env
    P = Pair(d, og, 1)
    getattr(m, name)(P.g)
    getattr(O, name)(P.o)
    P.check('var', 'var_i')


@pytest.mark.parametrize('name', ['interp_sc_x', 'interp_sc_y', 'interp_sc_z', 'interp_sc_xy',
                                  'interp_sc_xz', 'interp_sc_yz'])
def test_scalar_interpolations(env, name):
    d, g, m, og = env
    P = Pair(d, og, 2)
    out_g = torch.tensor(rnd(P.g.var.shape, 99), device='cuda')
    out_o = torch.tensor(rnd(P.g.var.shape, 99))
    getattr(m, name)(P.g.var, out_g)
    getattr(O, name)(og, P.o.var, out_o)
    torch.cuda.synchronize()
    assert np.array_equal(cpu(out_g), out_o.numpy())


@pytest.mark.parametrize('comp', ['u', 'v', 'w'])
def test_grad_vel_center(env, comp):
    d, g, m, og = env
    P = Pair(d, og, 3)
    m.grad_vel_center(P.g, comp)
    O.grad_vel_center(og, P.o, comp)
    P.check('var_i', 'grad_x', 'grad_y', 'grad_z')


def test_grad_vel_visc_and_grad_P(env):
    d, g, m, og = env
    P = Pair(d, og, 4)
    m.grad_vel_visc(P.g)
    O.grad_vel_visc(og, P.o)
    P.check('grad_x', 'grad_y', 'grad_z')
    Q = Pair(d, og, 5)
    m.grad_P(Q.g)
    O.grad_P(og, Q.o)
    Q.check('grad_x', 'grad_y', 'grad_z')


def test_div_visc(env):
    d, g, m, og = env
    P = (SHAPE[0] + 4, SHAPE[1] + 4, SHAPE[2] + 4)
    F = [rnd(P, 40 + q) for q in range(3)]
    r0 = rnd(SHAPE, 50)
    rg = torch.tensor(r0, device='cuda')
    ro = torch.tensor(r0)
    m.div_visc(*[torch.tensor(f, device='cuda') for f in F], rg)
    O.div_visc(og, *[torch.tensor(f) for f in F], ro)
    torch.cuda.synchronize()
    assert np.array_equal(cpu(rg), ro.numpy())


def test_div_vel_over_and_div_vel(env):
    d, g, m, og = env
    A, B, C = Pair(d, og, 6), Pair(d, og, 7), Pair(d, og, 8)
    Pshape = A.g.var.shape
    dg = torch.tensor(rnd(Pshape, 60), device='cuda')
    do = torch.tensor(rnd(Pshape, 60))
    m.div_vel_over(A.g, B.g, C.g, dg)
    O.div_vel_over(og, A.o, B.o, C.o, do)
    torch.cuda.synchronize()
    assert np.array_equal(cpu(dg), do.numpy())
    sg = torch.zeros(SHAPE, dtype=torch.float64, device='cuda')
    so = torch.zeros(SHAPE, dtype=torch.float64)
    m.div_vel(A.g, B.g, C.g, sg)
    O.div_vel(og, A.o, B.o, C.o, so)
    torch.cuda.synchronize()
    assert np.array_equal(cpu(sg), so.numpy())


@pytest.mark.parametrize('name', ['vel_conv_xx', 'vel_conv_yy', 'vel_conv_zz', 'vel_conv_x',
                                  'vel_conv_y', 'vel_conv_z'])
@pytest.mark.parametrize('sign', [1.0, -1.0])
def test_vel_conv(env, name, sign):
    d, g, m, og = env
    A, B = Pair(d, og, 9), Pair(d, og, 10)
    r0 = rnd(SHAPE, 70)
    rg = torch.tensor(r0, device='cuda')
    ro = torch.tensor(r0)
    getattr(m, name)(A.g, B.g, rg, sign)
    getattr(O, name)(og, A.o, B.o, ro, sign)
    torch.cuda.synchronize()
    assert np.array_equal(cpu(rg), ro.numpy())


def test_state_ops_and_halo(env):
    """state_P.ZAXPY / update / copy / vel_corr / update_border against the
    oracle's state.py restatement."""
    d, g, m, og = env
    A, B = Pair(d, og, 11), Pair(d, og, 12)
    X = rnd(SHAPE, 80)
    A.g.ZAXPY(0.37, torch.tensor(X, device='cuda'), B.g.interior())
    A.o.ZAXPY(0.37, torch.tensor(X), B.o.interior())
    A.check('var')
    A.g.vel_corr(B.g.grad_x, 0.21)
    A.o.vel_corr(B.o.grad_x, 0.21)
    A.check('var')
    C = Pair(d, og, 13)
    C.g.copy(torch.tensor(X, device='cuda'))
    C.o.interior().copy_(torch.tensor(X))
    C.check('var')
    C.g.update_border()
    O.halo_fill(C.o.var)
    C.check('var')
