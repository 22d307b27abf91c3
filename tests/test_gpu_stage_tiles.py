"""The tiled stage kernel on a box that no tile or chunk divides.

A block of the stage kernel owns a 32 (z) x 8 (y) tile and marches along x
over a chunk of planes, reading the stencil from an LDS ring of planes.  On
37 x 19 x 45 cells the last y and z tiles are partial (their threads past the
padding read clamped cells and store nothing), the x axis splits into
chunks of 10, 10, 10 and 7 planes, and the three spacings differ.  RK4, Euler
and the bare RHS, with a constant viscosity, a viscosity field and tensor
stresses, are compared with the oracle's restatement of velocity.py:221-383
and advancer.py:49-113, in FP64 and in FP32 (the oracle then issues the
same ops on float32 tensors).  Elementwise arithmetic only: bit-exact.
"""

import math

import numpy as np
import pytest
import torch

from helpers import cpu
from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (37, 19, 45)
L = (2 * math.pi, 1.1 * math.pi, 2.4 * math.pi)
MU, RHO, DT = 0.013, 1.3, 2e-3


def rnd(seed, shape=SHAPE):
    return np.random.Generator(np.random.PCG64(seed)).standard_normal(shape)


class Tensor6:
    modelType = 'tensor'


def make_case(visc_field, stresses, dtype=torch.float64):
    og = O.Grid(*SHAPE, Lx=L[0], Ly=L[1], Lz=L[2], dtype=dtype)
    u = [O.Field(og, rnd(s)) for s in (1, 2, 3)]
    if visc_field:
        VISC = og.from_interior(MU * (1.0 + 0.5 * np.abs(rnd(4))))
    else:
        VISC = torch.ones(og.nxo, og.nyo, og.nzo, dtype=dtype) * MU
    sfs = None
    if stresses:
        sfs = Tensor6()
        for k, n in enumerate(('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ')):
            setattr(sfs, n, og.from_interior(0.1 * rnd(10 + k)))
    return og, u, VISC, sfs


def gpu_box(og):
    from pyflow_amd import _native
    return _native.Box(*SHAPE, og.dx, og.dy, og.dz, dtype=og.dtype,
                       device=torch.device('cuda'))


def dev(t):
    return t.clone().to('cuda')


def sfs6(sfs):
    if sfs is None:
        return None
    return [dev(getattr(sfs, n)) for n in ('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ')]


CASES = [(False, False), (True, False), (False, True), (True, True)]
IDS = ['const-mu', 'visc-field', 'stresses', 'visc-field+stresses']
DTYPES = pytest.mark.parametrize('dtype', [torch.float64, torch.float32], ids=['f64', 'f32'])


@DTYPES
@pytest.mark.parametrize('visc_field,stresses', CASES, ids=IDS)
def test_rk4_step_bit_exact(visc_field, stresses, dtype):
    og, u, VISC, sfs = make_case(visc_field, stresses, dtype)
    box = gpu_box(og)
    U = [dev(f.var) for f in u]
    Ut = [torch.zeros_like(a) for a in U]
    S = sfs6(sfs)
    if visc_field:
        box.rk4_step_visc(*U, *Ut, DT, dev(VISC), RHO, S)
    else:
        box.rk4_step(*U, *Ut, DT, MU, RHO, S)
    t = [O.Field(og) for _ in range(3)]
    rhss = [O.NavierStokesRHS(og, RHO, VISC, sfs) for _ in range(4)]
    O.rk4_step(*u, *t, rhss, DT)
    torch.cuda.synchronize()
    for q in range(3):
        assert np.array_equal(cpu(U[q]), u[q].var.numpy()), ('u*', q)
        assert np.array_equal(cpu(Ut[q]), t[q].var.numpy()), ('Y4', q)


@DTYPES
@pytest.mark.parametrize('visc_field,stresses', CASES, ids=IDS)
def test_euler_step_and_rhs_bit_exact(visc_field, stresses, dtype):
    og, u, VISC, sfs = make_case(visc_field, stresses, dtype)
    box = gpu_box(og)
    U = [dev(f.var) for f in u]
    S = sfs6(sfs)
    R = [torch.empty(SHAPE, dtype=dtype, device='cuda') for _ in range(3)]
    if visc_field:
        box.rhs_ns_visc(*U, dev(VISC), RHO, *R, S)
        box.euler_step_visc(*U, DT, dev(VISC), RHO, S)
    else:
        box.rhs_ns(*U, MU, RHO, *R, S)
        box.euler_step(*U, DT, MU, RHO, S)
    rhs = O.NavierStokesRHS(og, RHO, VISC, sfs)
    rhs.evaluate(*u)
    ref_rhs = [rhs.rhs_u.clone(), rhs.rhs_v.clone(), rhs.rhs_w.clone()]
    O.euler_step(*u, rhs, DT)
    torch.cuda.synchronize()
    for q in range(3):
        assert np.array_equal(cpu(R[q]), ref_rhs[q].numpy()), ('rhs', q)
        assert np.array_equal(cpu(U[q]), u[q].var.numpy()), ('euler', q)
