"""The run loop's monitor values (pf_stats; core/PyFlow.py:378-395) and the
drop-in Pressure_Base.applyOperator (pf_apply_laplacian; solver/pressure.py:54-84)
on the GPU against the oracle, whose Box.stats and apply_operator are pinned by
closed forms (tests/test_oracle_stats.py) and golden vectors.

Bars: max|U|, max|V|, max|W| and max|div| bit-exact (abs and max are exact,
and the divergence is evaluated op for op); the interior divergence bit-exact;
TKE 1e-14 in FP64 (the kernel sums in FP64 in another order than torch.sum);
in FP32 the reference sums float32 squares with float32 accumulation while the
kernel accumulates in FP64, so TKE differs by float32 summation error: 1e-6.
applyOperator bit-exact in both precisions, including the overlap fill of A."""

import math

import numpy as np
import pytest
import torch

from helpers import gpu_env, cpu
from oracle import pyflow_oracle as O
import test_oracle_stats as S

pytestmark = pytest.mark.gpu


def gpu_stats(shape, L, dtype, fields):
    d, g, m = gpu_env(shape, L, dtype)
    U = [torch.tensor(f.numpy(), dtype=dtype, device='cuda') for f in fields]
    divg = torch.zeros(shape, dtype=dtype, device='cuda')
    st = d.box.stats(*U, divg)
    torch.cuda.synchronize()
    return dict(zip(('maxU', 'maxV', 'maxW', 'TKE', 'maxDiv'), st)), cpu(divg)


def compare(st, ref, divg, ref_div, tke_tol):
    for k in ('maxU', 'maxV', 'maxW', 'maxDiv'):
        assert st[k] == ref[k], (k, st[k], ref[k])
    assert abs(st['TKE'] - ref['TKE']) <= tke_tol * ref['TKE'], (st['TKE'], ref['TKE'])
    assert np.array_equal(divg, ref_div)


def test_stats_closed_form_field_with_overlap_junk():
    g, box = S.closed_form_box()
    ref = box.stats()
    st, divg = gpu_stats((S.NX, S.NY, S.NZ), (S.LX, S.LY, S.LZ), torch.float64,
                         [box.u.var, box.v.var, box.w.var])
    print('GPU', st, 'oracle', ref)
    compare(st, ref, divg, box.source.numpy(), 1e-14)
    S.check(st)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_stats_ragged_random_field(dtype):
    shape = (37, 19, 45)
    L = (2 * math.pi, 1.3 * math.pi, 2.2 * math.pi)
    og = O.Grid(*shape, *L, dtype=dtype)
    gen = torch.Generator().manual_seed(5)
    u, v, w = (torch.randn(shape, generator=gen, dtype=torch.float64) * s for s in (1.0, 0.6, 1.7))
    box = O.Box(og, u, v, w, solver='none')
    ref = box.stats()
    st, divg = gpu_stats(shape, L, dtype, [box.u.var, box.v.var, box.w.var])
    print(dtype, 'GPU', st, 'oracle', ref)
    compare(st, ref, divg, box.source.numpy(), 1e-14 if dtype == torch.float64 else 1e-6)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_apply_operator_bit_exact(dtype):
    from pyflow_amd import pressure
    shape = (37, 19, 45)
    L = (2 * math.pi, 1.3 * math.pi, 2.2 * math.pi)
    d, g, m = gpu_env(shape, L, dtype)
    og = O.Grid(*shape, *L, dtype=dtype)
    gen = torch.Generator().manual_seed(8)
    a = torch.randn(tuple(n + 4 for n in shape), generator=gen, dtype=torch.float64).to(dtype)
    # the overlap holds junk on entry: applyOperator fills it (communicate_border)
    A_ref = a.clone()
    B_ref = og.interior_zeros()
    O.apply_operator(og, O.laplace_coefficients(og), A_ref, B_ref)
    A = a.cuda()
    B = torch.full(shape, float('nan'), dtype=dtype, device='cuda')
    pressure.Pressure_Base(d, m).applyOperator(A, B)
    torch.cuda.synchronize()
    assert torch.equal(A.cpu(), A_ref)
    assert torch.equal(B.cpu(), B_ref)
