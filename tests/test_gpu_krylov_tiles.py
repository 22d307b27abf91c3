"""The tiled Krylov kernels on the decompositions the other solver tests
do not reach:

* 128 x 128 x 64: two z-neighbour cells per thread (64 x 8 tiles), here
  with x chunks of 16 planes in FP64 (128 tiles) and 4 in FP32 (512) - the
  one-plane-ahead load schedule of short chunks;
* 136 x 1024 x 64: the same tiles on chunks of 17 planes - the rotating
  load schedule and non-temporal stores of long chunks (512^3 takes it);
* 85 x 1024 x 45: one cell per thread (45 is no multiple of the 16-wide
  tile), chunks of 17 planes - the rotating schedule of that kernel;
* 91 x 256 x 64 (paired) and 38 x 19 x 45 (one cell per thread): chunks of
  5 planes, an odd count that leaves the two-fold unrolled loop of the
  one-plane-ahead schedule one plane to do on its own, and a last chunk of
  1 and 3 planes.

Ten BiCGStab iterations (pressure.py:229-290) on a random mean-free source
against the oracle's ten: the per-iteration relative residual and the
solution agree to 1e-10 in FP64 (only the inner products' summation order
differs, and BiCGStab amplifies that by a roughly constant factor per
iteration: ~1e-12 after ten).  FP32 storage with FP64 reductions against
the FP64 oracle: residuals and solution to 1e-4.
"""

import math

import numpy as np
import pytest
import torch

from helpers import cpu, rel_l2
from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = {'paired-short-chunks': ((128, 128, 64), (2 * math.pi, 2 * math.pi, math.pi)),
          'paired-long-chunks': ((136, 1024, 64), (2 * math.pi, 8 * math.pi, math.pi)),
          'one-cell-long-chunks': ((85, 1024, 45), (2 * math.pi, 8 * math.pi, 1.1 * math.pi)),
          'paired-odd-ragged-chunks': ((91, 256, 64), (2 * math.pi, 4 * math.pi, math.pi)),
          'one-cell-odd-ragged-chunks': ((38, 19, 45), (2 * math.pi, 1.1 * math.pi, 2.4 * math.pi))}
NIT = 10
RHO, DT = 1.0, 1e-3


def source(shape, seed=3):
    a = np.random.Generator(np.random.PCG64(seed)).standard_normal(shape)
    return a - a.mean()


@pytest.mark.parametrize('case', list(SHAPES))
@pytest.mark.parametrize('dtype,bar', [(torch.float64, 1e-10), (torch.float32, 1e-4)],
                         ids=['f64', 'f32'])
def test_tile_paths_against_oracle(case, dtype, bar):
    from pyflow_amd import _native
    SHAPE, L = SHAPES[case]
    og = O.Grid(*SHAPE, Lx=L[0], Ly=L[1], Lz=L[2])
    box = _native.Box(*SHAPE, og.dx, og.dy, og.dz, dtype=dtype, device=torch.device('cuda'))
    P = tuple(n + 4 for n in SHAPE)
    DP = torch.zeros(P, dtype=dtype, device='cuda')
    p = torch.zeros(P, dtype=dtype, device='cuda')
    src = torch.tensor(source(SHAPE), dtype=dtype, device='cuda')
    box.residual_history_enable(NIT + 4)
    res, its, status = box.bicgstab(DP, p, src, RHO, DT, 1e-30, NIT)
    hg = np.array(box.residual_history())

    torch.set_num_threads(max(1, min(16, torch.get_num_threads())))
    fDP, fp = O.Field(og), O.Field(og)
    B = O.BiCGStab(og, RHO, DT, 1e-30, NIT)
    B.solve(fDP, fp, torch.tensor(source(SHAPE)))
    ho = np.array(B.history)

    assert its == NIT == B.iterations
    rel = np.abs(hg - ho) / ho
    err = rel_l2(cpu(DP), fDP.var.numpy())
    print('%s %s: residual history max rel diff %.2e, solution rel L2 %.2e'
          % (case, dtype, rel.max(), err))
    assert rel.max() <= bar
    assert err <= bar
    assert res == pytest.approx(hg[-1], rel=0, abs=0)
