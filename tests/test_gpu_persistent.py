"""BiCGStab as one persistent cooperative kernel (pf_persist.hip) against the
launch-per-reduction kernels (KA, KB, KC) on the same box and source.

The persistent kernel keeps every block's cells on chip for the whole solve
and meets the other blocks through device-scope partial-sum slots and tile
faces; its blocks reproduce the launch grid's tiles, thread-to-cell map and
summation order, so a solve to convergence must give the same iteration
count, status, per-iteration residual history and solution bit for bit.  A
stale face or partial (a sync that lets a block run ahead) would change the
iterates within a few iterations, so bit identity over a few hundred
iterations is the test of the synchronisation.

Shapes: 128^3 (the headline box: 16-plane chunks in FP64, one per kernel
block, 256 blocks; 4-plane chunks in FP32, two per block at two blocks per
CU), 128 x 128 x 64 (16-plane chunks in FP64, 128 blocks) and 64^3 (4-plane
chunks, one per block: 128 blocks).  The pressure update p += DP (pressure.py:293-294) and the
velocity projection are covered through Domain.velCorr.
"""

import math

import numpy as np
import pytest
import torch

from helpers import cpu

pytestmark = pytest.mark.gpu

RHO, DT = 1.0, 1e-3
CASES = [((128, 128, 128), torch.float64, 1e-10),
         ((128, 128, 128), torch.float32, 1e-6),
         ((128, 128, 64), torch.float64, 1e-10),
         ((64, 64, 64), torch.float64, 1e-10),
         ((64, 64, 64), torch.float32, 1e-6)]


def source(shape, seed=5):
    a = np.random.Generator(np.random.PCG64(seed)).standard_normal(shape)
    return a - a.mean()


def solve(box, shape, dtype, tol, persistent, hist_cap=4096):
    P = tuple(n + 4 for n in shape)
    DP = torch.zeros(P, dtype=dtype, device='cuda')
    p = torch.full(P, 0.25, dtype=dtype, device='cuda')
    src = torch.tensor(source(shape), dtype=dtype, device='cuda')
    box.set_persistent(persistent)
    box.residual_history_enable(hist_cap)
    res, its, status = box.bicgstab(DP, p, src, RHO, DT, tol, 3000)
    torch.cuda.synchronize()
    return (cpu(DP), cpu(p), res, its, status, np.array(box.residual_history()),
            box.last_solve_persistent)


@pytest.mark.parametrize('shape,dtype,tol', CASES,
                         ids=['128c-f64', '128c-f32', '128x128x64-f64', '64c-f64', '64c-f32'])
def test_persistent_solve_bit_identical(shape, dtype, tol):
    from pyflow_amd import _native
    L = (2 * math.pi, 2 * math.pi, 2 * math.pi * shape[2] / shape[0])
    h = tuple(L[d] / shape[d] for d in range(3))
    box = _native.Box(*shape, *h, dtype=dtype, device=torch.device('cuda'))
    assert box.set_persistent(True), 'box has no persistent-kernel buffers'
    print('case %s %s tol %g: launch-per-reduction solve' % (shape, dtype, tol), flush=True)
    ref = solve(box, shape, dtype, tol, False)
    print('  %d iterations, status %d; persistent solve' % (ref[3], ref[4]), flush=True)
    got = solve(box, shape, dtype, tol, True)
    print('  %d iterations, status %d, ran persistent: %s' % (got[3], got[4], got[6]), flush=True)
    assert not ref[6]
    assert got[6], 'the persistent kernel did not run'
    assert got[3] == ref[3] and got[4] == ref[4]
    assert ref[4] == 0 and ref[3] > 50
    assert got[2] == ref[2]
    np.testing.assert_array_equal(got[5], ref[5])
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
    # and once more: the slot and face buffers are reused across solves
    again = solve(box, shape, dtype, tol, True)
    assert again[3] == ref[3]
    np.testing.assert_array_equal(again[0], ref[0])


def test_persistent_iteration_cap_and_history():
    """A solve capped at a few iterations stops there with PF_SOLVE_MAXIT on
    both paths, with the same partial history and iterate."""
    from pyflow_amd import _native
    shape, dtype = (128, 128, 128), torch.float64
    h = 2 * math.pi / 128
    box = _native.Box(*shape, h, h, h, dtype=dtype, device=torch.device('cuda'))
    P = tuple(n + 4 for n in shape)
    outs = []
    for persistent in (False, True):
        DP = torch.zeros(P, dtype=dtype, device='cuda')
        p = torch.zeros(P, dtype=dtype, device='cuda')
        src = torch.tensor(source(shape, seed=9), dtype=dtype, device='cuda')
        box.set_persistent(persistent)
        box.residual_history_enable(16)
        res, its, status = box.bicgstab(DP, p, src, RHO, DT, 1e-30, 7)
        outs.append((cpu(DP), res, its, status, np.array(box.residual_history()),
                     box.last_solve_persistent))
    (a, ra, ia, sa, ha, pa), (b, rb, ib, sb, hb, pb) = outs
    assert not pa and pb
    assert ia == ib == 7 and sa == sb == 1
    assert ra == rb
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(a, b)
