"""Host-side pieces of closure training (no GPU): the LES box filter, the NGA
config writer, the RMSprop step and the reference's gradient averaging."""

import math
import os
import sys

import numpy as np
import pytest
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'pyflow-public_amd'))

from pyflow_amd.les_targets import box_filter, divergence          # noqa: E402
from pyflow_amd.sfsmodel_mlp import rmsprop_step, GradientAverager  # noqa: E402
from pyflow_amd import dataReader as dr                               # noqa: E402


def solenoidal(n, seed):
    """Discrete curl of a random edge potential: staggered, periodic and
    divergence-free to rounding (the forward differences commute)."""
    g = torch.Generator().manual_seed(seed)
    h = 2 * math.pi / n
    Ax, Ay, Az = (torch.randn(n, n, n, generator=g, dtype=torch.float64) for _ in range(3))
    D = lambda f, d: (torch.roll(f, -1, d) - f) / h       # noqa: E731
    u = D(Az, 1) - D(Ay, 2)
    v = D(Ax, 2) - D(Az, 0)
    w = D(Ay, 0) - D(Ax, 1)
    return u, v, w, h


def test_box_filter_keeps_discrete_divergence_free():
    u, v, w, h = solenoidal(24, 1)
    assert float(divergence(u, v, w, h).abs().max()) < 1e-10
    r = 4
    uc, vc, wc = box_filter(u, v, w, r)
    assert uc.shape == (6, 6, 6)
    d = float(divergence(uc, vc, wc, r * h).abs().max())
    scale = float(uc.abs().max()) / (r * h)
    assert d < 1e-12 * scale, d
    # coarse divergence equals the block mean of the fine divergence for any field
    g = torch.Generator().manual_seed(2)
    f = [torch.randn(24, 24, 24, generator=g, dtype=torch.float64) for _ in range(3)]
    dfine = divergence(*f, h).reshape(6, 4, 6, 4, 6, 4).mean(dim=(1, 3, 5))
    dc = divergence(*box_filter(*f, r), r * h)
    assert torch.allclose(dc, dfine, rtol=0, atol=1e-12 * float(dfine.abs().max()))


def test_box_filter_planted_face_shift_breaks_divergence():
    """The filter must take the coarse face's own fine faces: averaging the
    next fine face instead leaves an O(1) divergence."""
    u, v, w, h = solenoidal(24, 3)
    r = 4
    m = 6
    uc = u[1::r].reshape(m, m, r, m, r).mean(dim=(2, 4))      # the mistake
    _, vc, wc = box_filter(u, v, w, r)
    d = float(divergence(uc, vc, wc, r * h).abs().max())
    assert d > 1e-3 * float(uc.abs().max()) / (r * h)


def test_box_filter_rejects_bad_ratio():
    u = torch.zeros(10, 10, 10)
    with pytest.raises(ValueError):
        box_filter(u, u, u, 4)


def test_nga_config_roundtrip(tmp_path):
    x = np.linspace(0.0, 2 * math.pi, 9)
    y = np.linspace(0.0, 1.0, 7)
    z = np.linspace(-1.0, 1.0, 5)
    f = str(tmp_path / 'config')
    dr.writeNGAconfig(f, x, y, z, configName='box', xper=1, yper=0, zper=1)
    xg, yg, zg, xp, yp, zp = dr.readNGAconfig(f)
    assert np.array_equal(xg, x) and np.array_equal(yg, y) and np.array_equal(zg, z)
    assert (xp, yp, zp) == (1, 0, 1)
    assert os.path.getsize(f) == 64 + 28 + 8 * (9 + 7 + 5) + 4 * 8 * 6


def test_rmsprop_step_matches_torch_optimizer():
    g = torch.Generator().manual_seed(4)
    p0 = torch.randn(746, generator=g)
    a = p0.clone()
    sq = torch.zeros_like(a)
    b = torch.nn.Parameter(p0.clone())
    opt = torch.optim.RMSprop([b], lr=0.01)
    for k in range(5):
        grad = torch.randn(746, generator=g) * (10.0 ** (k - 2))
        rmsprop_step(a, grad, sq, 0.01)
        b.grad = grad.clone()
        opt.step()
        assert torch.allclose(a, b.detach(), rtol=1e-6, atol=1e-7)


def test_gradient_averaging_cadence():
    """sfsmodel_ML.py:305-336: the step uses the mean of the last completed
    group of numSumDistr window gradients (zero before the first)."""
    av = GradientAverager(2, 3, 'cpu')
    seen = []
    for epoch in range(7):
        seen.append(av.push(torch.full((2,), float(epoch + 1), dtype=torch.float64), epoch).tolist())
    assert seen[0] == [0.0, 0.0] and seen[1] == [0.0, 0.0]
    assert seen[2] == [2.0, 2.0] and seen[3] == [2.0, 2.0] and seen[4] == [2.0, 2.0]
    assert seen[5] == [5.0, 5.0] and seen[6] == [5.0, 5.0]
