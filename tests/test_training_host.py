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


class _HostBox:
    """Stands in for the GPU box: the ML closure's constructor only reserves
    workspaces through it, and finalize() is host-side tensor arithmetic."""

    def __init__(self):
        self.reserved = []

    def sfs_ml_source_reserve(self):
        self.reserved.append('forward')

    def sfs_ml_source_train_reserve(self, norm):
        self.reserved.append(('train', len(norm)))


def _ml_closure(**cfg):
    import types
    from pyflow_amd import sfsmodel_ML
    d = types.SimpleNamespace(prec=torch.float64, device=torch.device('cpu'), box=_HostBox(),
                              nxo_=6, nyo_=6, nzo_=6)
    geo = types.SimpleNamespace(type='uniform', dx=0.5)
    return sfsmodel_ML.residual_stress(types.SimpleNamespace(**cfg), d, geo)


def test_ml_closure_finalize_matches_reference_bookkeeping():
    """sfsmodel_ML.py:284-352 with torch.optim.RMSprop over the model's twelve
    parameter tensors: grad *= -simDt, running sums, the mean every
    numSumDistr windows, the optimizer step, and the halving of the step size
    at epoch LR_decay (>= 4/5 LR_decay).  The reference builds its optimizer
    with lr 0.01 before it reads inputConfig.LR, so an LR override changes the
    reported rate only.  The window gradients keep one sign per parameter:
    this closure sums them in FP64 where the reference sums FP32 arrays, and
    sums that cancel would turn that rounding into a visible difference of
    the RMSprop step wherever the mean falls to the size of its eps."""
    from pyflow_amd import sfsmodel_ML
    simDt, nsum, decay = 2.5e-3, 2, 5
    ml = _ml_closure(simDt=simDt, numSumDistr=nsum, LR_decay=decay, LR=0.003,
                     mlSourceSeed=4, adjointTraining=True)
    assert ml.box.reserved == ['forward', ('train', 21)]
    sd = sfsmodel_ML.unpack_state_dict(ml.params.clone())
    params = [torch.nn.Parameter(t.clone()) for t in sd.values()]
    opt = torch.optim.RMSprop(params, lr=0.01)
    out = [torch.zeros_like(p) for p in params]
    acc = [torch.zeros_like(p) for p in params]
    gen = torch.Generator().manual_seed(9)
    sign = torch.sign(torch.randn(sfsmodel_ML.N_PARAMS, generator=gen))
    for epoch in range(7):
        mag = 0.5 + torch.rand(sfsmodel_ML.N_PARAMS, generator=gen)
        g32 = sign * mag * 10.0 ** (epoch % 3)
        ml.grad.copy_(g32.double())
        ml.finalize()
        # the reference's steps, on its per-tensor lists
        o = 0
        for i, p in enumerate(params):
            n = p.numel()
            gi = g32[o:o + n].reshape(p.shape) * -simDt
            o += n
            acc[i] += gi
            if (epoch + 1) % nsum == 0:
                out[i] = acc[i] / float(nsum)
                acc[i] = torch.zeros_like(p)
            p.grad = out[i].clone()
        opt.step()
        opt.zero_grad()
        if epoch + 1 >= 4 * decay // 5 and (epoch + 1) % decay == 0:
            for grp in opt.param_groups:
                grp['lr'] *= 0.5
        ref = torch.cat([p.detach().reshape(-1) for p in params])
        err = float((ml.params - ref).abs().max() / ref.abs().max())
        assert err < 1e-6, (epoch, err)
        assert float(ml.grad.abs().max()) == 0.0
    assert ml.epoch == 7
    assert ml.step_lr == opt.param_groups[0]['lr'] == 0.005
    assert ml.LR == 0.0015


def test_ml_closure_checkpoint_roundtrip(tmp_path):
    """save() writes the reference's keys (epoch, LR, model_state_dict,
    optimizer_state_dict, outParamsList, sumParamsList); loading it with
    weights_only=True restores the parameters, the RMSprop state, the
    averaging sums and the step size."""
    path = str(tmp_path / 'ml.pt')
    a = _ml_closure(simDt=1e-3, numSumDistr=3, LR_decay=2, mlSourceSeed=2, modelDictSave=path)
    gen = torch.Generator().manual_seed(1)
    for _ in range(4):
        a.grad.copy_(torch.randn(a.grad.numel(), generator=gen, dtype=torch.float64))
        a.finalize()
    a.save()
    d = torch.load(path, weights_only=True)
    assert set(d) == {'epoch', 'LR', 'model_state_dict', 'optimizer_state_dict',
                      'outParamsList', 'sumParamsList'}
    assert list(d['model_state_dict']) == ['fc1.weight', 'fc1.bias', 'fcG.weight', 'fcG.bias',
                                           'fcG2.weight', 'fcG2.bias', 'fc2.weight', 'fc2.bias',
                                           'fc3.weight', 'fc3.bias', 'fc4.weight', 'fc4.bias']
    b = _ml_closure(simDt=1e-3, numSumDistr=3, LR_decay=2, loadModel=True, modelDictName=path)
    assert b.epoch == a.epoch == 4 and b.step_lr == a.step_lr == 0.0025
    assert torch.equal(b.params, a.params)
    assert torch.equal(b.square_avg, a.square_avg)
    assert torch.allclose(b.averager.out, a.averager.out, rtol=1e-6, atol=0)
    assert torch.allclose(b.averager.sum, a.averager.sum, rtol=1e-6, atol=0)
