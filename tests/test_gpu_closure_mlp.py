"""Closure MLP kernel (pf_sfs_mlp, MFMA BF16 x BF16 -> FP32).

Two references on the same padded random velocity field:
* a torch-CPU emulation of the kernel's arithmetic (features, weights and
  hidden activations rounded to bfloat16, products accumulated in float32,
  float32 biases and tanh): checks the MFMA tiling/lane layout; bar relative
  L2 <= 2e-3 and max abs <= 2e-2 of the field scale (one-ulp BF16 flips of
  hidden activations are allowed);
* the FP64 oracle (features and MLP in float64): states the BF16 error,
  bar relative L2 <= 5e-2.
Features are the reference's cell-centred gradients (grad_vel_center),
bit-exact on the GPU, so both references see the same inputs.
"""

import math

import numpy as np
import pytest
import torch

from helpers import gpu_env, cpu, rel_l2
from oracle import pyflow_oracle as O

pytestmark = pytest.mark.gpu

SHAPE = (12, 10, 16)
L = (2 * math.pi, 2 * math.pi, 2 * math.pi)


def emulate(feats, params, cst):
    """feats [...,9] float64 -> [...,6] float32, kernel arithmetic on the CPU."""
    from pyflow_amd.sfsmodel_mlp import unpack
    W1, b1, W2, b2, W3, b3 = [torch.as_tensor(p) for p in unpack(params)]
    bf = lambda t: t.to(torch.bfloat16).to(torch.float32)
    x = bf(torch.as_tensor(feats).to(torch.float32))
    h1 = bf(torch.tanh(x @ bf(W1).T + b1))
    h2 = bf(torch.tanh(h1 @ bf(W2).T + b2))
    o = h2 @ bf(W3).T + b3
    return o * torch.tensor(cst, dtype=torch.float32)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
def test_mlp_stresses(dtype):
    import types
    from pyflow_amd import state
    from pyflow_amd.sfsmodel_mlp import residual_stress, init_params, unpack
    d, g, m = gpu_env(SHAPE, L, dtype)
    rng = np.random.Generator(np.random.PCG64(4))
    P = tuple(n + 4 for n in SHAPE)
    vel = [rng.standard_normal(SHAPE) for _ in range(3)]
    sg = [state.state_P(d, v) for v in vel]
    params = init_params(seed=3)
    # widen the weights a little so tanh leaves its linear range
    params = (params * 2.0).astype(np.float32)
    cfg = types.SimpleNamespace(mlpParams=params)
    mlp = residual_stress(cfg, d, g, m)
    mlp.update(*sg)
    torch.cuda.synchronize()
    og = O.Grid(*SHAPE, Lx=L[0], Ly=L[1], Lz=L[2], dtype=torch.float64)
    fo = [O.Field(og, cpu(s.interior())) for s in sg]
    for c, f in zip('uvw', fo):
        O.grad_vel_center(og, f, c)
    feats = torch.stack([fo[0].grad_x, fo[0].grad_y, fo[0].grad_z,
                         fo[1].grad_x, fo[1].grad_y, fo[1].grad_z,
                         fo[2].grad_x, fo[2].grad_y, fo[2].grad_z], dim=-1)
    if dtype == torch.float32:
        feats = feats.to(torch.float32).to(torch.float64)
    emu = emulate(feats.numpy(), params, np.float32(g.dx * g.dx / 12.0)).numpy()
    ref64 = O.mlp_closure_forward(og, *fo, unpack(params.astype(np.float64)))
    R = tuple(slice(0, n + 3) for n in SHAPE)   # the last overlap plane is an image
    names = ('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ')
    errs_e, errs_o = [], []
    for mm, n in enumerate(names):
        got = cpu(getattr(mlp, n))[R]
        e = emu[..., mm][R].astype(np.float64)
        scale = np.abs(e).max()
        errs_e.append(rel_l2(got, e))
        assert np.abs(got - e).max() <= 2e-2 * scale, (n, np.abs(got - e).max(), scale)
        errs_o.append(rel_l2(got, ref64[mm].numpy()[R]))
    print('MLP %s: vs BF16 emulation %s; vs FP64 %s' % (dtype, ['%.1e' % x for x in errs_e],
                                                      ['%.1e' % x for x in errs_o]))
    assert max(errs_e) <= 2e-3
    assert max(errs_o) <= 5e-2


def test_zero_weights_give_bias_only():
    """Closed form: with all weights zero the stresses are b3*dx^2/12 everywhere."""
    import types
    from pyflow_amd import state
    from pyflow_amd.sfsmodel_mlp import residual_stress, N_PARAMS
    d, g, m = gpu_env(SHAPE, L, torch.float64)
    params = np.zeros(N_PARAMS, dtype=np.float32)
    params[-6:] = np.array([1.0, -2.0, 0.5, 0.25, 3.0, -1.5], dtype=np.float32)
    cfg = types.SimpleNamespace(mlpParams=params)
    mlp = residual_stress(cfg, d, g, m)
    rng = np.random.Generator(np.random.PCG64(5))
    sg = [state.state_P(d, rng.standard_normal(SHAPE)) for _ in range(3)]
    mlp.update(*sg)
    torch.cuda.synchronize()
    cst = np.float32(g.dx * g.dx / 12.0)
    for mm, n in enumerate(('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ')):
        got = cpu(getattr(mlp, n))
        assert np.all(got == float(np.float32(params[-6 + mm] * cst)))
