"""Closure MLP with the 'tensor' contract of sfsmodel_gradient
(sfsmodel/sfsmodel_gradient.py:39-148): the nine cell-centred velocity
gradients of metric_uniform.grad_vel_center -> Linear(9,20) -> tanh ->
Linear(20,20) -> tanh -> Linear(20,6), times dx^2/12, written to FXX, FYY,
FZZ, FXY, FXZ, FYZ.  The reference has no such model (its ML closure is a
'source' model); this one is defined here and evaluated by one HIP kernel
(pf_sfs_mlp) on MFMA: BF16 operands, FP32 accumulation.

Parameters are one packed FP32 vector W1[20][9] b1[20] W2[20][20] b2[20]
W3[6][20] b3[6] with the seeded initialisation of torch.nn.Linear
(uniform in +-1/sqrt(fan_in)).
"""

import ctypes

import numpy as np
import torch

from . import _native

SHAPES = (('W1', (20, 9)), ('b1', (20,)), ('W2', (20, 20)), ('b2', (20,)),
          ('W3', (6, 20)), ('b3', (6,)))
FAN_IN = {'W1': 9, 'b1': 9, 'W2': 20, 'b2': 20, 'W3': 20, 'b3': 20}
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def init_params(seed=0):
    """Packed FP32 parameters with nn.Linear's default ranges, from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for name, shape in SHAPES:
        bound = 1.0 / np.sqrt(FAN_IN[name])
        parts.append(rng.uniform(-bound, bound, size=shape).astype(np.float32).ravel())
    return np.concatenate(parts)


def unpack(flat):
    """Packed vector -> [W1, b1, W2, b2, W3, b3] (numpy or torch, same kind)."""
    out = []
    o = 0
    for _, shape in SHAPES:
        n = int(np.prod(shape))
        out.append(flat[o:o + n].reshape(shape))
        o += n
    return out


class residual_stress:
    def __init__(self, inputConfig, decomp, geo, metric=None):
        self.modelType = 'tensor'
        self.box = decomp.box
        shape = (decomp.nxo_, decomp.nyo_, decomp.nzo_)
        for n in ('FXX', 'FYY', 'FZZ', 'FXY', 'FYZ', 'FXZ'):
            setattr(self, n, torch.zeros(shape, dtype=decomp.prec, device=decomp.device))
        p = getattr(inputConfig, 'mlpParams', None)
        if p is None:
            p = init_params(getattr(inputConfig, 'mlpSeed', 0))
        self.params = torch.as_tensor(np.asarray(p, dtype=np.float32), device=decomp.device)
        if self.params.numel() != N_PARAMS:
            raise ValueError('closure MLP needs %d parameters, got %d' % (N_PARAMS, self.params.numel()))
        if geo.type != 'uniform':
            raise Exception('sfsmodel_mlp: geometry type not implemented')
        self.const = geo.dx * geo.dx / 12.0

    def stresses(self):
        return (self.FXX, self.FYY, self.FZZ, self.FXY, self.FXZ, self.FYZ)

    def update(self, state_u, state_v, state_w, metric=None):
        self.box.sfs_mlp(state_u.var, state_v.var, state_w.var, self.params, self.stresses())
