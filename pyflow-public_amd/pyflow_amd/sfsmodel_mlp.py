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

Training follows the reference's ML closure (sfsmodel/sfsmodel_ML.py:123-349):
the adjoint step accumulates d(sum tau . taubar)/d theta into `grad` on every
adjoint RHS evaluation (the fused adjoint kernels write it); `finalize`
scales it by the time step, keeps the reference's running average over
`numSumDistr` windows, takes a torch-default RMSprop step (lr 0.01, alpha
0.99, eps 1e-8) with the averaged gradient, and halves the learning rate
every `LR_decay` windows after 4/5 of the first period.
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


def rmsprop_step(params, grad, square_avg, lr, alpha=0.99, eps=1e-8):
    """One step of torch.optim.RMSprop with its defaults (no momentum, not
    centred, no weight decay), in place on float32 tensors."""
    square_avg.mul_(alpha).addcmul_(grad, grad, value=1.0 - alpha)
    avg = square_avg.sqrt().add_(eps)
    params.addcdiv_(grad, avg, value=-lr)


class GradientAverager:
    """The reference's window-gradient bookkeeping (sfsmodel_ML.py:284-352):
    accumulate each window's gradient; every `numSumDistr` windows the mean of
    the last `numSumDistr` becomes the gradient every following step uses
    (zero until the first mean exists)."""

    def __init__(self, n, numSumDistr, device):
        self.numSumDistr = int(numSumDistr)
        self.out = torch.zeros(n, dtype=torch.float64, device=device)
        self.sum = torch.zeros(n, dtype=torch.float64, device=device)

    def push(self, g, epoch):
        self.sum += g
        if (epoch + 1) % self.numSumDistr == 0:
            self.out = self.sum / float(self.numSumDistr)
            self.sum = torch.zeros_like(self.sum)
        return self.out


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
        # training state (sfsmodel_ML.py:160-252)
        dev = self.params.device
        self.grad = torch.zeros(N_PARAMS, dtype=torch.float64, device=dev)
        self.simDt = getattr(inputConfig, 'simDt', None)
        self.LR = float(getattr(inputConfig, 'LR', 0.01))
        self.epoch = 0
        self.LR_decay = int(getattr(inputConfig, 'LR_decay', 2500))
        self.square_avg = torch.zeros(N_PARAMS, dtype=torch.float32, device=dev)
        self.averager = GradientAverager(N_PARAMS, getattr(inputConfig, 'numSumDistr', 8), dev)
        self.modelDictSave = getattr(inputConfig, 'modelDictSave', None)
        if getattr(inputConfig, 'loadModel', False):
            self.load(inputConfig.modelDictName)

    def param_grad(self):
        """FP64 accumulator the adjoint kernels add d(sum tau . taubar)/d theta to."""
        return self.grad

    def finalize(self, comms=None):
        """sfsmodel_ML.py:284-352 for this closure.  The reference scales its
        accumulated gradient by -simDt because its 'source' closure is
        subtracted from the RHS; this closure's stresses enter the RHS with a
        plus sign, so the same objective gradient is +simDt times the
        accumulation."""
        g = self.grad * self.simDt
        self.last_window_grad = g.clone()
        avg = self.averager.push(g, self.epoch)
        rmsprop_step(self.params, avg.to(torch.float32), self.square_avg, self.LR)
        self.grad.zero_()
        self.epoch += 1
        if self.epoch >= 4 * self.LR_decay // 5 and self.epoch % self.LR_decay == 0:
            self.LR *= 0.5

    def state_dict(self):
        return {'epoch': self.epoch, 'LR': self.LR, 'params': self.params.cpu(),
                'square_avg': self.square_avg.cpu(), 'outParams': self.averager.out.cpu(),
                'sumParams': self.averager.sum.cpu()}

    def save(self):
        """sfsmodel_ML.py:270-280 (tensors only)."""
        if self.modelDictSave:
            torch.save(self.state_dict(), self.modelDictSave)

    def load(self, path):
        d = torch.load(path, weights_only=True)
        dev = self.params.device
        self.epoch = int(d['epoch'])
        self.LR = float(d['LR'])
        self.params.copy_(d['params'].to(dev))
        self.square_avg.copy_(d['square_avg'].to(dev))
        self.averager.out = d['outParams'].to(dev)
        self.averager.sum = d['sumParams'].to(dev)

    def stresses(self):
        return (self.FXX, self.FYY, self.FZZ, self.FXY, self.FXZ, self.FYZ)

    def update(self, state_u, state_v, state_w, metric=None):
        self.box.sfs_mlp(state_u.var, state_v.var, state_w.var, self.params, self.stresses())
