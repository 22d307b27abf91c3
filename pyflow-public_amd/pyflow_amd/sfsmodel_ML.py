"""sfsmodel/sfsmodel_ML.py:119-593 on the GPU: the reference's 'source' ML
closure.  residual_stress.update(u_P, v_P, w_P) receives the var_i arrays
the RHS hands a source closure (velocity.py:365-371) and fills GX, GY, GZ
through one HIP call (pf_sfs_ml_source): 21 FP32 features per padded cell
(values, first and second differences, normalised by the reference's table),
the 273 inputs of the cell and its +-1, +-2 neighbours, NeuralNetworkModel2
(:76-116, 273 -> 200 gated tanh layers -> 18 outputs) on MFMA in FP32, and the
central-difference divergence of output channels 0-8 (:468-593).

Parameters: a checkpoint in the reference's format (loadModel,
modelDictName: a dict with 'model_state_dict', read with weights_only=True),
or `mlSourceParams` (packed vector, the layout of pack_state_dict), or a
seeded draw in torch.nn.Linear's default ranges (`mlSourceSeed`, default 0).

Training (sfsmodel_ML.py:160-352, adjoint.py:241-286).  The reference
differentiates the closure with autograd inside its adjoint RHS; here the
adjoint kernels do it by hand (closure kind 2 of pf_adj_closure_term, the
network's reverse pass on MFMA) and add d(sum G . u_A)/d theta to `grad`, an
FP64 accumulator, on every adjoint stage.  finalize() then follows the
reference step by step: scale by -simDt, add to the running sum, every
`numSumDistr` windows replace the distributed gradient by the mean of the last
`numSumDistr`, take one RMSprop step (torch's defaults) with the distributed
gradient, and halve the step size every `LR_decay` windows from 4/5 of
LR_decay on.  As in the reference, the RMSprop step size starts at 0.01 (or
the checkpoint's optimizer state) whatever `LR` says: the reference builds
its optimizer before it reads inputConfig.LR, which then only sets the
reported rate.
"""

import numpy as np
import torch

from .sfsmodel_mlp import GradientAverager, rmsprop_step

# (name, outputs, inputs) of NeuralNetworkModel2's layers (sfsmodel_ML.py:76-87),
# in the packed order pf_sfs_ml_source takes
LAYERS = (('fc1', 200, 273), ('fcG', 200, 273), ('fcG2', 200, 273),
          ('fc2', 200, 200), ('fc3', 200, 200), ('fc4', 18, 200))
N_PARAMS = sum(o * i + o for _, o, i in LAYERS)

# input normalisation (sfsmodel_ML.py:252-258), as FP32 like the reference's FloatTensor
NORM = (10.0 * np.array([1.0233662e+00, 1.0241578e+00, 1.0196122e+00,
                         5.0992206e+02, 5.2713013e+02, 5.2689368e+02, 5.2550342e+02,
                         5.1325964e+02, 5.2564642e+02, 5.3132422e+02, 5.2574353e+02,
                         5.1027972e+02, 9.5621462e+05, 1.2852805e+06, 1.2698791e+06,
                         1.2744681e+06, 9.6111162e+05, 1.2722050e+06, 1.2891108e+06,
                         1.2891108e+06, 9.6587112e+05])).astype(np.float32)


def init_params(seed=0):
    """Packed FP32 parameters drawn uniformly in +-1/sqrt(fan_in) (the ranges
    of torch.nn.Linear's default initialisation) from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for _, n_out, n_in in LAYERS:
        bound = 1.0 / np.sqrt(n_in)
        parts.append(rng.uniform(-bound, bound, size=n_out * n_in).astype(np.float32))
        parts.append(rng.uniform(-bound, bound, size=n_out).astype(np.float32))
    return np.concatenate(parts)


def pack_state_dict(sd):
    """The reference model's state_dict (fc1.weight, fc1.bias, ...) -> packed vector."""
    parts = []
    for name, n_out, n_in in LAYERS:
        w = torch.as_tensor(sd[name + '.weight']).to(torch.float32).cpu()
        b = torch.as_tensor(sd[name + '.bias']).to(torch.float32).cpu()
        if tuple(w.shape) != (n_out, n_in) or tuple(b.shape) != (n_out,):
            raise ValueError('%s: expected weight (%d, %d) and bias (%d,), got %s and %s'
                             % (name, n_out, n_in, n_out, tuple(w.shape), tuple(b.shape)))
        parts += [w.reshape(-1), b]
    return torch.cat(parts).numpy()


def unpack_state_dict(flat):
    """Packed vector -> state_dict with the reference model's names."""
    flat = torch.as_tensor(flat, dtype=torch.float32).cpu()
    sd, o = {}, 0
    for name, n_out, n_in in LAYERS:
        sd[name + '.weight'] = flat[o:o + n_out * n_in].reshape(n_out, n_in).clone()
        o += n_out * n_in
        sd[name + '.bias'] = flat[o:o + n_out].clone()
        o += n_out
    return sd


class residual_stress:
    def __init__(self, inputConfig, decomp, geo):
        self.prec = decomp.prec
        self.device = decomp.device
        self.box = decomp.box
        self.modelType = 'source'
        self.simDt = getattr(inputConfig, 'simDt', None)
        shape = (decomp.nxo_, decomp.nyo_, decomp.nzo_, 1)
        self.GX = torch.zeros(shape, dtype=self.prec, device=self.device)
        self.GY = torch.zeros(shape, dtype=self.prec, device=self.device)
        self.GZ = torch.zeros(shape, dtype=self.prec, device=self.device)
        # the kernels' feature and network-output workspace, reserved once
        self.box.sfs_ml_source_reserve()
        if geo.type != 'uniform':
            raise Exception('sfsmodel_ML: geometry type not implemented')
        self.dx = geo.dx
        self.const = geo.dx * geo.dx / 12.0
        self.epoch = 0
        self.LR = float(getattr(inputConfig, 'LR', 0.01))
        if getattr(inputConfig, 'loadModel', False):
            flat = self.load(inputConfig.modelDictName)
        else:
            flat = getattr(inputConfig, 'mlSourceParams', None)
            if flat is None:
                flat = init_params(getattr(inputConfig, 'mlSourceSeed', 0))
        self.params = torch.as_tensor(np.asarray(flat, dtype=np.float32)).to(self.device)
        if self.params.numel() != N_PARAMS:
            raise ValueError('sfsmodel_ML needs %d parameters, got %d'
                             % (N_PARAMS, self.params.numel()))
        self.norm = [float(x) for x in NORM]
        # training state (sfsmodel_ML.py:160-252)
        self.grad = torch.zeros(N_PARAMS, dtype=torch.float64, device=self.device)
        self.square_avg = torch.zeros(N_PARAMS, dtype=torch.float32, device=self.device)
        self.averager = GradientAverager(N_PARAMS, getattr(inputConfig, 'numSumDistr', 8),
                                         self.device)
        self.LR_decay = int(getattr(inputConfig, 'LR_decay', 2500))
        self.modelDictSave = getattr(inputConfig, 'modelDictSave', None)
        if self._ckpt is not None:
            self._load_training_state(self._ckpt)
        self._ckpt = None
        if hasattr(inputConfig, 'LR'):
            self.LR = float(inputConfig.LR)
        if getattr(inputConfig, 'adjointTraining', False):
            # the reverse pass's workspace (allocates once); its all-cell
            # cache lets each adjoint RK4 step evaluate the network once
            budget = getattr(inputConfig, 'mlSourceCacheBytes', None)
            if budget is not None:
                self.box.sfs_ml_source_cache_budget(budget)
            self.box.sfs_ml_source_train_reserve(self.norm)

    # the RMSprop step size of the reference's optimizer (sfsmodel_ML.py:164-168)
    step_lr = 0.01
    _ckpt = None

    def load(self, path):
        """A checkpoint in the reference's format (sfsmodel_ML.py:176-205),
        read with weights_only=True: epoch, LR, model_state_dict and, when
        present, the RMSprop state of optimizer_state_dict and the
        gradient-averaging lists (as tensors)."""
        d = torch.load(path, weights_only=True)
        self.epoch = int(d.get('epoch', 0))
        self.LR = float(d.get('LR', self.LR))
        self._ckpt = d
        return pack_state_dict(d['model_state_dict'])

    def _flat(self, tensors):
        return torch.cat([torch.as_tensor(t).reshape(-1).to(torch.float64) for t in tensors]) \
            .to(self.device)

    def _load_training_state(self, d):
        opt = d.get('optimizer_state_dict')
        if opt:
            groups = opt.get('param_groups') or [{}]
            self.step_lr = float(groups[0].get('lr', self.step_lr))
            st = opt.get('state') or {}
            if len(st) == 2 * len(LAYERS):
                self.square_avg.copy_(self._flat(st[i]['square_avg'] for i in range(len(st))))
        for key, attr in (('outParamsList', 'out'), ('sumParamsList', 'sum')):
            lst = d.get(key)
            if lst:
                setattr(self.averager, attr, self._flat(lst))

    def state_dict(self):
        """The model's parameters with the reference model's names."""
        return unpack_state_dict(self.params)

    def save(self):
        """sfsmodel_ML.py:270-280: the reference's checkpoint keys; the
        averaging lists are stored as tensors so weights_only loads read them."""
        if not self.modelDictSave:
            return
        sd = unpack_state_dict(self.params)
        names = list(sd)
        sq = unpack_state_dict(self.square_avg)
        state = {i: {'step': torch.tensor(float(self.epoch)), 'square_avg': sq[n]}
                 for i, n in enumerate(names)}
        group = {'lr': self.step_lr, 'alpha': 0.99, 'eps': 1e-08, 'weight_decay': 0,
                 'momentum': 0, 'centered': False, 'params': list(range(len(names)))}

        def split(v):
            return list(unpack_state_dict(v.to(torch.float32)).values())
        torch.save({'epoch': self.epoch, 'LR': self.LR, 'model_state_dict': sd,
                    'optimizer_state_dict': {'state': state, 'param_groups': [group]},
                    'outParamsList': split(self.averager.out),
                    'sumParamsList': split(self.averager.sum)}, self.modelDictSave)

    def param_grad(self):
        """FP64 accumulator the adjoint kernels add d(sum G . u_A)/d theta to."""
        return self.grad

    def update(self, u_P, v_P, w_P, requires_grad=False):
        """sfsmodel_ML.py:354-404: GX, GY, GZ from the RHS's var_i arrays.
        requires_grad only says that the caller will differentiate; the
        adjoint closure term does that by hand, so the evaluation is the same."""
        self.box.sfs_ml_source(self.params, self.norm, u_P, v_P, w_P, (self.GX, self.GY, self.GZ))

    def finalize(self, comms=None):
        """sfsmodel_ML.py:284-352 on one process: the window's gradient times
        -simDt, the running average over numSumDistr windows, one RMSprop step
        with the distributed average, and the step-size decay."""
        g = self.grad * (-self.simDt)
        self.last_window_grad = g.clone()
        avg = self.averager.push(g, self.epoch)
        rmsprop_step(self.params, avg.to(torch.float32), self.square_avg, self.step_lr)
        self.grad.zero_()
        self.epoch += 1
        if self.epoch >= 4 * self.LR_decay // 5 and self.epoch % self.LR_decay == 0:
            self.LR *= 0.5
            self.step_lr *= 0.5


def timings(N=64, dtype=torch.float32, reps=5,
            which=('forward', 'reverse', 'adjoint_step', 'four_terms')):
    """Milliseconds per call on random padded fields of an N^3 box: the
    closure's forward evaluation, its reverse pass (pf_sfs_ml_source_vjp:
    the forward recomputed, the input- and weight-gradient products, three
    times the forward's flops), the closure work of one adjoint RK4 step
    (closure kind 2 of pf_adj_rk4_step: one network forward and four stage
    reverses, BiCGStab-free) and four separate closure terms; plus the
    network's flops per forward evaluation."""
    import math
    import types
    from . import _native
    n = N
    h = 2 * math.pi / n
    box = _native.Box(n, n, n, h, h, h, dtype=dtype, device=torch.device('cuda'))
    d = types.SimpleNamespace(prec=dtype, device=torch.device('cuda'), box=box, nxo_=n + 4,
                              nyo_=n + 4, nzo_=n + 4)
    geo = types.SimpleNamespace(type='uniform', dx=h)
    ml = residual_stress(types.SimpleNamespace(simDt=1e-3, adjointTraining=True), d, geo)
    gen = torch.Generator(device='cuda').manual_seed(0)
    shape = (n + 4,) * 3
    U = [torch.randn(shape, dtype=dtype, device='cuda', generator=gen) for _ in range(3)]
    Gb = [torch.randn(shape, dtype=dtype, device='cuda', generator=gen) for _ in range(3)]
    dI = [torch.empty(shape, dtype=torch.float32, device='cuda') for _ in range(3)]
    A = [0.01 * torch.randn(shape, dtype=dtype, device='cuda', generator=gen) for _ in range(3)]
    At = [torch.zeros_like(x) for x in A]
    delta = [torch.zeros_like(x) for x in A]

    def forward():
        ml.update(*U)

    def reverse():
        box.sfs_ml_source_vjp(ml.params, *U, Gb, dI, ml.grad)

    def adjoint_step():
        box.adj_rk4_step(*U, *A, *At, 1e-3, 0.01, 1.0, 2, ml.params, None, delta, ml.grad)

    def four_terms():
        for _ in range(4):
            box.adj_closure_term(2, *U, ml.params, *A, None, delta, ml.grad)

    out = {}
    for name, fn in (('forward', forward), ('reverse', reverse), ('adjoint_step', adjoint_step),
                     ('four_terms', four_terms)):
        if name not in which:
            continue
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    out['flop_forward'] = 2.0 * (n + 4) ** 3 * sum(o * i for _, o, i in LAYERS)
    box.close()
    return out


def main(argv=None):
    """python -m pyflow_amd.sfsmodel_ML --N 64 --dtype f32 --reps 5
    [--reverse | --adjoint-step | --all]: the closure's forward evaluation
    (default), its reverse pass, or the closure work of one adjoint RK4 step
    beside four separate terms (timings()), as ms and TFLOP/s."""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument('--N', type=int, default=64)
    ap.add_argument('--dtype', choices=('f32', 'f64'), default='f32')
    ap.add_argument('--reps', type=int, default=5)
    ap.add_argument('--reverse', action='store_true')
    ap.add_argument('--adjoint-step', action='store_true')
    ap.add_argument('--all', action='store_true')
    a = ap.parse_args(argv)
    which = ('forward',)
    if a.reverse:
        which = ('reverse',)
    if a.adjoint_step:
        which = ('adjoint_step', 'four_terms')
    if a.all:
        which = ('forward', 'reverse', 'adjoint_step', 'four_terms')
    t = timings(a.N, torch.float32 if a.dtype == 'f32' else torch.float64, a.reps, which)
    f = t['flop_forward']
    head = 'ML source closure %d^3 %s:' % (a.N, a.dtype)
    if 'forward' in t:
        print('%s forward %.3f ms per evaluation, %.1f TFLOP/s of the network'
              % (head, t['forward'], f / t['forward'] / 1e9), flush=True)
    if 'reverse' in t:
        print('%s reverse %.3f ms per evaluation, %.1f TFLOP/s of the network'
              % (head, t['reverse'], 3 * f / t['reverse'] / 1e9), flush=True)
    if 'adjoint_step' in t:
        print('%s adjoint RK4 step (forward once, stage sums) %.3f ms; four separate closure '
              'terms %.3f ms' % (head, t['adjoint_step'], t['four_terms']), flush=True)
    return 0


if __name__ == '__main__':
    import sys
    sys.exit(main())
