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

Training through this closure (the reference differentiates it with
autograd, adjoint.py:241-286) is not part of this build: update() with
requires_grad=True and finalize() raise.
"""

import numpy as np
import torch

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

    def load(self, path):
        """A checkpoint in the reference's format (sfsmodel_ML.py:192-205)."""
        d = torch.load(path, weights_only=True)
        self.epoch = int(d.get('epoch', 0))
        self.LR = float(d.get('LR', self.LR))
        return pack_state_dict(d['model_state_dict'])

    def state_dict(self):
        return unpack_state_dict(self.params)

    def update(self, u_P, v_P, w_P, requires_grad=False):
        """sfsmodel_ML.py:354-404: GX, GY, GZ from the RHS's var_i arrays."""
        if requires_grad:
            raise NotImplementedError("gradients through the 'source' ML closure are not part "
                                      "of this build")
        self.box.sfs_ml_source(self.params, self.norm, u_P, v_P, w_P, (self.GX, self.GY, self.GZ))

    def finalize(self, comms=None):
        raise NotImplementedError("training the 'source' ML closure is not part of this build")


def main(argv=None):
    """Time one closure evaluation on random padded fields:
    python -m pyflow_amd.sfsmodel_ML --N 64 --dtype f32 --reps 5"""
    import argparse
    import math
    import types
    from . import _native
    ap = argparse.ArgumentParser()
    ap.add_argument('--N', type=int, default=64)
    ap.add_argument('--dtype', choices=('f32', 'f64'), default='f32')
    ap.add_argument('--reps', type=int, default=5)
    a = ap.parse_args(argv)
    dt = torch.float32 if a.dtype == 'f32' else torch.float64
    n = a.N
    h = 2 * math.pi / n
    box = _native.Box(n, n, n, h, h, h, dtype=dt, device=torch.device('cuda'))
    d = types.SimpleNamespace(prec=dt, device=torch.device('cuda'), box=box, nxo_=n + 4,
                              nyo_=n + 4, nzo_=n + 4)
    geo = types.SimpleNamespace(type='uniform', dx=h)
    ml = residual_stress(types.SimpleNamespace(simDt=1e-3), d, geo)
    gen = torch.Generator(device='cuda').manual_seed(0)
    U = [torch.randn((n + 4,) * 3, dtype=dt, device='cuda', generator=gen) for _ in range(3)]
    ml.update(*U)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.reps):
        ml.update(*U)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.reps
    flop = 2.0 * (n + 4) ** 3 * sum(o * i for _, o, i in LAYERS)
    print('ML source closure %d^3 %s: %.3f ms per evaluation, %.1f TFLOP/s of the network'
          % (n, a.dtype, ms, flop / ms / 1e9), flush=True)
    return 0


if __name__ == '__main__':
    import sys
    sys.exit(main())
