"""Closure training against filtered-DNS targets: the reference's
adjoint-training driver (examples/adjoint_training/
run_PyFlow_adjoint_training.py:43-141, which restarts from a target file and
trains its closure with adjointTraining, numCheckpointIt and
getTargetFileName), with the closure MLP or the reference's own 'source' ML
closure (--closure ml).

    python -m pyflow_amd.les_train --targets DIR --make --N 256 --ratio 4 \\
        --window 20 --windows 2 --dtype f32 [--closure ml]

--make first writes the targets (pyflow_amd.les_targets: a Passot-Pouquet
DNS on N^3, box-filtered ratio x onto the LES grid, one file every `window`
LES steps).  The run restarts from the first target and trains the MLP
closure (9 velocity gradients -> 20 -> 20 -> 6 stresses, tanh; its forward
on MFMA in BF16 with FP32 accumulation) over `windows` checkpointed windows
of `window` RK4 steps: forward steps, the adjoint seeded at the window's
target, the adjoint steps with the closure's parameter gradient, RMSprop.
Prints one JSON line: the per-window model error, seconds per window and
the forward + adjoint rate in Mcell-steps/s.
"""

import argparse
import json
import os
import sys
import time
import types

import numpy as np
import torch


def config(meta, window, windows, dtype, params, LR=0.01, tol=None, closure='mlp'):
    """The run's input configuration (attribute names of the reference's
    inputConfigClass)."""
    f32 = dtype == torch.float32
    base = meta['base']
    sfs = {'SFSmodel': 'MLP', 'mlpParams': params} if closure == 'mlp' else \
        {'SFSmodel': 'ML', 'mlSourceParams': params}
    cfg = types.SimpleNamespace(
        configName='restart', configFileStr=meta['config'], dataFileStr=meta['files'][0],
        dataFileType='restart', nproc_x=1, nproc_y=1, nproc_z=1, mu=meta['nu'], rho=1.0,
        simDt=meta['dt_les'], numIt=window * windows, **sfs,
        advancerName='RK4', equationMode='NS', pSolverMode='bicgstab',
        min_pressure_residual=tol if tol is not None else (1e-6 if f32 else 1e-12),
        max_pressure_iterations=3000, dtypeTorch=dtype,
        dtypeNumpy=np.float32 if f32 else np.float64, adjointTraining=True,
        numCheckpointIt=window, adjointVerification=False, startFileIt=0, LR=LR,
        numSumDistr=1, writeMonitor=False)
    cfg.getTargetFileName = lambda it: base + '{:08d}'.format(it)
    return cfg


def main(argv=None):
    from . import PyFlow, les_targets
    from .sfsmodel_mlp import init_params
    ap = argparse.ArgumentParser()
    ap.add_argument('--targets', required=True)
    ap.add_argument('--make', action='store_true', help='write the targets first')
    ap.add_argument('--N', type=int, default=256, help='DNS grid (with --make)')
    ap.add_argument('--ratio', type=int, default=4, help='filter ratio (with --make)')
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--nu', type=float, default=0.005)
    ap.add_argument('--window', type=int, default=20, help='LES steps per training window')
    ap.add_argument('--windows', type=int, default=2)
    ap.add_argument('--dtype', choices=('f32', 'f64'), default='f32')
    ap.add_argument('--mlp-seed', type=int, default=13, help='seed of the closure weights')
    ap.add_argument('--closure', choices=('mlp', 'ml'), default='mlp',
                    help="the closure MLP or the reference's 'source' ML closure")
    a = ap.parse_args(argv)
    dtype = torch.float32 if a.dtype == 'f32' else torch.float64
    if a.make:
        t = time.perf_counter()
        les_targets.generate(a.targets, N=a.N, ratio=a.ratio, steps=a.window * a.windows,
                             every=a.window, seed=a.seed, nu=a.nu, verbose=False)
        print('targets: %d^3 DNS filtered %dx, %.1f s' % (a.N, a.ratio, time.perf_counter() - t),
              flush=True)
    with open(os.path.join(a.targets, 'targets.json')) as f:
        meta = json.load(f)
    if meta['every'] != a.window or meta['steps'] < a.window * a.windows:
        raise SystemExit('the targets hold one file every %d steps up to %d; the run needs '
                         'every %d up to %d' % (meta['every'], meta['steps'], a.window,
                                                a.window * a.windows))
    if a.closure == 'ml':
        from .sfsmodel_ML import init_params as init_ml
        params = init_ml(a.mlp_seed)
    else:
        params = init_params(a.mlp_seed)
    cfg = config(meta, a.window, a.windows, dtype, params, closure=a.closure)
    torch.cuda.synchronize()
    res = PyFlow.simulate(cfg, verbose=False)
    torch.cuda.synchronize()
    m = meta['N'] // meta['ratio']
    out = {'closure': a.closure, 'les_grid': m, 'dns_grid': meta['N'], 'dtype': a.dtype,
           'window_steps': a.window,
           'windows': a.windows, 'wall_s': round(res.wall, 3),
           's_per_window': round(res.wall / a.windows, 3),
           'forward_plus_adjoint_mcell_steps_per_s':
               round(m ** 3 * a.window * a.windows / res.wall / 1e6, 2),
           'model_error': [round(h['model_error'], 6) for h in res.train_history],
           'norm_error': [round(h['norm_error'], 6) for h in res.train_history],
           'TKE': [round(h['TKE'], 6) for h in res.train_history],
           'poisson_iters_mean': round(float(np.mean([r['iters'] for r in res.history[1:]])), 1)}
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
