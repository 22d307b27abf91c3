"""Throughput of PyFlow's time step on one MI355X.

Workload (default): decaying isotropic turbulence, triply periodic box,
L = 2*pi, 128^3 staggered grid, FP64, seeded Passot-Pouquet field (k0 = 4,
u_rms = 1), nu = 0.005, rho = 1, fixed dt = 0.3*h/max|u0|, RK4 + BiCGStab to
a relative L-inf residual of 1e-10 -- the reference's forward loop
(Domain.forwardStep + the per-step monitor values) through this package's
PyFlow.setup.  A step is one full time step of that loop.

Prints one JSON line.  `value` is Mcell-steps/s of the timed steps; the
roofline object is for the dominant kernel (KA of BiCGStab) with its average
launch duration measured during the timed steps with start/stop hipEvents
recorded by the dispatch of every 32nd launch of each Krylov kernel (sampling
keeps the events' own cost out of `value`); `cpu_baseline` times the CPU oracle (a
torch-CPU FP64 restatement of the reference's op sequence) on a bounded
sample of the same workload on this host, and `oracle_error` is the relative
L2 error of the GPU velocity against the oracle's after those steps.
"""

import argparse
import json
import math
import os
import sys
import time
import types

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, 'pyflow-public_amd')):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0

# HBM-side bytes per KA launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on this bench's workload (one pass per counter), each scaled by
# the counter's reading for a plain 512 MiB device copy (FETCH 0.500,
# WRITE 1.000): (38403.6 KiB / 0.500 + 32099.1 KiB) per dispatch.  Keyed by
# (grid, dtype); other workloads report null.
KA_TRAFFIC_BYTES = {(128, 'f64'): int((38403.6 / 0.50003 + 32099.1) * 1024)}


def make_config(N, dtype, nu, tol, maxit, seed, solver='bicgstab'):
    import numpy as np
    import torch
    return types.SimpleNamespace(
        configName='passotPouquet', Nx=N, Ny=N, Nz=N, Lx=2 * math.pi, seed=seed, k0=4.0, urms=1.0,
        nproc_x=1, nproc_y=1, nproc_z=1, mu=nu, rho=1.0, simDt=None, numIt=0,
        SFSmodel='none', advancerName='RK4', equationMode='NS', pSolverMode=solver,
        min_pressure_residual=tol, max_pressure_iterations=maxit,
        dtypeTorch=dtype, dtypeNumpy=np.float64 if dtype == torch.float64 else np.float32,
        writeMonitor=False)


def gpu_run(N, dtype, nu, tol, maxit, steps, warmup, seed, profile=True, solver='bicgstab'):
    import torch
    from pyflow_amd import PyFlow
    from pyflow_amd.initial_conditions import passot_pouquet, cfl_dt
    u0, v0, w0 = passot_pouquet(N, k0=4.0, urms=1.0, seed=seed)
    dt = cfl_dt(u0, v0, w0, 2 * math.pi / N)
    cfg = make_config(N, dtype, nu, tol, maxit, seed, solver)
    cfg.simDt = dt
    comms, decomp, geometry, metric, D, t0 = PyFlow.setup(cfg)
    for _ in range(warmup):
        D.forwardStep(dt)
        D.stats()
    box = decomp.box
    if profile:
        box.profile_enable(profile)
    torch.cuda.synchronize()
    its = []
    status = []
    t = time.perf_counter()
    for _ in range(steps):
        D.forwardStep(dt)
        D.stats()
        its.append(D.poisson.iterations)
        status.append(D.poisson.status)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    prof = box.profile_read() if profile else None
    if profile:
        box.profile_enable(False)
    st = D.stats()
    return dict(wall=wall, its=its, prof=prof, dt=dt, stats=st, u0=(u0, v0, w0),
                res=D.max_resP, status={k: status.count(k) for k in set(status)})


def gpu_fields(N, dtype, nu, tol, maxit, steps, seed, dt, solver='bicgstab'):
    """Interior velocity after `steps` steps from the initial field (for the
    field error against the oracle)."""
    import torch
    from pyflow_amd import PyFlow
    cfg = make_config(N, dtype, nu, tol, maxit, seed, solver)
    cfg.simDt = dt
    comms, decomp, geometry, metric, D, t0 = PyFlow.setup(cfg)
    for _ in range(steps):
        D.forwardStep(dt)
    torch.cuda.synchronize()
    return [s.interior().double().cpu() for s in (D.state_u_P, D.state_v_P, D.state_w_P)]


def cpu_baseline(u0, v0, w0, N, nu, dt, tol, maxit, steps=1):
    """The CPU oracle (torch-CPU FP64) on `steps` steps of the same workload;
    returns its throughput and final interior velocity."""
    import torch
    from oracle import pyflow_oracle as O
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    torch.set_num_threads(cores)
    g = O.Grid(N, N, N)
    B = O.Box(g, u0, v0, w0, mu=nu, rho=1.0, dt=dt, tol=tol, maxit=maxit)
    t = time.perf_counter()
    for _ in range(steps):
        B.forward_step()
        B.stats()
    wall = time.perf_counter() - t
    return dict(value=N ** 3 * steps / wall / 1e6, wall=wall, cores=cores,
                iters=B.poisson.iterations,
                fields=[B.u.interior().clone(), B.v.interior().clone(), B.w.interior().clone()])


def calibrate_copy(mib):
    """A plain device copy of known size (the reference point for the HBM byte
    counters): 5 copies of `mib` MiB, timed with events (read + write bytes per
    second of the last 4, i.e. the achievable streaming bandwidth)."""
    import torch
    n = int(mib) << 20
    a = torch.empty(n, dtype=torch.uint8, device='cuda')
    b = torch.ones(n, dtype=torch.uint8, device='cuda')
    a.copy_(b)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(4):
        a.copy_(b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 4
    print(json.dumps({'copy_bytes': n, 'copies': 5, 'ms_per_copy': round(ms, 4),
                      'GB/s_read_plus_write': round(2 * n / ms / 1e6, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--N', type=int, default=128)
    ap.add_argument('--dtype', choices=['f64', 'f32'], default='f64')
    ap.add_argument('--nu', type=float, default=0.005)
    ap.add_argument('--tol', type=float, default=None)
    ap.add_argument('--maxit', type=int, default=3000)
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--solver', choices=['bicgstab', 'bicgstab_mpir'], default='bicgstab',
                    help='bicgstab_mpir: FP64 iterative refinement around BiCGStab (long '
                         'solves, FP32 storage at tight tolerances)')
    ap.add_argument('--cpu-steps', type=int, default=10)
    ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                    help='only run 5 device copies of MIB MiB (byte-counter reference) and exit')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-profile', action='store_true')
    ap.add_argument('--profile-period', type=int, default=32,
                    help='time every n-th launch of each Krylov kernel with dispatch events')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: this build runs one periodic box per GPU and has no '
              'multi-GPU path (requested --gpus %d)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    import torch
    dtype = torch.float64 if a.dtype == 'f64' else torch.float32
    w = 8 if a.dtype == 'f64' else 4
    tol = a.tol if a.tol is not None else (1e-10 if a.dtype == 'f64' else 1e-8)
    r = gpu_run(a.N, dtype, a.nu, tol, a.maxit, a.steps, a.warmup, a.seed,
                0 if a.no_profile else a.profile_period, a.solver)
    ncell = a.N ** 3
    K = a.steps
    value = ncell * K / r['wall'] / 1e6
    ms = r['wall'] / K * 1e3
    kbar = sum(r['its']) / len(r['its'])
    alg_bytes = sum(ncell * w * (65 + 15 * k) for k in r['its'])
    out = {
        'metric': 'Mcell-steps/s (RK4 + BiCGStab time step, %d^3 %s)' % (a.N, a.dtype.upper()),
        'value': round(value, 3), 'unit': 'Mcell-steps/s', 'n_gpus': 1, 'steps': K,
        'warmup': a.warmup, 'ms_per_step': round(ms, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': a.dtype, 'data': 'synthetic',
        'config': {'workload': 'decaying isotropic turbulence, periodic %d^3 staggered grid, '
                               'Passot-Pouquet k0=4 u_rms=1, nu=%g, RK4, %s tol %g'
                               % (a.N, a.nu, 'BiCGStab' if a.solver == 'bicgstab' else
                                  'BiCGStab in FP64 iterative refinement', tol),
                   'grid': a.N, 'nu': a.nu, 'tol': tol, 'dt': r['dt'],
                   'bicgstab_iters_mean': round(kbar, 2),
                   'bicgstab_last_residual': float('%.3e' % r['res']),
                   'bicgstab_status': r['status'], 'parallelism': 'single GPU'},
        'hbm_gbs_algorithmic': round(alg_bytes / r['wall'] / 1e9, 2),
        'step_roofline_frac': round(alg_bytes / r['wall'] / 1e9 / HBM_PEAK_GBS, 4),
    }
    if r['prof']:
        per = {}
        for name, wmul in (('KA', 6), ('KB', 2), ('KC', 7)):
            tot, n = r['prof'][name]
            if n:
                avg = tot / n / 1e3
                gbs = wmul * w * ncell / avg / 1e9
                per[name] = {'avg_us': round(avg * 1e6, 3), 'launches': n,
                             'bytes_per_launch': wmul * w * ncell, 'GB/s': round(gbs, 1)}
        out['kernels'] = per
        if 'KA' in per:
            out['roofline'] = {'kernel': 'KA: p update, v = A p, <v, r_hat> (bicg_tile2_kernel / bicg_tile_kernel)', 'bound': 'hbm',
                               'achieved': per['KA']['GB/s'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                               'frac': round(per['KA']['GB/s'] / HBM_PEAK_GBS, 4),
                               'traffic': KA_TRAFFIC_BYTES.get((a.N, a.dtype)),
                               'bytes_per_launch': per['KA']['bytes_per_launch']}
    if not a.no_cpu:
        u0, v0, w0 = r['u0']
        otol = 1e-10   # the oracle is the reference's FP64 path at the FP64 tolerance
        cb = cpu_baseline(u0, v0, w0, a.N, a.nu, r['dt'], otol, a.maxit, a.cpu_steps)
        out['cpu_baseline'] = {'value': round(cb['value'], 5), 'unit': 'Mcell-steps/s',
                               'cores': cb['cores'], 'kind': 'port',
                               'sample': '%d time step(s) of the same %d^3 workload in FP64 from '
                                         'the same initial field (CPU oracle, torch-CPU FP64, %d '
                                         'threads; %d BiCGStab iterations in its last step, '
                                         '%.1f s)' % (a.cpu_steps, a.N, cb['cores'], cb['iters'],
                                                      cb['wall'])}
        # field error of the GPU path against the oracle after the same steps
        gf = gpu_fields(a.N, dtype, a.nu, tol, a.maxit, a.cpu_steps, a.seed, r['dt'], a.solver)
        errs = [float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
                for x, y in zip(gf, cb['fields'])]
        out['oracle_error'] = {'steps': a.cpu_steps, 'rel_l2_uvw': [float('%.3e' % e) for e in errs],
                               'gpu_tol': tol, 'oracle_tol': otol}
    print(json.dumps(out))


if __name__ == '__main__':
    main()
