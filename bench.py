"""Throughput of PyFlow's time step on one MI355X.

Workload (default): decaying isotropic turbulence, triply periodic box,
L = 2*pi, 128^3 staggered grid, FP64, seeded Passot-Pouquet field (k0 = 4,
u_rms = 1), nu = 0.005, rho = 1, fixed dt = 0.3*h/max|u0|, RK4 + BiCGStab to
a relative L-inf residual of 1e-10 -- the reference's forward loop
(Domain.forwardStep + the per-step monitor values) through this package's
PyFlow.setup.  A step is one full time step of that loop.

Prints one JSON line.  `value` is Mcell-steps/s of the timed FP64 steps.
The Poisson solve of a step runs as one persistent cooperative kernel that
keeps the Krylov vectors on chip (KP, bicg_persist_kernel) when the box
fits; the roofline object is then for KP: the algorithmic bytes of the
iterations it ran over its duration (start/stop hipEvents around each
launch), which can exceed the HBM peak because the vectors never travel,
with its HBM traffic beside it (`roofline.traffic`: bytes per launch from two
rocprofv3 --pmc passes, FETCH_SIZE then WRITE_SIZE, over a short child run
before the timed legs, each counter scaled by its reading for a plain
device copy in the same pass; --no-traffic skips them).  `launch_path` is the same
workload with the launch-per-reduction kernels (KA, KB, KC once per
iteration each; their average durations come from start/stop hipEvents
recorded by the dispatch of every 32nd launch of each kernel during the
timed steps, a sample that keeps the events' own cost out of its value),
and its roofline object is for the one of them with the largest total.  `cpu_baseline` times the CPU oracle (a torch-CPU FP64
restatement of the reference's op sequence) on a bounded sample of the same
workload on this host, and `oracle_error` is the relative L2 error of the GPU
velocity against the oracle's after those steps.  The other legs of the
metric ride along: `f32` (the same workload with FP32 storage and FP64
Krylov reductions at tolerance 1e-8, its error against the FP64 oracle) and
`serving` (a bounded Poisson stream of mixed-size requests served with
continuous batching and one after another: throughput, p50/p99 step latency,
requests with bit-identical results).  `ml_closure` times the reference's
'source' ML closure at 64^3 FP32 (the network forward, its reverse pass, the
closure work of one adjoint RK4 step) against the FP32 MFMA peak.
"""

import argparse
import gc
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
INFINITY_CACHE_BYTES = 256 << 20   # MI355X last-level (MALL) cache

KERNEL_NAMES = {'KA': 'KA: p update, v = A p, <v, r_hat> (bicg_tile2_kernel mode 0)',
                'KB': 'KB: s = r - alpha v, t = A s, <s,t>, <t,t> (bicg_tile2_kernel mode 1)',
                'KC': 'KC: s, t = A s again, x and r updates, <r,r_hat>, max|r| '
                      '(bicg_tile2_kernel mode 2)'}
KERNEL_NAMES['KP'] = ('KP: the whole BiCGStab solve as one persistent cooperative kernel, '
                      'vectors on chip (bicg_persist_kernel)')
KERNEL_WORDS = {'KA': 6, 'KB': 2, 'KC': 7}   # algorithmic vector passes per launch
KERNEL_MATCH = {'KP': 'bicg_persist_kernel<%s,', 'KA': 'bicg_tile2_kernel<%s, 0,',
                'KB': 'bicg_tile2_kernel<%s, 1,', 'KC': 'bicg_tile2_kernel<%s, 2,'}
COPY_MIB = 512      # plain device copy that calibrates the byte counters


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


def gpu_run(N, dtype, nu, tol, maxit, steps, warmup, seed, profile=True, solver='bicgstab',
            persistent=True):
    import torch
    from pyflow_amd import PyFlow
    from pyflow_amd.initial_conditions import passot_pouquet, cfl_dt
    u0, v0, w0 = passot_pouquet(N, k0=4.0, urms=1.0, seed=seed)
    dt = cfl_dt(u0, v0, w0, 2 * math.pi / N)
    cfg = make_config(N, dtype, nu, tol, maxit, seed, solver)
    cfg.simDt = dt
    comms, decomp, geometry, metric, D, t0 = PyFlow.setup(cfg)
    box = decomp.box
    box.set_persistent(persistent)
    for _ in range(warmup):
        D.forwardStep(dt)
        D.stats()
    if profile:
        box.profile_enable(profile)
    # collect the previous leg's garbage here, not inside the timed steps (an
    # FP32 leg after the FP64 one otherwise ran 232-245 instead of 316
    # Mcell-steps/s)
    gc.collect()
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
                res=D.max_resP, status={k: status.count(k) for k in set(status)},
                persistent=box.last_solve_persistent)


def measure(N, dt_name, nu, tol, maxit, steps, warmup, seed, period, solver, persistent=True):
    """Timed steps of one workload: throughput, per-kernel bandwidth and the
    roofline of the Krylov kernel with the largest measured time."""
    import torch
    dtype = torch.float64 if dt_name == 'f64' else torch.float32
    w = 8 if dt_name == 'f64' else 4
    r = gpu_run(N, dtype, nu, tol, maxit, steps, warmup, seed, period, solver, persistent)
    ncell = N ** 3
    value = ncell * steps / r['wall'] / 1e6
    kbar = sum(r['its']) / len(r['its'])
    kt = sum(r['its'])
    # bytes per step: 65 vector passes outside BiCGStab plus 15 per iteration
    # (what the kernels must move; KC recomputes s and t) -- and the 17 per
    # iteration of a minimally fused step that stores s and t
    alg = sum(ncell * w * (65 + 15 * k) for k in r['its'])
    alg17 = sum(ncell * w * (65 + 17 * k) for k in r['its'])
    out = {'value': round(value, 3), 'ms_per_step': round(r['wall'] / steps * 1e3, 4),
           'bicgstab_iters_mean': round(kbar, 2),
           'hbm_gbs_algorithmic': round(alg / r['wall'] / 1e9, 2),
           'step_roofline_frac': round(alg / r['wall'] / 1e9 / HBM_PEAK_GBS, 4),
           'step_roofline_frac_17w': round(alg17 / r['wall'] / 1e9 / HBM_PEAK_GBS, 4),
           '_run': r}
    krylov_set = 9 * ncell * w
    out['infinity_cache'] = {'krylov_set_bytes': krylov_set,
                             'cache_bytes': INFINITY_CACHE_BYTES,
                             'set_fits': krylov_set <= INFINITY_CACHE_BYTES}
    out['krylov_path'] = 'persistent' if r['persistent'] else 'launch per reduction'
    if r['prof'] and r['prof']['KP'][1]:
        # the whole solve is one cooperative launch (bicg_persist_kernel): its
        # algorithmic bytes are those of the iterations it ran, 15 vector
        # passes each (what KA, KB and KC move), although the vectors stay on
        # chip and HBM sees only the tile faces
        ms, n, itp = r['prof']['KP']
        avg = ms / n / 1e3
        nbytes = 15 * w * ncell * itp / n
        gbs = nbytes / avg / 1e9
        out['kernels'] = {'KP': {'avg_us': round(avg * 1e6, 1), 'timed_launches': n,
                                 'iterations_per_launch': round(itp / n, 2),
                                 'bytes_per_launch': int(nbytes), 'GB/s': round(gbs, 1),
                                 'us_per_iteration': round(avg * 1e6 * n / itp, 3)}}
        out['roofline'] = {'kernel': KERNEL_NAMES['KP'], 'bound': 'hbm', 'achieved': round(gbs, 1),
                           'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                           'frac': round(gbs / HBM_PEAK_GBS, 4),
                           'traffic': None,   # measured by the --pmc passes at the end
                           'bytes_per_launch': int(nbytes),
                           'chosen_by': 'the solve is one launch: the largest kernel by time'}
        out['_top'] = 'KP'
    elif r['prof']:
        per = {}
        for name in ('KA', 'KB', 'KC'):
            tot, n = r['prof'][name]
            if n:
                avg = tot / n / 1e3
                nbytes = KERNEL_WORDS[name] * w * ncell
                per[name] = {'avg_us': round(avg * 1e6, 3), 'timed_launches': n,
                             'launches': kt, 'est_total_ms': round(avg * kt * 1e3, 3),
                             'bytes_per_launch': nbytes, 'GB/s': round(nbytes / avg / 1e9, 1)}
        out['kernels'] = per
        if per:
            # every kernel runs once per iteration: the largest total is the
            # largest average launch
            top = max(per, key=lambda k: per[k]['est_total_ms'])
            p = per[top]
            out['roofline'] = {'kernel': KERNEL_NAMES[top], 'bound': 'hbm',
                               'achieved': p['GB/s'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                               'frac': round(p['GB/s'] / HBM_PEAK_GBS, 4),
                               'traffic': None,
                               'bytes_per_launch': p['bytes_per_launch'],
                               'chosen_by': 'largest measured total time of KA, KB, KC'}
            out['_top'] = top
    return out


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


def host_cpu():
    """(model name, visible cores, threads to use): the threads are the
    process's CPU share -- OMP_NUM_THREADS when the host sets it, else every
    core the process may run on."""
    model = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0))
    share = os.environ.get('OMP_NUM_THREADS')
    threads = min(visible, int(share)) if share and share.isdigit() and int(share) > 0 \
        else visible
    return model, visible, max(1, threads)


def cpu_baseline(u0, v0, w0, N, nu, dt, tol, maxit, steps=1):
    """The CPU oracle (torch-CPU FP64) on `steps` steps of the same workload;
    returns its throughput and final interior velocity."""
    import torch
    from oracle import pyflow_oracle as O
    model, visible, cores = host_cpu()
    torch.set_num_threads(cores)
    g = O.Grid(N, N, N)
    B = O.Box(g, u0, v0, w0, mu=nu, rho=1.0, dt=dt, tol=tol, maxit=maxit)
    t = time.perf_counter()
    for _ in range(steps):
        B.forward_step()
        B.stats()
    wall = time.perf_counter() - t
    return dict(value=N ** 3 * steps / wall / 1e6, wall=wall, cores=cores, model=model,
                visible=visible, iters=B.poisson.iterations,
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


def traffic_child(a):
    """The program a --pmc pass profiles: a plain device copy of COPY_MIB MiB
    (5 dispatches: the counters' reference) and two steps of each leg named in
    --traffic-child ("f64:p" = FP64 with the persistent solve, "f64:l" = with
    the launch-per-reduction kernels, "f32:p"); prints the legs' BiCGStab
    iteration counts, one per dispatch of the persistent kernel."""
    import torch
    n = COPY_MIB << 20
    dst = torch.empty(n, dtype=torch.uint8, device='cuda')
    src = torch.ones(n, dtype=torch.uint8, device='cuda')
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    del dst, src
    legs = []
    for spec in a.traffic_child.split(','):
        dt_name, path = spec.split(':')
        dtype = torch.float64 if dt_name == 'f64' else torch.float32
        tol = 1e-10 if dt_name == 'f64' else 1e-8
        t0 = time.perf_counter()
        r = gpu_run(a.N, dtype, a.nu, tol, a.maxit, 2, 0, a.seed, 0, a.solver, path == 'p')
        legs.append({'leg': spec, 'its': r['its'], 'persistent': r['persistent']})
        print('traffic child: %s, 2 steps, %s iterations, %.1f s'
              % (spec, r['its'], time.perf_counter() - t0), flush=True)
    print('TRAFFIC_CHILD ' + json.dumps({'copy_bytes': n, 'legs': legs}), flush=True)


def pmc_pass(a, counter, specs, timeout):
    """One rocprofv3 --pmc pass of traffic_child: (its report, rows of
    (kernel name, counter value in KiB) in dispatch order)."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile
    import threading
    exe = shutil.which('rocprofv3')
    if exe is None:
        raise RuntimeError('rocprofv3 not found')
    d = tempfile.mkdtemp(prefix='pf_pmc_')
    try:
        cmd = [exe, '--pmc', counter, '--output-format', 'csv', '-d', d, '-o', 'run', '--',
               sys.executable, os.path.abspath(__file__), '--traffic-child', ','.join(specs),
               '--N', str(a.N), '--nu', repr(a.nu), '--maxit', str(a.maxit),
               '--seed', str(a.seed), '--solver', a.solver]
        proc = subprocess.Popen(cmd, cwd=d, env=dict(os.environ, TMPDIR=d),
                                stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                start_new_session=True)
        lines = []

        def pump():   # the child's progress lines go on to stderr as they come
            for raw in proc.stdout:
                ln = raw.decode(errors='replace').rstrip('\n')
                lines.append(ln)
                if ln.startswith('traffic child'):
                    print('%s pass: %s' % (counter, ln), file=sys.stderr, flush=True)

        th = threading.Thread(target=pump, daemon=True)
        th.start()
        try:
            proc.wait(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            raise RuntimeError('%s pass did not finish in %d s' % (counter, timeout))
        th.join(10)
        text = '\n'.join(lines)
        rep = [ln for ln in lines if ln.startswith('TRAFFIC_CHILD ')]
        path = os.path.join(d, 'run_counter_collection.csv')
        # the profiler may end with a signal after the program has finished
        # and its files are written: the report line and the file decide
        if not rep or not os.path.exists(path):
            raise RuntimeError('%s pass failed (exit %s): %s'
                               % (counter, proc.returncode, text[-300:]))
        rows = []
        with open(path) as f:
            for row in csv.DictReader(f):
                if row['Counter_Name'] == counter:
                    rows.append((row['Kernel_Name'], float(row['Counter_Value'])))
        return json.loads(rep[-1][len('TRAFFIC_CHILD '):]), rows
    finally:
        shutil.rmtree(d, ignore_errors=True)


def measure_traffic(a, specs, timeout=300):
    """HBM-side bytes of the Krylov kernels, measured in this run: two
    rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) over traffic_child,
    each counter scaled by what it reads for the plain copy of COPY_MIB MiB in
    the same pass (bytes = FETCH_KiB / f_fetch + WRITE_KiB / f_write, times
    1024).  Returns {(dtype, kernel): bytes per launch} for KA, KB, KC, and
    bytes per BiCGStab iteration for KP, plus the scale factors."""
    got = {}
    factors = {}
    for counter in ('FETCH_SIZE', 'WRITE_SIZE'):
        t0 = time.perf_counter()
        rep, rows = pmc_pass(a, counter, specs, timeout)
        n = rep['copy_bytes']
        copies = [v for k, v in rows if 'copyBuffer' in k and v * 1024 > 0.1 * n]
        if not copies:
            raise RuntimeError('%s pass: no device copy of %d bytes found' % (counter, n))
        copies.sort()
        f = copies[len(copies) // 2] * 1024 / n
        factors[counter] = round(f, 5)
        for leg in rep['legs']:
            dt_name, path = leg['leg'].split(':')
            tname = 'double' if dt_name == 'f64' else 'float'
            if path == 'p':
                if not leg['persistent']:
                    continue
                vals = [v for k, v in rows if KERNEL_MATCH['KP'] % tname in k]
                if len(vals) != len(leg['its']):
                    raise RuntimeError('%s pass: %d persistent dispatches for %d solves'
                                       % (counter, len(vals), len(leg['its'])))
                got.setdefault((dt_name, 'KP'), 0.0)
                got[(dt_name, 'KP')] += sum(vals) * 1024 / f / sum(leg['its'])
            else:
                for kn in ('KA', 'KB', 'KC'):
                    vals = [v for k, v in rows if KERNEL_MATCH[kn] % tname in k]
                    if vals:
                        got.setdefault((dt_name, kn), 0.0)
                        got[(dt_name, kn)] += sum(vals) / len(vals) * 1024 / f
        print('traffic: %s pass %.1f s' % (counter, time.perf_counter() - t0), file=sys.stderr,
              flush=True)
    return got, factors


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
    ap.add_argument('--no-f32', action='store_true',
                    help='skip the FP32-storage leg of the default FP64 run')
    ap.add_argument('--serving-requests', type=int, default=24,
                    help='requests of the serving leg (0 = skip it)')
    ap.add_argument('--no-ml-closure', action='store_true',
                    help="skip the leg that times the reference's ML closure at 64^3")
    ap.add_argument('--no-profile', action='store_true')
    ap.add_argument('--krylov', choices=['persistent', 'launches'], default='persistent',
                    help='Poisson solves of the main leg: one persistent cooperative kernel '
                         'when the box fits (default), or the launch-per-reduction kernels')
    ap.add_argument('--no-launch-path', action='store_true',
                    help='skip the leg that runs the same workload with the launch-per-'
                         'reduction Krylov kernels instead of the persistent solve')
    ap.add_argument('--profile-period', type=int, default=32,
                    help='time every n-th launch of each Krylov kernel with dispatch events')
    ap.add_argument('--no-traffic', action='store_true',
                    help='skip the two rocprofv3 --pmc passes that measure the Krylov kernels\' '
                         'HBM traffic (roofline.traffic is then null)')
    ap.add_argument('--traffic-child', default='', help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: this build runs one periodic box per GPU and has no '
              'multi-GPU path (requested --gpus %d)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.traffic_child:
        traffic_child(a)
        return
    # the --pmc passes run first, before this process opens the GPU: the
    # profiled child then has the device to itself
    traffic = None
    if not a.no_traffic and not a.no_profile:
        specs = ['%s:%s' % (a.dtype, 'p' if a.krylov == 'persistent' else 'l')]
        if not a.no_launch_path:
            specs.append('%s:l' % a.dtype)
        if a.dtype == 'f64' and not a.no_f32:
            specs.append('f32:%s' % ('p' if a.krylov == 'persistent' else 'l'))
        specs = sorted(set(specs))
        try:
            traffic = measure_traffic(a, specs) + (specs,)
        except Exception as e:   # the line still prints, with the reason
            traffic = str(e)[:400]
    import torch
    dtype = torch.float64 if a.dtype == 'f64' else torch.float32
    tol = a.tol if a.tol is not None else (1e-10 if a.dtype == 'f64' else 1e-8)
    metric = 'Mcell-steps/s + HBM GB/s at 128^3 FP64/FP32; served p50/p99 step latency'
    leg = measure(a.N, a.dtype, a.nu, tol, a.maxit, a.steps, a.warmup, a.seed,
                  0 if a.no_profile else a.profile_period, a.solver,
                  persistent=a.krylov == 'persistent')
    r = leg.pop('_run')
    legs = {'main': (a.dtype, leg.pop('_top', None), leg, leg.get('krylov_path') == 'persistent')}
    out = {
        'metric': metric, 'value': leg['value'], 'unit': 'Mcell-steps/s', 'n_gpus': 1,
        'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': leg['ms_per_step'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': a.dtype,
        'data': 'synthetic (seeded Passot-Pouquet field)',
        'config': {'workload': 'decaying isotropic turbulence, periodic %d^3 staggered grid, '
                               'Passot-Pouquet k0=4 u_rms=1, nu=%g, RK4, %s tol %g, %s'
                               % (a.N, a.nu, 'BiCGStab' if a.solver == 'bicgstab' else
                                  'BiCGStab in FP64 iterative refinement', tol, a.dtype.upper()),
                   'grid': a.N, 'nu': a.nu, 'tol': tol, 'dt': r['dt'],
                   'bicgstab_iters_mean': leg['bicgstab_iters_mean'],
                   'bicgstab_last_residual': float('%.3e' % r['res']),
                   'bicgstab_status': r['status'], 'parallelism': 'single GPU'},
    }
    for k in ('hbm_gbs_algorithmic', 'step_roofline_frac', 'step_roofline_frac_17w',
              'kernels', 'roofline', 'infinity_cache'):
        if k in leg:
            out[k] = leg[k]
    if 'krylov_path' in leg:
        out['config']['krylov_path'] = leg['krylov_path']
    if not a.no_launch_path:
        # the same workload with the launch-per-reduction kernels (KA, KB, KC)
        lp = measure(a.N, a.dtype, a.nu, tol, a.maxit, a.steps, a.warmup, a.seed,
                     0 if a.no_profile else a.profile_period, a.solver, persistent=False)
        lp.pop('_run')
        top = lp.pop('_top', None)
        out['launch_path'] = {k: lp[k] for k in ('value', 'ms_per_step', 'bicgstab_iters_mean',
                                                 'step_roofline_frac', 'kernels', 'roofline')
                              if k in lp}
        legs['launch_path'] = (a.dtype, top, out['launch_path'], False)
    cb = None
    if not a.no_cpu:
        u0, v0, w0 = r['u0']
        otol = 1e-10   # the oracle is the reference's FP64 path at the FP64 tolerance
        cb = cpu_baseline(u0, v0, w0, a.N, a.nu, r['dt'], otol, a.maxit, a.cpu_steps)
        out['cpu_baseline'] = {
            'value': round(cb['value'], 5), 'unit': 'Mcell-steps/s', 'cores': cb['cores'],
            'kind': 'port',
            'sample': '%d time step(s) of the same %d^3 workload in FP64 from the same initial '
                      'field: the CPU oracle (torch-CPU FP64 restatement of the reference\'s op '
                      'sequence) on %d threads (the process\'s CPU share; %d cores visible, %s); '
                      '%d BiCGStab iterations in its last step, %.1f s. The oracle multiplies by '
                      'the nine Laplace constants where the reference reads its 72 B/cell '
                      'coefficient tensor, so it is faster than the reference\'s own CPU path: '
                      'a conservative baseline'
                      % (a.cpu_steps, a.N, cb['cores'], cb['visible'], cb['model'], cb['iters'],
                         cb['wall'])}
        gf = gpu_fields(a.N, dtype, a.nu, tol, a.maxit, a.cpu_steps, a.seed, r['dt'], a.solver)
        errs = [float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
                for x, y in zip(gf, cb['fields'])]
        out['oracle_error'] = {'steps': a.cpu_steps, 'rel_l2_uvw': [float('%.3e' % e) for e in errs],
                               'gpu_tol': tol, 'oracle_tol': otol}
    if a.dtype == 'f64' and not a.no_f32:
        f32 = measure(a.N, 'f32', a.nu, 1e-8, a.maxit, a.steps, a.warmup, a.seed,
                      0 if a.no_profile else a.profile_period, a.solver,
                      persistent=a.krylov == 'persistent')
        f32.pop('_run')
        legs['f32'] = ('f32', f32.pop('_top', None), f32, f32.get('krylov_path') == 'persistent')
        f32['tol'] = 1e-8
        if cb is not None:
            gf = gpu_fields(a.N, torch.float32, a.nu, 1e-8, a.maxit, a.cpu_steps, a.seed,
                            r['dt'], a.solver)
            errs = [float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
                    for x, y in zip(gf, cb['fields'])]
            f32['oracle_error'] = {'steps': a.cpu_steps, 'vs': 'FP64 oracle, tol 1e-10',
                                   'rel_l2_uvw': [float('%.3e' % e) for e in errs]}
        out['f32'] = f32
    if a.serving_requests > 0:
        from pyflow_amd import serve_bench
        sv = serve_bench.run(requests=a.serving_requests)
        b, q = sv['batched'], sv['sequential']
        keys = ('mcell_steps_per_s', 'step_latency_p50_ms', 'step_latency_p99_ms', 'wall_s')
        out['serving'] = {
            'workload': sv['workload'],
            'batched': {k: round(b[k], 3) for k in keys},
            'sequential': {k: round(q[k], 3) for k in keys},
            'throughput_ratio': round(sv['throughput_ratio'], 3),
            'bit_identical_requests': sv['bit_identical_requests'],
            'max_resident_gb': round(b['max_resident_gb'], 2),
            'mean_batch': round(b['mean_batch'], 2)}
    if not a.no_ml_closure:
        from pyflow_amd import sfsmodel_ML
        t = sfsmodel_ML.timings(64, torch.float32, 5)
        f = t['flop_forward']
        out['ml_closure'] = {
            'workload': "the reference's 'source' ML closure (NeuralNetworkModel2, 0.495 MFLOP "
                        "per padded cell) on a 64^3 FP32 box (68^3 padded cells), random fields",
            'forward_ms': round(t['forward'], 3),
            'forward_tflops': round(f / t['forward'] / 1e9, 1),
            'reverse_ms': round(t['reverse'], 3),
            'reverse_tflops': round(3 * f / t['reverse'] / 1e9, 1),
            'adjoint_rk4_step_closure_ms': round(t['adjoint_step'], 3),
            'four_separate_terms_ms': round(t['four_terms'], 3),
            'peak_tflops_fp32_mfma': 157.3,
            'forward_frac': round(f / t['forward'] / 1e9 / 157.3, 4)}
    if isinstance(traffic, str):
        out['traffic_measurement'] = {'error': traffic}
    elif traffic is not None:
        got, factors, specs = traffic
        for dt, top, d, pers in legs.values():
            if top is None or (dt, top) not in got:
                continue
            per = got[(dt, top)]
            if top == 'KP':
                d['kernels']['KP']['hbm_bytes_per_iteration'] = int(per)
                per *= d['kernels']['KP']['iterations_per_launch']
            d['roofline']['traffic'] = int(per)
            d['roofline']['traffic_vs_algorithmic'] = round(
                per / d['roofline']['bytes_per_launch'], 4)
            for kn, kd in d['kernels'].items():
                if kn != 'KP' and (dt, kn) in got:
                    kd['hbm_bytes_measured'] = int(got[(dt, kn)])
        out['traffic_measurement'] = {
            'method': 'rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE (one pass each) over 2 steps '
                      'of each leg in a child run before the timed legs, KiB per dispatch '
                      'scaled by the same counter over a plain %d MiB device copy in that pass'
                      % COPY_MIB,
            'copy_scale': factors, 'legs': specs}
    print(json.dumps(out))


if __name__ == '__main__':
    main()
