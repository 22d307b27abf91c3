"""Decay curves of the periodic turbulence box: kinetic energy and
dissipation per step.

    python -m pyflow_amd.decay --N 128 --steps 200 --dtype f64 --out decay_f64.csv
    python -m pyflow_amd.decay --N 512 --steps 100 --dtype f32 --nu 0.001 \
        --solver bicgstab_mpir --tol 1e-8 --progress --out decay_512_f32.csv

Runs the forward loop of PyFlow.run (core/PyFlow.py:352-425) on a
Passot-Pouquet field (k0 = 4, u_rms = 1, L = 2 pi, fixed dt at CFL 0.3) and
writes per step: time, TKE = 0.5 <u.u> (the reference's monitor value,
PyFlow.py:388-390), dissipation eps = 2 nu <S_ij S_ij> with S the symmetric
part of the cell-centred velocity gradients of metric_uniform.grad_vel_center
(metric/metric_staggered.py:215-268), max|div u| and the BiCGStab
iterations.  The gradients come from the HIP metric kernels; the products and
means are a diagnostic computed with torch on the GPU.  Each row also carries
the solver's residual and stop reason and the step's wall time (the time
step alone, synchronized; the diagnostics are outside it).
"""

import argparse
import math
import sys
import time
import types

import numpy as np
import torch


def dissipation(metric, D, nu):
    """2 nu <S_ij S_ij> over the interior cells."""
    G = []
    for st, comp in ((D.state_u_P, 'u'), (D.state_v_P, 'v'), (D.state_w_P, 'w')):
        metric.grad_vel_center(st, comp)
        G.append([g[2:-2, 2:-2, 2:-2].double() for g in (st.grad_x, st.grad_y, st.grad_z)])
    s2 = torch.zeros_like(G[0][0])
    for i in range(3):
        for j in range(3):
            S = 0.5 * (G[i][j] + G[j][i])
            s2 = s2 + S * S
    return 2.0 * nu * float(s2.mean())


def run(N=128, steps=200, dtype='f64', nu=0.005, seed=2024, tol=None, maxit=3000, every=1,
        solver='bicgstab', progress=None):
    from . import PyFlow
    from .initial_conditions import passot_pouquet, cfl_dt
    tdt = torch.float64 if dtype == 'f64' else torch.float32
    tol = tol if tol is not None else (1e-10 if dtype == 'f64' else 1e-8)
    u0, v0, w0 = passot_pouquet(N, k0=4.0, urms=1.0, seed=seed)
    dt = cfl_dt(u0, v0, w0, 2 * math.pi / N)
    cfg = types.SimpleNamespace(
        configName='passotPouquet', Nx=N, Ny=N, Nz=N, Lx=2 * math.pi, seed=seed, k0=4.0, urms=1.0,
        nproc_x=1, nproc_y=1, nproc_z=1, mu=nu, rho=1.0, simDt=dt, numIt=steps,
        SFSmodel='none', advancerName='RK4', equationMode='NS', pSolverMode=solver,
        min_pressure_residual=tol, max_pressure_iterations=maxit, dtypeTorch=tdt,
        dtypeNumpy=np.float64 if tdt == torch.float64 else np.float32, writeMonitor=False)
    del u0, v0, w0
    comms, decomp, geometry, metric, D, t0 = PyFlow.setup(cfg)
    rows = []
    st = D.stats()
    rows.append((0, 0.0, st['TKE'], dissipation(metric, D, nu), st['maxDiv'], 0, 0.0, '-', 0.0))
    if progress:
        progress(rows[-1])
    for k in range(1, steps + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        D.forwardStep(dt)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        if k % every == 0 or k == steps:
            st = D.stats()
            rows.append((k, k * dt, st['TKE'], dissipation(metric, D, nu), st['maxDiv'],
                         D.poisson.iterations, D.max_resP,
                         str(D.poisson.status).replace(',', ';').replace(' ', '_'), wall))
            if progress:
                progress(rows[-1])
    return rows, dt


ROW = '%d,%.6e,%.10e,%.10e,%.3e,%d,%.3e,%s,%.4f\n'


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--N', type=int, default=128)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--dtype', choices=['f64', 'f32'], default='f64')
    ap.add_argument('--nu', type=float, default=0.005)
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--every', type=int, default=1)
    ap.add_argument('--tol', type=float, default=None)
    ap.add_argument('--maxit', type=int, default=3000)
    ap.add_argument('--solver', choices=['bicgstab', 'bicgstab_mpir'], default='bicgstab')
    ap.add_argument('--progress', action='store_true', help='print each row as it is computed')
    ap.add_argument('--out', default=None)
    a = ap.parse_args(argv)
    prog = (lambda r: (sys.stdout.write(ROW % r), sys.stdout.flush())) if a.progress else None
    rows, dt = run(a.N, a.steps, a.dtype, a.nu, a.seed, a.tol, a.maxit, a.every, a.solver, prog)
    out = open(a.out, 'w') if a.out else sys.stdout
    out.write('# N=%d dtype=%s nu=%g dt=%.6e solver=%s\n'
              'step,time,TKE,eps,maxdiv,iters,residual,status,step_s\n'
              % (a.N, a.dtype, a.nu, dt, a.solver))
    for r in rows:
        out.write(ROW % r)
    if a.out:
        out.close()
    return 0


if __name__ == '__main__':
    sys.exit(main())
