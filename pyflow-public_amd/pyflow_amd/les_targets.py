"""Training targets for the LES closure: box-filtered DNS snapshots.

    python -m pyflow_amd.les_targets --out targets/ --N 256 --ratio 4 --steps 20 --every 5

runs a DNS of decaying isotropic turbulence (Passot-Pouquet field, this
package's time step) on an N^3 grid, filters it onto the (N/ratio)^3 grid of
the LES and writes, as NGA files the run loop reads (dataReader):

* `<out>/config`: the LES grid (periodic, L = 2 pi);
* `<out>/target_<it:08d>`: the filtered velocity (U, V, W and a zero P) at
  LES iteration it = 0, every, 2*every, ... steps; the LES time step is
  ratio times the DNS one, so LES iteration it is DNS step ratio*it;
* `<out>/targets.json`: the time steps, sizes and file names.

The file for it = 0 doubles as the LES initial condition (configName
'restart', dataFileStr).  A run trains against these files with
adjointTraining, numCheckpointIt = every, targetFileBaseStr = '<out>/target_'
and startFileIt = 0.

The filter averages each staggered velocity component over the fine faces
that make up a coarse face (u on the coarse x-face at x = I*H is the mean of
the ratio^2 fine u values on that face).  This is the box filter whose
discrete divergence is the block mean of the fine divergence, so a
divergence-free DNS gives a divergence-free LES field.
"""

import argparse
import json
import math
import os
import sys
import types

import numpy as np
import torch


def box_filter(u, v, w, r):
    """Coarse faces of interior staggered fields (N, N, N) -> (N/r)^3 each."""
    n = u.shape[0]
    if n % r or u.shape != v.shape or u.shape != w.shape or len(set(u.shape)) != 1:
        raise ValueError('box_filter needs equal cubic fields with N divisible by the ratio')
    m = n // r
    uc = u[0::r].reshape(m, m, r, m, r).mean(dim=(2, 4))
    vc = v[:, 0::r].reshape(m, r, m, m, r).mean(dim=(1, 4))
    wc = w[:, :, 0::r].reshape(m, r, m, r, m).mean(dim=(1, 3))
    return uc, vc, wc


def divergence(u, v, w, h):
    """Cell-centred divergence of periodic interior staggered fields."""
    return ((torch.roll(u, -1, 0) - u) + (torch.roll(v, -1, 1) - v) + (torch.roll(w, -1, 2) - w)) / h


class _Fields:
    """The minimal host mirror the NGA writers need."""

    def __init__(self, fields, time, dt):
        self.fields = [f.detach().double().cpu().numpy() for f in fields]
        self.nx, self.ny, self.nz = self.fields[0].shape
        self.nx_, self.ny_, self.nz_ = self.nx, self.ny, self.nz
        self.names = ['U', 'V', 'W', 'P']
        self.nvar = 4
        self.ivar_read_start = 0
        self.nvar_read = 4
        self.time = time
        self.dt = dt

    def read(self, ivar):
        return self.fields[ivar]


def write_fields(fname, u, v, w, time, dt):
    from . import dataReader as dr
    data = _Fields((u, v, w, torch.zeros_like(u)), time, dt)
    dr.writeNGArestart(fname, data, True)
    dr.writeNGArestart_parallel(fname, data)


def generate(out, N=256, ratio=4, steps=20, every=5, seed=2024, k0=4.0, urms=1.0, nu=0.005,
             tol=1e-10, dtype=torch.float64, verbose=True):
    """Run the DNS and write the config and the filtered snapshots; returns
    dict(dt_les, dt_dns, config, base, files)."""
    from . import PyFlow
    from . import dataReader as dr
    from .initial_conditions import passot_pouquet, cfl_dt
    if N % ratio:
        raise ValueError('N must be divisible by ratio')
    os.makedirs(out, exist_ok=True)
    u0, v0, w0 = passot_pouquet(N, k0=k0, urms=urms, seed=seed)
    dt_dns = cfl_dt(u0, v0, w0, 2 * math.pi / N)
    del u0, v0, w0
    cfg = types.SimpleNamespace(
        configName='passotPouquet', Nx=N, Ny=N, Nz=N, Lx=2 * math.pi, seed=seed, k0=k0,
        urms=urms, nproc_x=1, nproc_y=1, nproc_z=1, mu=nu, rho=1.0, simDt=dt_dns, numIt=0,
        SFSmodel='none', advancerName='RK4', equationMode='NS', pSolverMode='bicgstab',
        min_pressure_residual=tol, max_pressure_iterations=5000, dtypeTorch=dtype,
        dtypeNumpy=np.float64 if dtype == torch.float64 else np.float32, writeMonitor=False)
    comms, decomp, geometry, metric, D, t0 = PyFlow.setup(cfg)
    m = N // ratio
    grid = np.linspace(0.0, 2 * math.pi, m + 1)
    config = os.path.join(out, 'config')
    dr.writeNGAconfig(config, grid, grid, grid, configName='filteredDNS')
    base = os.path.join(out, 'target_')
    dt_les = ratio * dt_dns
    files = []
    for it in range(0, steps + 1):
        if it > 0:
            for _ in range(ratio):
                D.forwardStep(dt_dns)
        if it % every == 0:
            U = [s.interior().double() for s in (D.state_u_P, D.state_v_P, D.state_w_P)]
            uc, vc, wc = box_filter(*U, ratio)
            fname = base + '{:08d}'.format(it)
            write_fields(fname, uc, vc, wc, it * dt_les, dt_les)
            files.append(fname)
            if verbose:
                st = D.stats()
                print('LES it %d (DNS step %d): DNS TKE %.6e, filtered TKE %.6e -> %s'
                      % (it, it * ratio, st['TKE'],
                         0.5 * float((uc ** 2 + vc ** 2 + wc ** 2).mean()), fname), flush=True)
    meta = dict(dt_les=dt_les, dt_dns=dt_dns, config=config, base=base, files=files, N=N,
                ratio=ratio, steps=steps, every=every, seed=seed, k0=k0, nu=nu)
    with open(os.path.join(out, 'targets.json'), 'w') as f:
        json.dump(meta, f, indent=1)
    return meta


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', required=True)
    ap.add_argument('--N', type=int, default=256)
    ap.add_argument('--ratio', type=int, default=4)
    ap.add_argument('--steps', type=int, default=20, help='LES steps covered')
    ap.add_argument('--every', type=int, default=5, help='LES steps between targets')
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--k0', type=float, default=4.0)
    ap.add_argument('--nu', type=float, default=0.005)
    a = ap.parse_args(argv)
    r = generate(a.out, a.N, a.ratio, a.steps, a.every, a.seed, a.k0, 1.0, a.nu)
    print('LES dt %.6e (DNS dt %.6e); config %s; %d target files'
          % (r['dt_les'], r['dt_dns'], r['config'], len(r['files'])))
    return 0


if __name__ == '__main__':
    sys.exit(main())
