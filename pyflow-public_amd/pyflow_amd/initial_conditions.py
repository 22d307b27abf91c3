"""Initial conditions (data/initial_conditions.py:39-143) plus the seeded
Passot-Pouquet decaying-turbulence field the benchmark configurations use.

The reference has no spectral initial condition; this one is defined here:

* host numpy, ``Generator(PCG64(seed))``;
* complex Gaussian modes with |u_hat(k)| proportional to sqrt(E(k)/(4 pi k^2)),
  E(k) = k^4 exp(-2 (k/k0)^2) (Passot & Pouquet 1987);
* each component lives on its staggered face grid (u on x-faces, v on y-faces,
  w on z-faces) and is made discretely solenoidal by projecting every mode
  onto the null space of the staggered divergence symbol
  sigma_d(k) = (exp(i k_d h_d) - 1)/h_d, so div_h u = 0 to rounding;
* scaled so that u_rms = sqrt(<u^2+v^2+w^2>/3) equals the requested value.
"""

import math

import numpy as np


def passot_pouquet(nx, ny=None, nz=None, Lx=2 * math.pi, Ly=None, Lz=None, k0=4.0, urms=1.0,
                   seed=0):
    """Return (u, v, w) float64 arrays of shape (nx, ny, nz) on the staggered faces."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    Ly = Lx if Ly is None else Ly
    Lz = Lx if Lz is None else Lz
    hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
    rng = np.random.Generator(np.random.PCG64(seed))
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=hx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=hy)
    kz = 2 * np.pi * np.fft.rfftfreq(nz, d=hz)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing='ij')
    kmag = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    with np.errstate(divide='ignore', invalid='ignore'):
        amp = np.where(kmag > 0, np.sqrt(kmag ** 4 * np.exp(-2.0 * (kmag / k0) ** 2)
                                         / (4 * np.pi * kmag ** 2)), 0.0)
    shape = KX.shape
    uh = []
    for _ in range(3):
        re = rng.standard_normal(shape)
        im = rng.standard_normal(shape)
        uh.append(amp * (re + 1j * im))
    sig = [(np.exp(1j * KX * hx) - 1.0) / hx,
           (np.exp(1j * KY * hy) - 1.0) / hy,
           (np.exp(1j * KZ * hz) - 1.0) / hz]
    div = sig[0] * uh[0] + sig[1] * uh[1] + sig[2] * uh[2]
    norm = np.abs(sig[0]) ** 2 + np.abs(sig[1]) ** 2 + np.abs(sig[2]) ** 2
    with np.errstate(divide='ignore', invalid='ignore'):
        fac = np.where(norm > 0, div / norm, 0.0)
    for d in range(3):
        uh[d] = uh[d] - np.conj(sig[d]) * fac
        uh[d][norm == 0] = 0.0
    fields = [np.fft.irfftn(uh[d], s=(nx, ny, nz), axes=(0, 1, 2)) for d in range(3)]
    cur = math.sqrt(sum(float(np.mean(f * f)) for f in fields) / 3.0)
    scale = urms / cur if cur > 0 else 0.0
    return tuple(np.ascontiguousarray(f * scale, dtype=np.float64) for f in fields)


def cfl_dt(u, v, w, h, cfl=0.3):
    """Fixed time step 0.3*h/max|u| from the initial field."""
    umax = max(float(np.abs(u).max()), float(np.abs(v).max()), float(np.abs(w).max()))
    return cfl * h / umax


def generate(IC, config, decomp):
    """data/initial_conditions.py:39-143.  Supported configNames:
    'passotPouquet' (this build's spectral field; IC.seed, IC.k0, IC.urms),
    'uniform' and 'periodicGaussian' (the reference's analytic fields), and
    'restart' (NGA restart files)."""
    names = ['U', 'V', 'W', 'P']
    startTime = 0.0
    nx_, ny_, nz_ = decomp.nx_, decomp.ny_, decomp.nz_
    name = IC.configName
    if name == 'restart':
        from . import dataReader as dr
        names, startTime = dr.readNGArestart(IC.dataFileStr)
        data_IC = np.zeros((nx_, ny_, nz_, 4), dtype=decomp.dtypeNumpy)
        return names, startTime, data_IC, config.xGrid, config.yGrid, config.zGrid
    xGrid, yGrid, zGrid = config.xGrid, config.yGrid, config.zGrid
    data_IC = np.zeros((nx_, ny_, nz_, 4), dtype=np.float64)
    if name == 'passotPouquet':
        u, v, w = passot_pouquet(nx_, ny_, nz_, config.Lx, config.Ly, config.Lz,
                                 k0=getattr(IC, 'k0', 4.0), urms=getattr(IC, 'urms', 1.0),
                                 seed=getattr(IC, 'seed', 0))
        data_IC[..., 0], data_IC[..., 1], data_IC[..., 2] = u, v, w
    elif name == 'uniform':                                  # initial_conditions.py:100-113
        data_IC[..., 0] = 2.0
        data_IC[..., 1] = 2.0
        data_IC[..., 2] = 2.0
    elif name == 'periodicGaussian':                         # initial_conditions.py:80-98
        xc = xGrid - 0.5 * (xGrid[0] + xGrid[-1])
        yc = yGrid - 0.5 * (yGrid[0] + yGrid[-1])
        zc = zGrid - 0.5 * (zGrid[0] + zGrid[-1])
        stdDev = 0.1
        bump = (np.exp(-0.5 * (xc[:nx_, None, None] / stdDev) ** 2) *
                np.exp(-0.5 * (yc[None, :ny_, None] / stdDev) ** 2) *
                np.exp(-0.5 * (zc[None, None, :nz_] / stdDev) ** 2))
        data_IC[..., 0] = 2.0 * bump
        data_IC[..., 1] = 2.0 * bump
        data_IC[..., 2] = 2.0 * bump
    else:
        raise NotImplementedError("configName '%s' is not supported" % name)
    return names, startTime, data_IC, xGrid, yGrid, zGrid
