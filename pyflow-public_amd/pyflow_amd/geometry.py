"""geometry/geometry.py:40-116 for a uniform periodic box."""

import math

import numpy as np


class config:
    """Grid sizes (geometry.py:40-70).

    configName 'restart' reads the NGA config file (dataReader.readNGAconfig);
    any other name takes Nx, Ny, Nz (and Lx, Ly, Lz, default 2*pi) from the
    input configuration.  Boxes are periodic in every direction.
    """

    def __init__(self, IC):
        if getattr(IC, 'configName', None) == 'restart':
            from . import dataReader as dr
            self.xGrid, self.yGrid, self.zGrid, xper, yper, zper = dr.readNGAconfig(IC.configFileStr)
            self.Nx = len(self.xGrid) - 1
            self.Ny = len(self.yGrid) - 1
            self.Nz = len(self.zGrid) - 1
            self.isper = [xper, yper, zper]
            if not all(self.isper):
                raise NotImplementedError('non-periodic NGA grids are outside this build')
        else:
            self.Nx, self.Ny, self.Nz = IC.Nx, IC.Ny, IC.Nz
            self.Lx = getattr(IC, 'Lx', 2 * math.pi)
            self.Ly = getattr(IC, 'Ly', self.Lx)
            self.Lz = getattr(IC, 'Lz', self.Lx)
            self.xGrid = np.linspace(0.0, self.Lx, self.Nx + 1)
            self.yGrid = np.linspace(0.0, self.Ly, self.Ny + 1)
            self.zGrid = np.linspace(0.0, self.Lz, self.Nz + 1)
            self.isper = [1, 1, 1]


class uniform:
    """geometry.py:76-116; also hands the spacing to the decomposition's box."""

    def __init__(self, xGrid, yGrid, zGrid, decomp):
        self.type = 'uniform'
        self.device = decomp.device
        self.prec = decomp.prec
        self.dtypeNumpy = decomp.dtypeNumpy
        self.Nx = len(xGrid) - 1
        self.Ny = len(yGrid) - 1
        self.Nz = len(zGrid) - 1
        self.Lx = xGrid[-1] - xGrid[0]
        self.Ly = yGrid[-1] - yGrid[0]
        self.Lz = zGrid[-1] - zGrid[0]
        self.dx = self.Lx / float(self.Nx); self.dxi = 1.0 / self.dx
        self.dy = self.Ly / float(self.Ny); self.dyi = 1.0 / self.dy
        self.dz = self.Lz / float(self.Nz); self.dzi = 1.0 / self.dz
        for a in ('nover', 'nxo_', 'nx_', 'nyo_', 'ny_', 'nzo_', 'nz_', 'imino_', 'jmino_',
                  'kmino_', 'imaxo_', 'jmaxo_', 'kmaxo_', 'imin_', 'imax_', 'jmin_', 'jmax_',
                  'kmin_', 'kmax_'):
            setattr(self, a, getattr(decomp, a))
        self.decomp = decomp
        decomp.box.set_spacing(self.dx, self.dy, self.dz)


def Geometry_Base(inputConfig, xGrid, yGrid, zGrid, decomp):
    """The factory core/PyFlow.py:103 calls (absent from the reference's
    geometry module): a uniform periodic box."""
    return uniform(xGrid, yGrid, zGrid, decomp)
