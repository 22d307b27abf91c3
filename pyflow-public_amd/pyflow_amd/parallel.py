"""Single-GPU replacement of library/parallel.py: one periodic box per
process, so the communicator is trivial and the overlap exchange is a HIP
kernel that writes periodic images.

Multi-rank decomposition is not part of this build (one GPU, one box); the
constructor rejects nproc_x*nproc_y*nproc_z != 1.
"""

import numpy as np
import torch

from . import _native


class comms:
    """library/parallel.py:40-112 on a single rank: reductions are identities."""

    def __init__(self, IC=None):
        self.rank = 0
        self.size = 1
        self.dtype = getattr(IC, 'dtypeNumpy', np.float32) if IC is not None else np.float32

    def finalize(self):
        pass

    def parallel_sum(self, sendBuf):
        return sendBuf

    def parallel_max(self, sendBuf):
        return sendBuf

    def parallel_min(self, sendBuf):
        return sendBuf


class decomp:
    """library/parallel.py:118-284 for one rank owning the whole periodic box.

    Same attribute names (nx_, nxo_, imin_, imax_, ...).  Owns the native box
    handle (grid sizes, dtype and all device workspace); geometry.uniform sets
    its spacing.
    """

    def __init__(self, IC, config):
        self.prec = getattr(IC, 'dtypeTorch', torch.float32)
        self.dtypeNumpy = getattr(IC, 'dtypeNumpy', np.float32)
        _native.require_gpu()
        self.device = torch.device('cuda', torch.cuda.current_device())
        self.rank = 0
        self.size = 1
        self.npx = getattr(IC, 'nproc_x', 1)
        self.npy = getattr(IC, 'nproc_y', 1)
        self.npz = getattr(IC, 'nproc_z', 1)
        if self.npx * self.npy * self.npz != 1:
            raise ValueError('this build runs one periodic box per GPU: nproc_x*nproc_y*nproc_z '
                             'must be 1 (parallel.py:149-150 requires it to match the task count)')
        self.iproc = self.jproc = self.kproc = 0
        self.N = config.Nx * config.Ny * config.Nz
        self.nx, self.ny, self.nz = config.Nx, config.Ny, config.Nz
        self.isper = list(getattr(config, 'isper', [1, 1, 1]))
        self.nover = 2
        self.nxo, self.nyo, self.nzo = self.nx + 4, self.ny + 4, self.nz + 4
        self.imino = self.jmino = self.kmino = 0
        self.imaxo, self.jmaxo, self.kmaxo = self.nxo - 1, self.nyo - 1, self.nzo - 1
        self.imin = self.jmin = self.kmin = 2
        self.imax, self.jmax, self.kmax = self.nx + 1, self.ny + 1, self.nz + 1
        self.nx_, self.ny_, self.nz_ = self.nx, self.ny, self.nz
        self.imin_loc = self.jmin_loc = self.kmin_loc = 0
        self.imax_loc, self.jmax_loc, self.kmax_loc = self.nx - 1, self.ny - 1, self.nz - 1
        self.nxo_, self.nyo_, self.nzo_ = self.nxo, self.nyo, self.nzo
        self.imino_ = self.jmino_ = self.kmino_ = 0
        self.imaxo_, self.jmaxo_, self.kmaxo_ = self.nxo - 1, self.nyo - 1, self.nzo - 1
        self.imin_ = self.jmin_ = self.kmin_ = 2
        self.imax_, self.jmax_, self.kmax_ = self.nx + 1, self.ny + 1, self.nz + 1
        self.box = _native.Box(self.nx, self.ny, self.nz, dtype=self.prec, device=self.device)

    def free(self):
        self.box.close()

    def communicate_border(self, A):
        """Periodic overlap fill of a padded field (parallel.py:301-414)."""
        self.box.fill_halo(A)
