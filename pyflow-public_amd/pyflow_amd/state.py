"""data/state.py on the GPU: halo-padded field containers whose updates run
in HIP kernels (pf_state_op, pf_fill_halo)."""

import ctypes

import numpy as np
import torch

from . import _native

LAYOUT_PADDED = 0
LAYOUT_INTERIOR = 1
EW_ZAXPY, EW_COPY, EW_SUBSCALED = 0, 1, 2


def field_arg(box, t, name='field'):
    """(pointer, layout) of a padded array, an interior array, or the interior
    view of a padded array (what state_P.interior() returns)."""
    if not isinstance(t, torch.Tensor):
        raise TypeError('%s must be a torch.Tensor' % name)
    if t.device.type != 'cuda' or t.dtype != box.dtype:
        raise ValueError('%s must be a %s GPU tensor' % (name, box.dtype))
    if tuple(t.shape) == box.padded_shape and t.is_contiguous():
        return ctypes.c_void_p(t.data_ptr()), LAYOUT_PADDED
    if tuple(t.shape) == box.interior_shape:
        if t.is_contiguous():
            return ctypes.c_void_p(t.data_ptr()), LAYOUT_INTERIOR
        nyo, nzo = box.padded_shape[1], box.padded_shape[2]
        if t.stride() == (nyo * nzo, nzo, 1):
            off = 2 * nyo * nzo + 2 * nzo + 2
            base = t.data_ptr() - off * t.element_size()
            return ctypes.c_void_p(base), LAYOUT_PADDED
        return field_arg(box, t.contiguous(), name)
    raise ValueError('%s has shape %s; expected %s or %s' % (name, tuple(t.shape),
                                                             box.padded_shape, box.interior_shape))


class data_all_CPU:
    """data/state.py:38-92: host mirror of a list of states (file I/O and abs-max)."""

    def __init__(self, decomp, time, dt, names, state_data_all):
        self.device = decomp.device
        self.nx, self.ny, self.nz = decomp.nx, decomp.ny, decomp.nz
        self.nx_, self.ny_, self.nz_ = decomp.nx_, decomp.ny_, decomp.nz_
        self.imin_loc, self.imax_loc = decomp.imin_loc, decomp.imax_loc
        self.jmin_loc, self.jmax_loc = decomp.jmin_loc, decomp.jmax_loc
        self.kmin_loc, self.kmax_loc = decomp.kmin_loc, decomp.kmax_loc
        self.time = time
        self.dt = dt
        self.names = names
        self.nvar = len(names)
        self.ivar_read_start = 0
        self.nvar_read = 0
        self.data = []
        for st in state_data_all:
            self.data.append(st.interior())
            self.nvar_read += 1
        self._states = list(state_data_all)

    def append(self, ivar, inData):
        if ivar < self.nvar_read:
            self.data[ivar].copy_(torch.from_numpy(np.asarray(inData)).to(self.device))

    def read(self, ivar):
        if ivar < self.nvar_read:
            return self.data[ivar].cpu().numpy()

    def absmax(self, ivar):
        """max|interior| through the monitor kernel (pf_stats on (f, f, f))."""
        if ivar < self.nvar_read:
            f = self._states[ivar].var
            return np.float64(self._states[ivar].box.stats(f, f, f)[0])


class state_P:
    """data/state.py:98-211.  `var` is the padded field; `var_i` and the
    gradient arrays are allocated on first use (the fused time step never
    touches them)."""

    def __init__(self, decomp, IC=None, need_gradients=True):
        self.decomp = decomp
        self.box = decomp.box
        prec = decomp.prec
        self.nover = decomp.nover
        self.imin_ = decomp.imin_; self.imax_ = decomp.imax_
        self.jmin_ = decomp.jmin_; self.jmax_ = decomp.jmax_
        self.kmin_ = decomp.kmin_; self.kmax_ = decomp.kmax_
        self.imino_ = decomp.imino_; self.imaxo_ = decomp.imaxo_
        self.jmino_ = decomp.jmino_; self.jmaxo_ = decomp.jmaxo_
        self.kmino_ = decomp.kmino_; self.kmaxo_ = decomp.kmaxo_
        self._shape = (decomp.nxo_, decomp.nyo_, decomp.nzo_)
        self._prec = prec
        self.need_gradients = need_gradients
        self.var = torch.zeros(self._shape, dtype=prec, device=decomp.device)
        self._var_i = None
        self._grad = [None, None, None]
        if IC is not None:
            IC = np.asarray(IC)
            if IC.any():
                self.interior().copy_(torch.from_numpy(IC).to(self.var.device, prec))
        self.update_border()

    def _alloc(self):
        return torch.zeros(self._shape, dtype=self._prec, device=self.var.device)

    @property
    def var_i(self):
        if self._var_i is None:
            self._var_i = self._alloc()
        return self._var_i

    @var_i.setter
    def var_i(self, t):
        self._var_i = t

    def _g(self, d):
        if self._grad[d] is None:
            self._grad[d] = self._alloc()
        return self._grad[d]

    grad_x = property(lambda self: self._g(0), lambda self, t: self._grad.__setitem__(0, t))
    grad_y = property(lambda self: self._g(1), lambda self, t: self._grad.__setitem__(1, t))
    grad_z = property(lambda self: self._g(2), lambda self, t: self._grad.__setitem__(2, t))

    def interior(self):
        return self.var[self.imin_:self.imax_ + 1, self.jmin_:self.jmax_ + 1,
                        self.kmin_:self.kmax_ + 1]

    def update_border(self):
        self.decomp.communicate_border(self.var)

    def update_border_i(self):
        self.decomp.communicate_border(self.var_i)

    def _op(self, op, a, X, Y, halo):
        box = self.box
        xp, xl = field_arg(box, X, 'X')
        yp, yl = field_arg(box, Y, 'Y') if Y is not None else (None, LAYOUT_INTERIOR)
        _native.check(_native.lib().pf_state_op(box.h, op, float(a), ctypes.c_void_p(self.var.data_ptr()),
                                                xp, xl, yp, yl, int(halo), box.stream()),
                      'pf_state_op')

    def ZAXPY(self, A, X, Y):
        """interior = A*X + Y, then the overlap fill (state.py:157-165)."""
        self._op(EW_ZAXPY, A, X, Y, True)

    def copy(self, input):
        """interior = input, overlap untouched (state.py:168-172)."""
        self._op(EW_COPY, 0.0, input, None, False)

    def update(self, input):
        """interior = input, then the overlap fill (state.py:190-195)."""
        self._op(EW_COPY, 0.0, input, None, True)

    def vel_corr(self, inGrad, inScalar):
        """interior -= inScalar*inGrad, then the overlap fill (state.py:198-211)."""
        self._op(EW_SUBSCALED, inScalar, inGrad, None, True)
