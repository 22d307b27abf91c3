"""metric/metric_staggered.py:38-526 (metric_uniform) on the GPU.

The fused time step (velocity/advancer/pressure entry points) never calls
these methods; they exist so that code written against the reference's
operator interface keeps working, each one a HIP kernel with the
reference's slice semantics on padded arrays.
"""

import ctypes

import torch

from . import _native
from .state import field_arg


class metric_uniform:
    def __init__(self, geo):
        self.geo = geo
        self.box = geo.decomp.box
        self.imin_ = geo.imin_; self.imax_ = geo.imax_
        self.jmin_ = geo.jmin_; self.jmax_ = geo.jmax_
        self.kmin_ = geo.kmin_; self.kmax_ = geo.kmax_
        self.imino_ = geo.imino_; self.imaxo_ = geo.imaxo_
        self.jmino_ = geo.jmino_; self.jmaxo_ = geo.jmaxo_
        self.kmino_ = geo.kmino_; self.kmaxo_ = geo.kmaxo_
        self.grid_geo = geo.type
        if self.grid_geo != 'uniform':
            raise NotImplementedError('grid type not implemented')
        self.div_x = self.grad_x = self.grad_xm = 1.0 * geo.dxi
        self.div_y = self.grad_y = self.grad_ym = 1.0 * geo.dyi
        self.div_z = self.grad_z = self.grad_zm = 1.0 * geo.dzi
        self.interp_x = self.interp_y = self.interp_z = 0.5
        self.interp_xm = self.interp_ym = self.interp_zm = 0.5
        self.vol = geo.dx * geo.dy * geo.dz
        self._prec = geo.prec
        self._device = geo.device
        self._shape = (geo.nxo_, geo.nyo_, geo.nzo_)
        self._laplace = None
        self._tmpv = None

    @property
    def Laplace(self):
        """The reference's constant coefficient array [nxo,nyo,nzo,3,3]
        (metric_staggered.py:85-102), materialised only if asked for: the
        kernels use its nine distinct values as scalars."""
        if self._laplace is None:
            g = self.geo
            L = torch.zeros(*self._shape, 3, 3, dtype=self._prec, device=self._device)
            for a, hi in enumerate((g.dxi, g.dyi, g.dzi)):
                L[:, :, :, a, 0].add_(hi * hi)
                L[:, :, :, a, 1].add_(-2.0 * hi * hi)
                L[:, :, :, a, 2].add_(hi * hi)
            L.mul_(-1.0 * self.vol)
            self._laplace = L
        return self._laplace

    @property
    def tmpv(self):
        if self._tmpv is None:
            self._tmpv = torch.zeros(self._shape, dtype=self._prec, device=self._device)
        return self._tmpv

    # ---- divergence (metric_staggered.py:438-453) ----
    def div_vel(self, state_u, state_v, state_w, divg):
        self.box.div_vel(state_u.var, state_v.var, state_w.var, divg)

    # ---- interpolations (metric_staggered.py:112-210) ----
    def _pair(self, kind, axis, upper, dst, src, coef):
        self.box.metric_pair(kind, axis, upper, dst, src, coef)

    def interp_sc_x(self, SC, interp_SC):
        self._pair(0, 0, False, interp_SC, SC, self.interp_x)

    def interp_sc_y(self, SC, interp_SC):
        self._pair(0, 1, False, interp_SC, SC, self.interp_y)

    def interp_sc_z(self, SC, interp_SC):
        self._pair(0, 2, False, interp_SC, SC, self.interp_z)

    def interp_sc_xy(self, SC, interp_SC):
        self.box.metric_edge4(0, interp_SC, SC)

    def interp_sc_xz(self, SC, interp_SC):
        self.box.metric_edge4(1, interp_SC, SC)

    def interp_sc_yz(self, SC, interp_SC):
        self.box.metric_edge4(2, interp_SC, SC)

    def interp_u_xm(self, state):
        self._pair(0, 0, True, state.var_i, state.var, 0.5)

    def interp_v_ym(self, state):
        self._pair(0, 1, True, state.var_i, state.var, 0.5)

    def interp_w_zm(self, state):
        self._pair(0, 2, True, state.var_i, state.var, 0.5)

    def interp_vw_x(self, state):
        self._pair(0, 0, False, state.var_i, state.var, self.interp_x)

    def interp_uw_y(self, state):
        self._pair(0, 1, False, state.var_i, state.var, self.interp_y)

    def interp_uv_z(self, state):
        self._pair(0, 2, False, state.var_i, state.var, self.interp_z)

    def interp_uvwi_x(self, state):
        self._pair(0, 0, False, state.var_i, state.var_i, self.interp_x)

    def interp_uvwi_y(self, state):
        self._pair(0, 1, False, state.var_i, state.var_i, self.interp_y)

    def interp_uvwi_z(self, state):
        self._pair(0, 2, False, state.var_i, state.var_i, self.interp_z)

    # ---- gradients (metric_staggered.py:215-401) ----
    def grad_vel_center(self, state, comp):
        """Cell-centred gradients (:215-268), same sequence of operations."""
        g = (state.grad_x, state.grad_y, state.grad_z)
        h = (self.grad_xm, self.grad_ym, self.grad_zm)
        ax = {'u': 0, 'v': 1, 'w': 2}[comp]
        interp = (self.interp_u_xm, self.interp_v_ym, self.interp_w_zm)[ax]
        uvwi = (self.interp_uvwi_x, self.interp_uvwi_y, self.interp_uvwi_z)
        for d in range(3):
            if d == ax:
                self._pair(1, d, True, g[d], state.var, h[d])
            else:
                interp(state)
                uvwi[d](state)
                self._pair(1, d, True, g[d], state.var_i, h[d])

    def grad_vel_visc(self, state):
        """:373-384"""
        self._pair(1, 0, True, state.grad_x, state.var, self.grad_x)
        self._pair(1, 1, True, state.grad_y, state.var, self.grad_y)
        self._pair(1, 2, True, state.grad_z, state.var, self.grad_z)

    def grad_P(self, state):
        """:390-401"""
        self._pair(1, 0, False, state.grad_x, state.var, self.grad_x)
        self._pair(1, 1, False, state.grad_y, state.var, self.grad_y)
        self._pair(1, 2, False, state.grad_z, state.var, self.grad_z)

    # ---- divergences (metric_staggered.py:406-433) ----
    def div_visc(self, FX, FY, FZ, rhs):
        self.box.div_visc(FX, FY, FZ, rhs)

    def div_vel_over(self, state_u, state_v, state_w, divg_o):
        self.box.div_vel_over(state_u.var, state_v.var, state_w.var, divg_o)

    # ---- convective fluxes (metric_staggered.py:458-526) ----
    def vel_conv_xx(self, state_u, state_v, grad_x, sign=1.0):
        self.box.vel_conv(0, True, state_u.var_i, state_v.var_i, grad_x, sign)

    def vel_conv_yy(self, state_u, state_v, grad_y, sign=1.0):
        self.box.vel_conv(1, True, state_u.var_i, state_v.var_i, grad_y, sign)

    def vel_conv_zz(self, state_u, state_w, grad_z, sign=1.0):
        self.box.vel_conv(2, True, state_u.var_i, state_w.var_i, grad_z, sign)

    def vel_conv_x(self, state_u, state_v, grad_x, sign=1.0):
        self.box.vel_conv(0, False, state_u.var_i, state_v.var_i, grad_x, sign)

    def vel_conv_y(self, state_u, state_v, grad_y, sign=1.0):
        self.box.vel_conv(1, False, state_u.var_i, state_v.var_i, grad_y, sign)

    def vel_conv_z(self, state_u, state_w, grad_z, sign=1.0):
        self.box.vel_conv(2, False, state_u.var_i, state_w.var_i, grad_z, sign)

    # ---- adjoint convective fluxes (metric_staggered.py:529-800) ----
    # Each is two subtractions, for s = 0 and s = e_s2:
    #   rhs -= ((a[m - e_d + s] - a[m + s]) * vi[m + s + vsh]) * coef
    # with vi the forward state's var_i left by the preceding interp_* call.
    def _adj(self, d, s2, centre, a, f, rhs, coef):
        self.box.adj_conv(d, s2, centre, a.var, f.var_i, rhs, coef)

    def adj_conv_xx(self, state_u_adj, state_u, rhs_u):
        self._adj(0, 0, True, state_u_adj, state_u, rhs_u, self.grad_x)

    def adj_conv_yy(self, state_v_adj, state_v, rhs_v):
        self._adj(1, 1, True, state_v_adj, state_v, rhs_v, self.grad_y)

    def adj_conv_zz(self, state_w_adj, state_w, rhs_w):
        self._adj(2, 2, True, state_w_adj, state_w, rhs_w, self.grad_z)

    def adj_conv_x(self, state_v_adj, state_u, rhs_v):
        self._adj(0, 0, False, state_v_adj, state_u, rhs_v, 0.5 * self.grad_x)

    def adj_conv_y(self, state_u_adj, state_v, rhs_u):
        self._adj(1, 1, False, state_u_adj, state_v, rhs_u, 0.5 * self.grad_y)

    def adj_conv_z(self, state_u_adj, state_w, rhs_u):
        self._adj(2, 2, False, state_u_adj, state_w, rhs_u, 0.5 * self.grad_z)

    def adj_conv_1d(self, state_v_adj, state_v, rhs_u):
        self._adj(0, 1, False, state_v_adj, state_v, rhs_u, 0.5 * self.grad_x)

    def adj_conv_1e(self, state_w_adj, state_w, rhs_u):
        self._adj(0, 2, False, state_w_adj, state_w, rhs_u, 0.5 * self.grad_x)

    def adj_conv_2d(self, state_u_adj, state_u, rhs_v):
        self._adj(1, 0, False, state_u_adj, state_u, rhs_v, 0.5 * self.grad_y)

    def adj_conv_2e(self, state_w_adj, state_w, rhs_v):
        self._adj(1, 2, False, state_w_adj, state_w, rhs_v, 0.5 * self.grad_y)

    def adj_conv_3d(self, state_u_adj, state_u, rhs_w):
        self._adj(2, 0, False, state_u_adj, state_u, rhs_w, 0.5 * self.grad_z)

    def adj_conv_3e(self, state_v_adj, state_v, rhs_w):
        self._adj(2, 1, False, state_v_adj, state_v, rhs_w, 0.5 * self.grad_z)
