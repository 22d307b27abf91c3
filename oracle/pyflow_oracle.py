"""CPU restatement of PyFlow's single-box time-step hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: only
the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  Nothing in the product package imports it, and
the product never falls back to it.

It restates, op for op, the reference's PyTorch sequence on halo-padded
arrays (2 overlap cells per side, z fastest) for a single triply periodic
box on one rank, so that its floating-point results are the reference's:

* the same elementwise operations in the same order (torch CPU rounds every
  op separately, except ``add_/sub_`` with an ``alpha`` which it fuses into
  one FMA; this module calls exactly those torch ops);
* the same reductions (``torch.sum`` / ``torch.max`` on CPU tensors).

Every function cites the reference file:line it follows (paths relative to
the reference's ``src/``).  How each function is pinned (golden vectors
written by importing the reference, closed forms, textbook routines) is
listed in the test modules that exercise it; functions without such a check
say "parity unpinned" in their docstring.
"""

import math

import numpy as np
import torch

NOVER = 2  # overlap depth, parallel.py:201


class Grid:
    """Uniform periodic box on one rank.

    Mirrors geometry.uniform (geometry/geometry.py:76-116) plus the single-rank
    part of parallel.decomp (library/parallel.py:194-284): padded sizes
    n+2*nover, interior indices [imin, imax] inclusive with imin = 2.
    """

    def __init__(self, nx, ny, nz, Lx=2 * math.pi, Ly=None, Lz=None,
                 dtype=torch.float64):
        Ly = Lx if Ly is None else Ly
        Lz = Lx if Lz is None else Lz
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = Lx, Ly, Lz
        self.dtype = dtype
        # geometry.py:97-99
        self.dx = Lx / float(nx); self.dxi = 1.0 / self.dx
        self.dy = Ly / float(ny); self.dyi = 1.0 / self.dy
        self.dz = Lz / float(nz); self.dzi = 1.0 / self.dz
        # metric_staggered.py:78
        self.vol = self.dx * self.dy * self.dz
        # parallel.py:268-284
        self.nxo, self.nyo, self.nzo = nx + 2 * NOVER, ny + 2 * NOVER, nz + 2 * NOVER
        self.imin = self.jmin = self.kmin = NOVER
        self.imax = NOVER + nx - 1
        self.jmax = NOVER + ny - 1
        self.kmax = NOVER + nz - 1

    def padded(self):
        return torch.zeros(self.nxo, self.nyo, self.nzo, dtype=self.dtype)

    def interior_zeros(self):
        return torch.zeros(self.nx, self.ny, self.nz, dtype=self.dtype)

    def interior(self, a):
        """state.py:142-146"""
        return a[self.imin:self.imax + 1, self.jmin:self.jmax + 1,
                 self.kmin:self.kmax + 1]

    def from_interior(self, x):
        """state.py:132-140: a padded array holding ``x`` with filled halos."""
        a = self.padded()
        self.interior(a).copy_(torch.as_tensor(x, dtype=self.dtype))
        halo_fill(a)
        return a


# ---------------------------------------------------------------------------
# Overlap cells
# ---------------------------------------------------------------------------
def halo_fill(a):
    """Periodic overlap fill on a single rank.

    parallel.py:301-414 with the Cartesian communicator periodic in every
    direction: each direction copies the first two interior planes into the
    far overlap and the last two into the near overlap, x then y then z, over
    the full extent of the other two directions (so edges and corners end up
    as periodic images too).
    """
    no = NOVER
    n1, n2, n3 = a.shape
    a[n1 - no:n1].copy_(a[no:2 * no].clone())
    a[0:no].copy_(a[n1 - 2 * no:n1 - no].clone())
    a[:, n2 - no:n2].copy_(a[:, no:2 * no].clone())
    a[:, 0:no].copy_(a[:, n2 - 2 * no:n2 - no].clone())
    a[:, :, n3 - no:n3].copy_(a[:, :, no:2 * no].clone())
    a[:, :, 0:no].copy_(a[:, :, n3 - 2 * no:n3 - no].clone())


# ---------------------------------------------------------------------------
# Field container
# ---------------------------------------------------------------------------
class Field:
    """state_P (data/state.py:98-211): var, var_i, grad_x/y/z, all padded."""

    def __init__(self, g, interior=None):
        self.g = g
        self.var = g.padded()
        self.var_i = g.padded()
        self.grad_x = g.padded()
        self.grad_y = g.padded()
        self.grad_z = g.padded()
        if interior is not None:
            g.interior(self.var).copy_(torch.as_tensor(interior, dtype=g.dtype))
        halo_fill(self.var)

    def interior(self):
        return self.g.interior(self.var)

    def update_border(self):
        halo_fill(self.var)

    def ZAXPY(self, A, X, Y):
        """state.py:157-165"""
        self.interior().copy_(A * X + Y)
        self.update_border()

    def update(self, x):
        """state.py:168-195"""
        self.interior().copy_(x)
        self.update_border()

    def vel_corr(self, grad, scalar):
        """state.py:198-211: interior -= scalar*grad (torch fuses this into one FMA)."""
        g = self.g
        self.interior().sub_(g.interior(grad), alpha=scalar)
        self.update_border()


# ---------------------------------------------------------------------------
# Staggered operators: metric_uniform (metric/metric_staggered.py:38-526)
# ---------------------------------------------------------------------------
def laplace_coefficients(g):
    """The constant 7-point coefficients (metric_staggered.py:85-102).

    Returns ((cx_m, cx_0, cx_p), (cy_m, cy_0, cy_p), (cz_m, cz_0, cz_p)) as
    0-d tensors of the grid dtype, i.e. -vol*(1/h^2, -2/h^2, 1/h^2) per axis
    rounded exactly as the reference's tensor ops round them.
    """
    out = []
    for hi in (g.dxi, g.dyi, g.dzi):
        c = []
        for v in (hi * hi, -2.0 * hi * hi, hi * hi):
            t = torch.zeros((), dtype=g.dtype)
            t.add_(v)
            t.mul_(-1.0 * g.vol)
            c.append(t)
        out.append(tuple(c))
    return tuple(out)


def interp_sc_x(g, SC, out):
    out[1:, :, :].copy_(0.5 * (SC[1:, :, :] + SC[:-1, :, :]))           # :112-113


def interp_sc_y(g, SC, out):
    out[:, 1:, :].copy_(0.5 * (SC[:, 1:, :] + SC[:, :-1, :]))           # :117-118


def interp_sc_z(g, SC, out):
    out[:, :, 1:].copy_(0.5 * (SC[:, :, 1:] + SC[:, :, :-1]))           # :122-123


def interp_sc_xy(g, SC, out):
    """metric_staggered.py:127-130 (0.5*0.5 is formed first, then scales the 4-sum)."""
    out[1:, 1:, :].copy_(0.5 * 0.5 * (SC[1:, 1:, :] + SC[:-1, 1:, :] +
                                      SC[1:, :-1, :] + SC[:-1, :-1, :]))


def interp_sc_xz(g, SC, out):
    out[1:, :, 1:].copy_(0.5 * 0.5 * (SC[1:, :, 1:] + SC[:-1, :, 1:] +   # :134-137
                                      SC[1:, :, :-1] + SC[:-1, :, :-1]))


def interp_sc_yz(g, SC, out):
    out[:, 1:, 1:].copy_(0.5 * 0.5 * (SC[:, 1:, 1:] + SC[:, :-1, 1:] +   # :141-144
                                      SC[:, 1:, :-1] + SC[:, :-1, :-1]))


def interp_u_xm(f):
    f.var_i[:-1, :, :] = 0.5 * (f.var[1:, :, :] + f.var[:-1, :, :])     # :149-151


def interp_v_ym(f):
    f.var_i[:, :-1, :] = 0.5 * (f.var[:, 1:, :] + f.var[:, :-1, :])     # :156-158


def interp_w_zm(f):
    f.var_i[:, :, :-1] = 0.5 * (f.var[:, :, 1:] + f.var[:, :, :-1])     # :163-165


def interp_vw_x(f):
    f.var_i[1:, :, :] = 0.5 * (f.var[1:, :, :] + f.var[:-1, :, :])      # :171-173


def interp_uw_y(f):
    f.var_i[:, 1:, :] = 0.5 * (f.var[:, 1:, :] + f.var[:, :-1, :])      # :178-180


def interp_uv_z(f):
    f.var_i[:, :, 1:] = 0.5 * (f.var[:, :, 1:] + f.var[:, :, :-1])      # :185-187


def interp_uvwi_x(f):
    f.var_i[1:, :, :] = 0.5 * (f.var_i[1:, :, :] + f.var_i[:-1, :, :])  # :193-195


def interp_uvwi_y(f):
    f.var_i[:, 1:, :] = 0.5 * (f.var_i[:, 1:, :] + f.var_i[:, :-1, :])  # :200-202


def interp_uvwi_z(f):
    f.var_i[:, :, 1:] = 0.5 * (f.var_i[:, :, 1:] + f.var_i[:, :, :-1])  # :207-209


def _dfwd(dst, src, axis, h):
    """dst[:-1] = (src[1:] - src[:-1]) * h along ``axis`` (copy, sub, mul)."""
    sl_hi = [slice(None)] * 3
    sl_lo = [slice(None)] * 3
    sl_hi[axis] = slice(1, None)
    sl_lo[axis] = slice(None, -1)
    d = dst[tuple(sl_lo)]
    d.copy_(src[tuple(sl_hi)])
    d.sub_(src[tuple(sl_lo)])
    d.mul_(h)


def _dbwd(dst, src, axis, h):
    """dst[1:] = (src[1:] - src[:-1]) * h along ``axis`` (copy, sub, mul)."""
    sl_hi = [slice(None)] * 3
    sl_lo = [slice(None)] * 3
    sl_hi[axis] = slice(1, None)
    sl_lo[axis] = slice(None, -1)
    d = dst[tuple(sl_hi)]
    d.copy_(src[tuple(sl_hi)])
    d.sub_(src[tuple(sl_lo)])
    d.mul_(h)


def grad_vel_center(g, f, comp):
    """Cell-centred velocity gradients, metric_staggered.py:215-268."""
    if comp == 'u':
        _dfwd(f.grad_x, f.var, 0, g.dxi)
        interp_u_xm(f); interp_uvwi_y(f)
        _dfwd(f.grad_y, f.var_i, 1, g.dyi)
        interp_u_xm(f); interp_uvwi_z(f)
        _dfwd(f.grad_z, f.var_i, 2, g.dzi)
    elif comp == 'v':
        interp_v_ym(f); interp_uvwi_x(f)
        _dfwd(f.grad_x, f.var_i, 0, g.dxi)
        _dfwd(f.grad_y, f.var, 1, g.dyi)
        interp_v_ym(f); interp_uvwi_z(f)
        _dfwd(f.grad_z, f.var_i, 2, g.dzi)
    elif comp == 'w':
        interp_w_zm(f); interp_uvwi_x(f)
        _dfwd(f.grad_x, f.var_i, 0, g.dxi)
        interp_w_zm(f); interp_uvwi_y(f)
        _dfwd(f.grad_y, f.var_i, 1, g.dyi)
        _dfwd(f.grad_z, f.var, 2, g.dzi)
    else:
        raise ValueError(comp)


def grad_vel_visc(g, f):
    """metric_staggered.py:373-384"""
    _dfwd(f.grad_x, f.var, 0, g.dxi)
    _dfwd(f.grad_y, f.var, 1, g.dyi)
    _dfwd(f.grad_z, f.var, 2, g.dzi)


def grad_P(g, f):
    """metric_staggered.py:390-401"""
    _dbwd(f.grad_x, f.var, 0, g.dxi)
    _dbwd(f.grad_y, f.var, 1, g.dyi)
    _dbwd(f.grad_z, f.var, 2, g.dzi)


def _ii(g, di=0, dj=0, dk=0):
    return (slice(g.imin + di, g.imax + 1 + di),
            slice(g.jmin + dj, g.jmax + 1 + dj),
            slice(g.kmin + dk, g.kmax + 1 + dk))


def div_visc(g, FX, FY, FZ, rhs):
    """metric_staggered.py:406-418 (each add_/sub_ with a scalar is one FMA)."""
    rhs.add_(FX[_ii(g)], alpha=g.dxi)
    rhs.sub_(FX[_ii(g, di=-1)], alpha=g.dxi)
    rhs.add_(FY[_ii(g)], alpha=g.dyi)
    rhs.sub_(FY[_ii(g, dj=-1)], alpha=g.dyi)
    rhs.add_(FZ[_ii(g)], alpha=g.dzi)
    rhs.sub_(FZ[_ii(g, dk=-1)], alpha=g.dzi)


def div_vel_over(g, fu, fv, fw, divg_o):
    """metric_staggered.py:424-433"""
    divg_o.zero_()
    d = divg_o[:-1, :, :]
    d.add_(fu.var[1:, :, :])
    d.sub_(fu.var[:-1, :, :])
    d.mul_(g.dxi)
    divg_o[:, :-1, :] += (fv.var[:, 1:, :] - fv.var[:, :-1, :]) * g.dyi
    divg_o[:, :, :-1] += (fw.var[:, :, 1:] - fw.var[:, :, :-1]) * g.dzi


def div_vel(g, fu, fv, fw, divg):
    """Interior cell-centred divergence, metric_staggered.py:438-453."""
    divg.zero_()
    divg.add_(fu.var[_ii(g, di=1)])
    divg.sub_(fu.var[_ii(g)])
    divg.mul_(g.dxi)
    divg += (fv.var[_ii(g, dj=1)] - fv.var[_ii(g)]) * g.dyi
    divg += (fw.var[_ii(g, dk=1)] - fw.var[_ii(g)]) * g.dzi


def vel_conv_xx(g, a, b, rhs, sign=1.0):
    """metric_staggered.py:458-466"""
    I = slice(g.imin, g.imax + 1); Im = slice(g.imin - 1, g.imax)
    J = slice(g.jmin, g.jmax + 1); K = slice(g.kmin, g.kmax + 1)
    rhs -= (a.var_i[I, J, K] * b.var_i[I, J, K] -
            a.var_i[Im, J, K] * b.var_i[Im, J, K]) * g.dxi * sign


def vel_conv_yy(g, a, b, rhs, sign=1.0):
    """metric_staggered.py:470-478"""
    I = slice(g.imin, g.imax + 1); K = slice(g.kmin, g.kmax + 1)
    J = slice(g.jmin, g.jmax + 1); Jm = slice(g.jmin - 1, g.jmax)
    rhs -= (a.var_i[I, J, K] * b.var_i[I, J, K] -
            a.var_i[I, Jm, K] * b.var_i[I, Jm, K]) * g.dyi * sign


def vel_conv_zz(g, a, b, rhs, sign=1.0):
    """metric_staggered.py:482-490"""
    I = slice(g.imin, g.imax + 1); J = slice(g.jmin, g.jmax + 1)
    K = slice(g.kmin, g.kmax + 1); Km = slice(g.kmin - 1, g.kmax)
    rhs -= (a.var_i[I, J, K] * b.var_i[I, J, K] -
            a.var_i[I, J, Km] * b.var_i[I, J, Km]) * g.dzi * sign


def vel_conv_x(g, a, b, rhs, sign=1.0):
    """metric_staggered.py:494-502"""
    Ip = slice(g.imin + 1, g.imax + 2); I = slice(g.imin, g.imax + 1)
    J = slice(g.jmin, g.jmax + 1); K = slice(g.kmin, g.kmax + 1)
    rhs -= (a.var_i[Ip, J, K] * b.var_i[Ip, J, K] -
            a.var_i[I, J, K] * b.var_i[I, J, K]) * g.dxi * sign


def vel_conv_y(g, a, b, rhs, sign=1.0):
    """metric_staggered.py:506-514"""
    I = slice(g.imin, g.imax + 1); K = slice(g.kmin, g.kmax + 1)
    Jp = slice(g.jmin + 1, g.jmax + 2); J = slice(g.jmin, g.jmax + 1)
    rhs -= (a.var_i[I, Jp, K] * b.var_i[I, Jp, K] -
            a.var_i[I, J, K] * b.var_i[I, J, K]) * g.dyi * sign


def vel_conv_z(g, a, b, rhs, sign=1.0):
    """metric_staggered.py:518-526"""
    I = slice(g.imin, g.imax + 1); J = slice(g.jmin, g.jmax + 1)
    Kp = slice(g.kmin + 1, g.kmax + 2); K = slice(g.kmin, g.kmax + 1)
    rhs -= (a.var_i[I, J, Kp] * b.var_i[I, J, Kp] -
            a.var_i[I, J, K] * b.var_i[I, J, K]) * g.dzi * sign


# ---------------------------------------------------------------------------
# Closures with the 'tensor' contract (sfsmodel_gradient.py:39-148)
# ---------------------------------------------------------------------------
class GradientModel:
    """Gradient SFS model, sfsmodel/sfsmodel_gradient.py:39-148."""

    modelType = 'tensor'

    def __init__(self, g):
        self.g = g
        self.const = g.dx * g.dx / 12.0                     # :63
        for n in ('FXX', 'FYY', 'FZZ', 'FXY', 'FYZ', 'FXZ', 'tmp'):
            setattr(self, n, g.padded())

    def update(self, fu, fv, fw):
        g = self.g
        grad_vel_center(g, fu, 'u')
        grad_vel_center(g, fv, 'v')
        grad_vel_center(g, fw, 'w')
        pairs = (('FXX', fu, fu), ('FYY', fv, fv), ('FZZ', fw, fw),
                 ('FXY', fu, fv), ('FXZ', fu, fw), ('FYZ', fv, fw))
        for name, a, b in pairs:                            # :77-148
            out = getattr(self, name)
            t = self.tmp
            t.copy_(a.grad_x); t.mul_(b.grad_x); out.copy_(t)
            t.copy_(a.grad_y); t.mul_(b.grad_y); out.add_(t)
            t.copy_(a.grad_z); t.mul_(b.grad_z); out.add_(t)
            out.mul_(self.const)


class Smagorinsky:
    """Constant-coefficient Smagorinsky eddy viscosity,
    sfsmodel/sfsmodel_smagorinsky.py:39-88 (modelType 'eddyVisc'; Box adds it
    to the molecular viscosity before each step, domain.py:348-352)."""

    modelType = 'eddyVisc'

    def __init__(self, g, Cs=0.18, expFilterFac=1.0):
        self.g = g
        self.Cs = Cs
        self.Delta = g.dx * expFilterFac                    # :55-57

    def eddyVisc(self, fu, fv, fw, rho):
        g = self.g
        grad_vel_center(g, fu, 'u')                          # :68-70
        grad_vel_center(g, fv, 'v')
        grad_vel_center(g, fw, 'w')
        divOver3 = (fu.grad_x + fv.grad_y + fw.grad_z) / 3.0    # :73
        S_11 = fu.grad_x - divOver3                          # :76-81
        S_22 = fv.grad_y - divOver3
        S_33 = fw.grad_z - divOver3
        S_12 = 0.5 * (fu.grad_y + fv.grad_x)
        S_13 = 0.5 * (fu.grad_z + fw.grad_x)
        S_23 = 0.5 * (fv.grad_z + fw.grad_y)
        S_mag = torch.sqrt(S_11 * S_11 + S_22 * S_22 + S_33 * S_33 +
                           2.0 * (S_12 * S_12 + S_13 * S_13 + S_23 * S_23))    # :82-83
        return rho * (self.Cs * self.Delta) ** 2 * S_mag    # :86


def mlp_closure_forward(g, fu, fv, fw, params, const=None):
    """Closure MLP with the 'tensor' contract: 9 cell-centred velocity
    gradients -> Linear(9,20) -> tanh -> Linear(20,20) -> tanh -> Linear(20,6)
    -> (FXX, FYY, FZZ, FXY, FXZ, FYZ), each scaled by dx^2/12 like the
    gradient model's stresses.

    The MLP has no reference implementation (it replaces the arithmetic of
    sfsmodel_gradient.py:77-148 behind the same interface); this FP64 torch
    version is its oracle.  Pinned by closed forms (tests/test_oracle_golden:
    zero weights with output biases; selection weights that fix the feature
    order, the orientation of each weight matrix and the dx^2/12 scale) and
    by autograd-consistency tests.
    """
    grad_vel_center(g, fu, 'u')
    grad_vel_center(g, fv, 'v')
    grad_vel_center(g, fw, 'w')
    feats = torch.stack([fu.grad_x, fu.grad_y, fu.grad_z,
                         fv.grad_x, fv.grad_y, fv.grad_z,
                         fw.grad_x, fw.grad_y, fw.grad_z], dim=-1)
    W1, b1, W2, b2, W3, b3 = [torch.as_tensor(p, dtype=torch.float64) for p in params]
    h = torch.tanh(feats.double() @ W1.T + b1)
    h = torch.tanh(h @ W2.T + b2)
    o = h @ W3.T + b3
    c = g.dx * g.dx / 12.0 if const is None else const
    o = o * c
    return [o[..., m].to(g.dtype).contiguous() for m in range(6)]


# ---------------------------------------------------------------------------
# Momentum right-hand side: rhs_NavierStokes (solver/velocity.py:189-383)
# ---------------------------------------------------------------------------
class MLSource:
    """The reference's 'source' ML closure, sfsmodel/sfsmodel_ML.py:354-593
    (update, Finite_Difference_ConCat, ML_Closure) with NeuralNetworkModel2
    (:76-116), restated with torch CPU ops in FP32 as the reference runs it.
    ``sd`` is the network's state_dict.  Shifts wrap over the padded extent
    (the reference's torch.cat of first and last planes; torch.roll moves the
    same values), and every direction uses dx (:408-465, :468-593)."""

    modelType = 'source'

    # sfsmodel_ML.py:252-258
    NORM = 10.0 * np.array([1.0233662e+00, 1.0241578e+00, 1.0196122e+00,
                            5.0992206e+02, 5.2713013e+02, 5.2689368e+02, 5.2550342e+02,
                            5.1325964e+02, 5.2564642e+02, 5.3132422e+02, 5.2574353e+02,
                            5.1027972e+02, 9.5621462e+05, 1.2852805e+06, 1.2698791e+06,
                            1.2744681e+06, 9.6111162e+05, 1.2722050e+06, 1.2891108e+06,
                            1.2891108e+06, 9.6587112e+05])

    def __init__(self, g, sd, trainable=False):
        self.g = g
        # trainable: the weights are autograd leaves, so a backward pass
        # through update() accumulates their gradients (the reference's model
        # parameters, whose .grad g.backward() fills in adjoint.py:265)
        self.sd = {k: v.to(torch.float32).clone().requires_grad_(trainable)
                   for k, v in sd.items()}
        self.norm = torch.tensor(self.NORM, dtype=torch.float32)
        self.last_output = None

    def packed_grad(self):
        """The weights' accumulated gradients in the packed order fc1, fcG,
        fcG2, fc2, fc3, fc4 (weight, then bias)."""
        parts = []
        for n in ('fc1', 'fcG', 'fcG2', 'fc2', 'fc3', 'fc4'):
            for k in ('.weight', '.bias'):
                t = self.sd[n + k]
                parts.append((t.grad if t.grad is not None else torch.zeros_like(t)).reshape(-1))
        return torch.cat(parts).detach()

    def _d1(self, f, dim):
        return (torch.roll(f, -1, dim) - torch.roll(f, 1, dim)) / (2 * self.g.dx)

    def _d2(self, f, dim):
        return (torch.roll(f, -1, dim) - 2 * f + torch.roll(f, 1, dim)) / (self.g.dx * self.g.dx)

    def features(self, fu, fv, fw):
        """The 21 inputs per padded cell (:372-378), float32."""
        f3 = [a.to(torch.float32) for a in (fu, fv, fw)]
        first = [self._d1(f, d) for f in f3 for d in range(3)]
        second = [self._d2(f, d) for f in f3 for d in range(3)]
        return torch.stack(f3 + first + second, 3)

    def network(self, x):
        """NeuralNetworkModel2.forward (:88-116)."""
        sd = self.sd
        lin = lambda n, a: torch.nn.functional.linear(a, sd[n + '.weight'], sd[n + '.bias'])
        H1 = torch.tanh(lin('fc1', x))
        H2 = torch.tanh(lin('fc2', H1))
        G = torch.tanh(lin('fcG', x))
        Gate2 = torch.tanh(lin('fcG2', x))
        H3 = torch.tanh(lin('fc3', G * H2))
        return lin('fc4', H3 * Gate2)

    def update(self, fu, fv, fw, requires_grad=False):
        x = self.features(fu, fv, fw)
        xn = x / self.norm
        parts = [xn]
        for d in (1, 2):                                     # :413-447 order
            for dim in range(3):
                parts.append(torch.roll(xn, -d, dim))
                parts.append(torch.roll(xn, d, dim))
        p = self.network(torch.cat(parts, 3))
        self.last_output = p
        c = [self._d1(p[..., q:q + 1], q % 3) for q in range(9)]
        G = [(c[3 * k] + c[3 * k + 1]) + c[3 * k + 2] for k in range(3)]   # :586-588
        self.GX, self.GY, self.GZ = (a.to(self.g.dtype) for a in G)


class NavierStokesRHS:
    """rhs_NavierStokes, solver/velocity.py:189-383 (viscous + tensor-SFS
    fluxes, then the nine advective terms)."""

    def __init__(self, g, rho, VISC, sfs=None):
        self.g = g
        self.rho = rho
        self.VISC = VISC                                    # padded tensor (domain.py:130-131)
        self.sfs = sfs
        self.rhs_u = g.interior_zeros()
        self.rhs_v = g.interior_zeros()
        self.rhs_w = g.interior_zeros()
        self.FX = g.padded(); self.FY = g.padded(); self.FZ = g.padded()
        self.div_vel = g.padded()
        self.interp_SC = g.padded()

    def evaluate(self, fu, fv, fw):
        g = self.g
        sfs = self.sfs
        self.rhs_u.zero_(); self.rhs_v.zero_(); self.rhs_w.zero_()   # :227-229
        self.viscous(fu, fv, fw, sfs is not None and sfs.modelType == 'tensor')

        # advection: :322-355
        interp_u_xm(fu); vel_conv_xx(g, fu, fu, self.rhs_u)
        interp_uw_y(fu); interp_vw_x(fv); vel_conv_y(g, fu, fv, self.rhs_u)
        interp_uv_z(fu); interp_vw_x(fw); vel_conv_z(g, fu, fw, self.rhs_u)

        interp_vw_x(fv); interp_uw_y(fu); vel_conv_x(g, fv, fu, self.rhs_v)
        interp_v_ym(fv); vel_conv_yy(g, fv, fv, self.rhs_v)
        interp_uv_z(fv); interp_uw_y(fw); vel_conv_z(g, fv, fw, self.rhs_v)

        interp_vw_x(fw); interp_uv_z(fu); vel_conv_x(g, fw, fu, self.rhs_w)
        interp_uw_y(fw); interp_uv_z(fv); vel_conv_y(g, fw, fv, self.rhs_w)
        interp_w_zm(fw); vel_conv_zz(g, fw, fw, self.rhs_w)

        # 'source' closures: :360-383 (inputs are the var_i arrays as the
        # interpolations above and these three leave them)
        if sfs is not None and sfs.modelType == 'source':
            interp_u_xm(fu); interp_v_ym(fv); interp_w_zm(fw)
            sfs.update(fu.var_i, fv.var_i, fw.var_i)
            SC = self.interp_SC
            for G, interp, rhs in ((sfs.GX, interp_sc_x, self.rhs_u),
                                   (sfs.GY, interp_sc_y, self.rhs_v),
                                   (sfs.GZ, interp_sc_z, self.rhs_w)):
                interp(g, G[..., 0] if G.dim() == 4 else G, SC)
                rhs.sub_(g.interior(SC))

    def viscous(self, fu, fv, fw, tensor):
        """Viscous (and tensor-closure) flux divergence accumulated into
        rhs_u/v/w, velocity.py:235-317."""
        g = self.g
        rho, VISC, sfs = self.rho, self.VISC, self.sfs
        FX, FY, FZ, SC = self.FX, self.FY, self.FZ, self.interp_SC

        grad_vel_visc(g, fu); grad_vel_visc(g, fv); grad_vel_visc(g, fw)   # :235-237
        div_vel_over(g, fu, fv, fw, self.div_vel)                        # :240
        self.div_vel.div_(3.0)                                           # :241

        # u: :248-269
        FX.copy_(fu.grad_x); FX.sub_(self.div_vel); FX.mul_(2.0 * VISC / rho)
        interp_sc_xy(g, VISC, SC)
        FY.copy_(fu.grad_y); FY.add_(fv.grad_x); FY.mul_(SC / rho)
        interp_sc_xz(g, VISC, SC)
        FZ.copy_(fu.grad_z); FZ.add_(fw.grad_x); FZ.mul_(SC / rho)
        if tensor:
            FX.add_(sfs.FXX)
            interp_sc_xy(g, sfs.FXY, SC); FY.add_(SC)
            interp_sc_xz(g, sfs.FXZ, SC); FZ.add_(SC)
        div_visc(g, FX, FY, FZ, self.rhs_u)

        # v: :272-293
        interp_sc_xy(g, VISC, SC)
        FX.copy_(fv.grad_x); FX.add_(fu.grad_y); FX.mul_(SC / rho)
        FY.copy_(fv.grad_y); FY.sub_(self.div_vel); FY.mul_(2.0 * VISC / rho)
        interp_sc_yz(g, VISC, SC)
        FZ.copy_(fv.grad_z); FZ.add_(fw.grad_y); FZ.mul_(SC / rho)
        if tensor:
            interp_sc_xy(g, sfs.FXY, SC); FX.add_(SC)
            FY.add_(sfs.FYY)
            interp_sc_yz(g, sfs.FYZ, SC); FZ.add_(SC)
        div_visc(g, FX, FY, FZ, self.rhs_v)

        # w: :296-317
        interp_sc_xz(g, VISC, SC)
        FX.copy_(fw.grad_x); FX.add_(fu.grad_z); FX.mul_(SC / rho)
        interp_sc_yz(g, VISC, SC)
        FY.copy_(fw.grad_y); FY.add_(fv.grad_z); FY.mul_(SC / rho)
        FZ.copy_(fw.grad_z); FZ.sub_(self.div_vel); FZ.mul_(2.0 * VISC / rho)
        if tensor:
            interp_sc_xz(g, sfs.FXZ, SC); FX.add_(SC)
            interp_sc_yz(g, sfs.FYZ, SC); FY.add_(SC)
            FZ.add_(sfs.FZZ)
        div_visc(g, FX, FY, FZ, self.rhs_w)


# ---------------------------------------------------------------------------
# Time advancers (solver/advancer.py:41-113)
# ---------------------------------------------------------------------------
def euler_step(fu, fv, fw, rhs, dt):
    """advancer.py:49-57"""
    rhs.evaluate(fu, fv, fw)
    fu.update(fu.interior() + rhs.rhs_u * dt)
    fv.update(fv.interior() + rhs.rhs_v * dt)
    fw.update(fw.interior() + rhs.rhs_w * dt)


def rk4_step(fu, fv, fw, tu, tv, tw, rhss, dt):
    """The reference's 'RK4', advancer.py:75-113.

    Stages 2-4 all start from u^n + (dt/2)*k_s (stage 4 too, advancer.py:93-95),
    then u = u^n + dt/6*(k1 + 2k2 + 2k3 + k4).  ``rhss`` holds four
    NavierStokesRHS objects (velocity.py:45-76 allocates four).
    """
    r1, r2, r3, r4 = rhss
    r1.evaluate(fu, fv, fw)
    tu.ZAXPY(0.5 * dt, r1.rhs_u, fu.interior())
    tv.ZAXPY(0.5 * dt, r1.rhs_v, fv.interior())
    tw.ZAXPY(0.5 * dt, r1.rhs_w, fw.interior())
    r2.evaluate(tu, tv, tw)
    tu.ZAXPY(0.5 * dt, r2.rhs_u, fu.interior())
    tv.ZAXPY(0.5 * dt, r2.rhs_v, fv.interior())
    tw.ZAXPY(0.5 * dt, r2.rhs_w, fw.interior())
    r3.evaluate(tu, tv, tw)
    tu.ZAXPY(0.5 * dt, r3.rhs_u, fu.interior())
    tv.ZAXPY(0.5 * dt, r3.rhs_v, fv.interior())
    tw.ZAXPY(0.5 * dt, r3.rhs_w, fw.interior())
    r4.evaluate(tu, tv, tw)
    fu.update(fu.interior() + dt / 6.0 * (r1.rhs_u + 2.0 * r2.rhs_u + 2.0 * r3.rhs_u + r4.rhs_u))
    fv.update(fv.interior() + dt / 6.0 * (r1.rhs_v + 2.0 * r2.rhs_v + 2.0 * r3.rhs_v + r4.rhs_v))
    fw.update(fw.interior() + dt / 6.0 * (r1.rhs_w + 2.0 * r2.rhs_w + 2.0 * r3.rhs_w + r4.rhs_w))


# ---------------------------------------------------------------------------
# Pressure Poisson (solver/pressure.py:44-297)
# ---------------------------------------------------------------------------
def apply_operator(g, lap, A, B):
    """Pressure_Base.applyOperator, pressure.py:54-84: fill A's overlap, then
    B = sum over axes of A[-1]*c0 + A[0]*c1 + A[+1]*c2 in the reference order."""
    halo_fill(A)
    B.zero_()
    (x0, x1, x2), (y0, y1, y2), (z0, z1, z2) = lap
    B.add_(A[_ii(g, di=-1)] * x0); B.add_(A[_ii(g)] * x1); B.add_(A[_ii(g, di=1)] * x2)
    B.add_(A[_ii(g, dj=-1)] * y0); B.add_(A[_ii(g)] * y1); B.add_(A[_ii(g, dj=1)] * y2)
    B.add_(A[_ii(g, dk=-1)] * z0); B.add_(A[_ii(g)] * z1); B.add_(A[_ii(g, dk=1)] * z2)


class BiCGStab:
    """solver_bicgstab, pressure.py:164-297 (unpreconditioned, x0 = 0,
    r_hat = r0, stop on max|r|/max|r0| <= tol or iter >= max_iterations,
    silent exit on the two breakdowns).  One deliberate difference: where
    rho2*omega == 0 the reference divides by zero (numpy scalars: NaN from
    then on), here the solve stops with the last finite iterate."""

    def __init__(self, g, rho, dt, tol, maxit):
        self.g, self.rho, self.dt = g, rho, dt
        self.tol, self.maxit = tol, maxit
        self.lap = laplace_coefficients(g)
        self.pp = g.padded(); self.s = g.padded()
        self.vv = g.interior_zeros(); self.t = g.interior_zeros()
        self.res = g.interior_zeros(); self.res0 = g.interior_zeros()
        self.history = []
        self.iterations = 0

    def solve(self, fDP, fp, source):
        g = self.g
        inner = _ii(g)
        fDP.var.zero_()                                     # :202
        source *= self.rho / self.dt                        # :205
        source *= -1.0 * g.vol                              # :208
        self.res.copy_(source); self.res0.copy_(self.res)   # :211-212
        max_res0 = torch.max(torch.abs(self.res)).item()    # :213
        self.vv.zero_(); self.pp.zero_()
        rho1 = 1.0; rho2 = 1.0; alpha = 0.0; omega = 1.0
        it = 0; done = False; max_res = 0.0
        self.history = []
        self.status = 'ok'
        while not done:                                     # :229-290
            rho1 = torch.sum(self.res * self.res0).item()
            if rho2 * omega == 0.0:
                # the reference's numpy scalars give beta = x/0 here and NaN
                # iterates from then on; stop at the last finite iterate
                self.status = 'breakdown_rho'
                break
            beta = alpha * rho1 / (rho2 * omega)
            pin = self.pp[inner]
            pin.sub_(self.vv, alpha=omega)
            pin.mul_(beta)
            pin.add_(self.res)
            apply_operator(g, self.lap, self.pp, self.vv)
            gamma = torch.sum(self.vv * self.res0).item()
            if gamma == 0.0:
                break
            alpha = rho1 / gamma
            fDP.var.add_(self.pp, alpha=alpha)
            sin = self.s[inner]
            sin.copy_(self.res)
            sin.sub_(self.vv, alpha=alpha)
            apply_operator(g, self.lap, self.s, self.t)
            buf1 = torch.sum(self.s[inner] * self.t).item()
            buf2 = torch.sum(self.t * self.t).item()
            if buf2 == 0:
                break
            omega = buf1 / buf2
            fDP.var.add_(self.s, alpha=omega)
            self.res.copy_(self.s[inner])
            self.res.sub_(self.t, alpha=omega)
            rho2 = rho1
            max_res = torch.max(torch.abs(self.res)).item()
            max_res /= max_res0
            self.history.append(max_res)
            it += 1
            if it >= self.maxit:
                done = True
            if max_res <= self.tol:
                done = True
        self.iterations = it
        fDP.update_border()                                 # :293
        fp.var.add_(fDP.var)                                # :294
        return max_res


class Jacobi:
    """solver_jacobi, pressure.py:92-161.  The pressure correction uses
    state_pOld (iterate N-1), and the reported residual mixes the vol-scaled
    operator with the unscaled source (pressure.py:153-158): both kept."""

    def __init__(self, g, rho, dt, nit):
        self.g, self.rho, self.dt, self.nit = g, rho, dt, nit
        self.lap = laplace_coefficients(g)
        self.res_P = g.interior_zeros()

    def solve(self, fpOld, fp, source):
        g = self.g
        fp.var.zero_()
        source *= self.rho / self.dt
        self.res_P.copy_(source)
        max_res0 = torch.max(torch.abs(self.res_P)).item()
        DInv = -1.0 / 6.0
        source *= g.dx ** 2
        for _ in range(self.nit):
            fpOld.update(fp.var[_ii(g)])
            po = fpOld.var
            fp.interior().copy_((-po[_ii(g, di=1)] - po[_ii(g, di=-1)]
                                 - po[_ii(g, dj=1)] - po[_ii(g, dj=-1)]
                                 - po[_ii(g, dk=1)] - po[_ii(g, dk=-1)]
                                 + source) * DInv)
        fp.update_border()
        source /= g.dx ** 2
        apply_operator(g, self.lap, fp.var, self.res_P)
        self.res_P.mul_(-1.0)
        self.res_P.add_(source)
        max_resP = torch.max(torch.abs(self.res_P)).item()
        return max_resP / max_res0


# ---------------------------------------------------------------------------
# Domain step (data/domain.py:342-402) and run-loop diagnostics
# ---------------------------------------------------------------------------
class Box:
    """The state of Domain (domain.py:56-336) for a periodic box with constant
    viscosity, the RK4 or Euler advancer and the BiCGStab or Jacobi solver."""

    def __init__(self, g, u, v, w, mu=0.01, rho=1.0, dt=1e-3, advancer='RK4',
                 solver='bicgstab', tol=1e-10, maxit=3000, sfs=None, p=None):
        self.g = g
        self.mu, self.rho, self.dt = mu, rho, dt
        self.u = Field(g, u); self.v = Field(g, v); self.w = Field(g, w)
        zeros = torch.zeros(g.nx, g.ny, g.nz, dtype=torch.float64)
        self.p = Field(g, zeros if p is None else p)
        self.DP = Field(g, zeros)
        self.ut = Field(g, u); self.vt = Field(g, v); self.wt = Field(g, w)
        self.source = g.interior_zeros()
        self.VISC = torch.ones(g.nxo, g.nyo, g.nzo, dtype=g.dtype)   # domain.py:130-131
        self.VISC.mul_(mu)
        self.advancer = advancer
        self.sfs = sfs
        nrhs = 4 if advancer == 'RK4' else 1
        self.rhss = [NavierStokesRHS(g, rho, self.VISC, sfs) for _ in range(nrhs)]
        if solver == 'bicgstab':
            self.poisson = BiCGStab(g, rho, dt, tol, maxit)
        elif solver == 'Jacobi':
            self.poisson = Jacobi(g, rho, dt, maxit)
        else:
            self.poisson = None
        self.max_resP = 0.0

    def forward_step(self):
        """domain.py:342-367"""
        if self.sfs is not None and self.sfs.modelType == 'eddyVisc':
            muEddy = self.sfs.eddyVisc(self.u, self.v, self.w, self.rho)
            self.VISC.copy_(self.mu + muEddy)                # domain.py:348-352
        if self.sfs is not None and self.sfs.modelType == 'tensor':
            self.sfs.update(self.u, self.v, self.w)
        if self.advancer == 'RK4':
            rk4_step(self.u, self.v, self.w, self.ut, self.vt, self.wt, self.rhss, self.dt)
        else:
            euler_step(self.u, self.v, self.w, self.rhss[0], self.dt)
        if self.poisson is not None:
            self.vel_corr()

    def vel_corr(self):
        """domain.py:373-402"""
        g = self.g
        div_vel(g, self.u, self.v, self.w, self.source)
        self.max_resP = self.poisson.solve(self.DP, self.p, self.source)
        grad_P(g, self.DP)
        self.u.vel_corr(self.DP.grad_x, self.dt / self.rho)
        self.v.vel_corr(self.DP.grad_y, self.dt / self.rho)
        self.w.vel_corr(self.DP.grad_z, self.dt / self.rho)

    def stats(self):
        """Per-step monitor values, core/PyFlow.py:378-395 (max|U|,|V|,|W|
        over the interior, TKE, max|div|, the solver residual)."""
        g = self.g
        ui, vi, wi = self.u.interior(), self.v.interior(), self.w.interior()
        maxU = torch.max(torch.abs(ui)).item()
        maxV = torch.max(torch.abs(vi)).item()
        maxW = torch.max(torch.abs(wi)).item()
        npts = float(g.nx * g.ny * g.nz)
        TKE = torch.sum(ui ** 2 + vi ** 2 + wi ** 2).item() / npts * 0.5
        div_vel(g, self.u, self.v, self.w, self.source)
        maxDiv = torch.max(torch.abs(self.source)).item()
        return dict(maxU=maxU, maxV=maxV, maxW=maxW, TKE=TKE, maxDiv=maxDiv,
                    max_resP=self.max_resP)
