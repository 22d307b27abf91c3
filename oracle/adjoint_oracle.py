"""CPU restatement of PyFlow's adjoint step (the reverse pass of closure
training) for a single periodic box.

TEST INFRASTRUCTURE ONLY, like pyflow_oracle: only the tests and
``__graft_entry__.smoke()`` may import it; the product never calls it.

Restated from the reference (paths relative to its ``src/``):

* the 15 discrete-exact adjoint convective terms ``adj_conv_*``
  (metric/metric_staggered.py:529-800);
* ``rhs_adjPredictor.evaluate`` (solver/adjoint.py:94-298): the viscous
  operator of the forward RHS applied to the adjoint field, then the
  convective transposes in the reference's order, then the closure term:
  for the reference's 'source' ML closure (adjoint.py:230-286) its own
  recipe, autograd through pyflow_oracle.MLSource (source_closure_term);
  for the 'tensor' closures of this build the exact transpose of the
  closure's contribution to the forward RHS, computed with torch autograd
  through this module's forward restatement;
* ``Domain.adjointStep`` (data/domain.py:409-438): adjoint projection with
  +dt/rho, then the advancer on the adjoint RHS with the forward state held
  at the checkpoint;
* the adjoint loop of ``PyFlow.run`` (core/PyFlow.py:312-373, 471-582) and
  the adjoint-verification protocol (core/PyFlow.py:147-166, 475-483;
  core/run_PyFlow_adjoint_verification.py:125-207).

How each part is pinned: golden vectors written by importing the reference
(tests/golden/make_golden.py: rhs_adjPredictor.evaluate and
Domain.adjointStep, and the evaluation with the 'source' ML closure: the
RHS and the parameter gradients its g.backward() leaves), and the transpose
identity against the forward RHS, whose restatement is itself pinned
bit-exactly (tests/test_oracle_adjoint.py).
"""

import torch

from . import pyflow_oracle as O

_E = ((1, 0, 0), (0, 1, 0), (0, 0, 1))


def _add(a, b):
    return tuple(x + y for x, y in zip(a, b))


def _neg(a):
    return tuple(-x for x in a)


def _adj_pair(g, rhs, ah, vi, d, s2, coef, vsh):
    """The two subtractions of one adj_conv_* method.

    For s in (0, s2): rhs -= ((ah[m - e_d + s] - ah[m + s]) * vi[m + s + vsh]) * coef
    at every interior cell m, each product and difference rounded on its own
    (tmpv copy_, sub_, mul_, mul_, then rhs.sub_).
    """
    ed = _E[d]
    for s in ((0, 0, 0), s2):
        lo = _add(_neg(ed), s)
        hi = s
        vp = _add(s, vsh)
        t = ah[O._ii(g, *lo)].clone()
        t.sub_(ah[O._ii(g, *hi)])
        t.mul_(vi[O._ii(g, *vp)])
        t.mul_(coef)
        rhs.sub_(t)


# metric_staggered.py:537-556 / :560-577 / :581-598 (A11, A22, A33)
def adj_conv_xx(g, au, fu, rhs):
    _adj_pair(g, rhs, au.var, fu.var_i, 0, _E[0], g.dxi, _neg(_E[0]))


def adj_conv_yy(g, av, fv, rhs):
    _adj_pair(g, rhs, av.var, fv.var_i, 1, _E[1], g.dyi, _neg(_E[1]))


def adj_conv_zz(g, aw, fw, rhs):
    _adj_pair(g, rhs, aw.var, fw.var_i, 2, _E[2], g.dzi, _neg(_E[2]))


# metric_staggered.py:604-621 / :627-644 / :650-667 (x, y, z)
def adj_conv_x(g, a, f, rhs):
    _adj_pair(g, rhs, a.var, f.var_i, 0, _E[0], 0.5 * g.dxi, (0, 0, 0))


def adj_conv_y(g, a, f, rhs):
    _adj_pair(g, rhs, a.var, f.var_i, 1, _E[1], 0.5 * g.dyi, (0, 0, 0))


def adj_conv_z(g, a, f, rhs):
    _adj_pair(g, rhs, a.var, f.var_i, 2, _E[2], 0.5 * g.dzi, (0, 0, 0))


# metric_staggered.py:671-688 / :692-709 (1d, 1e)
def adj_conv_1d(g, av, fv, rhs):
    _adj_pair(g, rhs, av.var, fv.var_i, 0, _E[1], 0.5 * g.dxi, (0, 0, 0))


def adj_conv_1e(g, aw, fw, rhs):
    _adj_pair(g, rhs, aw.var, fw.var_i, 0, _E[2], 0.5 * g.dxi, (0, 0, 0))


# metric_staggered.py:713-730 / :734-751 (2d, 2e)
def adj_conv_2d(g, au, fu, rhs):
    _adj_pair(g, rhs, au.var, fu.var_i, 1, _E[0], 0.5 * g.dyi, (0, 0, 0))


def adj_conv_2e(g, aw, fw, rhs):
    _adj_pair(g, rhs, aw.var, fw.var_i, 1, _E[2], 0.5 * g.dyi, (0, 0, 0))


# metric_staggered.py:755-772 / :776-793 (3d, 3e)
def adj_conv_3d(g, au, fu, rhs):
    _adj_pair(g, rhs, au.var, fu.var_i, 2, _E[0], 0.5 * g.dzi, (0, 0, 0))


def adj_conv_3e(g, av, fv, rhs):
    _adj_pair(g, rhs, av.var, fv.var_i, 2, _E[1], 0.5 * g.dzi, (0, 0, 0))


# ---------------------------------------------------------------------------
# Closure term for 'tensor' closures: exact transpose by autograd
# ---------------------------------------------------------------------------
def _periodic_padded(x):
    """Padded array holding the interior x with periodic overlap values, as a
    differentiable function of x (the same values halo_fill writes)."""
    p = O.NOVER
    return torch.nn.functional.pad(x[None, None], (p, p, p, p, p, p), mode='circular')[0, 0]


def closure_adjoint_term(g, rho, fu, fv, fw, au, av, aw, closure):
    """delta_c = d/du_c [ sum over interior cells of R_sfs(tau(u)) . u_adj ],
    with R_sfs the tensor-closure flux divergence of the forward RHS
    (velocity.py:262-315) and tau the closure evaluated on the forward state
    (sfsmodel_gradient.py:70-148, or the MLP).  ``closure(fu, fv, fw)``
    returns the six stresses as padded tensors.  The derivative is taken by
    torch autograd through this restatement's in-place op sequence, with the
    overlap cells as functions of the interior (so their adjoints fold onto
    the interior images).  Returns three interior arrays."""
    xs = [f.interior().detach().clone().requires_grad_(True) for f in (fu, fv, fw)]
    F = [O.Field(g) for _ in range(3)]
    for f, x in zip(F, xs):
        f.var = _periodic_padded(x)
    stresses = closure(*F)

    class _T:
        modelType = 'tensor'
    sfs = _T()
    for n, s in zip(('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ'), stresses):
        setattr(sfs, n, s)
    VISC0 = torch.zeros(g.nxo, g.nyo, g.nzo, dtype=g.dtype)
    r = O.NavierStokesRHS(g, rho, VISC0, sfs)
    r.rhs_u.zero_(); r.rhs_v.zero_(); r.rhs_w.zero_()
    # with VISC = 0 the viscous fluxes vanish and only the closure part remains
    Z = [O.Field(g) for _ in range(3)]
    r.viscous(*Z, True)
    L = (torch.sum(r.rhs_u * au.interior()) + torch.sum(r.rhs_v * av.interior())
         + torch.sum(r.rhs_w * aw.interior()))
    gr = torch.autograd.grad(L, xs)
    return [t.detach() for t in gr]


def source_closure_term(g, closure, fu, fv, fw, au, av, aw, rhs, interp_SC):
    """The closure term of rhs_adjPredictor.evaluate for a 'source' closure
    (adjoint.py:230-286), run after the convective transposes: the forward
    state's centre interpolants (overwriting var_i[:-1]; the last plane keeps
    what the transposes' interpolations left), the closure on them with
    autograd, the adjoint state's centre interpolants, g = sum of G . u_adj
    over the padded arrays, g.backward(), then rhs -= the face interpolant of
    dg/d var_i (taken to float32 first, :268-270).  The closure's trainable
    weights accumulate their gradients, as the reference's model does."""
    O.interp_u_xm(fu); O.interp_v_ym(fv); O.interp_w_zm(fw)               # :235-237
    V = [f.var_i.detach().clone().requires_grad_(True) for f in (fu, fv, fw)]   # :241-243
    closure.update(*V, requires_grad=True)                                 # :247
    O.interp_u_xm(au); O.interp_v_ym(av); O.interp_w_zm(aw)                 # :250-252
    A = [f.var_i.detach() for f in (au, av, aw)]                           # :254-256
    gsum = (torch.sum(closure.GX[:, :, :, 0] * A[0]) + torch.sum(closure.GY[:, :, :, 0] * A[1])
            + torch.sum(closure.GZ[:, :, :, 0] * A[2]))                    # :261-263
    gsum.backward()                                                        # :268
    gml = [v.grad.to(torch.float32) for v in V]                            # :271-273
    core = (slice(g.imin, g.imax + 1), slice(g.jmin, g.jmax + 1), slice(g.kmin, g.kmax + 1))
    for q, fn in enumerate((O.interp_sc_x, O.interp_sc_y, O.interp_sc_z)):  # :278-286
        fn(g, gml[q], interp_SC)
        rhs[q].sub_(interp_SC[core])
    return gml


def gradient_closure(g):
    """The gradient model as a closure callable for closure_adjoint_term."""
    def f(fu, fv, fw):
        m = O.GradientModel(g)
        m.update(fu, fv, fw)
        return [m.FXX, m.FYY, m.FZZ, m.FXY, m.FXZ, m.FYZ]
    return f


def mlp_closure(g, params):
    def f(fu, fv, fw):
        return O.mlp_closure_forward(g, fu, fv, fw, params)
    return f


# ---------------------------------------------------------------------------
# rhs_adjPredictor (adjoint.py:94-298) and the adjoint step
# ---------------------------------------------------------------------------
class AdjointRHS:
    """rhs_adjPredictor.  ``fu, fv, fw`` are the forward-state Fields (held
    fixed during an adjoint step); ``closure`` is None, a callable as for
    closure_adjoint_term ('tensor' closures), or a 'source' closure object
    such as pyflow_oracle.MLSource (source_closure_term)."""

    def __init__(self, g, rho, VISC, fu, fv, fw, closure=None):
        self.g = g
        self.rho = rho
        self.fu, self.fv, self.fw = fu, fv, fw
        self.closure = closure
        self.ns = O.NavierStokesRHS(g, rho, VISC, None)   # viscous operator and buffers
        self.rhs_u, self.rhs_v, self.rhs_w = self.ns.rhs_u, self.ns.rhs_v, self.ns.rhs_w

    def evaluate(self, au, av, aw):
        g = self.g
        fu, fv, fw = self.fu, self.fv, self.fw
        self.rhs_u.zero_(); self.rhs_v.zero_(); self.rhs_w.zero_()       # :109-111
        self.ns.viscous(au, av, aw, False)                                # :117-167
        ru, rv, rw = self.rhs_u, self.rhs_v, self.rhs_w
        # :170-230
        O.interp_u_xm(fu); adj_conv_xx(g, au, fu, ru)                     # 1a
        O.interp_vw_x(fv); adj_conv_y(g, au, fv, ru)                      # 1b
        O.interp_vw_x(fw); adj_conv_z(g, au, fw, ru)                      # 1c
        adj_conv_1d(g, av, fv, ru)                                        # 1d
        adj_conv_1e(g, aw, fw, ru)                                        # 1e
        O.interp_uw_y(fu); adj_conv_x(g, av, fu, rv)                      # 2a
        O.interp_v_ym(fv); adj_conv_yy(g, av, fv, rv)                     # 2b
        O.interp_uw_y(fw); adj_conv_z(g, av, fw, rv)                      # 2c
        adj_conv_2d(g, au, fu, rv)                                        # 2d
        adj_conv_2e(g, aw, fw, rv)                                        # 2e
        O.interp_uv_z(fu); adj_conv_x(g, aw, fu, rw)                      # 3a
        O.interp_uv_z(fv); adj_conv_y(g, aw, fv, rw)                      # 3b
        O.interp_w_zm(fw); adj_conv_zz(g, aw, fw, rw)                     # 3c
        adj_conv_3d(g, au, fu, rw)                                        # 3d
        adj_conv_3e(g, av, fv, rw)                                        # 3e
        if getattr(self.closure, 'modelType', None) == 'source':
            source_closure_term(g, self.closure, fu, fv, fw, au, av, aw, (ru, rv, rw),
                                self.ns.interp_SC)
        elif self.closure is not None:
            d = closure_adjoint_term(g, self.rho, fu, fv, fw, au, av, aw, self.closure)
            ru.add_(d[0]); rv.add_(d[1]); rw.add_(d[2])


class AdjointBox:
    """The adjoint part of Domain (domain.py:135-191, 263-288, 409-438) on top
    of an oracle Box: adjoint fields, temporaries, checkpoints."""

    def __init__(self, box, closure=None):
        self.box = box
        g = box.g
        self.au = O.Field(g); self.av = O.Field(g); self.aw = O.Field(g)
        self.aut = O.Field(g); self.avt = O.Field(g); self.awt = O.Field(g)
        n = 4 if box.advancer == 'RK4' else 1
        self.rhss = [AdjointRHS(g, box.rho, box.VISC, box.u, box.v, box.w, closure)
                     for _ in range(n)]

    def adjoint_step(self):
        """domain.py:409-438"""
        b = self.box
        g = b.g
        if b.poisson is not None:
            O.div_vel(g, self.au, self.av, self.aw, b.source)
            b.max_resP = b.poisson.solve(b.DP, b.p, b.source)
            O.grad_P(g, b.DP)
            self.au.vel_corr(b.DP.grad_x, +b.dt / b.rho)
            self.av.vel_corr(b.DP.grad_y, +b.dt / b.rho)
            self.aw.vel_corr(b.DP.grad_z, +b.dt / b.rho)
        if b.advancer == 'RK4':
            O.rk4_step(self.au, self.av, self.aw, self.aut, self.avt, self.awt, self.rhss, b.dt)
        else:
            O.euler_step(self.au, self.av, self.aw, self.rhss[0], b.dt)


def adjoint_verification_run(g, u, v, w, nsteps, perturbation=0.0, adjoint=True, closure=None,
                             **box_kw):
    """One PyFlow.run of the adjoint-verification protocol: perturb u at
    padded index nn = imin + nx//2 on all three axes (PyFlow.py:155-157, the
    overlap is not refreshed), nsteps forward steps with checkpoints
    (PyFlow.py:333-367), objective mean(u_interior**2) (PyFlow.py:476-477),
    adjoint seeded with 2u/numPoints (PyFlow.py:480-482, 528-530), nsteps
    adjoint steps each at the forward state of the start of its step
    (PyFlow.py:548-555).  Returns (objective, u_adj[nn, nn, nn])."""
    b = O.Box(g, u, v, w, **box_kw)
    nn = g.imin + g.nx // 2
    b.u.var[nn, nn, nn] += perturbation
    checks = [(b.u.var.clone(), b.v.var.clone(), b.w.var.clone())]
    for _ in range(nsteps):
        b.forward_step()
        checks.append((b.u.var.clone(), b.v.var.clone(), b.w.var.clone()))
    obj = torch.mean(b.u.interior() ** 2).item()
    if not adjoint:
        return obj, None
    A = AdjointBox(b, closure)
    A.au.var.copy_(2.0 * (b.u.var - 0.0))
    A.av.var.copy_(0.0 * (b.v.var - 0.0))
    A.aw.var.copy_(0.0 * (b.w.var - 0.0))
    npts = float(g.nx * g.ny * g.nz)
    A.au.var.div_(npts); A.av.var.div_(npts); A.aw.var.div_(npts)
    for it in range(nsteps, 0, -1):
        cu, cv, cw = checks[it - 1]
        b.u.var.copy_(cu); b.v.var.copy_(cv); b.w.var.copy_(cw)
        A.adjoint_step()
    return obj, A.au.var[nn, nn, nn].item()


# ---------------------------------------------------------------------------
# Differentiable forward step (the checker of the exact reverse pass)
# ---------------------------------------------------------------------------
def fft_project(g, u, v, w):
    """u - G L^-1 D u on interior arrays, with D = div_vel (forward
    differences, metric_staggered.py:438-453), G = grad_P (backward
    differences, :390-401) and L = D G the 7-point Laplacian, solved exactly
    through its Fourier symbol -(4/h^2) sin^2(k h/2) per axis (the mean of
    D u is zero).  This is the projection of Domain.velCorr
    (domain.py:373-402) with an exact Poisson solve, differentiable by
    autograd."""
    dxi, dyi, dzi = g.dxi, g.dyi, g.dzi
    D = ((torch.roll(u, -1, 0) - u) * dxi + (torch.roll(v, -1, 1) - v) * dyi
         + (torch.roll(w, -1, 2) - w) * dzi)
    def sym(n, hi):
        k = 2.0 * torch.pi * torch.fft.fftfreq(n, dtype=torch.float64)
        return -4.0 * hi * hi * torch.sin(0.5 * k) ** 2
    L = (sym(g.nx, dxi)[:, None, None] + sym(g.ny, dyi)[None, :, None]
         + sym(g.nz, dzi)[None, None, :])
    L[0, 0, 0] = 1.0
    ph = torch.fft.fftn(D) / L
    ph[0, 0, 0] = 0.0
    phi = torch.fft.ifftn(ph).real
    return (u - (phi - torch.roll(phi, 1, 0)) * dxi,
            v - (phi - torch.roll(phi, 1, 1)) * dyi,
            w - (phi - torch.roll(phi, 1, 2)) * dzi)


def functional_step(g, u, v, w, mu, rho, dt, closure=None, project=True):
    """One Domain.forwardStep (domain.py:342-402) as a differentiable
    function of the interior velocity: closure stresses from u^n (a callable
    as for closure_adjoint_term, or None), the reference's RK4
    (advancer.py:75-113, stage 4 from u^n + dt/2*k3), then the projection
    with an exact solve.  Returns interior arrays."""
    VISC = torch.ones(g.nxo, g.nyo, g.nzo, dtype=g.dtype) * mu

    def fields(x):
        F = [O.Field(g) for _ in range(3)]
        for f, t in zip(F, x):
            f.var = _periodic_padded(t)
        return F

    class _T:
        modelType = 'tensor'
    sfs = None
    if closure is not None:
        sfs = _T()
        with torch.autograd.graph.saved_tensors_hooks(_pack, _unpack):
            taus = closure(*fields((u, v, w)))
        for n, s in zip(('FXX', 'FYY', 'FZZ', 'FXY', 'FXZ', 'FYZ'), taus):
            setattr(sfs, n, s)

    def R(x):
        r = O.NavierStokesRHS(g, rho, VISC, sfs)
        r.evaluate(*fields(x))
        return r.rhs_u, r.rhs_v, r.rhs_w

    U = (u, v, w)
    with torch.autograd.graph.saved_tensors_hooks(_pack, _unpack):
        return _rk4_project(g, U, R, dt, project)


# The restated RHS reuses its var_i / flux buffers in place, as the reference
# does; autograd must see each saved operand as it was when it was used.
def _pack(t):
    return t.clone()


def _unpack(t):
    return t


def _rk4_project(g, U, R, dt, project):
    k1 = R(U)
    k2 = R(tuple(0.5 * dt * k + x for k, x in zip(k1, U)))
    k3 = R(tuple(0.5 * dt * k + x for k, x in zip(k2, U)))
    k4 = R(tuple(0.5 * dt * k + x for k, x in zip(k3, U)))
    us = tuple(x + dt / 6.0 * (a + 2.0 * b + 2.0 * c + d)
               for x, a, b, c, d in zip(U, k1, k2, k3, k4))
    return fft_project(g, *us) if project else us
