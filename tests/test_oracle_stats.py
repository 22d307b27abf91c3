"""Closed forms that pin the oracle's run-loop monitor values (Box.stats,
core/PyFlow.py:378-395 with state.data_all_CPU.absmax, data/state.py:90-92,
and metric.div_vel, metric/metric_staggered.py:438-453).

Field: u = A cos(2 pi i/nx), v = B cos(2 pi j/ny), w = C cos(2 pi k/nz) on
the interior indices, with nx, ny, nz = 2 (mod 4) so that every extremum below
lies on a grid point:

* max|u| = |A| (i = 0), likewise |B|, |C|;
* TKE = 1/2 mean(u^2 + v^2 + w^2) = (A^2 + B^2 + C^2)/4 (the mean of cos^2
  over a full period of n >= 3 points is 1/2);
* div = a(i) + b(j) + c(k) with a(i) = (u[i+1] - u[i])/dx
  = -2 A sin(pi/nx) sin(2 pi (i + 1/2)/nx)/dx, whose extremes +-2|A|sin(pi/nx)/dx
  sit at i + 1/2 = nx/4 and 3nx/4; the three terms reach their maxima at one
  cell, so max|div| = sum over axes of 2|A_d| sin(pi/n_d)/h_d.

The overlap cells that the reference never reads (the two near planes and
the outermost far plane of each axis) hold junk, so a reduction over the
padded array instead of the interior fails the maxima and the TKE, and the
planted mistake below shows it."""

import math

import pytest
import torch

from oracle import pyflow_oracle as O

NX, NY, NZ = 10, 14, 18
LX, LY, LZ = 2 * math.pi, 1.5 * math.pi, 2.5 * math.pi
A, B, C = 1.3, -0.7, 2.1
JUNK = 99.0


def closed_form_box():
    g = O.Grid(NX, NY, NZ, LX, LY, LZ)
    i = torch.arange(NX, dtype=torch.float64).view(-1, 1, 1)
    j = torch.arange(NY, dtype=torch.float64).view(1, -1, 1)
    k = torch.arange(NZ, dtype=torch.float64).view(1, 1, -1)
    u = (A * torch.cos(2 * math.pi * i / NX)).expand(NX, NY, NZ).clone()
    v = (B * torch.cos(2 * math.pi * j / NY)).expand(NX, NY, NZ).clone()
    w = (C * torch.cos(2 * math.pi * k / NZ)).expand(NX, NY, NZ).clone()
    box = O.Box(g, u, v, w, solver='none')
    for f in (box.u, box.v, box.w):
        a = f.var
        for ax in range(3):
            idx = [slice(None)] * 3
            for plane in (0, 1, a.shape[ax] - 1):
                idx[ax] = plane
                a[tuple(idx)] = JUNK
    return g, box


def expected():
    div = (2 * abs(A) * math.sin(math.pi / NX) / (LX / NX)
           + 2 * abs(B) * math.sin(math.pi / NY) / (LY / NY)
           + 2 * abs(C) * math.sin(math.pi / NZ) / (LZ / NZ))
    return dict(maxU=abs(A), maxV=abs(B), maxW=abs(C), TKE=(A * A + B * B + C * C) / 4.0,
                maxDiv=div)


def check(st):
    e = expected()
    assert st['maxU'] == e['maxU'] and st['maxV'] == e['maxV'] and st['maxW'] == e['maxW']
    assert abs(st['TKE'] - e['TKE']) <= 1e-14 * e['TKE']
    assert abs(st['maxDiv'] - e['maxDiv']) <= 1e-12 * e['maxDiv']


def test_stats_closed_forms():
    g, box = closed_form_box()
    st = box.stats()
    print(st, expected())
    check(st)


def test_stats_closed_forms_fail_a_padded_reduction(monkeypatch):
    """Planted mistake: reductions over the padded array (overlap junk
    included) instead of the interior."""
    g, box = closed_form_box()
    monkeypatch.setattr(O.Field, 'interior', lambda self: self.var)
    with pytest.raises(AssertionError):
        check(box.stats())


def test_stats_closed_forms_fail_a_wrong_normalisation(monkeypatch):
    """Planted mistake: TKE normalised by the padded point count."""
    g, box = closed_form_box()
    g.nx, g.ny, g.nz = g.nxo, g.nyo, g.nzo
    st = box.stats()
    g.nx, g.ny, g.nz = NX, NY, NZ
    with pytest.raises(AssertionError):
        check(st)


def test_apply_operator_single_mode_eigenvalue():
    """Pressure_Base.applyOperator (solver/pressure.py:54-84) on one Fourier
    mode: B = -vol * lambda * A with lambda = -sum_d (4/h_d^2) sin^2(k_d h_d/2),
    the 7-point Laplacian's eigenvalue (a sign or scale mistake in any axis
    fails it)."""
    g = O.Grid(NX, NY, NZ, LX, LY, LZ)
    m = (2, 3, 5)
    i = torch.arange(NX, dtype=torch.float64).view(-1, 1, 1)
    j = torch.arange(NY, dtype=torch.float64).view(1, -1, 1)
    k = torch.arange(NZ, dtype=torch.float64).view(1, 1, -1)
    ph = 2 * math.pi * (m[0] * i / NX + m[1] * j / NY + m[2] * k / NZ)
    a = torch.cos(ph + 0.3).expand(NX, NY, NZ).clone()
    Apad = g.from_interior(a)
    # halo junk: applyOperator refills the overlap itself
    Apad[0] = JUNK
    B = g.interior_zeros()
    O.apply_operator(g, O.laplace_coefficients(g), Apad, B)
    hs = (g.dx, g.dy, g.dz)
    ns = (NX, NY, NZ)
    lam = -sum(4.0 / h ** 2 * math.sin(math.pi * md / n) ** 2 for md, n, h in zip(m, ns, hs))
    ref = -g.vol * lam * a
    err = float((B - ref).abs().max() / ref.abs().max())
    print('eigenvalue check: max rel err %.2e' % err)
    assert err <= 1e-12
