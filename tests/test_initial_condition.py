"""The seeded Passot-Pouquet field (host numpy): discretely solenoidal on the
staggered grid, unit u_rms, reproducible from its seed, spectrum peaked at k0."""

import math

import numpy as np

from pyflow_amd.initial_conditions import passot_pouquet, cfl_dt


def div_h(u, v, w, h):
    return ((np.roll(u, -1, 0) - u) + (np.roll(v, -1, 1) - v) + (np.roll(w, -1, 2) - w)) / h


def test_solenoidal_and_normalised():
    N = 32
    u, v, w = passot_pouquet(N, k0=4.0, urms=1.0, seed=3)
    h = 2 * math.pi / N
    d = div_h(u, v, w, h)
    assert np.abs(d).max() < 1e-12 * np.abs(u).max() / h
    urms = math.sqrt((np.mean(u * u) + np.mean(v * v) + np.mean(w * w)) / 3)
    assert abs(urms - 1.0) < 1e-12


def test_seeded_and_distinct():
    a = passot_pouquet(16, seed=1)
    b = passot_pouquet(16, seed=1)
    c = passot_pouquet(16, seed=2)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])


def test_spectrum_peak_near_k0():
    N = 64
    u, v, w = passot_pouquet(N, k0=4.0, seed=0)
    E = np.zeros(N)
    k = np.fft.fftfreq(N, 1.0 / N)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing='ij')
    kk = np.rint(np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)).astype(int)
    for f in (u, v, w):
        F = np.abs(np.fft.fftn(f)) ** 2
        np.add.at(E, np.clip(kk, 0, N - 1).ravel(), F.ravel())
    # E(k) ~ k^4 exp(-2(k/k0)^2) peaks at k = k0
    assert 3 <= int(np.argmax(E[1:20])) + 1 <= 5


def test_cfl_dt():
    u, v, w = passot_pouquet(16, seed=0)
    h = 2 * math.pi / 16
    dt = cfl_dt(u, v, w, h)
    umax = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    assert abs(dt * umax / h - 0.3) < 1e-14
