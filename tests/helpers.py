"""Shared helpers for the tests: golden fixtures, set-up of the HIP-backed domain objects, error norms."""

import math
import os
import sys
import types

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, 'golden', 'reference_vectors.npz')


def golden():
    return np.load(GOLDEN)


def ml_source_params(seed=7):
    """Packed weights of the reference's ML closure network used by the golden
    vectors (tests/golden/source_stub.py), and the same as a state_dict."""
    d = os.path.join(HERE, 'golden')
    if d not in sys.path:
        sys.path.insert(0, d)
    from source_stub import ml_source_params as gen, ml_state_dict
    flat = gen(seed)
    return flat, ml_state_dict(flat)


def recording_source():
    """The stand-in 'source' closure the golden vectors were written with
    (tests/golden/source_stub.py)."""
    d = os.path.join(HERE, 'golden')
    if d not in sys.path:
        sys.path.insert(0, d)
    from source_stub import RecordingSource
    return RecordingSource()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def rel_max(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.abs(b).max()
    return float(np.abs(a - b).max() / (m if m > 0 else 1.0))


def rel_l2_mean_free(a, b, padded=True):
    """rel_l2 after removing each field's interior mean: the periodic pressure
    is defined up to a constant (the 7-point operator annihilates constants),
    so solvers that differ only in rounding can differ in that constant."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    core = (slice(2, -2),) * 3 if padded else (slice(None),) * 3
    return rel_l2(a - a[core].mean(), b - b[core].mean())


def gpu_env(shape, L=(2 * math.pi,) * 3, dtype=torch.float64):
    """(decomp, geometry, metric) of the GPU package for a periodic box."""
    from pyflow_amd import parallel, geometry, metric_staggered
    IC = types.SimpleNamespace(dtypeTorch=dtype,
                               dtypeNumpy=np.float64 if dtype == torch.float64 else np.float32,
                               nproc_x=1, nproc_y=1, nproc_z=1)
    cfg = types.SimpleNamespace(Nx=shape[0], Ny=shape[1], Nz=shape[2], isper=[1, 1, 1])
    d = parallel.decomp(IC, cfg)
    xg = np.linspace(0.0, L[0], shape[0] + 1)
    yg = np.linspace(0.0, L[1], shape[1] + 1)
    zg = np.linspace(0.0, L[2], shape[2] + 1)
    g = geometry.uniform(xg, yg, zg, d)
    m = metric_staggered.metric_uniform(g)
    return d, g, m


def cpu(t):
    return t.detach().to('cpu', torch.float64).numpy()
