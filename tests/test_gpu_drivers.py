"""The command-line drivers around the time step: decay curves
(pyflow_amd.decay) and filtered-DNS targets (pyflow_amd.les_targets)."""

import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_decay_curve_small(tmp_path):
    from pyflow_amd import decay
    out = str(tmp_path / 'decay.csv')
    assert decay.main(['--N', '32', '--steps', '4', '--nu', '0.01', '--out', out]) == 0
    rows = [l.strip().split(',') for l in open(out) if l[0].isdigit()]
    assert len(rows) == 5
    tke = [float(r[2]) for r in rows]
    eps = [float(r[3]) for r in rows]
    print('TKE', tke, 'eps', eps)
    # decaying turbulence: energy falls, and over a step by about dt*eps
    assert all(b < a for a, b in zip(tke, tke[1:]))
    dt = float(open(out).readline().split('dt=')[1].split()[0])
    for k in range(4):
        drop = tke[k] - tke[k + 1]
        assert 0.3 * dt * eps[k] < drop < 3.0 * dt * eps[k]
    assert all(r[7] == 'converged' for r in rows[1:])


def test_les_targets_cli(tmp_path):
    from pyflow_amd import les_targets, dataReader as dr
    out = str(tmp_path / 't')
    assert les_targets.main(['--out', out, '--N', '32', '--ratio', '4', '--steps', '2',
                             '--every', '1']) == 0
    x, y, z, *_ = dr.readNGAconfig(os.path.join(out, 'config'))
    assert len(x) == 9 and abs(x[-1] - 2 * math.pi) < 1e-14
    names, t, data = dr.readNGArestart(os.path.join(out, 'target_00000002'), headerOnly=False,
                                       printOut=False)
    assert names == ['U', 'V', 'W', 'P'] and data.shape == (8, 8, 8, 4) and t > 0
    assert np.isfinite(data).all() and np.abs(data[..., :3]).max() > 0
