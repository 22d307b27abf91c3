"""NGA file I/O (host only).  The two config files are the NGA grid fixtures
shipped with the reference (examples/*/restart/config_*), copied under
tests/golden/; their values are checked against what their headers encode.
Restart files round-trip through the writer and both readers."""

import os
import types

import numpy as np

from pyflow_amd import dataReader as dr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_config_dnsbox_128():
    x, y, z, xp, yp, zp = dr.readNGAconfig(os.path.join(GOLD, 'config_dnsbox_128_Lx0.0056'))
    assert (len(x), len(y), len(z)) == (129, 129, 129)
    assert (xp, yp, zp) == (1, 1, 1)
    assert x[0] == 0.0 and abs(x[-1] - 0.0056) < 1e-15
    assert np.allclose(np.diff(x), 4.375e-5, rtol=0, atol=1e-15)


def test_config_dnsbox_64():
    x, y, z, xp, yp, zp = dr.readNGAconfig(
        os.path.join(GOLD, 'config_dnsbox_1024_Lx0.045_NR_Delta16_Down16_0000'))
    assert (len(x), len(y), len(z)) == (65, 65, 65)
    assert (xp, yp, zp) == (1, 1, 1)
    assert np.allclose(np.diff(x), 7.03125e-4, rtol=0, atol=1e-14)
    assert abs(x[0] - 2.197265625e-05) < 1e-12


class HostData:
    """Stand-in for state.data_all_CPU with host arrays."""

    def __init__(self, arrays, names, time, dt):
        self.nx, self.ny, self.nz = arrays[0].shape
        self.nx_, self.ny_, self.nz_ = arrays[0].shape
        self.names = names
        self.nvar = len(names)
        self.ivar_read_start = 0
        self.nvar_read = len(arrays)
        self.time, self.dt = time, dt
        self.arrays = arrays

    def read(self, ivar):
        return self.arrays[ivar]

    def append(self, ivar, a):
        self.arrays[ivar][...] = a


def test_restart_round_trip(tmp_path):
    rng = np.random.Generator(np.random.PCG64(0))
    arrays = [rng.standard_normal((6, 5, 4)) for _ in range(4)]
    D = HostData(arrays, ['U', 'V', 'W', 'P'], 0.125, 1e-3)
    fn = str(tmp_path / 'restart_test')
    dr.writeNGArestart(fn, D, True)
    dr.writeNGArestart_parallel(fn, D)
    names, time, data = dr.readNGArestart(fn, headerOnly=False, printOut=False)
    assert names == ['U', 'V', 'W', 'P'] and time == 0.125
    for i in range(4):
        assert np.array_equal(data[..., i], arrays[i])
    # Fortran order on disk: the first block starts with u[0,0,0], u[1,0,0], ...
    raw = np.fromfile(fn, dtype='<f8', offset=16 + 4 * 8 + 16, count=3)
    assert np.array_equal(raw, arrays[0][:3, 0, 0])
    back = HostData([np.zeros((6, 5, 4)) for _ in range(4)], D.names, 0.0, 0.0)
    dr.readNGArestart_parallel(fn, back)
    for i in range(4):
        assert np.array_equal(back.arrays[i], arrays[i])
    # serial full write equals header + parallel write
    fn2 = str(tmp_path / 'restart_full')
    dr.writeNGArestart(fn2, D, False)
    assert open(fn, 'rb').read() == open(fn2, 'rb').read()
