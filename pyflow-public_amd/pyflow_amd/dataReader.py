"""NGA config/restart files (data/dataReader.py:60-390) for one rank.

Formats, as the reference reads and writes them (little endian):

* config: 64-byte name, int32 icyl, xper, yper, zper, nx, ny, nz, then the
  x, y, z grid-point coordinates (nx+1, ny+1, nz+1 float64), then a mask
  (not read).
* restart: int32 nx, ny, nz, nvar; float64 dt, time; nvar 8-character
  names; then one nx*ny*nz float64 block per variable in Fortran order.

The reference's *_parallel variants use MPI-IO subarray views; on one rank
the view is the whole array, so they reduce to plain reads and writes of the
same bytes.
"""

import struct

import numpy as np


def readNGAconfig(fName, verbose=False):
    """dataReader.py:60-112"""
    with open(fName, 'rb') as f:
        raw = f.read(64)
        configName = ''.join(chr(c) for c in raw if not chr(c).isspace())
        icyl, xper, yper, zper, nx, ny, nz = struct.unpack('<7i', f.read(28))
        xGrid = np.fromfile(f, dtype='<f8', count=nx + 1)
        yGrid = np.fromfile(f, dtype='<f8', count=ny + 1)
        zGrid = np.fromfile(f, dtype='<f8', count=nz + 1)
    if verbose:
        print(' --> Importing grid data from NGA config file')
        print('   Config file name:   {}'.format(fName))
        print('   Configuration name: {}'.format(configName))
        print('   xper,yper,zper:     {},{},{}'.format(xper, yper, zper))
        print('   Grid size:         nx={}, ny={}, nz={}'.format(nx, ny, nz))
    return xGrid, yGrid, zGrid, xper, yper, zper


def writeNGAconfig(fName, xGrid, yGrid, zGrid, configName='PyFlow', xper=1, yper=1, zper=1,
                   icyl=0):
    """The config layout readNGAconfig reads (the reference only reads it):
    name, flags, sizes, grid points, then an all-zero int32 nx*ny mask."""
    nx, ny, nz = len(xGrid) - 1, len(yGrid) - 1, len(zGrid) - 1
    with open(fName, 'wb') as f:
        f.write(configName.ljust(64)[:64].encode('UTF-8'))
        f.write(struct.pack('<7i', icyl, xper, yper, zper, nx, ny, nz))
        for g in (xGrid, yGrid, zGrid):
            f.write(np.asarray(g, dtype='<f8').tobytes())
        f.write(np.zeros(nx * ny, dtype='<i4').tobytes())


def _read_header(f):
    nx, ny, nz, nvar = struct.unpack('<4i', f.read(16))
    dt, time = struct.unpack('<2d', f.read(16))
    raw = f.read(8 * nvar)
    names = [''.join(chr(c) for c in raw[8 * i:8 * (i + 1)] if not chr(c).isspace())
             for i in range(nvar)]
    return nx, ny, nz, nvar, dt, time, names


def readNGArestart(fName, headerOnly=True, printOut=True):
    """dataReader.py:118-169: (names, time) or (names, time, data[nx,ny,nz,nvar])."""
    with open(fName, 'rb') as f:
        nx, ny, nz, nvar, dt, time, names = _read_header(f)
        if printOut:
            print(' --> Importing state data from NGA restart file')
            print('   Data file name:    {}'.format(fName))
            print('   Data file at time: {:10.4E}'.format(time))
            print('   Timestep size:     {:10.4E}'.format(dt))
            print('   Number of vars:    {}'.format(nvar))
            print('   Variables in file: {}'.format(names))
        if headerOnly:
            return names, time
        data = np.empty((nx, ny, nz, nvar))
        for ivar in range(nvar):
            block = np.fromfile(f, dtype='<f8', count=nx * ny * nz)
            data[:, :, :, ivar] = block.reshape((nx, ny, nz), order='F')
        return names, time, data


def writeNGArestart(fName, Data, headerOnly=True):
    """dataReader.py:179-208: header (and, if asked, the data) of a restart file."""
    with open(fName, 'wb') as f:
        f.write(struct.pack('<4i', Data.nx, Data.ny, Data.nz, Data.nvar))
        f.write(struct.pack('<2d', Data.dt, Data.time))
        for name in Data.names:
            f.write(name.ljust(8)[:8].encode('UTF-8'))
        if not headerOnly:
            for ivar in range(Data.nvar):
                a = np.asarray(Data.read(ivar), dtype='<f8')
                f.write(a.reshape(-1, order='F').tobytes())


def _offset(Data, ivar):
    # 4 int32 + 2 float64 + names, then whole variables (dataReader.py:282-283)
    return 4 * 4 + Data.nvar * 8 + 2 * 8 + Data.nx * Data.ny * Data.nz * ivar * 8


def readNGArestart_parallel(fName, data):
    """dataReader.py:214-300 on one rank: variables ivar_read_start ..
    ivar_read_start+nvar_read into the state mirror `data`."""
    start = data.ivar_read_start or 0
    count = data.nvar_read if data.nvar_read is not None else data.nvar - start
    n = data.nx * data.ny * data.nz
    with open(fName, 'rb') as f:
        for ivar in range(start, start + count):
            f.seek(_offset(data, ivar))
            block = np.fromfile(f, dtype='<f8', count=n)
            data.append(ivar, block.reshape((data.nx_, data.ny_, data.nz_), order='F'))


def writeNGArestart_parallel(fName, data):
    """dataReader.py:306-390 on one rank: the data blocks after a header
    written by writeNGArestart."""
    start = data.ivar_read_start or 0
    count = data.nvar_read if data.nvar_read is not None else data.nvar - start
    with open(fName, 'r+b') as f:
        for ivar in range(start, start + count):
            f.seek(_offset(data, ivar))
            a = np.asarray(data.read(ivar), dtype='<f8')
            f.write(a.reshape(-1, order='F').tobytes())
