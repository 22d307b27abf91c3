"""Build libpyflow_hip.so for gfx950 with hipcc (no torch extension machinery:
the library exposes a plain C-ABI and links only the HIP runtime)."""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, 'csrc')
LIB_DIR = os.path.join(HERE, '_lib')
LIB_NAME = 'libpyflow_hip.so'
LIB_PATH = os.path.join(LIB_DIR, LIB_NAME)

SOURCES = ['pf_runtime.hip', 'pf_step.hip', 'pf_krylov.hip', 'pf_persist.hip', 'pf_ops.hip', 'pf_closure.hip',
           'pf_adjoint.hip', 'pf_batch.hip', 'pf_source.hip', 'pf_source_train.hip']

ARCH = os.environ.get('PYFLOW_HIP_ARCH', 'gfx950')

FLAGS = [
    '-O3', '-std=c++17', '-fPIC',
    # no contractions beyond the fma() calls written in the sources: the
    # kernels reproduce the reference's per-op rounding
    '-ffp-contract=off',
    # keep the compiler from forming packed FP32 ops
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-mcode-object-version=5',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
]


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found (set HIPCC)')


def build(verbose=False, jobs=None):
    """Compile every source to an object (in parallel) and link the library."""
    os.makedirs(LIB_DIR, exist_ok=True)
    objdir = os.path.join(PKG_ROOT, '_build')
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    procs = []
    objs = []
    for s in srcs:
        obj = os.path.join(objdir, s.replace('.hip', '.o'))
        objs.append(obj)
        cmd = [cc, '--offload-arch=' + ARCH, *FLAGS, '-I', CSRC, '-c',
               os.path.join(CSRC, s), '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors='replace')
        # hipcc's host pass does not know the packed-fp32 feature; drop that note
        lines = [ln for ln in txt.splitlines() if 'packed-fp32-ops' not in ln]
        if p.returncode != 0:
            failed.append(s)
            sys.stderr.write('\n'.join(lines) + '\n')
        elif verbose and lines:
            print('\n'.join(lines))
    if failed:
        raise RuntimeError('hipcc failed for: ' + ', '.join(failed))
    tmp = LIB_PATH + '.tmp'
    cmd = [cc, '--offload-arch=' + ARCH, '-shared', '-fPIC', *objs, '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
