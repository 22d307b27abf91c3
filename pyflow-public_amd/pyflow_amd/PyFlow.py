"""core/PyFlow.py:69-464: the forward run loop on the GPU.

run(inputConfig) keeps the reference's signature and behaviour for forward
runs (set-up, per-step monitor line on stdout and in
monitor/velocity_<caseName>.txt, NGA restart output every numItDataOut
steps when fNameOut is set).  simulate(inputConfig) is the same loop
returning the Domain and the per-step monitor values for programmatic use.
"""

import os
import sys
import time

import numpy as np

from . import domain
from . import geometry as geo
from . import initial_conditions
from . import metric_staggered
from . import parallel

HEAD = "\n  {:10s}   {:9s}   {:9s}   {:9s}   {:9s}   {:9s}   {:9s}   {:9s}   {:9s}"
LINE = "  {:10d}   {:8.3E}   {:8.3E}   {:8.3E}   {:8.3E}   {:8.3E}   {:8.3E}   {:8.3E}   {:8.3E}"


class Result:
    def __init__(self, D, geometry, history, wall):
        self.domain = D
        self.geometry = geometry
        self.history = history
        self.wall = wall


def setup(inputConfig):
    """core/PyFlow.py:75-125: comms, grid, decomposition, initial condition,
    geometry, metric and Domain."""
    comms = parallel.comms(inputConfig)
    config = geo.config(inputConfig)
    decomp = parallel.decomp(inputConfig, config)
    names, startTime, data_IC, xGrid, yGrid, zGrid = initial_conditions.generate(inputConfig, config,
                                                                                 decomp)
    geometry = geo.Geometry_Base(inputConfig, xGrid, yGrid, zGrid, decomp)
    metric = metric_staggered.metric_uniform(geometry)
    D = domain.Domain(inputConfig, comms, decomp, data_IC, geometry, metric, names, startTime)
    if inputConfig.configName == 'restart' and getattr(inputConfig, 'dataFileType', '') == 'restart':
        from . import dataReader as dr
        dr.readNGArestart_parallel(inputConfig.dataFileStr, D.data_all_CPU)
    if getattr(inputConfig, 'adjointVerification', False):
        nn = geometry.imin_ + geometry.nx_ // 2                      # PyFlow.py:155-156
        D.state_u_P.var[nn, nn, nn] += inputConfig.perturbation
    return comms, decomp, geometry, metric, D, startTime


def _write_restart(inputConfig, D, simTime):
    fname = getattr(inputConfig, 'fNameOut', None)
    if not fname:
        return
    from . import dataReader as dr
    timeStr = "{:12.7E}".format(simTime)
    dr.writeNGArestart(fname + '_' + timeStr, D.data_all_CPU, True)
    dr.writeNGArestart_parallel(fname + '_' + timeStr, D.data_all_CPU)


def simulate(inputConfig, verbose=True):
    comms, decomp, geometry, metric, D, startTime = setup(inputConfig)
    simDt = inputConfig.simDt
    simTime = startTime
    stopTime = startTime + inputConfig.numIt * simDt
    for s in (D.state_u_P, D.state_v_P, D.state_w_P, D.state_p_P):   # PyFlow.py:204-207
        s.update_border()
    _write_restart(inputConfig, D, simTime)
    viscNu = inputConfig.mu / inputConfig.rho

    def cfl(st):
        c = max((st['maxU'] / geometry.dx, st['maxV'] / geometry.dy, st['maxW'] / geometry.dz)) * simDt
        vnn = max((viscNu / geometry.dx ** 2, viscNu / geometry.dy ** 2, viscNu / geometry.dz ** 2)) * simDt
        return max((c, vnn))

    st = D.stats()
    history = [dict(step=0, time=simTime, CFL=cfl(st), **st)]
    caseName = getattr(inputConfig, 'caseName', 'PyFlow')
    monitor = None
    if getattr(inputConfig, 'writeMonitor', True):
        os.makedirs('monitor', exist_ok=True)
        monitor = open(os.path.join('monitor', 'velocity_' + caseName + '.txt'), 'w')
        monitor.write((HEAD[1:] + " \n").format("Step", "Time", "max CFL", "max U", "max V",
                                                 "max W", "TKE", "divergence", "max res_P"))
    if verbose:
        print(HEAD.format("Step", "Time", "max CFL", "max U", "max V", "max W", "TKE",
                          "divergence", "max res_P"))
    t0 = time.time()
    itCount = 0
    numItDataOut = getattr(inputConfig, 'numItDataOut', 0) or 0
    while simTime < stopTime and itCount < inputConfig.numIt:
        D.forwardStep(simDt)
        itCount += 1
        simTime += simDt
        st = D.stats()
        rec = dict(step=itCount, time=simTime, CFL=cfl(st), **st)
        rec['iters'] = getattr(D.poisson, 'iterations', None) if D.doPressure else None
        history.append(rec)
        line = LINE.format(itCount, simTime, rec['CFL'], st['maxU'], st['maxV'], st['maxW'],
                           st['TKE'], st['maxDiv'], st['max_resP'])
        if verbose:
            print(line)
        if monitor:
            monitor.write(line + "\n")
        if numItDataOut and itCount % numItDataOut == 0:
            _write_restart(inputConfig, D, simTime)
    wall = time.time() - t0
    if monitor:
        monitor.close()
    return Result(D, geometry, history, wall)


def run(inputConfig):
    """core/PyFlow.py:69: forward run, returns None like the reference."""
    simulate(inputConfig)
    return None


if __name__ == '__main__':
    sys.exit('import pyflow_amd.PyFlow and call run(inputConfig)')
