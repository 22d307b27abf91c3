"""Continuous-batching server (pyflow_amd.serving) on the GPU.

* Batched steps equal the plain single-box path bit for bit (the batched
  kernels keep each box's block decomposition), FP32 and FP64.
* A request's result is the same bits alone and in a mix that it joins
  while other requests are already running.
* Admission holds the resident HBM footprint under the cap: with a cap of
  two slots, later arrivals wait and are admitted as earlier ones finish;
  the results kept from reused slots are those of a server with room for all.
"""

import numpy as np
import pytest
import torch

from helpers import cpu

pytestmark = pytest.mark.gpu


def _reqs(specs, tol=1e-8):
    from pyflow_amd.serving import Request
    return [Request(i, N, steps, seed, k0, nu, arrival=t, tol=tol, maxit=1500)
            for i, (N, steps, seed, k0, nu, t) in enumerate(specs)]


SPECS = [(32, 6, 11, 4.0, 0.01, 0.0), (48, 4, 12, 2.0, 0.005, 0.0),
         (64, 5, 13, 8.0, 0.02, 0.0), (32, 3, 14, 2.0, 0.002, 0.0)]


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64])
def test_batched_equals_single_box_path(dtype):
    from pyflow_amd.serving import BatchServer, SequentialRunner
    tol = 1e-8 if dtype == torch.float32 else 1e-11
    a = _reqs(SPECS, tol)
    b = _reqs(SPECS, tol)
    BatchServer(dtype=dtype, capacity=8).serve(a, keep_results=True, realtime=False)
    SequentialRunner(dtype=dtype).serve(b, keep_results=True, realtime=False)
    torch.cuda.synchronize()
    for ra, rb in zip(a, b):
        assert ra.iters == rb.iters, ra.rid
        for x, y in zip(ra.result, rb.result):
            assert np.array_equal(cpu(x), cpu(y)), ra.rid
        assert ra.stats['max_resP'] <= tol


def test_same_result_alone_and_in_a_mix():
    from pyflow_amd.serving import BatchServer
    alone = _reqs([(48, 8, 21, 4.0, 0.01, 0.0)])
    BatchServer(capacity=4).serve(alone, keep_results=True, realtime=False)
    # the same request arrives while three others are running
    mix = _reqs([(64, 12, 31, 2.0, 0.004, 0.0), (32, 20, 32, 8.0, 0.015, 0.0),
                 (96, 6, 33, 4.0, 0.01, 0.0), (48, 8, 21, 4.0, 0.01, 0.05)])
    srv = BatchServer(capacity=4)
    srv.serve(mix, keep_results=True, realtime=True)
    r = mix[3]
    assert r.admitted > min(m.admitted for m in mix[:3])
    assert max(srv.batch_sizes) >= 2
    for x, y in zip(alone[0].result, r.result):
        assert np.array_equal(cpu(x), cpu(y))


def test_admission_holds_the_hbm_cap():
    from pyflow_amd.serving import BatchServer, SlotPool
    pool = SlotPool(torch.float32, torch.device('cuda'))
    one = pool.take(32)
    cap = 2 * one.bytes + 1
    pool.give(one)
    reqs = _reqs([(32, 5, 41 + i, 4.0, 0.01, 0.0) for i in range(5)])
    srv = BatchServer(capacity=8, hbm_cap=cap)
    srv.serve(reqs, keep_results=False, realtime=False)
    assert srv.max_resident <= cap
    assert max(srv.batch_sizes) == 2
    order = sorted(reqs, key=lambda r: r.admitted)
    # the third request is admitted only after one of the first two finished
    assert order[2].admitted >= min(order[0].finished, order[1].finished)
    assert all(r.done_steps == r.steps for r in reqs)


def test_reused_slots_keep_each_result():
    """With room for two boxes, five requests of one size pass through two
    slots: a finished request's result is copied out while the loader may
    already write the next request's initial fields into the same slot.
    Every kept result equals the one from a server with room for all five."""
    from pyflow_amd.serving import BatchServer, SlotPool
    pool = SlotPool(torch.float32, torch.device('cuda'))
    one = pool.take(32)
    cap = 2 * one.bytes + 1
    pool.give(one)
    specs = [(32, 3 + i % 3, 51 + i, 4.0, 0.01, 0.0) for i in range(5)]
    tight, roomy = _reqs(specs), _reqs(specs)
    srv = BatchServer(capacity=8, hbm_cap=cap)
    srv.serve(tight, keep_results=True, realtime=False)
    assert max(srv.batch_sizes) == 2
    BatchServer(capacity=8).serve(roomy, keep_results=True, realtime=False)
    for a, b in zip(tight, roomy):
        for x, y in zip(a.result, b.result):
            assert torch.equal(x, y), a.rid


def test_cache_groups_do_not_change_results():
    """Grouping the boxes of a chunk for cache residency (and the store policy
    that goes with it) reorders independent boxes' launches only: the same
    bits with one group of all boxes, the default budget, and groups so small
    that every box is its own."""
    from pyflow_amd.serving import BatchServer
    specs = [(64, 5, 71, 4.0, 0.01, 0.0), (32, 7, 72, 2.0, 0.005, 0.0),
             (48, 6, 73, 8.0, 0.02, 0.0), (96, 3, 74, 4.0, 0.01, 0.0)]
    runs = []
    for gb in (0, None, 1):
        r = _reqs(specs, 1e-8)
        BatchServer(capacity=8, group_bytes=gb).serve(r, keep_results=True, realtime=False)
        runs.append(r)
    for a, b, c in zip(*runs):
        for x, y, z in zip(a.result, b.result, c.result):
            assert torch.equal(x, y) and torch.equal(x, z), a.rid
