"""Serving benchmark: a Poisson stream of decaying-turbulence requests served
by the continuous-batching server, and the same requests run one after
another through the plain single-box path.

Workload: requests arrive at `--rate` per second (default 50); box size
uniform over {32, 48, 64, 96, 128}^3, step count uniform in [10, 200],
k0 in {2, 4, 8}, nu uniform in [0.002, 0.02], each with its own seed; FP32
storage with FP64 Krylov reductions, BiCGStab to a relative L-inf residual
of 1e-6 (FP32 storage does not reach 1e-8 on every 128^3 field: such solves
stop at the iteration cap); admission holds at most `--hbm-gb` (default 200)
GB resident.

Prints one JSON line: aggregate Mcell-steps/s and p50/p99 per-step latency
of both runs, their ratio, and the number of requests whose final fields
are bit-identical in the two runs.

    python -m pyflow_amd.serve_bench --requests 48
"""

import argparse
import json
import sys

import numpy as np
import torch


def _same_bits(a, b):
    """Equal bit patterns (NaNs included)."""
    return torch.equal(a.contiguous().view(torch.uint8), b.contiguous().view(torch.uint8))


def run(requests=48, rate=50.0, seed=7, capacity=256, hbm_gb=200.0, tol=1e-6, dtype='f32',
        sequential=True, batched=True):
    """Serve one request stream batched and sequentially; returns the summary
    dict that main() prints."""
    from . import serving
    tdtype = torch.float32 if dtype == 'f32' else torch.float64
    reqs = serving.make_requests(requests, rate=rate, seed=seed, tol=tol)

    def warm_stream():
        """The stream's sizes with 2 steps each, queued at once: run untimed
        through the server that then serves the timed stream, it fills the
        server's slot pool and loads its kernels (a server that has been up
        for a while; both paths get the same pass)."""
        return [serving.Request(r.rid, r.N, 2, r.seed + 1, r.k0, r.nu, arrival=0.0, tol=tol)
                for r in reqs]

    out = {'workload': '%d requests, Poisson arrivals %.0f/s, N^3 in {32,48,64,96,128}, 10-200 '
                       'steps, k0 in {2,4,8}, nu in [0.002,0.02], %s, BiCGStab %g; each path '
                       'served by a server warmed with a 2-step pass of the same sizes'
                       % (requests, rate, dtype, tol)}
    if batched:
        srv = serving.BatchServer(dtype=tdtype, capacity=capacity,
                                  hbm_cap=int(hbm_gb * serving.GB))
        srv.serve(warm_stream(), realtime=False)
        srv.step_times, srv.batch_sizes, srv.max_resident = [], [], 0
        sb = srv.serve(reqs, keep_results=True, realtime=True)
        sb['max_resident_gb'] = srv.max_resident / serving.GB
        sb['mean_batch'] = float(np.mean(srv.batch_sizes))
        sb['server_steps'] = len(srv.step_times)
        out['batched'] = sb
    if sequential:
        seq = serving.make_requests(requests, rate=rate, seed=seed, tol=tol)
        runner = serving.SequentialRunner(dtype=tdtype)
        runner.serve(warm_stream(), realtime=False)
        ss = runner.serve(seq, keep_results=True, realtime=True)
        out['sequential'] = ss
        if batched:
            out['throughput_ratio'] = out['batched']['mcell_steps_per_s'] / ss['mcell_steps_per_s']
            same = 0
            for x, y in zip(reqs, seq):
                if all(_same_bits(p, q) for p, q in zip(x.result, y.result)):
                    same += 1
            out['bit_identical_requests'] = '%d/%d' % (same, len(reqs))
        out['nonfinite_requests'] = sum(
            1 for r in seq if not all(bool(torch.isfinite(t).all()) for t in r.result))
        for r in seq:
            r.result = None
    for r in reqs:
        r.result = None
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--requests', type=int, default=48)
    ap.add_argument('--rate', type=float, default=50.0)
    ap.add_argument('--seed', type=int, default=7)
    ap.add_argument('--capacity', type=int, default=256)
    ap.add_argument('--hbm-gb', type=float, default=200.0)
    ap.add_argument('--tol', type=float, default=1e-6)
    ap.add_argument('--dtype', choices=['f32', 'f64'], default='f32')
    ap.add_argument('--no-sequential', action='store_true')
    ap.add_argument('--no-batched', action='store_true')
    a = ap.parse_args(argv)
    out = run(a.requests, a.rate, a.seed, a.capacity, a.hbm_gb, a.tol, a.dtype,
              not a.no_sequential, not a.no_batched)
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
