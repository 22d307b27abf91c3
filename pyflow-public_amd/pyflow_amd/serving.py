"""Serving many independent box simulations on one GPU (continuous batching).

A request is one decaying-turbulence run: a periodic box of N^3 cells
(L = 2 pi), its own Passot-Pouquet field (seed, k0, u_rms = 1), viscosity nu,
a number of time steps, fixed dt at CFL 0.3, the reference's forward step
(RK4 + BiCGStab velCorr, core/PyFlow.py:352-425 / data/domain.py:342-402).

The server keeps every admitted request resident in HBM and advances all of
them together with pf_batch_advance: every kernel launch covers all the
boxes it concerns, the BiCGStab iterations of all resident boxes run from a
hipGraph captured for the current composition, and each box starts its next
time step as soon as its own solve is done (continuous batching at the
granularity of Krylov iterations: no request waits for a slower one).
Requests that arrive while others run wait in a FIFO queue and are admitted
as soon as their HBM footprint fits under the admission limit and a batch
slot is free; an admitted request's initial field is built by a loader
thread on its own stream while the batch keeps running (pf_batch_advance
releases the GIL), and it joins the batch once built; finished requests
leave when their last step ends.
Boxes and field tensors are pooled by size, so steady-state admission does
not allocate.

Per-request results do not depend on the mix: the batched kernels keep each
box's single-box block decomposition, so a request's final field is the
same bits alone, in any mix, and through the plain single-box path
(SequentialRunner).
"""

import math
import queue as _queue
import threading
import time

import numpy as np
import torch

from . import _native

GB = 1 << 30
FIELDS = ('u', 'v', 'w', 'ut', 'vt', 'wt', 'DP', 'p')


class Request:
    def __init__(self, rid, N, steps, seed, k0, nu, arrival=0.0, tol=1e-8, maxit=2000):
        self.rid = rid
        self.N = int(N)
        self.steps = int(steps)
        self.seed = int(seed)
        self.k0 = float(k0)
        self.nu = float(nu)
        self.arrival = float(arrival)
        self.tol = float(tol)
        self.maxit = int(maxit)
        # filled while served
        self.dt = None
        self.done_steps = 0
        self.admitted = None
        self.finished = None
        self.step_latency = []
        self.iters = []
        self.stats = None
        self.result = None

    @property
    def cells(self):
        return self.N ** 3


def initial_field(N, k0, seed, device, dtype, urms=1.0):
    """Passot-Pouquet field E(k) ~ k^4 exp(-2 (k/k0)^2) on the staggered faces,
    projected with the discrete divergence symbol (discretely solenoidal) and
    scaled to u_rms; random phases from the device generator seeded with
    `seed` (same construction as initial_conditions.passot_pouquet, which
    draws on the host)."""
    h = 2 * math.pi / N
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    f64 = torch.float64
    k = 2 * math.pi * torch.fft.fftfreq(N, d=h, device=device, dtype=f64)
    kr = 2 * math.pi * torch.fft.rfftfreq(N, d=h, device=device, dtype=f64)
    KX, KY, KZ = torch.meshgrid(k, k, kr, indexing='ij')
    kmag = torch.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp = torch.where(kmag > 0, torch.sqrt(kmag ** 4 * torch.exp(-2.0 * (kmag / k0) ** 2)
                                           / (4 * math.pi * kmag.clamp_min(1e-30) ** 2)),
                      torch.zeros_like(kmag))
    shape = KX.shape
    uh = []
    for _ in range(3):
        re = torch.randn(shape, generator=gen, device=device, dtype=f64)
        im = torch.randn(shape, generator=gen, device=device, dtype=f64)
        uh.append(amp * torch.complex(re, im))
    sig = [(torch.exp(1j * KX * h) - 1.0) / h, (torch.exp(1j * KY * h) - 1.0) / h,
           (torch.exp(1j * KZ * h) - 1.0) / h]
    div = sig[0] * uh[0] + sig[1] * uh[1] + sig[2] * uh[2]
    norm = sig[0].abs() ** 2 + sig[1].abs() ** 2 + sig[2].abs() ** 2
    fac = torch.where(norm > 0, div / norm.clamp_min(1e-300), torch.zeros_like(div))
    out = []
    for d in range(3):
        x = uh[d] - torch.conj(sig[d]) * fac
        x = torch.where(norm == 0, torch.zeros_like(x), x)
        out.append(torch.fft.irfftn(x, s=(N, N, N)))
    cur = math.sqrt(sum(float(torch.mean(f * f)) for f in out) / 3.0)
    scale = urms / cur if cur > 0 else 0.0
    return [(f * scale).to(dtype) for f in out]


class Slot:
    """A box and its fields (one resident request)."""

    def __init__(self, N, dtype, device):
        h = 2 * math.pi / N
        self.N = N
        self.box = _native.Box(N, N, N, h, h, h, dtype=dtype, device=device)
        P = self.box.padded_shape
        self.f = {n: torch.zeros(P, dtype=dtype, device=device) for n in FIELDS}
        self.f['src'] = torch.zeros(self.box.interior_shape, dtype=dtype, device=device)
        self.bytes = self.box.device_bytes + sum(t.numel() * t.element_size()
                                                 for t in self.f.values())
        # recorded on the serving stream after the last reads of a finished
        # request's fields; the next load of the slot waits for it
        self.released = None

    def load(self, req):
        """Initial state of `req` (fields, p = 0) and its fixed dt."""
        u = initial_field(req.N, req.k0, req.seed, self.box.device, self.box.dtype)
        for q, n in enumerate(('u', 'v', 'w')):
            self.f[n][2:-2, 2:-2, 2:-2].copy_(u[q])
            self.box.fill_halo(self.f[n])
            self.f[n + 't'].copy_(self.f[n])
        for n in ('DP', 'p'):
            self.f[n].zero_()
        umax = max(float(t.abs().max()) for t in u)
        req.dt = 0.3 * (2 * math.pi / req.N) / umax

    def params(self, req):
        return (req.dt, req.nu, 1.0, req.tol, req.maxit)


class SlotPool:
    """Boxes and fields by size, reused across requests; counts HBM bytes."""

    def __init__(self, dtype, device):
        self.dtype, self.device = dtype, device
        self.free = {}
        self.bytes_resident = 0     # slots in use
        self.bytes_pooled = 0       # idle slots kept for reuse
        self.known = {}             # measured slot bytes by size

    def footprint(self, N):
        """Device bytes of a slot of size N: measured once a slot of that size
        exists, else an upper estimate (8 padded fields, the source, and the
        box's 12 interior vectors, 3 padded buffers and reduction space)."""
        if N in self.known:
            return self.known[N]
        e = torch.tensor([], dtype=self.dtype).element_size()
        P = (N + 4) ** 3
        I = N ** 3
        return (11 * P + 13 * I) * e + (1 << 20)

    def take(self, N):
        lst = self.free.get(N)
        if lst:
            s = lst.pop()
            self.bytes_pooled -= s.bytes
        else:
            s = Slot(N, self.dtype, self.device)
            self.known[N] = s.bytes
        self.bytes_resident += s.bytes
        return s

    def give(self, s):
        self.bytes_resident -= s.bytes
        self.free.setdefault(s.N, []).append(s)
        self.bytes_pooled += s.bytes

    def evict(self, need, cap):
        """Drop idle slots until `need` more bytes fit under `cap`."""
        for N in sorted(self.free, reverse=True):
            while self.free[N] and self.bytes_resident + self.bytes_pooled + need > cap:
                s = self.free[N].pop()
                self.bytes_pooled -= s.bytes
                s.box.close()
                del s


def make_requests(n, rate=50.0, sizes=(32, 48, 64, 96, 128), steps=(10, 200), seed=0,
                  tol=1e-8, maxit=2000):
    """n requests with Poisson arrivals (mean `rate` per second), size uniform
    over `sizes`, step count uniform in [steps[0], steps[1]], k0 in {2,4,8},
    nu uniform in [0.002, 0.02], each with its own field seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(Request(i, rng.choice(sizes), rng.integers(steps[0], steps[1] + 1),
                           rng.integers(1, 2 ** 31 - 1), rng.choice([2.0, 4.0, 8.0]),
                           rng.uniform(0.002, 0.02), arrival=t, tol=tol, maxit=maxit))
    return out


def _finish(req, slot, keep_result):
    req.finished = time.perf_counter()
    if keep_result:
        req.result = [slot.f[n][2:-2, 2:-2, 2:-2].clone() for n in ('u', 'v', 'w', 'p')]
    # the clones are queued on this thread's stream; a later load of the slot
    # (on the loader's stream) must not overwrite the fields before they ran
    ev = torch.cuda.Event()
    ev.record()
    slot.released = ev


def _record(req, row, npts):
    req.stats = dict(maxU=row[0], maxV=row[1], maxW=row[2], TKE=row[3] / npts * 0.5,
                     maxDiv=row[4], max_resP=row[5])
    req.iters.append(int(row[6]))


class _Loader(threading.Thread):
    """Builds admitted requests' initial states (Slot.load) on a stream of
    its own, so that the batch's kernels keep the GPU busy meanwhile."""

    def __init__(self, device):
        super().__init__(daemon=True)
        self.device = device
        self.stream = torch.cuda.Stream(device=device)
        self.todo = _queue.Queue()
        self.done = _queue.Queue()

    def run(self):
        torch.cuda.set_device(self.device)
        while True:
            item = self.todo.get()
            if item is None:
                return
            req, slot = item
            try:
                if slot.released is not None:
                    self.stream.wait_event(slot.released)
                with torch.cuda.stream(self.stream):
                    slot.load(req)
                self.stream.synchronize()
                self.done.put((req, slot, None))
            except Exception as e:   # handed to the serving loop, which raises it
                self.done.put((req, slot, e))


class BatchServer:
    """Continuous-batching server on one GPU.  hbm_cap: admission limit in
    bytes for resident slots (idle pooled slots are evicted to make room)."""

    def __init__(self, dtype=torch.float32, device=None, capacity=256, hbm_cap=200 * GB,
                 group_bytes=None, persistent=True):
        _native.require_gpu()
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.dtype = dtype
        self.capacity = int(capacity)
        self.hbm_cap = int(hbm_cap)
        self.pool = SlotPool(dtype, self.device)
        self.batch = _native.Batch(self.capacity, dtype)
        if group_bytes is not None:
            self.batch.set_group_bytes(group_bytes)
        # big boxes (128^3) solve each step as one persistent launch; an int
        # caps such solves per advance call (True: one)
        self.batch.set_persistent(persistent)
        self.step_times = []
        self.batch_sizes = []
        self.max_resident = 0

    def serve(self, requests, keep_results=False, realtime=True):
        """Serve `requests` (arrival times in seconds from now when realtime,
        else all queued at once).  Returns a summary dict."""
        loader = _Loader(self.device)
        loader.start()
        try:
            return self._serve(requests, keep_results, realtime, loader)
        finally:
            loader.todo.put(None)
            loader.join()

    def _serve(self, requests, keep_results, realtime, loader):
        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []
        active = []          # [req, slot, inside a step, step start time]
        loading = 0          # admitted requests whose initial state is being built
        t0 = time.perf_counter()
        i = 0
        dirty = True
        while i < len(pending) or queue or active or loading:
            now = time.perf_counter() - t0
            while i < len(pending) and (not realtime or pending[i].arrival <= now):
                queue.append(pending[i])
                i += 1
            # admission: FIFO while a slot is free and the footprint fits
            while queue and len(active) + loading < self.capacity:
                req = queue[0]
                need = self.pool.footprint(req.N)
                if self.pool.bytes_resident + need > self.hbm_cap:
                    break
                if not self.pool.free.get(req.N):
                    self.pool.evict(need, self.hbm_cap)
                slot = self.pool.take(req.N)
                loader.todo.put((req, slot))
                loading += 1
                queue.pop(0)
            self.max_resident = max(self.max_resident, self.pool.bytes_resident)
            # built requests join the batch; wait for one only when nothing runs
            while loading:
                try:
                    req, slot, err = loader.done.get(block=not active, timeout=0.01)
                except _queue.Empty:
                    break
                if err is not None:
                    raise err
                loading -= 1
                req.admitted = time.perf_counter()
                active.append([req, slot, False, None])
                dirty = True
            if not active:
                if i < len(pending) and not loading:
                    time.sleep(max(0.0, min(0.01, pending[i].arrival - now)))
                continue
            if dirty:
                self.batch.set([(e[1].box, e[1].f, e[1].params(e[0])) for e in active])
                dirty = False
            # every idle resident request starts its next step now
            ts = time.perf_counter()
            start = []
            for e in active:
                start.append(not e[2])
                if not e[2]:
                    e[2], e[3] = True, ts
            fin, rows, _ = self.batch.advance(start, self.device)
            te = time.perf_counter()
            self.step_times.append(te - ts)
            self.batch_sizes.append(len(active))
            still = []
            for idx, e in enumerate(active):
                req, slot = e[0], e[1]
                if idx in rows:
                    e[2] = False
                    req.done_steps += 1
                    req.step_latency.append(te - e[3])
                    _record(req, rows[idx], float(req.cells))
                    if req.done_steps >= req.steps:
                        _finish(req, slot, keep_results)
                        self.pool.give(slot)
                        dirty = True
                        continue
                still.append(e)
            active = still
        wall = time.perf_counter() - t0
        return summarize(requests, wall)


class SequentialRunner:
    """The same requests one after another, each through the plain
    single-box entry points (pf_rk4_step, pf_vel_corr, pf_stats), in arrival
    order: the unbatched baseline."""

    def __init__(self, dtype=torch.float32, device=None):
        _native.require_gpu()
        self.device = torch.device('cuda', torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.dtype = dtype
        self.pool = SlotPool(dtype, self.device)

    def run_one(self, req, keep_result=False):
        slot = self.pool.take(req.N)
        slot.load(req)
        req.admitted = time.perf_counter()
        f, box = slot.f, slot.box
        for _ in range(req.steps):
            ts = time.perf_counter()
            box.rk4_step(f['u'], f['v'], f['w'], f['ut'], f['vt'], f['wt'], req.dt, req.nu, 1.0)
            res, its, st = box.vel_corr(f['u'], f['v'], f['w'], f['DP'], f['p'], f['src'],
                                        _native.PF_POISSON_BICGSTAB, 1.0 / req.dt, req.dt / 1.0,
                                        req.tol, req.maxit)
            m = box.stats(f['u'], f['v'], f['w'])
            req.step_latency.append(time.perf_counter() - ts)
            req.done_steps += 1
            req.iters.append(its)
            req.stats = dict(maxU=m[0], maxV=m[1], maxW=m[2], TKE=m[3], maxDiv=m[4],
                             max_resP=res)
        _finish(req, slot, keep_result)
        self.pool.give(slot)

    def serve(self, requests, keep_results=False, realtime=True):
        t0 = time.perf_counter()
        for req in sorted(requests, key=lambda r: r.arrival):
            if realtime:
                wait = req.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            self.run_one(req, keep_results)
        return summarize(requests, time.perf_counter() - t0)


def summarize(requests, wall):
    lat = np.array([x for r in requests for x in r.step_latency]) if requests else np.zeros(1)
    cell_steps = sum(r.cells * r.steps for r in requests)
    its = {}
    for r in requests:
        its.setdefault(r.N, []).extend(r.iters)
    capped = sum(1 for r in requests for k in r.iters if k >= r.maxit)
    return dict(requests=len(requests), wall_s=wall, cell_steps=cell_steps,
                mcell_steps_per_s=cell_steps / wall / 1e6,
                step_latency_p50_ms=float(np.percentile(lat, 50)) * 1e3,
                step_latency_p99_ms=float(np.percentile(lat, 99)) * 1e3,
                steps=int(lat.size),
                iters_by_size={int(N): [int(np.mean(v)), int(np.max(v))] for N, v in its.items()},
                capped_solves=capped)
