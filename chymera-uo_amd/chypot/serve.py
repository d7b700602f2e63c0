"""Serving loop for potential solves on one GPU.

Requests arrive as a Poisson process over a mix of grids.  Each request names
a grid shape and a device-resident density; the server

  * keeps a cache of prepared grids (SETBDY state + workspace), least recently
    used first out, and admits a request only when its grid and its output
    buffers fit the HBM budget (bytes reported by chy_grid_device_bytes plus
    the request's PHI and COEF(K=2) buffers);
  * batches whatever is admitted (up to max_batch requests per call) into one
    chy_solve_batch call -- batches are formed continuously, a new one as soon
    as requests are admitted, while earlier batches still run;
  * records per-request latency from arrival to the host seeing completion.

Solves are deterministic and independent, so a request's potential does not
depend on what it was batched with (tests/test_gpu_serve.py checks this bit
for bit).
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import COEF_K2, Context, PreparedGrid, fields, to_device


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds from the start of the run
    shape: tuple            # (JMAX, KMAX, LMAX)
    rho: object             # device tensor (LMAX, KMAX+2, JMAX+2), float64
    phi: object = None      # output (allocated at admission)
    coef: object = None
    admitted: float = -1.0
    done: float = -1.0


def field_bytes(shape) -> int:
    j, k, l = shape
    return (j + 2) * (k + 2) * l * 8


class Server:
    def __init__(self, ctx: Context, hbm_budget: int, max_batch: int = 16, workers: int = 4,
                 keep_outputs: bool = False):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.keep_outputs = keep_outputs  # else a finished request's buffers are released
        ctx.workers(workers)
        # the context's stream: batches fork from and join back to it
        self.stream = torch.cuda.Stream()
        ctx.set_stream(self.stream)
        self.grids = collections.OrderedDict()   # shape -> PreparedGrid (LRU order)
        self.in_use = collections.Counter()      # shape -> admitted requests
        self.used = 0                            # bytes of grids + admitted buffers
        self.queue = collections.deque()
        self.inflight = []                       # (event, [requests])
        self.finished = []

    # -- memory accounting --------------------------------------------------
    def _request_bytes(self, shape) -> int:
        j, k, l = shape
        return field_bytes(shape) + 2 * (l // 2) * (k + 2) * (j + 2) * 8

    def _grid_for(self, shape):
        if shape in self.grids:
            self.grids.move_to_end(shape)
            return self.grids[shape]
        g = fields.unit_grid(*shape)
        pg = PreparedGrid(self.ctx, g)
        self.grids[shape] = pg
        self.used += pg.device_bytes
        return pg

    def _evict_until(self, need: int) -> bool:
        """Free idle prepared grids (least recently used first) until `need`
        more bytes fit the budget."""
        for shape in list(self.grids):
            if self.used + need <= self.budget:
                break
            if self.in_use[shape] == 0:
                pg = self.grids.pop(shape)
                self.used -= pg.device_bytes
                pg.close()
        return self.used + need <= self.budget

    def _admit(self, req: Request, now: float) -> bool:
        need = self._request_bytes(req.shape)
        if req.shape not in self.grids:
            # a grid's workspace is two spectra plus tables: bound it before
            # preparing (the exact figure is booked once it exists)
            need += 3 * field_bytes(req.shape)
        if not self._evict_until(need):
            return False
        pg = self._grid_for(req.shape)
        t = self.torch
        j, k, l = req.shape
        with t.cuda.stream(self.stream):  # ordered before the batch that uses them
            req.phi = t.zeros((l, k + 2, j + 2), dtype=t.float64, device="cuda")
            req.coef = t.zeros((2, l // 2, k + 2, j + 2), dtype=t.float64, device="cuda")
        self.used += self._request_bytes(req.shape)
        self.in_use[req.shape] += 1
        req.admitted = now
        req.grid = pg
        return True

    def _release(self, req: Request):
        self.used -= self._request_bytes(req.shape)
        self.in_use[req.shape] -= 1
        if not self.keep_outputs:
            req.phi = req.coef = None

    def prepare(self, shapes):
        """Prepare grids ahead of traffic (SETBDY work kept off request latency)."""
        for shape in shapes:
            self._grid_for(tuple(shape))

    # -- loop -----------------------------------------------------------------
    def poll(self, now: float):
        still = []
        for ev, reqs in self.inflight:
            if ev.query():
                for r in reqs:
                    r.done = now
                    self._release(r)
                    self.finished.append(r)
            else:
                still.append((ev, reqs))
        self.inflight = still

    def step(self, now: float):
        self.poll(now)
        batch = []
        for _ in range(len(self.queue)):
            req = self.queue.popleft()
            if len(batch) < self.max_batch and self._admit(req, now):
                batch.append(req)
            else:
                self.queue.append(req)
        if batch:
            self.ctx.solve_batch([dict(grid=r.grid, rho=r.rho, phi=r.phi, coef=r.coef,
                                       coef_mode=COEF_K2) for r in batch])
            ev = self.torch.cuda.Event()
            ev.record(self.stream)
            self.inflight.append((ev, batch))

    def run(self, requests):
        """Serve `requests` in real time (arrival times in seconds from now);
        returns the finished requests."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        self.torch.cuda.synchronize()  # inputs made on other streams are complete
        t0 = time.perf_counter()
        while pending or self.queue or self.inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                self.queue.append(pending.popleft())
            self.step(now)
        return self.finished

    def close(self):
        self.torch.cuda.synchronize()
        for pg in self.grids.values():
            pg.close()
        self.grids.clear()


def poisson_requests(rate: float, duration: float, mix, densities, seed: int = 0):
    """Arrival times of a Poisson process of `rate` per second over
    `duration` seconds; each request draws its grid from `mix` (list of
    (shape, probability)) and a density from `densities[shape]` (a list of
    device tensors) in turn."""
    rng = np.random.default_rng(seed)
    shapes = [m[0] for m in mix]
    probs = np.array([m[1] for m in mix], dtype=float)
    probs /= probs.sum()
    t, out, used = 0.0, [], collections.Counter()
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        shape = shapes[rng.choice(len(shapes), p=probs)]
        pool = densities[shape]
        out.append(Request(len(out), t, shape, pool[used[shape] % len(pool)]))
        used[shape] += 1


def density_pool(shape, n: int, seed: int = 0, on_device: bool = False):
    """n device-resident tori on a unit grid of `shape` (random centre,
    radius, index and log-normal noise as in fields.random_field; generated
    on the GPU with on_device, for large grids)."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed + 7919 * i)
        g = fields.unit_grid(*shape)
        rc = float(rng.uniform(0.3, 0.6))
        a = float(rng.uniform(0.15, 0.35))
        nidx = float(rng.choice([1.0, 1.5, 2.5]))
        if on_device:
            out.append(fields.torus_device(g, rc, a, nidx, 0.05, seed + 7919 * i))
            continue
        lognorm = np.exp(0.05 * rng.standard_normal(size=g.shape))
        rho = fields._place(g, fields.torus(g, rc, a, nidx) * lognorm)
        out.append(to_device(rho))
    return out


def latency_summary(finished):
    lat = np.array([r.done - r.arrival for r in finished]) * 1e3
    wait = np.array([r.admitted - r.arrival for r in finished]) * 1e3
    return dict(requests=len(finished), p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), mean_admit_wait_ms=float(wait.mean()))
