"""Serving loop for potential solves on one GPU.

Requests arrive as a Poisson process over a mix of grids.  Each request names
a grid shape and a device-resident density; the server

  * keeps, per grid shape, a pool of slots: a prepared grid (SETBDY state and
    workspace) with its PHI and COEF(K=2) output buffers allocated once.  A
    request is admitted into a free slot of its shape; a new slot is made
    only when its bytes (chy_grid_device_bytes plus the output buffers, and
    the refinement buffers when refining) fit the HBM budget, after evicting
    idle slots of other shapes, least recently used first.  Because each
    slot is its own grid object, same-shape requests run together in one
    batched launch per kernel;
  * batches continuously: whatever is admitted (up to max_batch requests) is
    launched as one chy_solve_batch (or chy_solve_batch_refined) call as soon
    as it is admitted, while earlier batches still run on the GPU;
  * waits without spinning: a helper thread blocks on each batch's HIP event
    (blocking synchronisation) and wakes the loop, which otherwise sleeps
    until the next arrival;
  * records per-request latency from arrival to the host seeing completion.

refine_steps > 0 serves the refined solve (binary32 correction solves on the
FP32 matrix cores, binary64 residuals) on grids it supports (KMAX <= 256);
other grids get the plain binary64 solve.  Each request records its path.
A chypot.autotune table (tune_table) sets the launch shapes of every grid
the server prepares.

Solves are deterministic and independent, so a request's potential does not
depend on what it was batched with (tests/test_gpu_serve.py checks this bit
for bit).
"""
from __future__ import annotations

import collections
import dataclasses
import queue as queue_mod
import threading
import time

import numpy as np

from . import COEF_K2, Context, PreparedGrid, autotune, fields, to_device


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds from the start of the run
    shape: tuple            # (JMAX, KMAX, LMAX)
    rho: object             # device tensor (LMAX, KMAX+2, JMAX+2), float64
    phi: object = None      # the result (a copy taken at completion when the server keeps outputs)
    admitted: float = -1.0
    done: float = -1.0
    path: str = ""          # "refined" or "plain"


def field_bytes(shape) -> int:
    j, k, l = shape
    return (j + 2) * (k + 2) * l * 8


def refine_ok(shape) -> bool:
    """The refined solve's z transforms keep a slot's weights in registers
    for KMAX-1 <= 255."""
    return shape[1] <= 256


class Slot:
    """A prepared grid and the output buffers of the request it serves."""

    def __init__(self, ctx, shape, torch):
        j, k, l = shape
        self.shape = shape
        self.grid = PreparedGrid(ctx, fields.unit_grid(*shape))
        self.phi = torch.zeros((l, k + 2, j + 2), dtype=torch.float64, device="cuda")
        self.coef = torch.zeros((2, l // 2, k + 2, j + 2), dtype=torch.float64, device="cuda")
        self.busy = False
        self.bytes = 0

    def close(self):
        self.grid.close()
        self.phi = self.coef = None


class Server:
    def __init__(self, ctx: Context, hbm_budget: int, max_batch: int = 16, workers: int = 4,
                 keep_outputs: bool = False, refine_steps: int = 0, slots_per_shape: int = 16,
                 tune_table: dict | None = None):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.keep_outputs = keep_outputs  # copy each result out of its slot at completion
        self.refine_steps = int(refine_steps)
        self.slots_per_shape = slots_per_shape
        self.tune_table = tune_table or {}  # chypot.autotune table applied to every new slot's grid
        ctx.workers(workers)
        # eager launches: each request brings its own density buffer, so a
        # captured single-field graph (keyed on its pointers) would rarely replay
        ctx.graphs(False)
        # the context's stream: batches fork from and join back to it
        self.stream = torch.cuda.Stream()
        ctx.set_stream(self.stream)
        self.pools = collections.OrderedDict()  # shape -> [Slot] (LRU order of shapes)
        self.slot_bytes = {}                    # shape -> bytes of one slot, once one exists
        self.used = 0                           # bytes of all slots
        self.queue = collections.deque()
        self.inflight = 0
        self.finished = []
        # completion: a helper thread blocks on each batch's event
        self._cv = threading.Condition()
        self._completed = collections.deque()
        self._waits = queue_mod.Queue()
        self._waiter = threading.Thread(target=self._wait_loop, daemon=True)
        self._waiter.start()

    # -- slots and memory accounting -----------------------------------------
    def _refined(self, shape) -> bool:
        return self.refine_steps > 0 and refine_ok(shape)

    def _estimate(self, shape) -> int:
        """Bytes of a new slot: exact once a slot of the shape has existed,
        otherwise an upper bound (two binary64 spectra, the binary32 ones,
        tables well under one field, PHI and COEF)."""
        if shape in self.slot_bytes:
            return self.slot_bytes[shape]
        return 6 * field_bytes(shape)

    def _evict_until(self, need: int, keep) -> bool:
        """Free idle slots of other shapes (least recently used shape first)
        until `need` more bytes fit the budget."""
        for shape in list(self.pools):
            if self.used + need <= self.budget:
                break
            if shape == keep:
                continue
            pool = self.pools[shape]
            for s in [s for s in pool if not s.busy]:
                if self.used + need <= self.budget:
                    break
                pool.remove(s)
                self.used -= s.bytes
                s.close()
            if not pool:
                del self.pools[shape]
        return self.used + need <= self.budget

    def _new_slot(self, shape):
        need = self._estimate(shape)
        if not self._evict_until(need, shape):
            return None
        s = Slot(self.ctx, shape, self.torch)
        if self.tune_table:
            autotune.apply(s.grid, self.tune_table)
        s.bytes = s.grid.device_bytes + field_bytes(shape) + 2 * (shape[2] // 2) * (shape[1] + 2) * (shape[0] + 2) * 8
        if self._refined(shape):
            s.bytes += field_bytes(shape)  # the binary32 spectrum and eigen coordinates (allocated on first use)
        self.slot_bytes[shape] = s.bytes
        self.used += s.bytes
        self.pools.setdefault(shape, []).append(s)
        if self.used > self.budget:  # the estimate was low: give the slot back
            self.pools[shape].remove(s)
            self.used -= s.bytes
            s.close()
            return None
        return s

    def _slot_for(self, shape):
        pool = self.pools.get(shape, [])
        if shape in self.pools:
            self.pools.move_to_end(shape)
        for s in pool:
            if not s.busy:
                return s
        if len(pool) < self.slots_per_shape:
            return self._new_slot(shape)
        return None

    def prepare(self, shapes, per_shape: int = 1):
        """Make slots ahead of traffic (SETBDY work kept off request latency)."""
        for shape in shapes:
            for _ in range(per_shape):
                if self._new_slot(tuple(shape)) is None:
                    break

    # -- completion ------------------------------------------------------------
    def _wait_loop(self):
        while True:
            item = self._waits.get()
            if item is None:
                return
            ev, batch = item
            ev.synchronize()
            with self._cv:
                self._completed.append(batch)
                self._cv.notify()

    def _finish(self, batch, now: float):
        for r, s in batch:
            r.done = now
            s.busy = False
            self.finished.append(r)
        self.inflight -= 1

    # -- loop -----------------------------------------------------------------
    def step(self, now: float):
        """Admit what fits and launch it as one batch (per path)."""
        groups = {"refined": [], "plain": []}
        n = 0
        for _ in range(len(self.queue)):
            req = self.queue.popleft()
            slot = self._slot_for(req.shape) if n < self.max_batch else None
            if slot is None:
                self.queue.append(req)
                continue
            slot.busy = True
            req.admitted = now
            req.path = "refined" if self._refined(req.shape) else "plain"
            groups[req.path].append((req, slot))
            n += 1
        for path, batch in groups.items():
            if not batch:
                continue
            jobs = [dict(grid=s.grid, rho=r.rho, phi=s.phi, coef=s.coef, coef_mode=COEF_K2) for r, s in batch]
            self.ctx.solve_batch(jobs, refine_steps=self.refine_steps if path == "refined" else 0)
            if self.keep_outputs:
                with self.torch.cuda.stream(self.stream):
                    for r, s in batch:
                        r.phi = s.phi.clone()
            ev = self.torch.cuda.Event(blocking=True)
            ev.record(self.stream)
            self.inflight += 1
            self._waits.put((ev, batch))

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
            if self.queue and not self.inflight:
                raise RuntimeError("a request does not fit the HBM budget on an idle GPU")
            # sleep until the next arrival or a batch completes
            timeout = None if not pending else max(0.0, pending[0].arrival - (time.perf_counter() - t0))
            with self._cv:
                if not self._completed:
                    if timeout is None and self.inflight:
                        self._cv.wait()
                    elif timeout:
                        self._cv.wait(timeout)
                done = list(self._completed)
                self._completed.clear()
            now = time.perf_counter() - t0
            for batch in done:
                self._finish(batch, now)
        return self.finished

    def close(self):
        self._waits.put(None)
        self._waiter.join()
        self.torch.cuda.synchronize()
        self.ctx.graphs(True)
        for pool in self.pools.values():
            for s in pool:
                s.close()
        self.pools.clear()
        self.used = 0


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
    paths = collections.Counter(r.path for r in finished)
    return dict(requests=len(finished), p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), mean_admit_wait_ms=float(wait.mean()),
                paths=dict(paths))
