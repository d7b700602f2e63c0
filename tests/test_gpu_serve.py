"""The serving loop (chypot.serve): Poisson arrivals over a grid mix, HBM-budget
admission, continuous batching.

A request's potential must not depend on what it is served with: solved
alone and solved inside a mixed stream of requests it is bit for bit the same.
The HBM budget must hold at every moment.
"""
import numpy as np
import pytest

import chypot
from chypot import serve

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MIX = [((32, 16, 32), 0.6), ((64, 32, 64), 0.3), ((32, 16, 64), 0.1)]


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return chypot.Context(0)


def test_alone_equals_mixed(ctx):
    dens = {shape: serve.density_pool(shape, 3, seed=11) for shape, _ in MIX}
    reqs = serve.poisson_requests(rate=2000.0, duration=0.05, mix=MIX, densities=dens, seed=3)
    assert len(reqs) > 40
    print(f"[gpu] serving {len(reqs)} requests", flush=True)
    srv = serve.Server(ctx, hbm_budget=8 << 30, max_batch=8, workers=4, keep_outputs=True)
    done = srv.run(reqs)
    torch.cuda.synchronize()
    assert len(done) == len(reqs)
    mixed = {r.rid: r.phi.cpu() for r in done}
    summary = serve.latency_summary(done)
    print(f"[gpu] serving: {summary}", flush=True)
    srv.close()
    # every distinct (grid, density) alone, through the plain entry point
    for r in reqs[:12]:
        g = chypot.fields.unit_grid(*r.shape)
        pg = chypot.PreparedGrid(ctx, g)
        phi = torch.zeros_like(r.rho)
        pg.solve(r.rho, phi)
        ctx.sync()
        assert torch.equal(phi.cpu(), mixed[r.rid]), r.rid
        pg.close()


def test_budget_admission_and_eviction(ctx):
    """A budget that holds one prepared 64x32x64 grid plus a few requests:
    requests of the other shape wait until the first grid can be evicted,
    and the books balance at the end."""
    mix = [((64, 32, 64), 0.5), ((32, 16, 32), 0.5)]
    dens = {shape: serve.density_pool(shape, 2, seed=5) for shape, _ in mix}
    reqs = serve.poisson_requests(rate=3000.0, duration=0.02, mix=mix, densities=dens, seed=9)
    g64 = chypot.PreparedGrid(ctx, chypot.fields.unit_grid(64, 32, 64))
    budget = g64.device_bytes + 6 * serve.field_bytes((64, 32, 64))
    g64.close()
    srv = serve.Server(ctx, hbm_budget=budget, max_batch=4, workers=2)
    peak = 0
    orig_step = srv.step

    def step(now):
        nonlocal peak
        orig_step(now)
        peak = max(peak, srv.used)
    srv.step = step
    done = srv.run(reqs)
    assert len(done) == len(reqs)
    assert peak <= budget
    assert all(v == 0 for v in srv.in_use.values())
    srv.close()
