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


def _alone(ctx, r, refine_steps):
    g = chypot.fields.unit_grid(*r.shape)
    pg = chypot.PreparedGrid(ctx, g)
    phi = torch.zeros_like(r.rho)
    if refine_steps and serve.refine_ok(r.shape):
        pg.solve_refined(r.rho, phi, steps=refine_steps)
    else:
        pg.solve(r.rho, phi)
    ctx.sync()
    pg.close()
    return phi.cpu()


@pytest.mark.parametrize("refine_steps", [0, 2])
def test_alone_equals_mixed(ctx, refine_steps):
    dens = {shape: serve.density_pool(shape, 3, seed=11) for shape, _ in MIX}
    reqs = serve.poisson_requests(rate=2000.0, duration=0.05, mix=MIX, densities=dens, seed=3)
    assert len(reqs) > 40
    print(f"[gpu] serving {len(reqs)} requests, refine_steps={refine_steps}", flush=True)
    srv = serve.Server(ctx, hbm_budget=8 << 30, max_batch=8, workers=4, keep_outputs=True,
                       refine_steps=refine_steps)
    done = srv.run(reqs)
    torch.cuda.synchronize()
    assert len(done) == len(reqs)
    mixed = {r.rid: r.phi.cpu() for r in done}
    summary = serve.latency_summary(done)
    print(f"[gpu] serving: {summary}", flush=True)
    assert summary["paths"] == {("refined" if refine_steps else "plain"): len(reqs)}
    srv.close()
    # every request checked alone, through the single-field entry points
    for r in reqs[:12]:
        assert torch.equal(_alone(ctx, r, refine_steps), mixed[r.rid]), r.rid


def test_same_shape_requests_share_a_launch(ctx):
    """Two requests of one shape admitted together take two slots (two grid
    objects) and so run as one batched launch."""
    shape = (32, 16, 32)
    dens = serve.density_pool(shape, 2, seed=1)
    srv = serve.Server(ctx, hbm_budget=1 << 30, max_batch=8, workers=2, keep_outputs=True)
    reqs = [serve.Request(0, 0.0, shape, dens[0]), serve.Request(1, 0.0, shape, dens[1])]
    srv.queue.extend(reqs)
    srv.step(0.0)
    assert len(srv.pools[shape]) == 2 and all(s.busy for s in srv.pools[shape])
    assert srv.inflight == 1  # one batch
    srv.queue.clear()
    done = srv.run([])
    assert {r.rid for r in done} == {0, 1}
    srv.close()


def test_served_mix_under_a_budget(ctx):
    """The served mix: Poisson arrivals over 60%
    128x64x128, 30% 256x128x256 and 10% 512x256x512 under an HBM budget,
    every request by the refined solve (KMAX <= 256); every request alone
    equals the same request in the mix, bit for bit, and the books balance."""
    mix = [((128, 64, 128), 0.6), ((256, 128, 256), 0.3), ((512, 256, 512), 0.1)]
    dens = {shape: serve.density_pool(shape, 2, seed=17, on_device=True) for shape, _ in mix}
    torch.cuda.synchronize()
    reqs = serve.poisson_requests(rate=400.0, duration=0.15, mix=mix, densities=dens, seed=5)
    assert any(r.shape == (512, 256, 512) for r in reqs)
    print(f"[gpu] served mix: {len(reqs)} requests", flush=True)
    srv = serve.Server(ctx, hbm_budget=24 << 30, max_batch=16, workers=8, keep_outputs=True, refine_steps=2)
    peak = 0
    orig_step = srv.step

    def step(now):
        nonlocal peak
        orig_step(now)
        peak = max(peak, srv.used)
    srv.step = step
    done = srv.run(reqs)
    torch.cuda.synchronize()
    summary = serve.latency_summary(done)
    print(f"[gpu] served mix: {summary}, peak {peak / 2**30:.2f} GiB", flush=True)
    assert len(done) == len(reqs)
    assert peak <= 24 << 30
    assert summary["paths"].get("refined", 0) == len(reqs)
    mixed = {r.rid: r.phi for r in done}
    seen = set()
    for r in reqs:
        if r.shape in seen:
            continue
        seen.add(r.shape)
        assert torch.equal(_alone(ctx, r, 2), mixed[r.rid].cpu()), r.shape
    srv.close()


def test_budget_admission_and_eviction(ctx):
    """A budget that holds one slot of 64x32x64 plus less than a slot of
    32x16x32 beside it: requests of the other shape wait until the first
    shape's slot can be evicted, and the books balance at the end."""
    mix = [((64, 32, 64), 0.5), ((32, 16, 32), 0.5)]
    dens = {shape: serve.density_pool(shape, 2, seed=5) for shape, _ in mix}
    reqs = serve.poisson_requests(rate=3000.0, duration=0.02, mix=mix, densities=dens, seed=9)
    probe = serve.Server(ctx, hbm_budget=1 << 40, max_batch=1, workers=1)
    probe.prepare([(64, 32, 64), (32, 16, 32)])
    big, small = probe.slot_bytes[(64, 32, 64)], probe.slot_bytes[(32, 16, 32)]
    probe.close()
    budget = big + small // 2
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
    assert all(not s.busy for pool in srv.pools.values() for s in pool)
    srv.close()
