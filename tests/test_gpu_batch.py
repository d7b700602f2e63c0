"""Batched solves of independent heterogeneous fields (chy_solve_batch).

Every field of a batch must come out bit for bit as when it is solved alone
with chy_solve, and agree with the oracle; a bad item stops the whole batch
before anything runs.
"""
import numpy as np
import pytest

import chypot
from chypot import fields
from helpers import rel_max

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SMALL_GRIDS = ((32, 64), (16, 32), (32, 64))   # JMAX, KMAX, LMAX choices


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return chypot.Context(0)


def make_jobs(ctx, seeds, grids=SMALL_GRIDS, coef_mode=chypot.COEF_K2):
    prepared = {}
    jobs = []
    for s in seeds:
        g, rho = fields.random_field(s, grids)
        key = (g.jmax, g.kmax, g.lmax)
        if key not in prepared:
            prepared[key] = chypot.PreparedGrid(ctx, g)
        coef = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64, device="cuda")
        jobs.append(dict(grid=prepared[key], g=g, rho_h=rho, rho=chypot.to_device(rho),
                         phi=torch.zeros((g.lmax, g.kmax + 2, g.jmax + 2), dtype=torch.float64,
                                         device="cuda"),
                         coef=coef, coef_mode=coef_mode, maxtrm=10, npoint=8, redge=0.0))
    return jobs, prepared


@pytest.mark.parametrize("workers", [1, 4])
def test_batch_equals_single_solves(ctx, workers):
    print(f"[gpu] batch of 12 heterogeneous fields, {workers} worker streams", flush=True)
    jobs, prepared = make_jobs(ctx, range(100, 112))
    alone = []
    for jb in jobs:
        phi = torch.zeros_like(jb["phi"])
        coef = torch.zeros_like(jb["coef"])
        d = jb["grid"].solve(jb["rho"], phi, coef=coef, coef_mode=jb["coef_mode"], diag=True)
        alone.append((phi.cpu(), coef.cpu(), (d.tmass, d.phichk, d.klocat)))
    ctx.workers(workers)
    try:
        diags = ctx.solve_batch(jobs, diag=True)
        ctx.sync()
        for jb, (phi, coef, d), db in zip(jobs, alone, diags):
            assert torch.equal(jb["phi"].cpu(), phi)
            assert torch.equal(jb["coef"].cpu(), coef)
            assert (db.tmass, db.phichk, db.klocat) == d
        # the same batch again (graph replays) gives the same numbers
        for jb in jobs:
            jb["phi"].zero_()
        ctx.solve_batch(jobs)
        ctx.sync()
        for jb, (phi, _, _) in zip(jobs, alone):
            assert torch.equal(jb["phi"].cpu(), phi)
    finally:
        ctx.workers(4)
        for pg in prepared.values():
            pg.close()


def test_batch_matches_oracle(ctx):
    import pyoracle as po
    po.use_extended(False)
    jobs, prepared = make_jobs(ctx, [7, 8, 9])
    print("[gpu] batch of 3 fields vs the oracle", flush=True)
    diags = ctx.solve_batch(jobs, diag=True)
    for jb, d in zip(jobs, diags):
        want = po.solve(jb["g"], jb["rho_h"])
        assert rel_max(chypot.to_host(jb["phi"]), want["phi"]) < 1e-12
        assert abs(d.tmass - want["tmass"]) <= 1e-13 * abs(want["tmass"])
        assert d.klocat == want["klocat"]
    for pg in prepared.values():
        pg.close()


def test_batch_rejects_bad_item_before_running(ctx):
    jobs, prepared = make_jobs(ctx, [1, 2])
    jobs[1]["npoint"] = 0
    with pytest.raises(chypot.ChyError) as ei:
        ctx.solve_batch(jobs)
    assert ei.value.code == chypot.CHY_EINVAL
    ctx.sync()
    assert float(jobs[0]["phi"].abs().max()) == 0.0   # nothing ran
    for pg in prepared.values():
        pg.close()
