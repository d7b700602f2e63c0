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


def _own_grid_job(ctx, seed, shape, coef_mode=chypot.COEF_K2, maxtrm=10, npoint=8, pg=None):
    """A job on a grid object of its own (or on `pg`), prepared from the
    description of `shape`; fields on such objects share launches."""
    g, rho = fields.random_field(seed, tuple((n,) for n in shape))
    pg = pg if pg is not None else chypot.PreparedGrid(ctx, g)
    z = dict(dtype=torch.float64, device="cuda")
    coef = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), **z) if coef_mode else None
    return dict(grid=pg, g=g, rho_h=rho, rho=chypot.to_device(rho),
                phi=torch.zeros((g.lmax, g.kmax + 2, g.jmax + 2), **z), coef=coef,
                coef_mode=coef_mode, maxtrm=maxtrm, npoint=npoint, redge=0.0)


def test_batched_launches_equal_single_solves(ctx):
    """Fields on separate grid objects with one description run as one launch
    per kernel (up to 32 fields): 34 such fields (two launch groups), fields of
    a second shape, a repeat on a grid object already in the batch (a later
    round), fields whose arguments keep them out of the groups (no COEF,
    other MAXTRM) and two fields whose NPOINT agree only once lowered to 10 --
    every one equal bit for bit to its single chy_solve, PHAC included."""
    print("[gpu] batched launches: 34+3 fields on own grid objects, repeats, odd arguments", flush=True)
    a, b = (32, 16, 64), (16, 32, 32)
    jobs = [_own_grid_job(ctx, 300 + i, a) for i in range(34)]
    jobs += [_own_grid_job(ctx, 400 + i, b) for i in range(3)]
    jobs.insert(5, _own_grid_job(ctx, 500, a, pg=jobs[0]["grid"]))          # same object: next round
    jobs.insert(9, _own_grid_job(ctx, 501, a, coef_mode=chypot.COEF_NONE))
    jobs.insert(13, _own_grid_job(ctx, 502, a, maxtrm=6))
    jobs.insert(2, _own_grid_job(ctx, 503, a, npoint=12))                   # lowered to 10 ...
    jobs.insert(17, _own_grid_job(ctx, 504, a, npoint=10))                  # ... so it shares this launch
    grids = {id(jb["grid"]): jb["grid"] for jb in jobs}
    try:
        alone = []
        for jb in jobs:
            phi = torch.zeros_like(jb["phi"])
            coef = torch.zeros_like(jb["coef"]) if jb["coef"] is not None else None
            d = jb["grid"].solve(jb["rho"], phi, maxtrm=jb["maxtrm"], npoint=jb["npoint"], coef=coef,
                                 coef_mode=jb["coef_mode"], diag=True)
            alone.append((phi.cpu(), None if coef is None else coef.cpu(), (d.tmass, d.phichk, d.klocat),
                          jb["grid"].phac()))
        for workers in (4, 1):
            ctx.workers(workers)
            for jb in jobs:
                jb["phi"].zero_()
                if jb["coef"] is not None:
                    jb["coef"].zero_()
            diags = ctx.solve_batch(jobs, diag=True)
            for i, (jb, (phi, coef, d, phac), db) in enumerate(zip(jobs, alone, diags)):
                assert torch.equal(jb["phi"].cpu(), phi), f"PHI of item {i}, {workers} workers"
                if coef is not None:
                    assert torch.equal(jb["coef"].cpu(), coef), f"COEF of item {i}"
                assert (db.tmass, db.phichk, db.klocat) == d, f"diagnostics of item {i}"
                if jb["grid"] is not jobs[0]["grid"]:     # PHAC is the grid object's last solve's
                    assert np.array_equal(jb["grid"].phac(), phac), f"PHAC of item {i}"
            # asynchronous form: the same numbers once the context's stream is joined
            for jb in jobs:
                jb["phi"].zero_()
            ctx.solve_batch(jobs)
            ctx.sync()
            for i, (jb, (phi, _, _, _)) in enumerate(zip(jobs, alone)):
                assert torch.equal(jb["phi"].cpu(), phi), f"PHI of item {i} (async), {workers} workers"
        # the launches the batch made: the 34 fields of shape a with NPOINT 8
        # take two launches (32 + 2), the two with NPOINT 12 (lowered to 10)
        # and 10 one, the fields of shape b one, the fields without COEF and
        # with MAXTRM 6 one each, and the repeat on the first grid object one
        ctx.profile(True)
        try:
            for jb in jobs:
                jb["phi"].zero_()
            ctx.solve_batch(jobs)
            ctx.sync()
            launches = ctx.stage_times()
        finally:
            ctx.profile(False)
        assert launches["fft_fwd"][1] == 7 and launches["ghost"][1] == 7, launches
        for i, (jb, (phi, _, _, _)) in enumerate(zip(jobs, alone)):
            assert torch.equal(jb["phi"].cpu(), phi), f"PHI of item {i} (stage timing on)"
    finally:
        ctx.workers(4)
        for pg in grids.values():
            pg.close()


def test_batched_launches_match_oracle(ctx):
    import pyoracle as po
    po.use_extended(False)
    print("[gpu] batched launches: 4 fields on own grid objects vs the oracle", flush=True)
    jobs = [_own_grid_job(ctx, 600 + i, (32, 16, 64)) for i in range(4)]
    try:
        diags = ctx.solve_batch(jobs, diag=True)
        for jb, d in zip(jobs, diags):
            want = po.solve(jb["g"], jb["rho_h"])
            assert rel_max(chypot.to_host(jb["phi"]), want["phi"]) < 1e-12
            assert abs(d.tmass - want["tmass"]) <= 1e-13 * abs(want["tmass"])
            assert d.klocat == want["klocat"]
    finally:
        for jb in jobs:
            jb["grid"].close()


def test_64_heterogeneous_fields_refined(ctx):
    """The heterogeneous batch: 64 independent fields drawn from a
    seed (JMAX in {64,128,256}, KMAX in {32,64,128}, LMAX in {64,128,256},
    tori of random centre, radius and index with log-normal noise), each on
    its own grid object, solved in one chy_solve_batch_refined call (fields
    on identical grids share launches).  Every field against its own plain
    FP64 solve (1e-10), a sample bitwise against its own refined solve, and
    the smallest fields against the oracle's solve."""
    import pyoracle as po
    jobs = []
    for i in range(64):
        g, rho_d = fields.random_field_device(1000 + i)
        pg = chypot.PreparedGrid(ctx, g)
        jobs.append(dict(grid=pg, g=g, rho=rho_d, phi=torch.zeros_like(rho_d),
                         coef=torch.zeros(pg.coef_shape, dtype=torch.float64, device="cuda"),
                         coef_mode=chypot.COEF_K2))
    shapes = {(j["g"].jmax, j["g"].kmax, j["g"].lmax) for j in jobs}
    print(f"[gpu] heterogeneous batch: 64 fields on {len(shapes)} distinct grids", flush=True)
    ctx.solve_batch(jobs, refine_steps=2)
    ctx.sync()
    worst = 0.0
    for i, j in enumerate(jobs):
        plain = torch.zeros_like(j["rho"])
        j["grid"].solve(j["rho"], plain)
        ctx.sync()
        e = float((j["phi"] - plain).abs().max() / plain.abs().max())
        worst = max(worst, e)
        if i % 8 == 0:
            alone = torch.zeros_like(j["rho"])
            j["grid"].solve_refined(j["rho"], alone, steps=2)
            ctx.sync()
            assert torch.equal(alone, j["phi"]), i
    print(f"[gpu] heterogeneous batch: worst refined vs plain FP64 {worst:.2e}", flush=True)
    assert worst < 1e-10
    small = sorted(range(64), key=lambda i: jobs[i]["g"].jmax * jobs[i]["g"].kmax * jobs[i]["g"].lmax)[:3]
    for i in small:
        g = jobs[i]["g"]
        rho = chypot.to_host(jobs[i]["rho"])
        want = po.solve(g, rho)
        assert rel_max(chypot.to_host(jobs[i]["phi"]), want["phi"]) < 1e-10, g.name
    for j in jobs:
        j["grid"].close()


@pytest.mark.parametrize("steps", [0, 2])
def test_group_graph_replays_equal_eager(ctx, steps):
    """Multi-field groups replay from captured HIP graphs: four calls of the
    same batch with graphs on (capture, then three replays), with PHI cleared
    before each and an eager batch into other buffers after each, must each
    equal the eager batch bit for bit; a group that changes an array is
    captured anew."""
    g = fields.unit_grid(64, 32, 64)
    pgs = [chypot.PreparedGrid(ctx, g) for _ in range(4)] + [chypot.PreparedGrid(ctx, fields.unit_grid(32, 16, 64))]
    rhos = [chypot.to_device(fields.density_spiral_torus(pg.grid, seed=60 + i)) for i, pg in enumerate(pgs)]

    def batch(phis):
        return [dict(grid=pg, rho=r, phi=p, coef=torch.zeros(pg.coef_shape, dtype=torch.float64, device="cuda"),
                     coef_mode=chypot.COEF_K2) for pg, r, p in zip(pgs, rhos, phis)]

    ctx.graphs(False)
    want = [torch.zeros_like(r) for r in rhos]
    ctx.solve_batch(batch(want), refine_steps=steps)
    ctx.sync()
    ctx.graphs(True)
    phis = [torch.zeros_like(r) for r in rhos]
    jobs = batch(phis)
    for rep in range(4):
        for p in phis:
            p.zero_()
        ctx.solve_batch(jobs, refine_steps=steps)
        ctx.sync()
        for i, (p, w) in enumerate(zip(phis, want)):
            assert torch.equal(p, w), (rep, i)
        # eager work between the replays
        ctx.graphs(False)
        other = [torch.zeros_like(r) for r in rhos]
        ctx.solve_batch(batch(other), refine_steps=steps)
        ctx.sync()
        ctx.graphs(True)
    # another PHI array for one field: a new capture, still the eager result
    phis2 = list(phis)
    phis2[1] = torch.zeros_like(rhos[1])
    ctx.solve_batch(batch(phis2), refine_steps=steps)
    ctx.sync()
    assert torch.equal(phis2[1], want[1])
    for pg in pgs:
        pg.close()


@pytest.mark.parametrize("steps", [0, 2])
def test_batched_mixed_radix_fields_equal_single_solves(ctx, steps):
    """Fields whose LMAX/2 is not a power of two (the mixed-radix transforms)
    in batched launches beside a power-of-two shape: five fields at LMAX 40
    (20 = 4 x 5) and three at LMAX 48 (24 = 8 x 3) on their own grid objects,
    two at LMAX 64, in one batch, plain and refined; each equal bit for bit to
    the same field solved alone."""
    print(f"[gpu] batched mixed-radix fields, refine steps {steps}", flush=True)
    jobs = [_own_grid_job(ctx, 700 + i, (24, 16, 40)) for i in range(5)]
    jobs += [_own_grid_job(ctx, 710 + i, (32, 16, 48)) for i in range(3)]
    jobs += [_own_grid_job(ctx, 720 + i, (32, 16, 64)) for i in range(2)]
    try:
        alone = []
        for jb in jobs:
            phi = torch.zeros_like(jb["phi"])
            coef = torch.zeros_like(jb["coef"])
            if steps:
                jb["grid"].solve_refined(jb["rho"], phi, steps=steps, coef=coef, coef_mode=jb["coef_mode"])
            else:
                jb["grid"].solve(jb["rho"], phi, coef=coef, coef_mode=jb["coef_mode"])
            ctx.sync()
            alone.append((phi.cpu(), coef.cpu()))
        ctx.solve_batch(jobs, refine_steps=steps)
        ctx.sync()
        for i, (jb, (phi, coef)) in enumerate(zip(jobs, alone)):
            assert torch.equal(jb["phi"].cpu(), phi), f"PHI of item {i}"
            assert torch.equal(jb["coef"].cpu(), coef), f"COEF of item {i}"
    finally:
        for jb in jobs:
            jb["grid"].close()


def test_worker_counts_graph_caches_and_teardown(ctx):
    """Changing the number of worker streams between batches, with graphs of
    single fields and of multi-field groups cached on the streams used before,
    a batch with stage timing, then closing every grid and the context: every
    batch equals the eager one, and the caller's next GPU call finds no error
    left behind (a failed wait in the cache clean-up must not surface there)."""
    print("[gpu] worker counts, graph caches and teardown", flush=True)
    c = chypot.Context(0)
    jobs = [_own_grid_job(c, 800 + i, (32, 16, 64)) for i in range(3)]
    jobs += [_own_grid_job(c, 810, (16, 32, 32)), _own_grid_job(c, 811, (24, 16, 40))]
    try:
        for steps in (2, 0):
            c.graphs(False)
            c.solve_batch(jobs, refine_steps=steps)
            c.sync()
            want = [jb["phi"].clone() for jb in jobs]
            c.graphs(True)
            for w in (8, 4, 16, 8, 1, 8):
                c.workers(w)
                for rep in range(2):
                    for jb in jobs:
                        jb["phi"].zero_()
                    c.solve_batch(jobs, refine_steps=steps)
                    c.sync()
                    for i, (jb, wv) in enumerate(zip(jobs, want)):
                        assert torch.equal(jb["phi"], wv), (steps, w, rep, i)
            c.profile(True)
            c.solve_batch(jobs, refine_steps=steps)
            c.sync()
            assert c.stage_times()["fft_fwd"][1] > 0
            c.profile(False)
    finally:
        for jb in jobs:
            jb["grid"].close()
        c.close()
    x = torch.ones(4096, dtype=torch.float64, device="cuda")
    assert float(x.sum()) == 4096.0
    torch.cuda.synchronize()


def test_batch_graph_replays_with_diagnostics(ctx):
    """A batch with diagnostics replayed from its captured graph several
    times (the per-item diagnostics go through a copy node of the graph),
    with eager work between the replays: every replay's PHI and diagnostics
    equal the eager batch's."""
    print("[gpu] batch graph replays with diagnostics", flush=True)
    jobs = [_own_grid_job(ctx, 900 + i, (32, 16, 64)) for i in range(3)]
    jobs += [_own_grid_job(ctx, 910, (24, 16, 40)), _own_grid_job(ctx, 911, (16, 32, 32))]
    try:
        ctx.graphs(False)
        want_d = ctx.solve_batch(jobs, diag=True)
        want = [jb["phi"].clone() for jb in jobs]
        ctx.graphs(True)
        for rep in range(4):
            for jb in jobs:
                jb["phi"].zero_()
            got_d = ctx.solve_batch(jobs, diag=True)
            for i, (jb, w, dw, dg) in enumerate(zip(jobs, want, want_d, got_d)):
                assert torch.equal(jb["phi"], w), (rep, i)
                assert (dg.tmass, dg.phichk, dg.klocat) == (dw.tmass, dw.phichk, dw.klocat), (rep, i)
            # eager work between the replays, into other buffers
            ctx.graphs(False)
            other = [dict(jb, phi=torch.zeros_like(jb["phi"])) for jb in jobs]
            ctx.solve_batch(other)
            ctx.sync()
            ctx.graphs(True)
    finally:
        ctx.graphs(True)
        for jb in jobs:
            jb["grid"].close()
