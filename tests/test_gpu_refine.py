"""Refined solve (chy_solve_refined, chy_solve_batch_refined): iterative
refinement of the mode systems in the azimuthal spectrum -- binary64
residuals, binary32 correction solves on the FP32 matrix cores with binary64
r solves for the orders m <= 4 -- must reproduce the plain binary64 solve to
1e-10 (max relative PHI error), with the reference's boundary values, COEF
and diagnostics; the error is reported per refinement step.
"""
import numpy as np
import pytest

import chypot
from chypot import fields
from helpers import GOLDEN, MIXED_CASES, SMALL_CASES, assert_coef_close, load_case, rel_max

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return chypot.Context(0)


def refined(ctx, pg, g, rho_d, params, steps, phi_init=0.0):
    phi = torch.full_like(rho_d, phi_init)
    coef = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64, device="cuda")
    d, rel = pg.solve_refined(rho_d, phi, steps=steps, maxtrm=params["maxtrm"], isym=params["isym"],
                              redge=params["redge"], npoint=params["npoint"], coef=coef,
                              coef_mode=chypot.COEF_FULL, diag=True, report=True)
    ctx.sync()
    return chypot.to_host(phi), np.asfortranarray(coef.cpu().numpy().transpose(3, 2, 1, 0)), d, rel


@pytest.mark.parametrize("name", SMALL_CASES + MIXED_CASES)
def test_refined_matches_reference(ctx, name):
    g, rho, params, f = load_case(name)
    pg = chypot.PreparedGrid(ctx, g, isym=params["isym"])
    rho_d = chypot.to_device(rho)
    # garbage in the interior of PHI must not matter
    phi, coef, d, rel = refined(ctx, pg, g, rho_d, params, steps=3, phi_init=float("nan"))
    print(f"[gpu] refined {name}: relative corrections {rel}", flush=True)
    pg.close()
    phi[:, g.kmax + 1, :] = 0.0          # cells POT3 leaves alone (NaN here)
    phi[g.jmax + 1, 0, :] = 0.0
    assert rel_max(phi, f["phi"]) < 1e-10
    assert rel[1] < 1e-5 and rel[2] < 1e-9
    assert abs(d.tmass - float(f["tmass_b"])) <= 1e-13 * abs(float(f["tmass_b"]))
    assert abs(d.phichk - float(f["phichk"])) < 1e-8
    assert_coef_close(coef[:, 1], f["coef_k2"])


@pytest.mark.parametrize("name", SMALL_CASES + MIXED_CASES)
def test_refined_two_steps_per_stage_matches_reference(ctx, name):
    """Two steps by the kernel-per-stage form, whose one stored iterate is
    kept in binary32 (rows of JMAX-1 not a multiple of 4 end in partial
    vectors), against the reference's whole solve, COEF and diagnostics."""
    g, rho, params, f = load_case(name)
    pg = chypot.PreparedGrid(ctx, g, isym=params["isym"])
    pg.refine_fused = False
    rho_d = chypot.to_device(rho)
    phi, coef, d, rel = refined(ctx, pg, g, rho_d, params, steps=2, phi_init=float("nan"))
    print(f"[gpu] refined 2 steps per stage {name}: relative corrections {rel}", flush=True)
    pg.close()
    phi[:, g.kmax + 1, :] = 0.0          # cells POT3 leaves alone (NaN here)
    phi[g.jmax + 1, 0, :] = 0.0
    assert rel_max(phi, f["phi"]) < 1e-10
    assert rel[1] < 1e-5
    assert abs(d.tmass - float(f["tmass_b"])) <= 1e-13 * abs(float(f["tmass_b"]))
    assert abs(d.phichk - float(f["phichk"])) < 1e-8
    assert_coef_close(coef[:, 1], f["coef_k2"])


@pytest.mark.parametrize("name", SMALL_CASES + MIXED_CASES)
def test_refined_two_steps_matches_reference(ctx, name):
    """Two steps by the one-kernel-per-plane form against the reference's
    whole solve; the interior of PHI starts as NaN."""
    g, rho, params, f = load_case(name)
    pg = chypot.PreparedGrid(ctx, g, isym=params["isym"])
    pg.refine_fused = True
    rho_d = chypot.to_device(rho)
    phi = torch.full_like(rho_d, float("nan"))
    pg.solve_refined(rho_d, phi, steps=2, maxtrm=params["maxtrm"], isym=params["isym"], redge=params["redge"],
                     npoint=params["npoint"])
    ctx.sync()
    pg.close()
    phi = chypot.to_host(phi)
    phi[:, g.kmax + 1, :] = 0.0          # cells POT3 leaves alone (NaN here)
    phi[g.jmax + 1, 0, :] = 0.0
    err = rel_max(phi, f["phi"])
    print(f"[gpu] refined 2 steps {name}: vs reference {err:.2e}", flush=True)
    assert err < 1e-10


@pytest.mark.parametrize("steps", [2, 3])
def test_refined_256_against_plain_and_reference(ctx, steps):
    """The benchmark grid: the refined solve against the plain binary64 GPU
    solve (the 1e-10 criterion) and against the reference's planes."""
    g = fields.unit_grid(256, 128, 256)
    rho = fields.density_spiral_torus(g, seed=2)
    f = np.load(GOLDEN / "solve_large_256x128x256.npz")
    pg = chypot.PreparedGrid(ctx, g)
    rho_d = chypot.to_device(rho)
    plain = torch.zeros_like(rho_d)
    pg.solve(rho_d, plain)
    ctx.sync()
    plain = chypot.to_host(plain)
    phi, _, d, rel = refined(ctx, pg, g, rho_d, dict(maxtrm=10, isym=2, redge=0.0, npoint=8), steps)
    pg.close()
    e_plain = rel_max(phi, plain)
    scale = float(f["phi_absmax"])
    e_ref = max(rel_max(phi[:, 1, :], f["phi_k2"], scale), rel_max(phi[:, :, 0], f["phi_l1"], scale))
    print(f"[gpu] refined 256x128x256, {steps} steps: corrections {rel}, vs plain {e_plain:.2e}, "
          f"vs reference {e_ref:.2e}", flush=True)
    assert e_plain < 1e-10
    assert e_ref < 1e-10


def _plain(ctx, pg, rho_d):
    out = torch.zeros_like(rho_d)
    pg.solve(rho_d, out)
    ctx.sync()
    return out


@pytest.mark.parametrize("density", ["small_torus", "spiral"])
def test_refined_error_per_step_256(ctx, density):
    """The small polytropic torus (and the spiral-arm torus) at
    256x128x256.  The error against the plain binary64 solve after each
    refinement step: the first step (binary32 solves) is ~1e-6, the second
    below 1e-10; then it settles at the distance between the exact solution
    of the mode systems and the plain path's cyclic reduction (its
    reference-faithful root tables), ~2e-12."""
    g = fields.unit_grid(256, 128, 256)
    rho = fields.density_small_torus(g, 1) if density == "small_torus" else fields.density_spiral_torus(g, 2)
    pg = chypot.PreparedGrid(ctx, g)
    assert pg.refine_m64 == 4
    rho_d = chypot.to_device(rho)
    plain = _plain(ctx, pg, rho_d)
    scale = float(plain.abs().max())
    errs = []
    for steps in range(1, 5):
        phi = torch.zeros_like(rho_d)
        pg.solve_refined(rho_d, phi, steps=steps)
        ctx.sync()
        errs.append(float((phi - plain).abs().max()) / scale)
    pg.close()
    print(f"[gpu] refined {density} 256x128x256: error vs plain FP64 per step " +
          " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert errs[0] < 1e-5
    assert errs[1] < 1e-10
    assert max(errs[2:]) < 1e-11


@pytest.mark.parametrize("jkl", [(256, 128, 256), (64, 32, 64), (100, 64, 32), (200, 128, 64), (40, 16, 16),
                                 (257, 8, 16)])
@pytest.mark.parametrize("steps", [1, 2])
def test_fused_planes_match_stage_kernels(ctx, jkl, steps):
    """The one-kernel-per-mode-plane form of the refined solve (the binary32
    plane in LDS) against the kernel-per-stage form on the same grid: the
    same products and residual; the binary32 r solves differ in method (one
    lane per system with the grid's LU factors, against parallel cyclic
    reduction across a wavefront), which moves the ~1e-7 first step, so one
    step agrees to that and two steps to rounding.  Shapes: every (PP, JP)
    the fused kernel is built for, JMAX-1 = JP (the r system filling every
    lane's rows) and KMAX = 8."""
    g = fields.unit_grid(*jkl)
    rho_d = chypot.to_device(fields.density_spiral_torus(g, seed=21))
    pg = chypot.PreparedGrid(ctx, g)
    out = {}
    for fused in (True, False):
        pg.refine_fused = fused
        assert pg.refine_fused == (32 if fused else 0)
        phi = torch.zeros_like(rho_d)
        pg.solve_refined(rho_d, phi, steps=steps)
        ctx.sync()
        out[fused] = phi
    pg.close()
    scale = float(out[False].abs().max())
    err = float((out[True] - out[False]).abs().max()) / scale
    print(f"[gpu] fused planes {g.name} steps {steps}: vs stage kernels {err:.2e}", flush=True)
    assert err < (1e-5 if steps == 1 else 2e-11)


def test_batched_refined_equals_single(ctx):
    """chy_solve_batch_refined over fields on identical grids (one launch per
    kernel) and on another grid: each field bitwise its own refined solve."""
    gs = [fields.unit_grid(64, 32, 64)] * 3 + [fields.unit_grid(32, 16, 64)]
    pgs = [chypot.PreparedGrid(ctx, g) for g in gs]
    rhos = [chypot.to_device(fields.density_spiral_torus(g, seed=10 + i)) for i, g in enumerate(gs)]
    jobs = [dict(grid=pg, rho=r, phi=torch.zeros_like(r),
                 coef=torch.zeros(pg.coef_shape, dtype=torch.float64, device="cuda"), coef_mode=chypot.COEF_K2)
            for pg, r in zip(pgs, rhos)]
    diags = ctx.solve_batch(jobs, diag=True, refine_steps=2)
    for i, (pg, j) in enumerate(zip(pgs, jobs)):
        phi = torch.zeros_like(j["rho"])
        coef = torch.zeros_like(j["coef"])
        d, _ = pg.solve_refined(j["rho"], phi, steps=2, coef=coef, coef_mode=chypot.COEF_K2, diag=True)
        ctx.sync()
        assert torch.equal(phi, j["phi"]), i
        assert torch.equal(coef, j["coef"]), i
        assert d.tmass == diags[i].tmass and d.klocat == diags[i].klocat
    for pg in pgs:
        pg.close()


@pytest.mark.parametrize("jkl", [(300, 16, 32), (600, 8, 24), (1500, 8, 16)])
def test_refined_wide_grids_against_oracle(ctx, jkl):
    """JMAX above 257: the residual's columns span two or three blocks of
    256 (the last one partial), the r solves take more rows per lane; LMAX 24
    adds a mixed-radix transform.  Two steps against the oracle's binary64
    solve."""
    import pyoracle as po
    po.use_extended(False)
    g = fields.unit_grid(*jkl)
    rho = fields.density_spiral_torus(g, seed=9)
    print(f"[gpu] refined wide grid {g.name}", flush=True)
    want = po.solve(g, rho)["phi"]
    pg = chypot.PreparedGrid(ctx, g)
    rho_d = chypot.to_device(rho)
    phi, _, _, rel = refined(ctx, pg, g, rho_d, dict(maxtrm=10, isym=2, redge=0.0, npoint=8), 2)
    pg.close()
    err = rel_max(phi, want)
    print(f"[gpu] refined wide grid {g.name}: corrections {rel}, vs oracle {err:.2e}", flush=True)
    assert err < 1e-10


@pytest.mark.parametrize("jkl,steps", [((64, 256, 32), 2), ((64, 256, 32), 3), ((130, 256, 48), 2)])
def test_refined_kmax256_against_oracle(ctx, jkl, steps):
    """KMAX = 256: the binary32 z products on 16 wavefronts of 16 rows each
    (16-column tiles); 255 z rows in the residual.  Against the oracle's
    binary64 solve, every step's correction printed."""
    import pyoracle as po
    po.use_extended(False)
    g = fields.unit_grid(*jkl)
    rho = fields.density_spiral_torus(g, seed=11)
    print(f"[gpu] refined KMAX=256 {g.name}, {steps} steps", flush=True)
    want = po.solve(g, rho)["phi"]
    pg = chypot.PreparedGrid(ctx, g)
    rho_d = chypot.to_device(rho)
    phi, _, _, rel = refined(ctx, pg, g, rho_d, dict(maxtrm=10, isym=2, redge=0.0, npoint=8), steps)
    pg.close()
    err = rel_max(phi, want)
    print(f"[gpu] refined KMAX=256 {g.name}: corrections {rel}, vs oracle {err:.2e}", flush=True)
    assert err < 1e-10


def test_refined_needs_kmax_at_most_256(ctx):
    g = fields.unit_grid(16, 512, 16)
    pg = chypot.PreparedGrid(ctx, g)
    rho = torch.zeros(pg.field_shape, dtype=torch.float64, device="cuda")
    with pytest.raises(chypot.ChyError) as ei:
        pg.solve_refined(rho, torch.zeros_like(rho), steps=2)
    assert ei.value.code == chypot.CHY_EUNSUPPORTED
    pg.close()


def test_refined_rejects_bad_steps(ctx):
    g = fields.unit_grid(16, 8, 16)
    pg = chypot.PreparedGrid(ctx, g)
    rho = torch.zeros((16, 10, 18), dtype=torch.float64, device="cuda")
    for steps in (0, 9):
        with pytest.raises(chypot.ChyError) as ei:
            pg.solve_refined(rho, torch.zeros_like(rho), steps=steps)
        assert ei.value.code == chypot.CHY_EINVAL
    pg.close()


def test_autotune_small_grid(ctx):
    """The autotuner on a small grid: it picks candidates the grid accepts,
    applies them, and the tuned refined solve still meets the 1e-10
    criterion against the plain solve for one field and a batch."""
    from chypot import autotune
    g = fields.unit_grid(64, 32, 64)
    table = autotune.tune(ctx, [g], batch_sizes=(1, 4), refine_steps=2, reps=2)
    e = table["64x32x64"]
    print(f"[gpu] autotune 64x32x64: {e}", flush=True)
    assert e["cr_lanes"] in (64, 32) and 0 <= e["refine_fused"] <= 4
    pgs = [chypot.PreparedGrid(ctx, g) for _ in range(4)]
    for pg in pgs:
        assert autotune.apply(pg, table)
        assert pg.cr_lanes == e["cr_lanes"] and pg.refine_fused == e["refine_fused"]
    rhos = [chypot.to_device(fields.density_spiral_torus(g, seed=40 + i)) for i in range(4)]
    jobs = [dict(grid=pg, rho=r, phi=torch.zeros_like(r)) for pg, r in zip(pgs, rhos)]
    ctx.solve_batch(jobs[:1], refine_steps=2)
    ctx.solve_batch(jobs[1:], refine_steps=2)
    for j in jobs:
        plain = torch.zeros_like(j["rho"])
        j["grid"].solve(j["rho"], plain)
        ctx.sync()
        assert float((j["phi"] - plain).abs().max()) / float(plain.abs().max()) < 1e-10
    for pg in pgs:
        pg.close()
