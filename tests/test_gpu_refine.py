"""Refined solve (chy_solve_refined): binary32 correction solves on a binary64
residual must reproduce the plain binary64 solve to 1e-10 (max relative
PHI error), with the reference's boundary values, COEF and diagnostics.
"""
import numpy as np
import pytest

import chypot
from chypot import fields
from helpers import GOLDEN, SMALL_CASES, load_case, phase_diff, rel_max

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


@pytest.mark.parametrize("name", SMALL_CASES)
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
    ck, rk = coef[:, 1], f["coef_k2"]
    assert np.abs(ck[..., 0] - rk[..., 0]).max() < 1e-11
    sig = rk[..., 0] > 1e-8
    assert np.abs(phase_diff(ck[..., 1], rk[..., 1])[sig]).max() < 1e-6


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


def test_refined_rejects_bad_steps(ctx):
    g = fields.unit_grid(16, 8, 16)
    pg = chypot.PreparedGrid(ctx, g)
    rho = torch.zeros((16, 10, 18), dtype=torch.float64, device="cuda")
    for steps in (0, 9):
        with pytest.raises(chypot.ChyError) as ei:
            pg.solve_refined(rho, torch.zeros_like(rho), steps=steps)
        assert ei.value.code == chypot.CHY_EINVAL
    pg.close()
