"""GPU parity of the potential solve against the reference's golden vectors.

Every test calls the HIP library through its C-ABI (chypot ctypes bindings)
and compares with fixtures written by the compiled reference
(tests/golden/make_golden.py) or with the oracle under oracle/.
"""
import numpy as np
import pytest

import chypot
from chypot import fields
from helpers import GOLDEN, MIXED_CASES, SMALL_CASES, assert_coef_close, load_case, phase_diff, rel_max

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return chypot.Context(0)


SOLVERS = {"cr": chypot.MODES_CR, "eigen": chypot.MODES_EIGEN}
# PHI against the reference's FP64 solve: the cyclic reduction follows the
# reference's own arithmetic (BLKTRI's root tables), the z eigenbasis is
# closer to the exact solution and within 1e-12 on these grids
PHI_BAR = {"cr": 1e-13, "eigen": 1e-12}


def run_solve(ctx, g, rho, params, coef_mode=chypot.COEF_FULL, phi_init=0.0, solver="cr"):
    pg = chypot.PreparedGrid(ctx, g, isym=params["isym"])
    pg.mode_solver = SOLVERS[solver]
    rho_d = chypot.to_device(rho)
    phi_d = torch.full_like(rho_d, phi_init)
    coef = None
    if coef_mode != chypot.COEF_NONE:
        coef = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64, device="cuda")
    pg.bdygen(rho_d, phi_d, maxtrm=params["maxtrm"], isym=params["isym"], redge=params["redge"])
    ctx.sync()
    phib = chypot.to_host(phi_d)
    d_b = pg.bdygen(rho_d, phi_d, maxtrm=params["maxtrm"], isym=params["isym"], redge=params["redge"],
                    diag=True)
    d = pg.pot3(rho_d, phi_d, npoint=params["npoint"], coef=coef, coef_mode=coef_mode, diag=True)
    ctx.sync()
    out = dict(phi=chypot.to_host(phi_d), phi_b=phib, tmass=d_b.tmass, phichk=d.phichk, klocat=d.klocat)
    if coef is not None:
        out["coef"] = np.asfortranarray(coef.cpu().numpy().transpose(3, 2, 1, 0))
    pg.close()
    return out


@pytest.mark.parametrize("solver", list(SOLVERS))
@pytest.mark.parametrize("name", SMALL_CASES + MIXED_CASES)
def test_solve_matches_reference(ctx, name, solver):
    g, rho, params, f = load_case(name)
    print(f"[gpu] solve {name} {g.name} {params} {solver}", flush=True)
    res = run_solve(ctx, g, rho, params, solver=solver)
    # boundary values from BDYGEN
    top = res["phi_b"][1:g.jmax + 1, g.kmax, :]
    out = res["phi_b"][g.jmax, 1:g.kmax, :]
    assert rel_max(top, f["phi_b_top"]) < 1e-13
    assert rel_max(out, f["phi_b_out"]) < 1e-13
    assert abs(res["tmass"] - float(f["tmass_b"])) <= 1e-13 * abs(float(f["tmass_b"]))
    # the potential, every cell POT3 writes (and the ones it leaves alone)
    assert rel_max(res["phi"], f["phi"]) < PHI_BAR[solver]
    # ZAXPHI diagnostics: PHICHK is a difference of nearly equal numbers
    assert abs(res["phichk"] - float(f["phichk"])) < 1e-10
    assert res["klocat"] == int(f["klocat"])
    assert_coef_close(res["coef"][:, 1, :, :], f["coef_k2"])
    if "coef" in f.files:
        assert_coef_close(res["coef"], f["coef"])


@pytest.mark.parametrize("npoint", range(1, 11))
def test_every_npoint_against_oracle(ctx, npoint):
    """ZAXPHI's extrapolation is compiled once per NPOINT (1..10): each against
    the oracle's POT3 with the same NPOINT on a 24x8x16 spiral torus -- PHI
    everywhere POT3 writes it (the axis column PHI(JMAX+2,K,L) is ZAXPHI's),
    PHICHK and KLOCAT."""
    import pyoracle as po
    po.use_extended(False)
    g = fields.unit_grid(24, 8, 16)
    rho = fields.density_spiral_torus(g, seed=5)
    print(f"[gpu] npoint {npoint} {g.name}", flush=True)
    res = run_solve(ctx, g, rho, dict(isym=2, maxtrm=10, redge=0.0, npoint=npoint), coef_mode=chypot.COEF_NONE)
    want = po.solve(g, rho, 10, 2, 0.0, npoint)
    assert rel_max(res["phi"], want["phi"]) < 1e-12
    assert abs(res["phichk"] - want["phichk"]) < 1e-10
    assert res["klocat"] == want["klocat"]


def test_zero_density_keeps_klocat(ctx):
    """Zero density: PHI stays zero (no NaN from ZAXPHI's 0/0), PHICHK is zero
    and KLOCAT keeps the value it had, as in the oracle (pot/pot3.f:380-399)."""
    g = fields.unit_grid(24, 8, 16)
    print(f"[gpu] zero density {g.name}", flush=True)
    pg = chypot.PreparedGrid(ctx, g)
    rho_d = chypot.to_device(np.zeros(g.shape, order="F"))
    phi_d = torch.zeros_like(rho_d)
    pg.bdygen(rho_d, phi_d, maxtrm=10, isym=2, redge=0.0)
    d = pg.pot3(rho_d, phi_d, npoint=8, diag=chypot.Diag(0.0, 0.0, -3))
    ctx.sync()
    phi = chypot.to_host(phi_d)
    pg.close()
    assert not np.abs(phi).any()
    assert (d.phichk, d.klocat) == (0.0, -3)


@pytest.mark.parametrize("path", ["pot3", "solve"])
@pytest.mark.parametrize("mode", ["full", "k2"])
def test_coef_where_a0_is_zero_against_oracle(ctx, path, mode):
    """COEF on a torus with no density on J=2..9 and COEF prefilled with 5:
    where a0 = 0 the amplitudes keep the old 5, the phases are written (720
    degrees where a = b = 0) and COEF(J,K,N,2) = a0 = 0, as the oracle does
    (pot/pot3.f:105-132); in K2 mode only the K=2 plane is written."""
    import pyoracle as po
    po.use_extended(False)
    g = fields.unit_grid(24, 8, 16)
    rho = fields.density_spiral_torus(g, seed=5)
    rho[:9] = 0.0
    rho = np.asfortranarray(rho)
    print(f"[gpu] COEF a0=0 {g.name} {path} {mode}", flush=True)
    phi_b, _ = po.bdygen(g, rho)
    _, want, _, _ = po.pot3(g, rho, phi_b, 8, coef=np.full((g.jmax + 2, g.kmax + 2, g.lmax // 2, 2), 5.0,
                                                           order="F"))
    cm = chypot.COEF_FULL if mode == "full" else chypot.COEF_K2
    pg = chypot.PreparedGrid(ctx, g)
    rho_d = chypot.to_device(rho)
    phi_d = torch.zeros_like(rho_d)
    coef = torch.full((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), 5.0, dtype=torch.float64, device="cuda")
    if path == "pot3":
        pg.bdygen(rho_d, phi_d, maxtrm=10, isym=2, redge=0.0)
        pg.pot3(rho_d, phi_d, npoint=8, coef=coef, coef_mode=cm)
    else:
        pg.solve(rho_d, phi_d, maxtrm=10, isym=2, redge=0.0, npoint=8, coef=coef, coef_mode=cm)
    ctx.sync()
    got = np.asfortranarray(coef.cpu().numpy().transpose(3, 2, 1, 0))
    pg.close()
    if mode == "k2":
        rest = np.ones(g.kmax + 2, dtype=bool)
        rest[1] = False
        assert (got[:, rest] == 5.0).all()
        got, want = got[:, 1:2], want[:, 1:2]
    assert np.abs(got[..., 0] - want[..., 0]).max() < 1e-11
    exact = (want[..., 1] == 720.0) | (want[..., 0] == 5.0)
    assert exact[1:9].any()
    n = g.lmax // 2
    assert (got[..., :n - 1, 1][exact[..., :n - 1]] == want[..., :n - 1, 1][exact[..., :n - 1]]).all()
    assert np.abs(got[..., n - 1, 1] - want[..., n - 1, 1]).max() <= 1e-12 * np.abs(want[..., n - 1, 1]).max()
    sig = (want[..., :n - 1, 0] > 1e-8) & ~exact[..., :n - 1]
    assert np.abs(phase_diff(got[..., :n - 1, 1], want[..., :n - 1, 1])[sig]).max() < 1e-6


def test_untouched_cells_and_determinism(ctx):
    g, rho, params, f = load_case("g24x16x32_spiral")
    a = run_solve(ctx, g, rho, params, phi_init=7.25)
    b = run_solve(ctx, g, rho, params, phi_init=7.25)
    assert np.array_equal(a["phi"], b["phi"])
    # POT3 leaves the K=KMAX+2 row and PHI(JMAX+2,1,L) alone (pot3.f:287-298,377)
    assert np.all(a["phi"][:, g.kmax + 1, :] == 7.25)
    assert np.all(a["phi"][g.jmax + 1, 0, :] == 7.25)


def test_host_entry_points_match_device(ctx):
    g, rho, params, f = load_case("g32_redge")
    dev = run_solve(ctx, g, rho, params)
    pg = chypot.PreparedGrid(ctx, g, isym=params["isym"])
    phi = np.zeros(g.shape, order="F")
    coef = np.zeros((g.jmax + 2, g.kmax + 2, g.lmax // 2, 2), order="F")
    pg.bdygen_host(rho, phi, maxtrm=params["maxtrm"], redge=params["redge"])
    d = pg.pot3_host(rho, phi, npoint=params["npoint"], coef=coef, coef_mode=chypot.COEF_FULL)
    assert np.array_equal(phi, dev["phi"])
    assert np.array_equal(coef, dev["coef"])
    assert d.klocat == dev["klocat"]


def _parity_256(ctx, solver):
    """The 256x128x256 spiral-arm torus: (GPU result, fixture, reference's max |PHI|)."""
    f = np.load(GOLDEN / "solve_large_256x128x256.npz")
    g = fields.unit_grid(256, 128, 256)
    rho = fields.density_spiral_torus(g, seed=2)
    res = run_solve(ctx, g, rho, dict(isym=2, maxtrm=10, npoint=8, redge=0.0), coef_mode=chypot.COEF_K2,
                    solver=solver)
    return g, rho, res, f, float(f["phi_absmax"])


def test_256x128x256_against_reference(ctx):
    """256x128x256, the spiral-arm torus: the cyclic reduction
    (default solver) within 1e-12 of the reference's FP64 solve, on the
    fixture's K=2 and L=1 planes and the whole-field statistics, and within
    5e-13 of the oracle's POT3 by the same cyclic reduction (bitwise BLKTRI)
    over the whole field (the GPU's r solves -- partitioned elimination and
    parallel cyclic reduction over the lanes -- round differently from
    BLKTRI's serial ones: ~2e-13 here)."""
    import pyoracle as po
    g, rho, res, f, scale = _parity_256(ctx, "cr")
    phi = res["phi"]
    e_ref = max(rel_max(phi[:, 1, :], f["phi_k2"], scale), rel_max(phi[:, :, 0], f["phi_l1"], scale))
    stats = {}
    # per-K and per-L max |PHI| and sums over every cell POT3 writes (K <= KMAX+1)
    wr = phi[:, :g.kmax + 1, :]
    stats["kmax"] = rel_max(np.abs(wr).max(axis=(0, 2)), f["phi_kmax"], scale)
    stats["lmax"] = rel_max(np.abs(wr).max(axis=(0, 1)), f["phi_lmax"], scale)
    stats["ksum"] = rel_max(wr.sum(axis=(0, 2)), f["phi_ksum"], scale * (g.jmax + 2) * g.lmax)
    stats["lsum"] = rel_max(wr.sum(axis=(0, 1)), f["phi_lsum"], scale * (g.jmax + 2) * (g.kmax + 1))
    print("[gpu] 256x128x256 oracle (cyclic reduction)", flush=True)
    po.use_cyclic_reduction(True)
    try:
        want = po.solve(g, rho)
    finally:
        po.use_cyclic_reduction(False)
    e_or = rel_max(phi, want["phi"], scale)
    print(f"[gpu] 256x128x256 cyclic reduction: vs reference planes {e_ref:.3e}, vs oracle (whole field) "
          f"{e_or:.3e}, stats {stats}", flush=True)
    assert e_ref < 1e-12
    assert e_or < 5e-13
    for k, v in stats.items():
        assert v < 1e-12, k


@pytest.mark.parametrize("lanes", [64, 32])
@pytest.mark.parametrize("jkl", [(128, 64, 128), (64, 32, 64), (200, 16, 32)])
def test_cr_lane_shapes_against_oracle(ctx, jkl, lanes):
    """Both r-system shapes of the cyclic reduction (64 lanes x R rows, or two
    orders per wavefront at 32 lanes x 2R rows) against the oracle's POT3 by
    BLKTRI's reduction."""
    import pyoracle as po
    g = fields.unit_grid(*jkl)
    rho = fields.density_spiral_torus(g, seed=3)
    pg = chypot.PreparedGrid(ctx, g)
    pg.cr_lanes = lanes
    assert pg.cr_lanes == lanes
    rho_d = chypot.to_device(rho)
    phi_d = torch.zeros_like(rho_d)
    pg.solve(rho_d, phi_d)
    ctx.sync()
    phi = chypot.to_host(phi_d)
    pg.close()
    po.use_cyclic_reduction(True)
    try:
        want = po.solve(g, rho)
    finally:
        po.use_cyclic_reduction(False)
    err = rel_max(phi, want["phi"])
    print(f"[gpu] cr lanes {lanes} {g.name}: vs oracle's reduction {err:.2e}", flush=True)
    assert err < 5e-13


def test_256x128x256_eigenbasis_against_exact(ctx):
    """256x128x256 by the z eigenbasis (opt-in fast solver): within 1.5e-12
    of the 80-bit solve of the discrete system; the reference's own FP64
    result is 2.3e-12 from that solve (its TQLRAT root tables), so the
    eigenbasis is 2-4e-12 from the reference."""
    g, rho, res, f, scale = _parity_256(ctx, "eigen")
    phi = res["phi"]
    e_ref = max(rel_max(phi[:, 1, :], f["phi_k2"], scale), rel_max(phi[:, :, 0], f["phi_l1"], scale))
    e_ext = max(rel_max(phi[:, 1, :], f["ext_phi_k2"], scale), rel_max(phi[:, :, 0], f["ext_phi_l1"], scale))
    print(f"[gpu] 256x128x256 eigenbasis: vs reference {e_ref:.3e}, vs 80-bit {e_ext:.3e}, "
          f"reference vs 80-bit {float(f['ext_ref_maxdiff']):.3e}", flush=True)
    assert e_ext < 1.5e-12
    assert e_ref < 5e-12


@pytest.mark.parametrize("jkl,code", [((16, 6, 16), chypot.CHY_EUNSUPPORTED),
                                      ((16, 8, 58), chypot.CHY_EUNSUPPORTED),
                                      ((16, 8, 25), chypot.CHY_EUNSUPPORTED),
                                      ((16, 8, 2048), chypot.CHY_EUNSUPPORTED),
                                      ((16, 8, 1026), chypot.CHY_EUNSUPPORTED),
                                      ((4100, 8, 16), chypot.CHY_EUNSUPPORTED),
                                      ((2, 8, 16), chypot.CHY_EINVAL)])
def test_rejected_grids(ctx, jkl, code):
    """Grids the drop-in refuses with a code instead of solving wrongly:
    KMAX not 2^k (BLKTRI's own precondition, pot/blktri.f:21-31); LMAX/2
    with a prime factor above 23 (58 = 2 x 29: the reference's FFT stops on
    it too, pot/fft.f:111-126); odd LMAX (REALTR needs LMAX = 2N); LMAX above
    1024 (the transforms keep a lane's spectrum in LDS); JMAX above 4096
    (the r solves' pivot table and padding grow with the next power of two);
    JMAX < 3."""
    g = fields.unit_grid(*jkl)
    with pytest.raises(chypot.ChyError) as ei:
        chypot.PreparedGrid(ctx, g)
    assert ei.value.code == code


def test_bad_arguments(ctx):
    g = fields.unit_grid(16, 8, 16)
    pg = chypot.PreparedGrid(ctx, g)
    rho = torch.zeros((16, 10, 18), dtype=torch.float64, device="cuda")
    phi = torch.zeros_like(rho)
    with pytest.raises(chypot.ChyError) as ei:
        pg.pot3(rho, phi, npoint=0)
    assert ei.value.code == chypot.CHY_EINVAL
    with pytest.raises(chypot.ChyError) as ei:
        pg.bdygen(rho, phi, maxtrm=11)
    assert ei.value.code == chypot.CHY_EINVAL
    with pytest.raises(chypot.ChyError) as ei:
        pg.bdygen(rho, phi, isym=1)
    assert ei.value.code == chypot.CHY_EINVAL


def fused_solve(ctx, g, rho_d, params, coef_mode=chypot.COEF_FULL, phi_init=0.0, pg=None, solver="cr"):
    own = pg is None
    if own:
        pg = chypot.PreparedGrid(ctx, g, isym=params["isym"])
        pg.mode_solver = SOLVERS[solver]
    phi_d = torch.full_like(rho_d, phi_init)
    coef = None
    if coef_mode != chypot.COEF_NONE:
        coef = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64, device="cuda")
    d = pg.solve(rho_d, phi_d, maxtrm=params["maxtrm"], isym=params["isym"], redge=params["redge"],
                 npoint=params["npoint"], coef=coef, coef_mode=coef_mode, diag=True)
    ctx.sync()
    out = dict(phi=chypot.to_host(phi_d), tmass=d.tmass, phichk=d.phichk, klocat=d.klocat)
    if coef is not None:
        out["coef"] = np.asfortranarray(coef.cpu().numpy().transpose(3, 2, 1, 0))
    if own:
        pg.close()
    return out


@pytest.mark.parametrize("solver", list(SOLVERS))
@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("name", SMALL_CASES)
def test_fused_solve_matches_reference(ctx, name, graphs, solver):
    """chy_solve (BDYGEN evaluated in the azimuthal spectrum when LMAX/2 > 10,
    through PHI otherwise) against the reference's BDYGEN+POT3."""
    g, rho, params, f = load_case(name)
    print(f"[gpu] fused solve {name} {g.name} graphs={graphs} {solver}", flush=True)
    ctx.graphs(graphs)
    try:
        res = fused_solve(ctx, g, chypot.to_device(rho), params, phi_init=3.5, solver=solver)
    finally:
        ctx.graphs(True)
    # the cells POT3 does not write keep their contents (the fixture holds 0 there)
    phi = res["phi"]
    assert np.all(phi[:, g.kmax + 1, :] == 3.5)
    assert np.all(phi[g.jmax + 1, 0, :] == 3.5)
    phi[:, g.kmax + 1, :] = 0.0
    phi[g.jmax + 1, 0, :] = 0.0
    assert rel_max(phi, f["phi"]) < PHI_BAR[solver]
    assert abs(res["tmass"] - float(f["tmass_b"])) <= 1e-13 * abs(float(f["tmass_b"]))
    assert abs(res["phichk"] - float(f["phichk"])) < 1e-10
    assert res["klocat"] == int(f["klocat"])
    assert_coef_close(res["coef"][:, 1], f["coef_k2"])
    if "coef" in f.files:
        assert_coef_close(res["coef"], f["coef"])


def test_solve_graph_replays_match_eager(ctx):
    """Replays of the captured solve graph, with eager solves, other GPU work
    and host waits between them, each equal bit for bit to the eager solve of
    the density the buffer holds at replay time; a call with another output
    buffer re-captures."""
    g, rho, params, f = load_case("c1_64x32x64")
    rho2 = fields.density_spiral_torus(g, seed=5)
    kw = dict(maxtrm=params["maxtrm"], isym=params["isym"], redge=params["redge"], npoint=params["npoint"])
    pg = chypot.PreparedGrid(ctx, g, isym=params["isym"])
    eager = {}
    ctx.graphs(False)
    for key, r in (("a", rho), ("b", rho2)):
        rd = chypot.to_device(r)
        ph = torch.zeros_like(rd)
        cf = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64, device="cuda")
        pg.solve(rd, ph, coef=cf, coef_mode=chypot.COEF_FULL, **kw)
        ctx.sync()
        eager[key] = (ph.cpu(), cf.cpu())
    ctx.graphs(True)
    rho_d = chypot.to_device(rho)
    phi_d = torch.zeros_like(rho_d)
    coef = torch.zeros((2, g.lmax // 2, g.kmax + 2, g.jmax + 2), dtype=torch.float64, device="cuda")
    try:
        for it, key in enumerate(("a", "a", "b", "a", "b", "b")):
            print(f"[gpu] graph replay {it} density {key}", flush=True)
            rho_d.copy_(chypot.to_device(rho if key == "a" else rho2))
            phi_d.fill_(0.0)
            coef.fill_(0.0)
            pg.solve(rho_d, phi_d, coef=coef, coef_mode=chypot.COEF_FULL, **kw)
            ctx.sync()
            assert float(phi_d.abs().max()) > 0.0, it
            assert torch.equal(phi_d.cpu(), eager[key][0]), it
            assert torch.equal(coef.cpu(), eager[key][1]), it
            # eager work between replays: an un-captured solve on the same grid
            # and a torch reduction
            ctx.graphs(False)
            tmp = torch.zeros_like(phi_d)
            pg.solve(rho_d, tmp, **kw)
            ctx.graphs(True)
            ctx.sync()
            assert float(tmp.abs().sum()) > 0.0
        # another output buffer: a fresh capture, same numbers
        phi2 = torch.zeros_like(rho_d)
        pg.solve(rho_d, phi2, coef=coef, coef_mode=chypot.COEF_FULL, **kw)
        ctx.sync()
        assert torch.equal(phi2.cpu(), eager["b"][0])
    finally:
        ctx.graphs(True)
        pg.close()
    assert rel_max(chypot.to_host(eager["a"][0]), f["phi"]) < 1e-12


def test_128x64x128_against_reference(ctx):
    """The 128x64x128 spiral solve against the reference's planes (K=2 and
    L=1), its boundary values, TMASS and COEF(K=2)."""
    f = np.load(GOLDEN / "solve_large_128x64x128.npz")
    g = fields.unit_grid(128, 64, 128)
    rho = fields.density_spiral_torus(g, seed=2)
    res = run_solve(ctx, g, rho, dict(isym=2, maxtrm=10, npoint=8, redge=0.0), coef_mode=chypot.COEF_K2)
    phi, phib = res["phi"], res["phi_b"]
    scale = float(f["phi_absmax"])
    err = max(rel_max(phi[:, 1, :], f["phi_k2"], scale), rel_max(phi[:, :, 0], f["phi_l1"], scale))
    print(f"[gpu] 128x64x128: vs reference {err:.3e}", flush=True)
    assert err < 1e-12
    assert rel_max(phib[1:g.jmax + 1, g.kmax, :], f["phi_b_top"]) < 1e-13
    assert rel_max(phib[g.jmax, 1:g.kmax, :], f["phi_b_out"]) < 1e-13
    assert abs(res["tmass"] - float(f["tmass"])) <= 1e-13 * abs(float(f["tmass"]))
    nm = f["coef_k2"].shape[1]  # the fixture keeps the first orders
    assert np.abs(res["coef"][:, 1, :nm, 0] - f["coef_k2"][..., 0]).max() < 1e-11


@pytest.mark.parametrize("jkl", [(32, 256, 32), (512, 16, 32), (16, 8, 512), (40, 8, 16), (200, 32, 64),
                                 (20, 4, 8), (64, 64, 2), (1024, 8, 16), (16, 8, 1024), (16, 512, 16),
                                 (128, 64, 192), (64, 32, 768), (16, 8, 1000), (16, 8, 6), (24, 8, 18),
                                 (1100, 8, 16), (2050, 16, 8)])
def test_grid_shapes_against_oracle(ctx, jkl):
    """Shapes that select the other kernel variants: KMAX=256 (general z
    GEMM, eight reduction levels), JMAX=512 (eight r rows per lane), LMAX=512
    and LMAX<=16 transforms, JMAX not a multiple of 64, KMAX=4, LMAX=2
    (boundary values through PHI), and the largest grids the library takes in
    each direction: JMAX=1024 (sixteen r rows per lane: z eigenbasis only),
    LMAX=1024, KMAX=512; mixed-radix transforms: LMAX/2 = 96 (one pass of
    16, 2 and 3), 384 and 500 (4 lanes per block), 3 and 9 (no spectral
    boundary lanes: LMAX/2 <= MAXTRM); JMAX 1100 and 2050, whose r systems
    are solved by the Thomas algorithm (one thread per system).

    Cyclic reduction (default where the grid allows it): against the oracle's
    POT3 by the same reduction, which is the reference's BLKTRI bit for bit;
    the GPU's r solves round differently (partitioned elimination and
    parallel cyclic reduction across lanes), which moves PHI by ~1e-14.

    z eigenbasis: two oracle solves bound the error, the 80-bit one (the
    discrete system solved closer to exactly) and the binary64 one.  The
    binary64 oracle's own distance from the 80-bit solve grows with the grid
    through the conditioning of the z and r operators (6e-12 at KMAX=256,
    5e-11 at KMAX=512, 4e-12 at JMAX=512), so the bar is relative to it: the
    GPU may be no further from the 80-bit solve than twice what the binary64
    oracle is, plus 1e-13."""
    import pyoracle as po
    g = fields.unit_grid(*jkl)
    rho = fields.density_spiral_torus(g, seed=4)
    print(f"[gpu] shape {g.name}", flush=True)
    po.use_extended(True)
    try:
        exact = po.solve(g, rho)
    finally:
        po.use_extended(False)
    want = po.solve(g, rho)
    pg = chypot.PreparedGrid(ctx, g)
    rho_d = chypot.to_device(rho)
    cr_ok = g.jmax - 1 <= 512
    assert pg.mode_solver == (chypot.MODES_CR if cr_ok else chypot.MODES_EIGEN)
    if not cr_ok:
        with pytest.raises(chypot.ChyError) as ei:
            pg.mode_solver = chypot.MODES_CR
        assert ei.value.code == chypot.CHY_EUNSUPPORTED
    results = {}
    for solver in (["cr"] if cr_ok else []) + ["eigen"]:
        pg.mode_solver = SOLVERS[solver]
        phi_d = torch.zeros_like(rho_d)
        d = pg.solve(rho_d, phi_d, diag=True)
        ctx.sync()
        results[solver] = (chypot.to_host(phi_d), d)
    pg.close()
    if cr_ok:
        po.use_cyclic_reduction(True)
        try:
            want_cr = po.solve(g, rho)
        finally:
            po.use_cyclic_reduction(False)
        phi, d = results["cr"]
        err_cr = rel_max(phi, want_cr["phi"])
        print(f"[gpu] shape {g.name}: cyclic reduction vs oracle's {err_cr:.2e} (oracle's reduction vs 80-bit "
              f"{rel_max(want_cr['phi'], exact['phi']):.2e})", flush=True)
        assert err_cr < 1e-12
        assert abs(d.tmass - want["tmass"]) <= 1e-12 * abs(want["tmass"])
    phi, d = results["eigen"]
    err_x, err_64 = rel_max(phi, exact["phi"]), rel_max(phi, want["phi"])
    err_o = rel_max(want["phi"], exact["phi"])
    print(f"[gpu] shape {g.name}: eigenbasis vs 80-bit oracle {err_x:.2e}, vs binary64 oracle {err_64:.2e}"
          f" (binary64 oracle vs 80-bit {err_o:.2e})", flush=True)
    assert err_x <= 2 * err_o + 1e-13
    assert err_64 < 1e-10
    assert abs(d.tmass - want["tmass"]) <= 1e-12 * abs(want["tmass"])


def test_512x256x512_against_oracle(ctx):
    """512x256x512 (KMAX=256: eight reduction levels, eight rows per lane in
    the r solves, the 512-point transform, 767 boundary points) by the cyclic
    reduction, against the oracle's POT3 by the same reduction."""
    import pyoracle as po
    po.use_extended(False)
    g = fields.unit_grid(512, 256, 512)
    rho = fields.density_spiral_torus(g, seed=2)
    print("[gpu] 512x256x512 solve", flush=True)
    pg = chypot.PreparedGrid(ctx, g)
    assert pg.mode_solver == chypot.MODES_CR
    rho_d = chypot.to_device(rho)
    phi_d = torch.zeros_like(rho_d)
    d = pg.solve(rho_d, phi_d, diag=True)
    ctx.sync()
    phi = chypot.to_host(phi_d)
    del rho_d, phi_d
    pg.close()
    print("[gpu] 512x256x512 oracle", flush=True)
    po.use_cyclic_reduction(True)
    try:
        want = po.solve(g, rho)
    finally:
        po.use_cyclic_reduction(False)
    err = rel_max(phi, want["phi"])
    print(f"[gpu] 512x256x512: max rel error vs oracle {err:.2e}", flush=True)
    assert err < 1e-12
    assert abs(d.tmass - want["tmass"]) <= 1e-12 * abs(want["tmass"])
    assert d.klocat == want["klocat"]
