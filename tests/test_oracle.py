"""The oracle (CPU restatement in oracle/chyoracle.c) pinned before it is trusted.

Pins, in order of strength:
  * golden vectors written by the reference itself (pot/fft.f, pot/blktri.f,
    whole SETBDY+BDYGEN+POT3 solves; tests/golden/make_golden.py);
  * closed forms and independent library routines: numpy's FFT, scipy's sparse
    direct solver on the same 2-D system, associated Legendre functions for the
    multipole table, the discrete operator's residual;
  * the oracle's own 80-bit build on the same inputs.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from scipy.special import lpmv

import make_golden as mg
import pyoracle as po
from chypot import fields
from helpers import GOLDEN, MIXED_CASES, SMALL_CASES, assert_coef_close, load_case, rel_max


@pytest.fixture(autouse=True)
def fp64_oracle():
    po.use_extended(False)
    yield
    po.use_extended(False)


# --------------------------------------------------------------------------- FFT

def test_fft_matches_reference_fft_realtr():
    d = np.load(GOLDEN / "fft_realtr.npz")
    ns = sorted({int(k.split("_")[1]) for k in d.files if k.startswith("x_")})
    assert ns, "no FFT golden vectors"
    for n in ns:
        for v in range(d[f"x_{n}"].shape[0]):
            A, B = po.rfft_fwd(d[f"x_{n}"][v])
            sc = np.abs(d[f"a_{n}"][v]).max()
            assert rel_max(A, d[f"a_{n}"][v], sc) < 1e-14, n
            assert rel_max(B, d[f"b_{n}"][v], sc) < 1e-14, n
            ref = np.empty(2 * n)
            ref[0::2], ref[1::2] = d[f"ia_{n}"][v], d[f"ib_{n}"][v]
            assert rel_max(po.rfft_inv(d[f"a_{n}"][v], d[f"b_{n}"][v]), ref) < 1e-14, n


@pytest.mark.parametrize("L", [2, 6, 16, 48, 64, 256, 512])
def test_fft_matches_numpy(L):
    x = np.random.default_rng(L).standard_normal(L)
    A, B = po.rfft_fwd(x)
    f = np.fft.fft(x)  # numpy: exp(-i theta); the reference's FFT(+1): exp(+i theta)
    N = L // 2
    assert rel_max(A, 2 * f.real[:N + 1], np.abs(f).max()) < 1e-14
    assert rel_max(B, -2 * f.imag[:N + 1], np.abs(f).max()) < 1e-14
    # inverse: 0.5 * raw inverse of the SF1-scaled coefficients is the identity (pot3.f:60-61,276-279)
    B2 = B.copy()
    B2[0] = B2[N] = 0.0
    assert rel_max(0.5 * po.rfft_inv(A / L, B2 / L), x) < 1e-14


# ------------------------------------------------------------------- mode solves

def test_mode_solve_matches_reference_blktri():
    e = np.load(GOLDEN / "blktri_modes.npz")
    ncase = len([k for k in e.files if k.endswith("_grid")])
    assert ncase >= 6
    for i in range(ncase):
        jk = e[f"case{i}_grid"]
        g = fields.Grid(int(jk[0]), int(jk[1]), int(jk[2]), float(jk[3]), float(jk[4]))
        assert int(e[f"case{i}_ierr"]) == 0
        x = po.mode_solve(*mg.pot3_coeffs(g, int(jk[5])), e[f"case{i}_y"])
        assert rel_max(x, e[f"case{i}_x"]) < 1e-13, (i, g.name)


def test_cyclic_reduction_is_the_reference_blktri_bitwise():
    """The oracle's restatement of BLKTRI's generalised cyclic reduction
    (pot/blktri.f:60-423: COMPB/TQLRAT root tables, BLKTR1's reduction and
    back substitution, PRDCT's root order and multiplier choice) gives the
    reference's own BLKTRI results bit for bit on every golden system.  A
    root table converged to another tolerance, another root order or another
    elimination order all change the last bits."""
    e = np.load(GOLDEN / "blktri_modes.npz")
    ncase = len([k for k in e.files if k.endswith("_grid")])
    for i in range(ncase):
        jk = e[f"case{i}_grid"]
        g = fields.Grid(int(jk[0]), int(jk[1]), int(jk[2]), float(jk[3]), float(jk[4]))
        x = po.cr_solve(*mg.pot3_coeffs(g, int(jk[5])), e[f"case{i}_y"])
        assert np.array_equal(x, e[f"case{i}_x"]), (i, g.name, rel_max(x, e[f"case{i}_x"]))
        # the variants are not: eliminating the r systems from the first row
        # down moves the result by rounding only
        x2 = po.cr_solve(*mg.pot3_coeffs(g, int(jk[5])), e[f"case{i}_y"], order="other")
        assert rel_max(x2, e[f"case{i}_x"]) < 1e-13


@pytest.mark.parametrize("grid,m", [((16, 8, 16), 0), ((24, 16, 32), 7), ((40, 32, 64), 1),
                                    ((20, 16, 64), 31), ((64, 64, 16), 2)])
def test_cyclic_reduction_solves_the_mode_system(grid, m):
    """Both root tables solve the same 2-D system as a sparse direct solve."""
    g = fields.Grid(grid[0], grid[1], grid[2], 0.9 / grid[0], 0.55 / grid[1])
    co = mg.pot3_coeffs(g, m)
    A = _mode_matrix(*co)
    y = np.random.default_rng(m).standard_normal((grid[1] - 1, grid[0] - 1))
    x_sp = spla.spsolve(A, y.ravel()).reshape(y.shape)
    assert rel_max(po.cr_solve(*co, y), x_sp) < 1e-12
    assert rel_max(po.cr_solve(*co, y, roots="exact"), x_sp) < 1e-12


def test_reference_error_at_kmax128_comes_from_its_root_table():
    """At KMAX=128 BLKTRI's result is ~1e-12 from the exact solution of its
    system.  Its root table (TQLRAT's convergence rule) makes that error: the
    same reduction with exact roots is 10x closer to the exact solution, and
    with the reference's roots but another elimination order it stays within
    ~1e-14 of the reference's result."""
    g = fields.unit_grid(256, 128, 16)
    co = mg.pot3_coeffs(g, 1)
    y = np.random.default_rng(5).uniform(-1, 1, (127, 255))
    po.use_extended(True)
    try:
        exact = po.mode_solve(*co, y)
    finally:
        po.use_extended(False)
    ref_roots = po.cr_solve(*co, y)
    other = po.cr_solve(*co, y, order="other")
    ex_roots = po.cr_solve(*co, y, roots="exact")
    e_ref = rel_max(ref_roots, exact)
    assert e_ref > 5e-13
    assert rel_max(ex_roots, exact) < 0.1 * e_ref
    assert rel_max(other, ref_roots) < 2e-14


def test_pot3_by_cyclic_reduction_matches_reference():
    """The oracle's POT3 with the cyclic reduction: the reference's whole solve
    on the golden grids to well inside the 1e-12 bar."""
    po.use_cyclic_reduction(True)
    try:
        for name in SMALL_CASES:
            g, rho, params, f = load_case(name)
            phi, _ = po.bdygen(g, rho, None, params["maxtrm"], params["isym"], params["redge"])
            phi, coef, phichk, klocat = po.pot3(g, rho, phi, params["npoint"])
            assert rel_max(phi, f["phi"]) < 1e-13, name
    finally:
        po.use_cyclic_reduction(False)


def _mode_matrix(an, bn, cn, am, bm, cm):
    """The 2-D system BLKTRI solves (pot/blktri.f:3-56), unknown (i,k) at k*M+i."""
    M, N = len(am), len(an)
    rows, cols, vals = [], [], []

    def add(r, c, v):
        if v != 0.0:
            rows.append(r); cols.append(c); vals.append(v)
    for k in range(N):
        for i in range(M):
            r = k * M + i
            add(r, r, bn[k] + bm[i])
            if i > 0: add(r, r - 1, am[i])
            if i < M - 1: add(r, r + 1, cm[i])
            if k > 0: add(r, r - M, an[k])
            if k < N - 1: add(r, r + M, cn[k])
    return sp.csc_matrix((vals, (rows, cols)), shape=(M * N, M * N))


@pytest.mark.parametrize("grid,m", [((16, 8, 16), 0), ((24, 16, 32), 7), ((40, 32, 64), 1),
                                    ((20, 16, 64), 31)])
def test_mode_solve_matches_sparse_direct(grid, m):
    g = fields.Grid(grid[0], grid[1], grid[2], 0.9 / grid[0], 0.55 / grid[1])
    co = mg.pot3_coeffs(g, m)
    A = _mode_matrix(*co)
    y = np.random.default_rng(m).standard_normal((grid[1] - 1, grid[0] - 1))
    x_sp = spla.spsolve(A, y.ravel()).reshape(y.shape)
    assert rel_max(po.mode_solve(*co, y), x_sp) < 1e-12


# --------------------------------------------------------------- multipole table

def test_multipole_terms_are_associated_legendre():
    """DLM*BLM(c)*BLM(c') = (l-m)!/(l+m)! P_l^m(c) P_l^m(c') (Legendre addition
    theorem; boundary.f:151-175).  The reference's DLM use decimal TW15..TW18
    (boundary.f:184-185), exact to 5e-10 relative, so that is the tolerance."""
    rng = np.random.default_rng(3)
    for l, m, dlm, poly in po.terms():
        assert (l + m) % 2 == 0 and l <= 10 and m <= l
        for _ in range(5):
            th1, th2 = rng.uniform(0.05, math.pi / 2, 2)
            c1, s1, c2, s2 = math.cos(th1), math.sin(th1), math.cos(th2), math.sin(th2)
            b1 = sum(p * c1 ** (2 * i) for i, p in enumerate(poly)) * s1 ** m
            b2 = sum(p * c2 ** (2 * i) for i, p in enumerate(poly)) * s2 ** m
            want = math.factorial(l - m) / math.factorial(l + m) * lpmv(m, l, c1) * lpmv(m, l, c2)
            assert abs(dlm * b1 * b2 - want) <= 5e-10 * abs(want) + 1e-15, (l, m)
    assert len(po.terms()) == 36


def test_ring_moments_are_rotated_fourier_coefficients():
    """SUMASS, CM(m), SM(m) over cell-centre angles (boundary.f:526-538) equal
    the FFT coefficients rotated by m*DTHETA/2 -- the identity the GPU uses to
    take the moments from its forward transform."""
    L = 64
    rho = np.random.default_rng(7).uniform(0.5, 1.5, L)
    dth = 2 * math.pi / L
    psi = (np.arange(1, L + 1) - 0.5) * dth
    A, B = po.rfft_fwd(rho)
    for m in range(0, 11):
        X = 0.5 * complex(A[m], B[m])  # sum rho_n exp(+i m n dtheta)
        rot = X * complex(math.cos(m * dth / 2), math.sin(m * dth / 2))
        assert abs(rot.real - np.sum(rho * np.cos(m * psi))) < 1e-12
        assert abs(rot.imag - np.sum(rho * np.sin(m * psi))) < 1e-12


# ----------------------------------------------------------------- whole solves

@pytest.mark.parametrize("name", SMALL_CASES + MIXED_CASES)
def test_solve_matches_reference(name):
    g, rho, p, f = load_case(name)
    res = po.solve(g, rho, p["maxtrm"], p["isym"], p["redge"], p["npoint"])
    assert rel_max(res["phi_b"][1:g.jmax + 1, g.kmax, :], f["phi_b_top"]) < 1e-13
    assert rel_max(res["phi_b"][g.jmax, 1:g.kmax, :], f["phi_b_out"]) < 1e-13
    assert abs(res["tmass"] - float(f["tmass_b"])) <= 1e-13 * abs(float(f["tmass_b"]))
    assert rel_max(res["phi"], f["phi"]) < 1e-12
    assert abs(res["phichk"] - float(f["phichk"])) < 1e-10
    assert res["klocat"] == int(f["klocat"])
    assert_coef_close(res["coef"][:, 1], f["coef_k2"])
    if "coef" in f.files:
        assert_coef_close(res["coef"], f["coef"])


def test_solution_satisfies_discrete_operator():
    """The residual of the 7-point operator of pot3.f:158-231 is at the rounding
    floor of its own terms (the terms are differences of nearly equal
    potentials, so the floor is a few hundred ulps of each term); a wrong
    coefficient or sign leaves an O(1) residual."""
    g, rho, p, f = load_case("g24x16x32_spiral")
    inner = (slice(1, g.jmax), slice(1, g.kmax), slice(None))
    for phi in (po.solve(g, rho)["phi"], f["phi"]):
        res, scale = po.residual(g, phi, rho)
        assert np.abs(res[inner] / scale[inner]).max() < 5e-11
    # the same check fails for a potential off by a plausible mistake
    bad = f["phi"].copy(order="F")
    bad[1:g.jmax + 1, 1:g.kmax + 1, :] *= 1.0 + 1e-6
    res, scale = po.residual(g, bad, rho)
    assert np.abs(res[inner] / scale[inner]).max() > 1e-8


# --------------------------------------------------------------------- COEF

def coef_case():
    """24x8x16 spiral torus with no density on J=2..9, COEF prefilled with 5."""
    g = fields.unit_grid(24, 8, 16)
    rho = fields.density_spiral_torus(g, seed=5)
    rho[:9] = 0.0
    return g, np.asfortranarray(rho), np.full((g.jmax + 2, g.kmax + 2, g.lmax // 2, 2), 5.0, order="F")


def test_coef_keeps_amplitudes_where_a0_is_zero():
    """COEF (pot/pot3.f:105-132) from the transform of PHI on every lane: where
    a0 = 0 the amplitudes keep their old values while the phases are written
    (720 degrees where a = b = 0) and COEF(J,K,N,2) = a0; elsewhere the
    amplitudes are 2 sqrt(a^2+b^2)/|a0| and |aN|/|a0|, checked against numpy's
    FFT of the density lane (interior lanes carry RHO, pot3.f:71-81)."""
    g, rho, coef0 = coef_case()
    phi_b, _ = po.bdygen(g, rho)
    _, coef, _, _ = po.pot3(g, rho, phi_b, 8, coef=coef0)
    n = g.lmax // 2
    dead = coef[1:9, 1:g.kmax, :, :]  # J=2..9, K=2..KMAX: RHO = 0, a0 = 0
    assert (dead[..., 0] == 5.0).all()
    assert (dead[..., :n - 1, 1] == 720.0).all() and (dead[..., n - 1, 1] == 0.0).all()
    live = rho[9:g.jmax, 1:g.kmax, :]  # J=10..JMAX, K=2..KMAX
    f = np.fft.fft(live, axis=2)
    amp = 2 * np.abs(f[..., 1:n]) / np.abs(f[..., :1])
    got = coef[9:g.jmax, 1:g.kmax, :, 0]
    assert np.abs(got[..., :n - 1] - amp).max() < 1e-12 * amp.max()
    assert np.abs(got[..., n - 1] - np.abs(f[..., n]) / np.abs(f[..., 0])).max() < 1e-12


# ------------------------------------------------------------------- ZAXPHI

def axis_value_lagrange(g, phi, npoint):
    """PHI(JMAX+2,K,L), L=1..N, as ZAXPHI defines it (pot/pot3.f:336-379): the
    value at r=0 of the polynomial through NPOINT samples, PHI(J,K,L) at
    x=+RHF(J) and PHI(J,K,L+N) at x=-RHF(J) for J=2..NUM+1, and for odd NPOINT
    also PHI(NUM+2,K,L) at x=+RHF(NUM+2).  Lagrange form in numpy: the Aitken
    recurrence of the reference is Neville's scheme for the same value."""
    rhf, num, n = g.rhf(), npoint // 2, g.lmax // 2
    xs = [rhf[j] for j in range(1, num + 1)] + [-rhf[j] for j in range(1, num + 1)]
    ys = [phi[j, 1:g.kmax + 1, :n] for j in range(1, num + 1)] + \
         [phi[j, 1:g.kmax + 1, n:] for j in range(1, num + 1)]
    if npoint % 2:
        xs.append(rhf[num + 1])
        ys.append(phi[num + 1, 1:g.kmax + 1, :n])
    w = [math.prod(-xs[m] / (xs[i] - xs[m]) for m in range(len(xs)) if m != i) for i in range(len(xs))]
    return sum(wi * y for wi, y in zip(w, ys))


def phichk_klocat(ph, klocat_in):
    """PHAC(K) = the deviation 1 - PH/PHMAX of largest magnitude over L, PHMAX
    the PH of largest magnitude (first L wins ties); PHICHK = the PHAC of
    largest magnitude (first K wins), KLOCAT its K, or KLOCAT unchanged when
    every PHAC is zero (pot/pot3.f:380-399).  ph is PH(K=2.., L=1..N)."""
    phac = np.zeros(ph.shape[0])
    for k in range(ph.shape[0]):
        phmax = ph[k, np.argmax(np.abs(ph[k]))]
        err = 1.0 - ph[k] / phmax
        phac[k] = err[np.argmax(np.abs(err))]
    i = int(np.argmax(np.abs(phac)))
    return phac[i], (i + 2 if phac[i] != 0.0 else klocat_in)


@pytest.mark.parametrize("npoint", range(1, 11))
def test_zaxphi_is_the_interpolant_at_the_axis(npoint):
    """Every NPOINT the solve accepts (1..10): the axis column POT3 writes is the
    interpolating polynomial's value at r=0, the same value in both halves
    L and L+N, and PHICHK/KLOCAT follow the reference's scans.  Planted in a
    scratch copy of the oracle, reading L for the far side fails it for
    NPOINT >= 2 and RHF(J+1) for RHF(J) for NPOINT >= 3."""
    g = fields.unit_grid(24, 8, 16)
    rho = fields.density_spiral_torus(g, seed=5)
    phi_b, _ = po.bdygen(g, rho)
    phi, _, phichk, klocat = po.pot3(g, rho, phi_b, npoint, klocat=-3)
    n = g.lmax // 2
    axis = phi[g.jmax + 1, 1:g.kmax + 1, :]
    want = axis_value_lagrange(g, phi, npoint)
    assert np.abs(axis[:, :n] - want).max() <= 1e-13 * np.abs(phi).max()
    assert np.array_equal(axis[:, n:], axis[:, :n])
    pc, kl = phichk_klocat(axis[:, :n], -3)
    assert phichk == pytest.approx(pc, rel=1e-14, abs=0.0)
    assert klocat == kl


def test_zaxphi_keeps_klocat_for_a_zero_density():
    """Zero density: PHI is zero, every PH is zero, 1 - PH/PHMAX is NaN and the
    strict comparisons of pot/pot3.f:380-399 keep PHAC at zero, so PHICHK is
    zero and KLOCAT keeps the value it had."""
    g = fields.unit_grid(24, 8, 16)
    rho = np.zeros(g.shape, order="F")
    phi_b, _ = po.bdygen(g, rho)
    phi, _, phichk, klocat = po.pot3(g, rho, phi_b, 8, klocat=-3)
    assert not np.abs(phi).any()
    assert (phichk, klocat) == (0.0, -3)


def test_extended_build_matches_fp64_build():
    g, rho, p, f = load_case("c1_64x32x64")
    a = po.solve(g, rho)["phi"]
    po.use_extended(True)
    b = po.solve(g, rho)["phi"]
    assert rel_max(a, b) < 2e-13
    # and the reference's own FP64 result is as close to the 80-bit one
    assert rel_max(f["phi"], b) < 2e-13


def test_large_fixture_records_reference_error():
    """At 256x128x256 the reference's FP64 result is 2.3e-12 from the 80-bit
    solve of the same discrete system (make_golden.py stores both)."""
    f = np.load(GOLDEN / "solve_large_256x128x256.npz")
    assert 1e-12 < float(f["ext_ref_maxdiff"]) < 4e-12
    f2 = np.load(GOLDEN / "solve_large_128x64x128.npz")
    assert float(f2["ext_ref_maxdiff"]) < 1e-12
