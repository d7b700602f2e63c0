"""Momentum sources that read the potential: the first loop of SOURCE
(radhydro/source.f:41-79), which the hydro step runs right after
BDYGEN(MAXTRM,ISYM,R(JMAX1)) + POT3 (radhydro/3dhyd-main.f:467-477).

CPU: the oracle's restatement against closed forms -- a potential or pressure
linear in J or K has an exact one-sided difference, sin(psi) an exact central
difference identity, constant RHO and OMEGA an exact centrifugal term; every
one uses dyadic values (exact arithmetic) or a stated tolerance, and the
potential grid is larger than the hydro grid so that a pitch mix-up shows.
GPU: chy_source_grav bitwise equal to the oracle, and the timestep loop (PHI
zeroed, BDYGEN with REDGE=R(JMAX1) + POT3, the star potential of ExternalPot,
then the sources) step by step.
"""
import numpy as np
import pytest

import pyoracle as po

# hydro grid 8 x 4 x 8 inside a potential grid 10 x 6 x 8
JM, KM, LM, PJ, PK = 8, 4, 8, 10, 6
DELT, DR, DZ = 0.25, 0.25, 0.125


def hyd(fill=0.0):
    return np.full((JM + 2, KM + 2, LM), fill, order="F")


def pot(fill=0.0):
    return np.full((PJ + 2, PK + 2, LM), fill, order="F")


def start_values():
    """S, T, A before the update: distinct dyadic values per cell."""
    rng = np.random.default_rng(5)
    return [np.asfortranarray(rng.integers(-64, 64, size=(JM + 2, KM + 2, LM)) / 16.0) for _ in range(3)]


def run(rho, omega, phi, p, r=None, jstart=2, dtheta=None):
    s, t, a = start_values()
    s0, t0, a0 = s.copy(), t.copy(), a.copy()
    r = np.arange(1, PJ + 3, dtype=float) * 0.5 if r is None else r
    dtheta = 2 * np.pi / LM if dtheta is None else dtheta
    po.source_grav(JM, KM, LM, PJ, PK, jstart, DELT, DR, DZ, dtheta, r, rho, omega, phi, p, s, t, a)
    return (s0, t0, a0), (s, t, a)


def updated(jstart=2):
    """Masks (Fortran order) of the cells each of S, T, A may change."""
    ms, mt = np.zeros((JM + 2, KM + 2, LM), bool), np.zeros((JM + 2, KM + 2, LM), bool)
    ms[jstart:JM + 1, 1:KM + 1, :] = True      # J = JSTART+1..JMAX+1, K = 2..KMAX+1
    mt[jstart - 1:JM + 1, 1:KM + 1, :] = True  # J = JSTART..JMAX+1
    return ms, mt, mt.copy()


def psi():
    return (np.arange(1, LM + 1) - 0.5) * (2 * np.pi / LM)


def test_radial_gravity_of_linear_potential_is_exact():
    # PHI = 0.5 J, RHO = 2: (RHO+RHO')(PHI-PHI')/(2 DR) = 4 -> S -= DELT*4
    phi = pot()
    phi[:] = 0.5 * np.arange(1, PJ + 3)[:, None, None]
    (s0, t0, a0), (s, t, a) = run(pot(2.0), hyd(), phi, hyd())
    ms, _, _ = updated()
    assert np.array_equal(s[ms], s0[ms] - 1.0)
    assert np.array_equal(s[~ms], s0[~ms])
    assert np.array_equal(t, t0) and np.array_equal(a, a0)


def test_vertical_gravity_of_linear_potential_is_exact():
    # PHI = 0.5 K, RHO = 2: 0.5 (RHO+RHO')(PHI-PHI')/DZ = 8 -> T -= DELT*8
    phi = pot()
    phi[:] = 0.5 * np.arange(1, PK + 3)[None, :, None]
    (s0, t0, a0), (s, t, a) = run(pot(2.0), hyd(), phi, hyd())
    _, mt, _ = updated()
    assert np.array_equal(t[mt], t0[mt] - 2.0)
    assert np.array_equal(t[~mt], t0[~mt])
    assert np.array_equal(s, s0) and np.array_equal(a, a0)


def test_azimuthal_gravity_of_sine_potential_wraps_periodically():
    # PHI = 0.75 sin(psi): (PHI(L+1)-PHI(L-1))/(2 dtheta) = 0.75 cos(psi) sin(dtheta)/dtheta,
    # including L=1 and L=LMAX, where the neighbours wrap round
    dth = 2 * np.pi / LM
    phi = pot()
    phi[:] = 0.75 * np.sin(psi())[None, None, :]
    (s0, t0, a0), (s, t, a) = run(pot(1.5), hyd(), phi, hyd())
    _, _, ma = updated()
    want = a0 - DELT * 1.5 * 0.75 * (np.cos(psi()) * np.sin(dth) / dth)[None, None, :]
    assert np.allclose(a[ma], want[ma], rtol=0, atol=1e-14)
    assert np.array_equal(a[~ma], a0[~ma])
    assert np.array_equal(s, s0) and np.array_equal(t, t0)


def test_pressure_gradients_are_exact():
    # P = 0.5 J + 0.25 K + 0.5 sin(psi), no potential: S -= DELT*0.5/DR, T -= DELT*0.25/DZ,
    # A -= DELT*0.5 cos(psi) sin(dtheta)/dtheta
    dth = 2 * np.pi / LM
    p = hyd()
    p[:] = (0.5 * np.arange(1, JM + 3)[:, None, None] + 0.25 * np.arange(1, KM + 3)[None, :, None]
            + 0.5 * np.sin(psi())[None, None, :])
    (s0, t0, a0), (s, t, a) = run(pot(3.0), hyd(), pot(), p)
    ms, mt, ma = updated()
    assert np.allclose(s[ms], s0[ms] - DELT * 0.5 / DR, rtol=0, atol=1e-14)
    assert np.allclose(t[mt], t0[mt] - DELT * 0.25 / DZ, rtol=0, atol=1e-14)
    want = a0 - DELT * 0.5 * (np.cos(psi()) * np.sin(dth) / dth)[None, None, :]
    assert np.allclose(a[ma], want[ma], rtol=0, atol=1e-14)


def test_centrifugal_term_uses_r_at_the_face():
    # RHO = 2, OMEGA = 0.5: R(J)(RHO+RHO')(OMEGA+OMEGA')^2/8 = R(J)*RHO*OMEGA^2 = R(J)/2,
    # with R(J) = J/2 (not the cell centre RHF)
    (s0, t0, a0), (s, t, a) = run(pot(2.0), hyd(0.5), pot(), hyd())
    ms, _, _ = updated()
    rj = 0.5 * np.arange(1, JM + 3)[:, None, None] * np.ones((1, KM + 2, LM))
    assert np.array_equal(s[ms], (s0 + DELT * rj * 0.5)[ms])
    assert np.array_equal(t, t0) and np.array_equal(a, a0)


def test_cells_outside_the_loops_are_untouched():
    rng = np.random.default_rng(11)
    fields = [np.asfortranarray(rng.standard_normal(x().shape)) for x in (pot, hyd, pot, hyd)]
    (s0, t0, a0), (s, t, a) = run(*fields, jstart=4)
    ms, mt, ma = updated(jstart=4)
    for new, old, m in ((s, s0, ms), (t, t0, mt), (a, a0, ma)):
        assert np.array_equal(new[~m], old[~m])
        assert np.all(new[m] != old[m])


def test_oracle_rejects_bad_arguments():
    s, t, a = start_values()
    r = np.ones(PJ + 2)
    with pytest.raises(ValueError):
        po.source_grav(JM, KM, LM, PJ, PK, 1, DELT, DR, DZ, 0.5, r, pot(), hyd(), pot(), hyd(), s, t, a)
    with pytest.raises(ValueError):  # potential grid smaller than the hydro grid
        po.source_grav(JM, KM, LM, JM - 1, PK, 2, DELT, DR, DZ, 0.5, r, np.zeros((JM + 1, PK + 2, LM), order="F"),
                       hyd(), np.zeros((JM + 1, PK + 2, LM), order="F"), hyd(), s, t, a)


# ---------------------------------------------------------------- GPU parity

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import chypot
    return chypot.Context(0)


def dev(x):
    """A Fortran-ordered array as a device tensor of the same byte layout."""
    return torch.from_numpy(np.ascontiguousarray(x.transpose(2, 1, 0))).to("cuda")


def host(t):
    return np.asfortranarray(t.cpu().numpy().transpose(2, 1, 0))


CASES = {
    "64x32x64": (64, 32, 64, 64, 32, 2),
    "40x20x32_in_48x24": (40, 20, 32, 48, 24, 5),
    "256x128x256": (256, 128, 256, 256, 128, 2),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_source_grav_bitwise_equals_oracle(ctx, case):
    import chypot
    jm, km, lm, pj, pk, js = CASES[case]
    print(f"[gpu] chy_source_grav {case} jstart={js}", flush=True)
    rng = np.random.default_rng(jm + km + lm)
    hs, ps = (jm + 2, km + 2, lm), (pj + 2, pk + 2, lm)
    rho, phi = (np.asfortranarray(rng.standard_normal(ps)) for _ in range(2))
    omega, p, s, t, a = (np.asfortranarray(rng.standard_normal(hs)) for _ in range(5))
    r = (np.arange(1, pj + 3) - 2) / pj
    d = chypot.SourceDesc(jm, km, lm, pj, pk, js, 1.0e-3, 1.0 / pj, 0.5 / pk, 2 * np.pi / lm)
    sd, td, ad = dev(s), dev(t), dev(a)
    ctx.source_grav(d, torch.from_numpy(r).to("cuda"), dev(rho), dev(omega), dev(phi), dev(p), sd, td, ad)
    ctx.sync()
    po.source_grav(jm, km, lm, pj, pk, js, d.delt, d.dr, d.dz, d.dtheta, r, rho, omega, phi, p, s, t, a)
    for name, got, want in (("S", sd, s), ("T", td, t), ("A", ad, a)):
        assert np.array_equal(host(got), want), name


@pytest.mark.gpu
def test_source_grav_rejects_bad_arguments(ctx):
    import chypot
    z = torch.zeros((8, 6, 10), dtype=torch.float64, device="cuda")
    r = torch.zeros(10, dtype=torch.float64, device="cuda")
    for bad in (dict(jstart=1), dict(jstart=10), dict(dr=0.0)):
        args = dict(jmax=8, kmax=4, lmax=8, pot_jmax=8, pot_kmax=4, jstart=2, delt=1.0, dr=1.0, dz=1.0,
                    dtheta=1.0)
        args.update(bad)
        d = chypot.SourceDesc(*args.values())
        with pytest.raises(chypot.ChyError) as ei:
            ctx.source_grav(d, r, z, z, z, z, z, z, z)
        assert ei.value.code == chypot.CHY_EINVAL


def wiggle_step(star, x, y, f, delt):
    """The star's leapfrog step of wiggle (pot/wiggle.F:55-90) with the pull f."""
    star["vx"] += f[0] * star["odelt"] * 0.5
    star["vy"] += f[1] * star["odelt"] * 0.5
    star["vx"] += f[0] * delt * 0.5
    star["vy"] += f[1] * delt * 0.5
    cx, cy = x + star["vx"] * delt, y + star["vy"] * delt
    star["odelt"] = delt
    star["r"] = float(np.sqrt(cx * cx + cy * cy))
    ph = float(np.arctan2(cy, cx))
    star["phi"] = ph + 2 * np.pi if ph < 0 else ph


@pytest.mark.gpu
def test_timestep_loop_with_device_resident_fields(ctx):
    """Three steps of the hydro step's gravity part at 32x16x32 in the
    reference's order: PHI zeroed (3dhyd-main.f:457-465), BDYGEN(10,2,R(JMAX1))
    + POT3(8) on the device density (:474-477), the star moved by wiggle with
    the disk's pull from find_force (:479-486; the leapfrog itself is scalar
    host arithmetic, done here as wiggle.F:55-90 does it), the star potential
    (ExternalPot, :511), then the sources (SOURCE, :524); the density changes
    between steps.  PHI is checked against the oracle's solve at 1e-12, the
    pull against the oracle's to 1e-13 of the sum of |terms|, PHI after
    ExternalPot and S/T/A bitwise against the oracle's ExternalPot and sources
    applied to the same arrays."""
    import chypot
    from chypot import fields
    from helpers import rel_max
    po.use_extended(False)
    g, rho_h = fields.random_field(21, ((32,), (16,), (32,)))
    pg = chypot.PreparedGrid(ctx, g)
    redge = float(g.r()[g.jmax])  # R(JMAX1)
    rng = np.random.default_rng(3)
    omega, p = (np.asfortranarray(rng.standard_normal(g.shape)) for _ in range(2))
    s, t, a = (np.zeros(g.shape, order="F") for _ in range(3))
    rho_d, om_d, p_d = dev(rho_h), dev(omega), dev(p)
    s_d, t_d, a_d = dev(s), dev(t), dev(a)
    phi_d = torch.empty_like(rho_d)
    r_d = torch.from_numpy(g.r()).to("cuda")
    d = chypot.SourceDesc(g.jmax, g.kmax, g.lmax, g.jmax, g.kmax, 2, 1.0e-3, g.delr, g.delz, g.dtheta)
    rhf_d, zhf_d = torch.from_numpy(g.rhf()).to("cuda"), torch.from_numpy(g.zhf()).to("cuda")
    sp_d = torch.empty_like(rho_d)
    sp = np.zeros(g.shape, order="F")
    fd = chypot.ForceDesc(g.jmax, g.kmax, g.lmax, 4, g.delr, g.delz, g.dtheta, 0.01)
    star = dict(r=0.02, phi=0.4, vx=0.0, vy=0.0, odelt=d.delt)  # wiggle's state
    try:
        for step in range(3):
            print(f"[gpu] timestep {step}", flush=True)
            phi_d.zero_()
            diag = pg.solve(rho_d, phi_d, maxtrm=10, isym=2, redge=redge, npoint=8, diag=True)
            phi = host(phi_d)
            want = po.solve(g, rho_h, maxtrm=10, isym=2, redge=redge, npoint=8)
            assert rel_max(phi, want["phi"]) < 1e-12
            assert abs(diag.tmass - want["tmass"]) <= 1e-13 * abs(want["tmass"])
            x, y = star["r"] * np.cos(star["phi"]), star["r"] * np.sin(star["phi"])
            f = ctx.find_force(fd, rhf_d, zhf_d, rho_d, x, y)
            want_f, scale = po.find_force(g.jmax, g.kmax, g.lmax, 4, g.delr, g.delz, g.dtheta, 0.01, g.rhf(),
                                          g.zhf(), rho_h, x, y)
            assert all(abs(f[i] - want_f[i]) <= 1e-13 * scale[i] for i in range(2)), (f, want_f)
            wiggle_step(star, x, y, f, d.delt)
            e = chypot.ExtPotDesc(g.jmax, g.kmax, g.lmax, g.jmax, g.kmax, g.dtheta, 1.0, 1.0e-3 * step,
                                  star["r"], star["phi"])
            ctx.external_pot(e, rhf_d, zhf_d, phi_d, sp_d)
            po.external_pot(g.jmax, g.kmax, g.lmax, g.jmax, g.kmax, e.dtheta, e.mass_star, e.tmassacc, e.rpstar,
                            e.phi_star, g.rhf(), g.zhf(), phi, sp)
            assert np.array_equal(host(phi_d), phi), f"PHI after ExternalPot at step {step}"
            assert np.array_equal(host(sp_d), sp), f"STARPHI at step {step}"
            ctx.source_grav(d, r_d, rho_d, om_d, phi_d, p_d, s_d, t_d, a_d)
            po.source_grav(g.jmax, g.kmax, g.lmax, g.jmax, g.kmax, 2, d.delt, d.dr, d.dz, d.dtheta, g.r(),
                           rho_h, omega, phi, p, s, t, a)
            for name, got, w in (("S", s_d, s), ("T", t_d, t), ("A", a_d, a)):
                assert np.array_equal(host(got), w), f"{name} at step {step}"
            # the next step's density: the same factors multiply both copies
            fac = np.asfortranarray(1.0 + 0.01 * np.cos(0.7 * (step + 1) + g.psi())[None, None, :]
                                    * np.ones(g.shape))
            rho_h = np.asfortranarray(rho_h * fac)
            rho_d = rho_d * dev(fac)
    finally:
        pg.close()
