"""The central star's potential added after the solve: ExternalPot
(pot/ExternalPot.f:1-88) with the default build's flags (EXTERNAL_POT=1,
INDIRECT=0, BINARY=0), called by the hydro step right after POT3
(radhydro/3dhyd-main.f:494-511).

CPU: the oracle against the point-mass potential -M/|x - x*| computed from
Cartesian coordinates (x = RHF cos psi, RHF sin psi, ZHF; x* on the midplane),
its axisymmetry when the star sits at the origin, and the cells it may touch.
GPU: chy_external_pot bitwise equal to the oracle.
"""
import numpy as np
import pytest

import pyoracle as po

JM, KM, LM, PJ, PK = 8, 4, 8, 10, 6
DTH = 2 * np.pi / LM


def centres():
    rhf = (np.arange(1, PJ + 3) - 1.5) / PJ
    zhf = (np.arange(1, PK + 3) - 1.5) * 0.5 / PK
    return rhf, zhf


def run(mass=1.5, macc=0.25, rp=0.3, ph=0.7, phi=None):
    rhf, zhf = centres()
    phi = np.zeros((PJ + 2, PK + 2, LM), order="F") if phi is None else phi
    sp = np.zeros((JM + 2, KM + 2, LM), order="F")
    po.external_pot(JM, KM, LM, PJ, PK, DTH, mass, macc, rp, ph, rhf, zhf, phi, sp)
    return phi, sp


def test_star_potential_is_the_point_mass_potential():
    rhf, zhf = centres()
    _, sp = run()
    psi = (np.arange(1, LM + 1) - 0.5) * DTH
    R, Z, P = np.meshgrid(rhf[:JM + 2], zhf[:KM + 2], psi, indexing="ij")
    star = np.array([0.3 * np.cos(0.7), 0.3 * np.sin(0.7), 0.0])
    d = np.sqrt((R * np.cos(P) - star[0]) ** 2 + (R * np.sin(P) - star[1]) ** 2 + (Z - star[2]) ** 2)
    assert np.allclose(sp, -(1.5 + 0.25) / d, rtol=2e-15, atol=0)


def test_star_at_the_origin_is_axisymmetric():
    rhf, zhf = centres()
    _, sp = run(rp=0.0, ph=1.1)
    want = -1.75 / np.sqrt(rhf[:JM + 2, None] ** 2 + zhf[None, :KM + 2] ** 2)
    for L in range(LM):
        assert np.array_equal(sp[:, :, L], sp[:, :, 0])
    assert np.allclose(sp[:, :, 0], want, rtol=1e-15, atol=0)


def test_phi_changes_on_the_hydro_grid_only():
    rng = np.random.default_rng(4)
    phi0 = np.asfortranarray(rng.standard_normal((PJ + 2, PK + 2, LM)))
    phi, sp = run(phi=phi0.copy(order="F"))
    inside = np.zeros(phi.shape, bool)
    inside[:JM + 2, :KM + 2, :] = True
    assert np.array_equal(phi[~inside], phi0[~inside])
    assert np.array_equal(phi[:JM + 2, :KM + 2, :], phi0[:JM + 2, :KM + 2, :] + sp)


def test_oracle_rejects_a_hydro_grid_larger_than_the_potential_grid():
    rhf, zhf = centres()
    with pytest.raises(ValueError):
        po.external_pot(JM, KM, LM, JM - 1, PK, DTH, 1.0, 0.0, 0.0, 0.0, rhf, zhf,
                        np.zeros((JM + 1, PK + 2, LM), order="F"), np.zeros((JM + 2, KM + 2, LM), order="F"))


# ---------------------------------------------------------------- GPU parity

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import chypot
    return chypot.Context(0)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x.transpose(2, 1, 0))).to("cuda")


def host(t):
    return np.asfortranarray(t.cpu().numpy().transpose(2, 1, 0))


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(64, 32, 64, 80, 40), (256, 128, 256, 256, 128)])
def test_external_pot_bitwise_equals_oracle(ctx, case):
    import chypot
    jm, km, lm, pj, pk = case
    print(f"[gpu] chy_external_pot {jm}x{km}x{lm} in {pj}x{pk}", flush=True)
    rng = np.random.default_rng(jm + lm)
    phi = np.asfortranarray(rng.standard_normal((pj + 2, pk + 2, lm)))
    rhf = (np.arange(1, pj + 3) - 1.5) / pj
    zhf = (np.arange(1, pk + 3) - 1.5) * 0.5 / pk
    d = chypot.ExtPotDesc(jm, km, lm, pj, pk, 2 * np.pi / lm, 1.0, 0.0123, 0.05, 2.5)
    phi_d = dev(phi)
    sp_d = torch.zeros((lm, km + 2, jm + 2), dtype=torch.float64, device="cuda")
    for _ in range(2):  # the second call uses the cached cos/sin table
        ctx.external_pot(d, torch.from_numpy(rhf).to("cuda"), torch.from_numpy(zhf).to("cuda"), phi_d, sp_d)
    ctx.sync()
    sp = np.zeros((jm + 2, km + 2, lm), order="F")
    for _ in range(2):
        po.external_pot(jm, km, lm, pj, pk, d.dtheta, d.mass_star, d.tmassacc, d.rpstar, d.phi_star, rhf, zhf,
                        phi, sp)
    assert np.array_equal(host(sp_d), sp)
    assert np.array_equal(host(phi_d), phi)


@pytest.mark.gpu
def test_external_pot_rejects_bad_arguments(ctx):
    import chypot
    z = torch.zeros((8, 6, 10), dtype=torch.float64, device="cuda")
    v = torch.zeros(10, dtype=torch.float64, device="cuda")
    with pytest.raises(chypot.ChyError) as ei:
        ctx.external_pot(chypot.ExtPotDesc(8, 4, 8, 8, 4, 0.0, 1.0, 0.0, 0.0, 0.0), v, v, z, z)
    assert ei.value.code == chypot.CHY_EINVAL
