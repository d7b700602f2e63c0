"""The disk's pull on the central star: find_force (pot/wiggle.F:99-135),
which wiggle calls every hydro step (radhydro/3dhyd-main.f:479-486).

CPU: the oracle against the point-mass force of a single cell computed from
Cartesian coordinates, the cancellation of an axisymmetric disk's pull on a
star at the origin, linearity in RHO, and the index ranges the sum covers.
GPU: chy_find_force against the oracle's 80-bit sum of the same binary64
terms to 1e-13 of the sum of |terms| (the GPU adds 16 planes per thread,
then in trees and strided runs: a few dozen roundings deep, well under 1e-14
of the sum of |terms|; the reference's own OpenMP order depends on the thread
count), and repeatable bit for bit.
"""
import numpy as np
import pytest

import pyoracle as po

JM, KM, LM, JMIN = 12, 6, 8, 4
DR, DZ, GS = 1.0 / 12, 0.5 / 6, 0.01
DPHI = 2 * np.pi / LM


def grid():
    rhf = (np.arange(1, JM + 3) - 1.5) * DR
    zhf = (np.arange(1, KM + 3) - 1.5) * DZ
    return rhf, zhf


def force(rho, x=0.13, y=-0.07, jmin=JMIN):
    rhf, zhf = grid()
    return po.find_force(JM, KM, LM, jmin, DR, DZ, DPHI, GS, rhf, zhf, np.asfortranarray(rho), x, y)


def test_single_cell_is_a_softened_point_mass():
    rhf, zhf = grid()
    x, y = 0.13, -0.07
    for (j, k, l) in ((JMIN - 3, 2, 1), (7, 4, 5), (JM + 1, KM + 1, LM)):
        rho = np.zeros((JM + 2, KM + 2, LM), order="F")
        rho[j - 1, k - 1, l - 1] = 2.5
        (fx, fy), _ = force(rho, x, y)
        ph = (l - 0.5) * DPHI
        dx, dy, dz = rhf[j - 1] * np.cos(ph) - x, rhf[j - 1] * np.sin(ph) - y, zhf[k - 1]
        m = 2.5 * rhf[j - 1] * DR * DZ * DPHI * 2
        d3 = (dx * dx + dy * dy + dz * dz + GS * GS) ** 1.5
        assert fx == pytest.approx(m * dx / d3, rel=1e-14)
        assert fy == pytest.approx(m * dy / d3, rel=1e-14)


def test_axisymmetric_disk_pulls_a_centred_star_nowhere():
    rng = np.random.default_rng(2)
    ring = rng.uniform(0.5, 1.5, size=(JM + 2, KM + 2, 1))
    rho = np.asfortranarray(np.repeat(ring, LM, axis=2))
    (fx, fy), (ax, ay) = force(rho, 0.0, 0.0)
    assert abs(fx) <= 1e-14 * ax and abs(fy) <= 1e-14 * ay


def test_force_is_linear_in_the_density():
    rng = np.random.default_rng(3)
    r1, r2 = (rng.uniform(0, 1, size=(JM + 2, KM + 2, LM)) for _ in range(2))
    f1, f2 = force(r1)[0], force(r2)[0]
    f12, a12 = force(r1 + r2)
    for i in range(2):
        assert abs(f12[i] - f1[i] - f2[i]) <= 1e-14 * a12[i]


def test_sum_covers_jmin_minus_3_to_jmax_plus_1_and_k_2_to_kmax_plus_1():
    def single(j, k):
        rho = np.zeros((JM + 2, KM + 2, LM), order="F")
        rho[j - 1, k - 1, :] = 1.0
        return force(rho)[0]
    assert single(JMIN - 4, 3) == (0.0, 0.0)
    assert single(JM + 2, 3) == (0.0, 0.0)
    assert single(5, 1) == (0.0, 0.0)
    assert single(5, KM + 2) == (0.0, 0.0)
    assert single(JMIN - 3, 3) != (0.0, 0.0) and single(JM + 1, KM + 1) != (0.0, 0.0)


def test_oracle_rejects_jmin_below_4():
    with pytest.raises(ValueError):
        force(np.zeros((JM + 2, KM + 2, LM)), jmin=3)


# ---------------------------------------------------------------- GPU parity

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import chypot
    return chypot.Context(0)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(64, 32, 64, 4), (256, 128, 256, 5)])
def test_find_force_matches_oracle(ctx, case):
    import chypot
    from chypot import fields
    jm, km, lm, jmin = case
    print(f"[gpu] chy_find_force {jm}x{km}x{lm} jmin={jmin}", flush=True)
    g = fields.unit_grid(jm, km, lm)
    rho = fields.density_spiral_torus(g, seed=4)
    d = chypot.ForceDesc(jm, km, lm, jmin, g.delr, g.delz, g.dtheta, 0.01)
    rho_d = torch.from_numpy(np.ascontiguousarray(rho.transpose(2, 1, 0))).to("cuda")
    rhf_d, zhf_d = torch.from_numpy(g.rhf()).to("cuda"), torch.from_numpy(g.zhf()).to("cuda")
    for (x, y) in ((0.0, 0.0), (0.031, -0.012)):
        f1 = ctx.find_force(d, rhf_d, zhf_d, rho_d, x, y)
        f2 = ctx.find_force(d, rhf_d, zhf_d, rho_d, x, y)
        assert f1 == f2  # fixed summation order
        po.use_extended(True)
        try:
            want, scale = po.find_force(jm, km, lm, jmin, g.delr, g.delz, g.dtheta, 0.01, g.rhf(), g.zhf(), rho,
                                        x, y)
        finally:
            po.use_extended(False)
        for i in range(2):
            assert abs(f1[i] - want[i]) <= 1e-13 * scale[i], (f1, want, scale)


@pytest.mark.gpu
def test_find_force_rejects_bad_arguments(ctx):
    import chypot
    rho = torch.zeros((8, 6, 14), dtype=torch.float64, device="cuda")
    v = torch.zeros(14, dtype=torch.float64, device="cuda")
    with pytest.raises(chypot.ChyError) as ei:
        ctx.find_force(chypot.ForceDesc(12, 4, 8, 3, 1.0, 1.0, 1.0, 0.0), v, v, rho, 0.0, 0.0)
    assert ei.value.code == chypot.CHY_EINVAL
