"""chy_find_force_noalloc: find_force (pot/wiggle.F:99-135) without any
allocation, on the library's fixed per-GPU scratch.  Two calls in a row
(the second on a smaller grid, so the scratch holds the first call's
partials beyond the second's range) each against chy_find_force on the same
arguments, to 1e-13 of the sum of |terms| (the bar of chy_find_force against
the oracle; the no-allocation form builds its cos/sin table on the GPU, so
the two agree to rounding, not bit for bit), and a grid beyond the scratch
refused on the host before anything is launched."""
import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import chypot
    return chypot.Context(0)


def _case(jm, km, lm, jmin, seed):
    import chypot
    from chypot import fields
    g = fields.unit_grid(jm, km, lm)
    rho = fields.density_spiral_torus(g, seed=seed)
    d = chypot.ForceDesc(jm, km, lm, jmin, g.delr, g.delz, g.dtheta, 0.01)
    rho_d = torch.from_numpy(np.ascontiguousarray(rho.transpose(2, 1, 0))).to("cuda")
    rhf_d, zhf_d = torch.from_numpy(g.rhf()).to("cuda"), torch.from_numpy(g.zhf()).to("cuda")
    return g, rho, d, rho_d, rhf_d, zhf_d


def test_noalloc_matches_find_force_twice_smaller_second(ctx):
    f_d = torch.zeros(2, dtype=torch.float64, device="cuda")
    for (jm, km, lm, jmin, seed, x, y) in ((256, 128, 256, 4, 4, 0.031, -0.012), (64, 32, 64, 5, 6, 0.0, 0.02)):
        print(f"[gpu] chy_find_force_noalloc {jm}x{km}x{lm} jmin={jmin}", flush=True)
        g, rho, d, rho_d, rhf_d, zhf_d = _case(jm, km, lm, jmin, seed)
        f_d.fill_(float("nan"))
        ctx.find_force_noalloc(d, rhf_d, zhf_d, rho_d, x, y, f_d)
        ctx.sync()
        got = f_d.cpu().numpy()
        want = ctx.find_force(d, rhf_d, zhf_d, rho_d, x, y)
        _, scale = po.find_force(jm, km, lm, jmin, g.delr, g.delz, g.dtheta, 0.01, g.rhf(), g.zhf(), rho, x, y)
        for i in range(2):
            assert abs(got[i] - want[i]) <= 1e-13 * scale[i], (got, want, scale)
        # repeatable: the same call again gives the same bits
        ctx.find_force_noalloc(d, rhf_d, zhf_d, rho_d, x, y, f_d)
        ctx.sync()
        assert np.array_equal(f_d.cpu().numpy(), got)


def test_noalloc_refuses_grids_beyond_its_scratch(ctx):
    import chypot
    # 1025x512 (J,K) range at 1024 planes: 2052 x 64 block partials > 65536
    d = chypot.ForceDesc(1020, 512, 1024, 4, 1e-3, 1e-3, 2 * np.pi / 1024, 0.01)
    rho = torch.zeros((1, 1, 1), dtype=torch.float64, device="cuda")
    v = torch.zeros(1100, dtype=torch.float64, device="cuda")
    f_d = torch.zeros(2, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):  # shape checked first on the host
        ctx.find_force_noalloc(d, v, v, rho, 0.0, 0.0, f_d)
    lib = chypot.lib()
    import ctypes as C
    rc = lib.chy_find_force_noalloc(ctx.handle, C.byref(d), C.c_void_p(v.data_ptr()), C.c_void_p(v.data_ptr()),
                                    C.c_void_p(v.data_ptr()), 0.0, 0.0, C.c_void_p(f_d.data_ptr()))
    assert rc == chypot.CHY_ELIMIT
    assert "exceed its scratch" in lib.chy_last_device_error().decode()
