"""Fixture loading shared by the CPU and GPU tests."""
import json
import pathlib

import numpy as np

from chypot import fields

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
SMALL_CASES = ["g16_small", "g16_m4_np5", "g16_isymneg_redge", "g24x16x32_spiral", "g32_redge",
               "g32x16x64_aniso", "c1_64x32x64"]
# LMAX/2 not a power of two (the reference's mixed-radix FFT, pot/fft.f:111-126):
# factors 3, 5, 7, 11 x 13, 17, 19, 23 and 3 x 5 with powers of two
MIXED_CASES = ["g16x8x24_radix3", "g24x16x40_radix5", "g32x16x84_radix7", "g16x8x286_radix11x13",
               "g16x8x68_radix17", "g16x8x76_radix19", "g16x8x46_radix23", "g32x16x120_radix3x5"]


def meta():
    return json.loads((GOLDEN / "solve_meta.json").read_text())


def load_case(name):
    """(grid, rho, params dict, fixture npz) for a golden solve."""
    f = np.load(GOLDEN / f"solve_{name}.npz")
    md = meta()[name]
    J, K, L = (int(v) for v in f["grid"])
    g = fields.Grid(J, K, L, float(f["spacing"][0]), float(f["spacing"][1]))
    if "rho" in f.files:
        rho = np.asfortranarray(f["rho"])
    elif md["density"] == "small":
        rho = fields.density_small_torus(g, 1)
    else:
        rho = fields.density_spiral_torus(g, 2)
    if "params" in f.files:
        isym, maxtrm, npoint = (int(v) for v in f["params"])
    else:
        isym, maxtrm, npoint = md["isym"], md["maxtrm"], md["npoint"]
    params = dict(isym=isym, maxtrm=maxtrm, npoint=npoint, redge=float(md["redge"]))
    return g, rho, params, f


def rel_max(a, b, scale=None):
    s = np.abs(b).max() if scale is None else scale
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / s)


def phase_diff(p1, p2):
    """Difference of two COEF phases in degrees, modulo 360, in [-180,180)."""
    return ((np.asarray(p1) - np.asarray(p2)) + 180.0) % 360.0 - 180.0


def assert_coef_close(c, ref, amp_tol=1e-11, tang_tol=1e-13, floor=1e-8):
    """COEF arrays (..., LMAX/2, 2) of pot/pot3.f:105-132 against a reference.

    Orders m < LMAX/2: the amplitude (relative to |A0|) within amp_tol, and
    where the reference amplitude exceeds floor, the phase within
    tang_tol / amplitude radians.  A phase is the angle of a coefficient whose
    components carry rounding errors of ~1e-16 |A0|, so its error grows as
    the amplitude falls: the product amplitude x phase error is the
    coefficient's error across its direction, bounded like the amplitude's.
    Order m = LMAX/2 keeps |A_N|/|A0| and A0 itself in the phase slot."""
    c, ref = np.asarray(c), np.asarray(ref)
    n = c.shape[-2]
    assert np.abs(c[..., 0] - ref[..., 0]).max() < amp_tol
    amp = ref[..., :n - 1, 0]
    sig = amp > floor
    if sig.any():
        dphi = np.radians(np.abs(phase_diff(c[..., :n - 1, 1], ref[..., :n - 1, 1])))
        assert (amp * dphi)[sig].max() < tang_tol
    a0 = ref[..., n - 1, 1]
    assert np.abs(c[..., n - 1, 1] - a0).max() <= 1e-12 * max(np.abs(a0).max(), 1e-300)

