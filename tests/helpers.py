"""Fixture loading shared by the CPU and GPU tests."""
import json
import pathlib

import numpy as np

from chypot import fields

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
SMALL_CASES = ["g16_small", "g16_m4_np5", "g16_isymneg_redge", "g24x16x32_spiral", "g32_redge",
               "g32x16x64_aniso", "c1_64x32x64"]


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
