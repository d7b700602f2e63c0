"""The Fortran drop-in: SETBDY/BDYGEN/POT3/ZAXPHI over the C-ABI.

chymera-uo_amd/fortran/potsolve.F exports the reference's four subroutine
names with its COMMON-block data; potdiag.F is the ITYPE=71 potential-only
run of radhydro/io.f:151-229 built on them.  The GPU tests run potdiag on the
golden cases and compare its phi3d file and diagnostics with the reference's
own results.
"""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from helpers import assert_coef_close, load_case, rel_max

ROOT = pathlib.Path(__file__).resolve().parents[1]
FDIR = ROOT / "chymera-uo_amd" / "fortran"
REFERENCE_HEADERS = pathlib.Path("/root/reference/headers")
FLANG = shutil.which("flang")

DROPIN_CASES = ["g16_small", "g16_m4_np5", "g16_isymneg_redge", "g32_redge", "c1_64x32x64",
                "g24x16x40_radix5", "g16x8x46_radix23"]


def potdiag(grid_name):
    exe = FDIR / "_build" / grid_name / "potdiag"
    if not exe.exists():
        pytest.skip(f"potdiag not built for {grid_name} (make -C chymera-uo_amd/fortran)")
    return exe


def nml(path, **kw):
    body = ", ".join(f"{k.upper()}={v!r}" if isinstance(v, str) else f"{k.upper()}={v}"
                     for k, v in kw.items())
    path.write_text(f"&POTDIAG {body} /\n")


def read_sequential(path, shape):
    """The two records of phi3d.NNNNNN (4-byte record markers): PHI, TIME."""
    raw = path.read_bytes()
    n = int(np.prod(shape)) * 8
    assert int(np.frombuffer(raw[:4], "<i4")[0]) == n
    phi = np.frombuffer(raw[4:4 + n], "<f8").reshape(shape, order="F")
    assert int(np.frombuffer(raw[4 + n:8 + n], "<i4")[0]) == n
    off = 8 + n
    assert int(np.frombuffer(raw[off:off + 4], "<i4")[0]) == 8
    time = float(np.frombuffer(raw[off + 4:off + 12], "<f8")[0])
    return phi, time


def read_out(path, g):
    raw = np.fromfile(path, "<f8")
    n = int(np.prod(g.shape))
    tmass, phichk, klocat = raw[:3]
    phi = raw[3:3 + n].reshape(g.shape, order="F")
    coef = raw[3 + n:].reshape((g.jmax + 2, g.kmax + 2, g.lmax // 2, 2), order="F")
    return float(tmass), float(phichk), int(klocat), phi, coef


# ------------------------------------------------------------------ CPU checks

@pytest.mark.skipif(FLANG is None, reason="flang not installed")
def test_dropin_exports_reference_symbols():
    obj = FDIR / "_build" / "16x8x16" / "potsolve.o"
    if not obj.exists():
        pytest.skip("fortran layer not built")
    out = subprocess.run(["nm", str(obj)], capture_output=True, text=True, check=True).stdout
    defined = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert {"setbdy_", "bdygen_", "pot3_", "zaxphi_"} <= defined


@pytest.mark.skipif(FLANG is None or not REFERENCE_HEADERS.exists(),
                    reason="needs flang and a CHYMERA tree's headers")
def test_dropin_compiles_against_reference_headers(tmp_path):
    """potsolve.F builds with the reference's own hydroparam.h/globals.h, i.e.
    it uses only names and COMMON blocks those headers define."""
    mod = FDIR / "_build" / "mod"
    if not (mod / "chypot_c.mod").exists():
        pytest.skip("fortran layer not built")
    subprocess.run([FLANG, "-O0", "-fopenmp", "-cpp", "-ffixed-form", f"-I{REFERENCE_HEADERS}",
                    f"-I{mod}", "-c", str(FDIR / "potsolve.F"), "-o", str(tmp_path / "p.o")],
                   check=True, capture_output=True)


# ------------------------------------------------------------------ GPU runs

def _have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.gpu
@pytest.mark.parametrize("name", DROPIN_CASES)
def test_potdiag_matches_reference(tmp_path, name):
    if not _have_gpu():
        pytest.skip("no HIP device")
    g, rho, p, f = load_case(name)
    exe = potdiag(g.name)
    with open(tmp_path / "rho3d.bin", "wb") as fh:
        fh.write(np.array([g.delr, g.delz], "<f8").tobytes())
        fh.write(np.asfortranarray(rho, "<f8").tobytes(order="F"))
    nml(tmp_path / "potdiag.nml", input="rho3d", rhofile="rho3d.bin", isym=p["isym"],
        maxtrm=p["maxtrm"], npoint=p["npoint"], iprint=0, redge=p["redge"], itstop=7, time=1.25,
        outfile="out.bin")
    print(f"[gpu] potdiag {name} {g.name}", flush=True)
    r = subprocess.run([str(exe), "potdiag.nml"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    phi3d, time = read_sequential(tmp_path / "phi3d.000007", g.shape)
    assert time == 1.25
    tmass, phichk, klocat, phi, coef = read_out(tmp_path / "out.bin", g)
    assert np.array_equal(phi3d, phi)
    assert rel_max(phi, f["phi"]) < 1e-12
    assert abs(tmass - float(f["tmass_b"])) <= 1e-13 * abs(float(f["tmass_b"]))
    assert abs(phichk - float(f["phichk"])) < 1e-10
    assert klocat == int(f["klocat"])
    assert_coef_close(coef[:, 1], f["coef_k2"])
    if "coef" in f.files:
        assert np.abs(coef[..., 0] - f["coef"][..., 0]).max() < 1e-11


@pytest.mark.gpu
def test_potdiag_fort2_model_and_report(tmp_path):
    """The ITYPE=71 input path: an axisymmetric model in fort.2 (header line +
    DENNY list-directed, floored at GRIDLIM*DENCEN), IPRINT=1 report on unit 3;
    the potential against the oracle's solve of the same density."""
    if not _have_gpu():
        pytest.skip("no HIP device")
    import pyoracle as po
    from chypot import fields
    g = fields.unit_grid(32, 16, 32)
    exe = potdiag(g.name)
    r, z = g.rhf()[:, None], g.zhf()[None, :]
    denny = np.maximum(0.0, 1.0 - ((r - 0.5) ** 2 + z ** 2) / 0.09) ** 1.5
    dencen = float(denny.max())
    lines = [" ".join(f"{v:.17e}" for v in (1.5, 0.0, 0.0, 0.0, dencen, 0.0, g.delr, g.delz, 0.0))
             + " 1 1"]
    flat = np.asfortranarray(denny).ravel(order="F")
    lines += [" ".join(f"{v:.17e}" for v in flat[i:i + 4]) for i in range(0, flat.size, 4)]
    (tmp_path / "fort.2").write_text("\n".join(lines) + "\n")
    nml(tmp_path / "potdiag.nml", input="fort2", iprint=1, outfile="out.bin")
    print("[gpu] potdiag fort.2 32x16x32", flush=True)
    res = subprocess.run([str(exe), "potdiag.nml"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    tmass, phichk, klocat, phi, coef = read_out(tmp_path / "out.bin", g)
    rho = np.zeros(g.shape, order="F")
    rho[1:g.jmax + 1, 1:g.kmax + 1, :] = np.maximum(denny[1:g.jmax + 1, 1:g.kmax + 1, None],
                                                    1e-10 * dencen)
    po.use_extended(False)
    want = po.solve(g, rho)
    assert rel_max(phi, want["phi"]) < 1e-12
    assert abs(tmass - want["tmass"]) <= 1e-13 * abs(want["tmass"])
    # the unit-3 report (pot3.f:402-419): PHAC line, then COEF(J,2,1:8) per J
    rep = (tmp_path / "fort.3").read_text().splitlines()
    i = rep.index("PHAC(1:KMAX1):")
    k = rep.index("Fourier amplitudes and phases:")
    assert k > i + 1
    first = rep[k + 1]
    assert int(first[20:25]) == 2 and first[25:32] == " COEFS="
    vals = [float(first[32 + 12 * m:44 + 12 * m]) for m in range(8)]
    np.testing.assert_allclose(vals, coef[1, 1, :8, 0], rtol=6e-3, atol=1e-300)
