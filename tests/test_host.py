"""CPU-side checks of the product: the C-ABI library loads and exports every
declared entry point, host logic (grids, synthetic fields) matches what the
fixtures were made from, and the entry points fail cleanly without a device."""
import hashlib
import pathlib
import re

import numpy as np
import pytest

import chypot
from chypot import fields
from helpers import load_case, meta

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "chypot.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*\*?\s*(chy_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = chypot.lib()
    syms = declared_symbols()
    assert len(syms) >= 16
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(chypot.EXPORTS) == syms


def test_strerror_covers_codes():
    for code in range(7):
        assert chypot.lib().chy_strerror(code)
    assert b"unknown" in chypot.lib().chy_strerror(99)


def test_context_without_device_fails_cleanly():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(chypot.ChyError) as ei:
        chypot.Context(0)
    assert ei.value.code == chypot.CHY_EINVAL


def test_grid_arrays_follow_io_f():
    g = fields.unit_grid(16, 8, 16)
    assert np.allclose(g.r()[:3], [-1 / 16, 0.0, 1 / 16])
    assert np.allclose(g.rhf()[1], 0.5 / 16)
    assert np.allclose(g.zhf()[1], 0.25 / 8)
    assert abs(g.dtheta - 2 * np.pi / 16) < 1e-15


@pytest.mark.parametrize("name", ["c1_64x32x64", "large_128x64x128", "large_256x128x256"])
def test_regenerated_densities_match_fixtures(name):
    md = meta()[name]
    J, K, L = (int(v) for v in md["grid"].split("x"))
    g = fields.Grid(J, K, L, md["delr"], md["delz"])
    rho = fields.density_small_torus(g, 1) if md["density"] == "small" else fields.density_spiral_torus(g, 2)
    digest = hashlib.sha256(np.asarray(rho, dtype="<f8").tobytes(order="F")).hexdigest()
    assert digest == md["rho_sha256"]


def test_density_floor_and_support():
    g, rho, p, f = load_case("g16_small")
    inner = rho[1:g.jmax + 1, 1:g.kmax + 1, :]
    assert inner.min() >= 1e-10 * inner.max() * (1 - 1e-15)
    assert np.all(rho[0] == 0) and np.all(rho[:, 0] == 0)
    assert np.all(rho[g.jmax + 1] == 0) and np.all(rho[:, g.kmax + 1] == 0)


def test_random_field_draws_documented_ranges():
    seen = set()
    for s in range(12):
        g, rho = fields.random_field(s)
        assert g.jmax in (64, 128, 256) and g.kmax in (32, 64, 128) and g.lmax in (64, 128, 256)
        assert rho.shape == g.shape and rho.flags.f_contiguous
        seen.add(g.name)
    assert len(seen) > 3


def test_host_arrays_must_be_fortran_float64_of_the_grid_shape():
    with pytest.raises(ValueError):
        chypot._host_ok((3, 3, 3), rho=np.zeros((3, 3, 3)))
    with pytest.raises(ValueError, match="shape"):
        chypot._host_ok((3, 3, 4), phi=np.zeros((3, 3, 3), order="F"))
    with pytest.raises(ValueError):
        chypot._host_ok((3, 3, 3), coef=np.zeros((3, 3, 3), dtype=np.float32, order="F"))
    chypot._host_ok((3, 3, 3), rho=np.zeros((3, 3, 3), order="F"))


class _FakeCudaTensor:
    """Stands in for a torch CUDA tensor in the argument checks (no GPU here)."""

    def __init__(self, shape, device=0, dtype="torch.float64", contiguous=True):
        self.shape, self.is_cuda = tuple(shape), True
        self.dtype = type("D", (), {"__repr__": lambda _s: dtype})()
        self.device = type("Dev", (), {"index": device})()
        self._c = contiguous

    def is_contiguous(self):
        return self._c

    def data_ptr(self):
        return 4096


def test_device_arguments_are_checked_against_grid_and_device():
    shape = (16, 10, 18)
    assert chypot._dptr(_FakeCudaTensor(shape), shape, 0, "phi").value == 4096
    with pytest.raises(ValueError, match="shape"):
        chypot._dptr(_FakeCudaTensor((16, 10, 17)), shape, 0, "phi")
    with pytest.raises(ValueError, match="cuda:1"):
        chypot._dptr(_FakeCudaTensor(shape, device=1), shape, 0, "rho")
    with pytest.raises(ValueError, match="float64"):
        chypot._dptr(_FakeCudaTensor(shape, dtype="torch.float32"), shape, 0, "rho")
    with pytest.raises(ValueError):
        chypot._dptr(_FakeCudaTensor(shape, contiguous=False), shape, 0, "rho")
    # COEF(JMAX+2, KMAX+2, LMAX/2, 2) in Fortran order is a (2, LMAX/2, KMAX+2, JMAX+2) tensor
    with pytest.raises(ValueError, match="shape"):
        chypot._dptr(_FakeCudaTensor((18, 10, 8, 2)), (2, 8, 10, 18), 0, "coef")


def test_autotune_fused_threshold():
    """The one-kernel-per-plane threshold is the largest batch size up to
    which that form won at every timed size."""
    from chypot import autotune
    assert autotune.fused_threshold({1: 0.3, 8: 2.0, 32: 9.0}, {1: 0.4, 8: 2.5, 32: 8.0}) == 8
    assert autotune.fused_threshold({1: 0.5, 32: 7.0}, {1: 0.4, 32: 8.0}) == 0
    assert autotune.fused_threshold({1: 0.3, 32: 7.0}, {1: 0.4, 32: 8.0}) == 32
    assert autotune.fused_threshold({}, {}) == 0


def test_autotune_table_round_trip(tmp_path):
    from chypot import autotune
    t = {"256x128x256": {"cr_lanes": 64, "refine_fused": 1, "timings_ms": {"cr_lanes": {"64": 15.0}}}}
    autotune.save(t, tmp_path / "t.json")
    assert autotune.load(tmp_path / "t.json") == t


def test_autotune_measured_table():
    """The shipped MI355X table: every entry names candidates the grid class
    takes (32 lanes only where JMAX-1 <= 256; the one-kernel refined form only
    for KMAX <= 128) and its choice is the fastest timed candidate."""
    from chypot import autotune
    t = autotune.measured_table()
    assert "256x128x256" in t
    for key, e in t.items():
        j, k, l = map(int, key.split("x"))
        lanes = e["timings_ms"]["cr_lanes"]
        assert e["cr_lanes"] == int(min(lanes, key=lanes.get))
        if j - 1 > 256:
            assert set(lanes) == {"64"}
        if "refine_fused" in e["timings_ms"]:
            assert k <= 128
            tf = {int(b): v["one_kernel"] for b, v in e["timings_ms"]["refine_fused"].items()}
            ts = {int(b): v["per_stage"] for b, v in e["timings_ms"]["refine_fused"].items()}
            assert e["refine_fused"] == autotune.fused_threshold(tf, ts)
        else:
            assert k > 128 or e["refine_fused"] == 0
