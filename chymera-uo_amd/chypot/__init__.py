"""Python host mirror of the MI355X potential solve (ctypes over libchypot.so).

Mirrors the reference's operator interface for the self-gravity solve --
``SETBDY(NCALL,ISYM)`` (pot/boundary.f:2), ``BDYGEN(MAXTRM,ISYM,REDGE)``
(pot/boundary.f:88) and ``POT3(NPOINT,IPRINT)`` (pot/pot3.f:15) -- with explicit
arrays in place of the COMMON blocks.  Arrays keep the reference's Fortran
layout ``PHI(JMAX+2,KMAX+2,LMAX)``: a torch tensor of shape ``(LMAX, KMAX+2,
JMAX+2)`` (C order) or a numpy array of shape ``(JMAX+2, KMAX+2, LMAX)`` with
``order='F'`` has exactly that byte layout.

Errors: every call raises :class:`ChyError` (carrying the library's code) where
the reference would STOP or silently continue; nothing falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

from . import fields  # noqa: F401  (re-exported helper module)

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE.parent / "_lib" / "libchypot.so"

CHY_OK, CHY_EINVAL, CHY_EUNSUPPORTED, CHY_ENOMEM, CHY_EDEVICE, CHY_ESTATE, CHY_ENUMERIC, CHY_ELIMIT = range(8)
COEF_NONE, COEF_K2, COEF_FULL = 0, 1, 2
STAGES = ("bdygen", "fft_fwd", "zfwd", "rsolve", "zinv", "fft_inv", "ghost", "fft_bdy", "cr", "resid", "planes")
# per-mode solver of a grid (chy_grid_set_mode_solver)
MODES_CR, MODES_EIGEN = 0, 1
# tunable kernel shapes and precision split (chy_grid_set_tune)
TUNE_CR_LANES, TUNE_REFINE_M64, TUNE_REFINE_FUSED = 0, 1, 2

# every entry point declared in include/chypot.h
EXPORTS = (
    "chy_ctx_create", "chy_ctx_destroy", "chy_ctx_set_stream", "chy_ctx_sync",
    "chy_setbdy", "chy_grid_destroy", "chy_grid_device_bytes",
    "chy_grid_set_mode_solver", "chy_grid_mode_solver", "chy_grid_set_tune", "chy_grid_tune",
    "chy_bdygen", "chy_pot3", "chy_solve", "chy_bdygen_host", "chy_pot3_host",
    "chy_ctx_profile", "chy_ctx_stage_times", "chy_ctx_graphs",
    "chy_zaxphi", "chy_zaxphi_host", "chy_phac", "chy_solve_batch", "chy_solve_batch_refined", "chy_ctx_workers",
    "chy_solve_refined", "chy_source_grav", "chy_external_pot", "chy_find_force", "chy_find_force_noalloc",
    "chy_strerror", "chy_last_device_error",
)


class ChyError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().chy_strerror(code).decode()
        if code == CHY_EDEVICE:
            msg += f" ({lib().chy_last_device_error().decode()})"
        super().__init__(f"{where}: {msg} [code {code}]")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("jmax", C.c_int), ("kmax", C.c_int), ("lmax", C.c_int),
                ("r", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
                ("rhf", C.POINTER(C.c_double)), ("zhf", C.POINTER(C.c_double)),
                ("dtheta", C.c_double), ("pi", C.c_double), ("grav", C.c_double)]


class SourceDesc(C.Structure):
    _fields_ = [("jmax", C.c_int), ("kmax", C.c_int), ("lmax", C.c_int),
                ("pot_jmax", C.c_int), ("pot_kmax", C.c_int), ("jstart", C.c_int),
                ("delt", C.c_double), ("dr", C.c_double), ("dz", C.c_double), ("dtheta", C.c_double)]


class ExtPotDesc(C.Structure):
    _fields_ = [("jmax", C.c_int), ("kmax", C.c_int), ("lmax", C.c_int),
                ("pot_jmax", C.c_int), ("pot_kmax", C.c_int), ("dtheta", C.c_double),
                ("mass_star", C.c_double), ("tmassacc", C.c_double), ("rpstar", C.c_double),
                ("phi_star", C.c_double)]


class ForceDesc(C.Structure):
    _fields_ = [("jmax", C.c_int), ("kmax", C.c_int), ("lmax", C.c_int), ("jmin", C.c_int),
                ("dr", C.c_double), ("dz", C.c_double), ("dphi", C.c_double), ("gsoft", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("tmass", C.c_double), ("phichk", C.c_double), ("klocat", C.c_int)]


class SolveItem(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("rho_d", C.c_void_p), ("phi_d", C.c_void_p),
                ("coef_d", C.c_void_p), ("coef_mode", C.c_int), ("maxtrm", C.c_int),
                ("isym", C.c_int), ("npoint", C.c_int), ("redge", C.c_double),
                ("diag", C.POINTER(Diag))]


_lib = None


def lib():
    """Load libchypot.so (raises if it was not built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libchypot.so not built: {LIB_PATH} (run __graft_entry__.build())")
        try:
            # torch first: the library then binds to the HIP runtime torch
            # loaded, so both see the device (the other order leaves torch
            # with a runtime that finds no GPU)
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.chy_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.chy_ctx_destroy.argtypes = [vp]
        L.chy_ctx_set_stream.argtypes = [vp, vp]
        L.chy_ctx_sync.argtypes = [vp]
        L.chy_setbdy.argtypes = [vp, C.POINTER(GridDesc), C.c_int, C.POINTER(vp)]
        L.chy_grid_destroy.argtypes = [vp]
        L.chy_grid_device_bytes.argtypes = [vp]
        L.chy_grid_device_bytes.restype = C.c_size_t
        L.chy_grid_set_mode_solver.argtypes = [vp, C.c_int]
        L.chy_grid_mode_solver.argtypes = [vp]
        L.chy_grid_set_tune.argtypes = [vp, C.c_int, C.c_int]
        L.chy_grid_tune.argtypes = [vp, C.c_int]
        L.chy_bdygen.argtypes = [vp, C.c_int, C.c_int, C.c_double, vp, vp, C.POINTER(Diag)]
        L.chy_pot3.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int, C.POINTER(Diag)]
        L.chy_solve.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp, C.c_int,
                                C.POINTER(Diag)]
        L.chy_bdygen_host.argtypes = [vp, C.c_int, C.c_int, C.c_double, dp, dp, C.POINTER(Diag)]
        L.chy_pot3_host.argtypes = [vp, C.c_int, dp, dp, dp, C.c_int, C.POINTER(Diag)]
        L.chy_ctx_profile.argtypes = [vp, C.c_int]
        L.chy_ctx_graphs.argtypes = [vp, C.c_int]
        L.chy_zaxphi.argtypes = [vp, C.c_int, vp, C.POINTER(Diag)]
        L.chy_zaxphi_host.argtypes = [vp, C.c_int, dp, C.POINTER(Diag)]
        L.chy_phac.argtypes = [vp, dp, C.c_int]
        L.chy_solve_batch.argtypes = [vp, C.POINTER(SolveItem), C.c_int]
        L.chy_solve_batch_refined.argtypes = [vp, C.POINTER(SolveItem), C.c_int, C.c_int]
        L.chy_ctx_workers.argtypes = [vp, C.c_int]
        L.chy_solve_refined.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, vp, vp, vp, C.c_int,
                                        C.c_int, C.POINTER(Diag), dp]
        L.chy_ctx_stage_times.argtypes = [vp, dp, C.POINTER(C.c_int), C.c_int]
        L.chy_source_grav.argtypes = [vp, C.POINTER(SourceDesc)] + [vp] * 8
        L.chy_external_pot.argtypes = [vp, C.POINTER(ExtPotDesc)] + [vp] * 4
        L.chy_find_force.argtypes = [vp, C.POINTER(ForceDesc), vp, vp, vp, C.c_double, C.c_double, dp]
        L.chy_find_force_noalloc.argtypes = [vp, C.POINTER(ForceDesc), vp, vp, vp, C.c_double, C.c_double, vp]
        L.chy_strerror.argtypes = [C.c_int]
        L.chy_strerror.restype = C.c_char_p
        L.chy_last_device_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int, where: str):
    if rc != CHY_OK:
        raise ChyError(rc, where)


def _dptr(t, shape=None, device=None, name="tensor"):
    """Device pointer of a torch CUDA tensor (float64, contiguous) or None.
    shape/device: what the kernels will index (checked, so that a wrong-sized
    or misplaced tensor is an error here rather than an out-of-bounds access)."""
    if t is None:
        return None
    if not t.is_cuda or t.dtype.__repr__() != "torch.float64" or not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous float64 CUDA tensor")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if device is not None and t.device.index != device:
        raise ValueError(f"{name}: on cuda:{t.device.index}, the context is on cuda:{device}")
    return C.c_void_p(t.data_ptr())


class Context:
    """One HIP device and the stream all work is issued on."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().chy_ctx_create(device, C.byref(h)), "chy_ctx_create")
        self.handle = h
        self.device = device

    def set_stream(self, stream=None):
        """Issue on a torch.cuda.Stream (or its raw handle); None = own stream.
        The null (legacy default) stream is refused rather than replaced by
        the context's own stream, which would not be ordered with it."""
        raw = None
        if stream is not None:
            raw = getattr(stream, "cuda_stream", stream)
            if not raw:
                raise ValueError("set_stream: the null stream cannot be used; pass a created stream "
                                 "(torch.cuda.Stream()) or None for the context's own")
        _check(lib().chy_ctx_set_stream(self.handle, C.c_void_p(raw) if raw else None),
               "chy_ctx_set_stream")

    def sync(self):
        _check(lib().chy_ctx_sync(self.handle), "chy_ctx_sync")

    def profile(self, enable: bool = True):
        """Bracket every solver stage with HIP events (resets the totals)."""
        _check(lib().chy_ctx_profile(self.handle, int(enable)), "chy_ctx_profile")

    def graphs(self, enable: bool = True):
        """Replay chy_solve from captured HIP graphs (default on)."""
        _check(lib().chy_ctx_graphs(self.handle, int(enable)), "chy_ctx_graphs")

    def workers(self, n: int):
        """Worker streams used by solve_batch (1..16)."""
        _check(lib().chy_ctx_workers(self.handle, int(n)), "chy_ctx_workers")

    def solve_batch(self, jobs, diag=False, refine_steps=0):
        """Solve independent fields in one call.  jobs: iterable of dicts with
        keys grid (PreparedGrid), rho, phi and optionally coef, coef_mode,
        maxtrm, isym, redge, npoint.  refine_steps > 0: every field by the
        refined solve with that many correction steps.  Returns a list of Diag
        (diag=True, waits) or None (asynchronous on this context's stream)."""
        jobs = list(jobs)
        items = (SolveItem * len(jobs))()
        diags = [Diag() for _ in jobs] if diag else None
        for i, jb in enumerate(jobs):
            pg = jb["grid"]
            if pg.ctx.device != self.device:
                raise ValueError(f"job {i}: grid prepared on cuda:{pg.ctx.device}, batch on cuda:{self.device}")
            items[i] = SolveItem(pg.handle.value, pg._f(jb["rho"], "rho").value, pg._f(jb["phi"], "phi").value,
                                 pg._c(jb.get("coef")).value if jb.get("coef") is not None else None,
                                 jb.get("coef_mode", COEF_NONE), jb.get("maxtrm", 10),
                                 jb.get("isym", pg.isym), jb.get("npoint", 8), jb.get("redge", 0.0),
                                 C.pointer(diags[i]) if diag else None)
        if refine_steps > 0:
            _check(lib().chy_solve_batch_refined(self.handle, items, len(jobs), int(refine_steps)),
                   "chy_solve_batch_refined")
        else:
            _check(lib().chy_solve_batch(self.handle, items, len(jobs)), "chy_solve_batch")
        return diags

    def source_grav(self, desc: SourceDesc, r, rho, omega, phi, p, s, t, a):
        """The momentum sources of radhydro/source.f:41-79 (asynchronous on this
        context's stream).  Device tensors in the Fortran layout: rho, phi of
        shape (LMAX, POT3KMAX+2, POT3JMAX+2); omega, p, s, t, a of shape
        (LMAX, KMAX+2, JMAX+2); r of length POT3JMAX+2.  s, t, a are updated."""
        hyd = (desc.lmax, desc.kmax + 2, desc.jmax + 2)
        pot = (desc.lmax, desc.pot_kmax + 2, desc.pot_jmax + 2)
        for name, x, shape in (("rho", rho, pot), ("phi", phi, pot), ("omega", omega, hyd), ("p", p, hyd),
                               ("s", s, hyd), ("t", t, hyd), ("a", a, hyd)):
            if tuple(x.shape) != shape:
                raise ValueError(f"{name}: shape {tuple(x.shape)}, expected {shape}")
        if r.numel() < desc.pot_jmax + 2:
            raise ValueError("r: fewer than POT3JMAX+2 values")
        ptr = [_dptr(x, device=self.device, name=n)
               for n, x in (("r", r), ("rho", rho), ("omega", omega), ("phi", phi), ("p", p), ("s", s), ("t", t),
                            ("a", a))]
        _check(lib().chy_source_grav(self.handle, C.byref(desc), *ptr), "chy_source_grav")

    def external_pot(self, desc: ExtPotDesc, rhf, zhf, phi, starphi):
        """The star potential of pot/ExternalPot.f (INDIRECT=0, BINARY=0), asynchronous on
        this context's stream: starphi (LMAX, KMAX+2, JMAX+2) is written and added to
        phi (LMAX, POT3KMAX+2, POT3JMAX+2); rhf, zhf are device vectors of RHF, ZHF."""
        if tuple(phi.shape) != (desc.lmax, desc.pot_kmax + 2, desc.pot_jmax + 2):
            raise ValueError(f"phi: shape {tuple(phi.shape)}")
        if tuple(starphi.shape) != (desc.lmax, desc.kmax + 2, desc.jmax + 2):
            raise ValueError(f"starphi: shape {tuple(starphi.shape)}")
        if rhf.numel() < desc.jmax + 2 or zhf.numel() < desc.kmax + 2:
            raise ValueError("rhf/zhf: too few values")
        ptr = [_dptr(x, device=self.device, name=n)
               for n, x in (("rhf", rhf), ("zhf", zhf), ("phi", phi), ("starphi", starphi))]
        _check(lib().chy_external_pot(self.handle, C.byref(desc), *ptr), "chy_external_pot")

    def find_force(self, desc: ForceDesc, rhf, zhf, rho, x: float, y: float):
        """find_force of pot/wiggle.F: the disk's pull (fx, fy) on a star at (x, y, 0);
        rho is a device tensor (LMAX, KMAX+2, JMAX+2).  Waits."""
        if tuple(rho.shape) != (desc.lmax, desc.kmax + 2, desc.jmax + 2):
            raise ValueError(f"rho: shape {tuple(rho.shape)}")
        if rhf.numel() < desc.jmax + 2 or zhf.numel() < desc.kmax + 2:
            raise ValueError("rhf/zhf: too few values")
        f = (C.c_double * 2)()
        ptr = [_dptr(x, device=self.device, name=n) for n, x in (("rhf", rhf), ("zhf", zhf), ("rho", rho))]
        _check(lib().chy_find_force(self.handle, C.byref(desc), *ptr, float(x), float(y), f), "chy_find_force")
        return float(f[0]), float(f[1])

    def find_force_noalloc(self, desc: ForceDesc, rhf, zhf, rho, x: float, y: float, f):
        """find_force into the device tensor f (2 float64: fx, fy) without any
        allocation (the library's fixed per-GPU scratch); asynchronous on this
        context's stream."""
        if tuple(rho.shape) != (desc.lmax, desc.kmax + 2, desc.jmax + 2):
            raise ValueError(f"rho: shape {tuple(rho.shape)}")
        if rhf.numel() < desc.jmax + 2 or zhf.numel() < desc.kmax + 2:
            raise ValueError("rhf/zhf: too few values")
        ptr = [_dptr(v, device=self.device, name=n) for n, v in (("rhf", rhf), ("zhf", zhf), ("rho", rho))]
        _check(lib().chy_find_force_noalloc(self.handle, C.byref(desc), *ptr, float(x), float(y),
                                            _dptr(f, (2,), self.device, "f")), "chy_find_force_noalloc")

    def stage_times(self):
        """{stage: (total_ms, launches)} since profiling was enabled (waits)."""
        ms = (C.c_double * len(STAGES))()
        cnt = (C.c_int * len(STAGES))()
        _check(lib().chy_ctx_stage_times(self.handle, ms, cnt, len(STAGES)), "chy_ctx_stage_times")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(STAGES)}

    def close(self):
        if self.handle:
            lib().chy_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PreparedGrid:
    """SETBDY: a grid with its device-resident tables and workspace."""

    def __init__(self, ctx: Context, grid, isym: int = 2, pi: float = fields.PI, grav: float = 1.0,
                 r=None, z=None, rhf=None, zhf=None, dtheta=None):
        self.ctx = ctx
        self.grid = grid
        self._arrs = [np.ascontiguousarray(a if a is not None else d(), dtype=np.float64)
                      for a, d in ((r, grid.r), (z, grid.z), (rhf, grid.rhf), (zhf, grid.zhf))]
        dp = C.POINTER(C.c_double)
        desc = GridDesc(grid.jmax, grid.kmax, grid.lmax,
                        *[a.ctypes.data_as(dp) for a in self._arrs],
                        grid.dtheta if dtheta is None else dtheta, pi, grav)
        h = C.c_void_p()
        _check(lib().chy_setbdy(ctx.handle, C.byref(desc), isym, C.byref(h)), "chy_setbdy")
        self.handle = h
        self.isym = isym

    @property
    def device_bytes(self) -> int:
        return int(lib().chy_grid_device_bytes(self.handle))

    @property
    def mode_solver(self) -> int:
        """MODES_CR (BLKTRI's cyclic reduction, the reference's results) or
        MODES_EIGEN (z eigenbasis on the matrix cores)."""
        return int(lib().chy_grid_mode_solver(self.handle))

    @mode_solver.setter
    def mode_solver(self, kind: int):
        _check(lib().chy_grid_set_mode_solver(self.handle, int(kind)), "chy_grid_set_mode_solver")

    @property
    def cr_lanes(self) -> int:
        """Lanes per r system in the cyclic reduction (64 or 32)."""
        return int(lib().chy_grid_tune(self.handle, TUNE_CR_LANES))

    @cr_lanes.setter
    def cr_lanes(self, lanes: int):
        _check(lib().chy_grid_set_tune(self.handle, TUNE_CR_LANES, int(lanes)), "chy_grid_set_tune")

    @property
    def refine_m64(self) -> int:
        """Refined solves: orders m <= this value are solved in binary64."""
        return int(lib().chy_grid_tune(self.handle, TUNE_REFINE_M64))

    @refine_m64.setter
    def refine_m64(self, m: int):
        _check(lib().chy_grid_set_tune(self.handle, TUNE_REFINE_M64, int(m)), "chy_grid_set_tune")

    @property
    def refine_fused(self) -> int:
        """Refined solves of 1 or 2 steps launched for at most this many
        fields run all mode work of a plane in one kernel (0: never).  True
        and False set 32 and 0."""
        return int(lib().chy_grid_tune(self.handle, TUNE_REFINE_FUSED))

    @refine_fused.setter
    def refine_fused(self, n):
        n = 32 if n is True else (0 if n is False else int(n))
        _check(lib().chy_grid_set_tune(self.handle, TUNE_REFINE_FUSED, n), "chy_grid_set_tune")

    # device tensor shapes: PHI/RHO (LMAX, KMAX+2, JMAX+2); COEF (2, LMAX/2, KMAX+2, JMAX+2)
    @property
    def field_shape(self):
        g = self.grid
        return (g.lmax, g.kmax + 2, g.jmax + 2)

    @property
    def coef_shape(self):
        g = self.grid
        return (2, g.lmax // 2, g.kmax + 2, g.jmax + 2)

    def _f(self, t, name):
        return _dptr(t, self.field_shape, self.ctx.device, name)

    def _c(self, t):
        return _dptr(t, self.coef_shape, self.ctx.device, "coef")

    # -- device-resident forms (torch CUDA float64 tensors) --
    def bdygen(self, rho, phi, maxtrm=10, isym=None, redge=0.0, diag=False):
        d = Diag() if diag else None
        _check(lib().chy_bdygen(self.handle, maxtrm, self.isym if isym is None else isym, redge,
                                self._f(rho, "rho"), self._f(phi, "phi"), C.byref(d) if d else None), "chy_bdygen")
        return d

    def pot3(self, rho, phi, npoint=8, coef=None, coef_mode=COEF_NONE, diag=None):
        d = diag if isinstance(diag, Diag) else (Diag() if diag else None)
        _check(lib().chy_pot3(self.handle, npoint, self._f(rho, "rho"), self._f(phi, "phi"), self._c(coef),
                              coef_mode, C.byref(d) if d else None), "chy_pot3")
        return d

    def solve(self, rho, phi, maxtrm=10, isym=None, redge=0.0, npoint=8, coef=None,
              coef_mode=COEF_NONE, diag=None):
        d = diag if isinstance(diag, Diag) else (Diag() if diag else None)
        _check(lib().chy_solve(self.handle, maxtrm, self.isym if isym is None else isym, redge, npoint,
                               self._f(rho, "rho"), self._f(phi, "phi"), self._c(coef), coef_mode,
                               C.byref(d) if d else None), "chy_solve")
        return d

    def solve_refined(self, rho, phi, steps=2, maxtrm=10, isym=None, redge=0.0, npoint=8, coef=None,
                      coef_mode=COEF_NONE, diag=None, report=False):
        """BDYGEN + POT3 by refinement with binary32 correction solves.
        Returns (Diag or None, per-step relative corrections or None)."""
        d = diag if isinstance(diag, Diag) else (Diag() if diag else None)
        rel = (C.c_double * steps)() if report else None
        _check(lib().chy_solve_refined(self.handle, maxtrm, self.isym if isym is None else isym, redge, npoint,
                                       self._f(rho, "rho"), self._f(phi, "phi"), self._c(coef), coef_mode, steps,
                                       C.byref(d) if d else None, rel), "chy_solve_refined")
        return d, (list(rel) if report else None)

    def zaxphi(self, phi, npoint=8, klocat=0):
        """ZAXPHI alone on a device PHI (pot3.f:311-399); returns the diagnostics."""
        d = Diag(0.0, 0.0, klocat)
        _check(lib().chy_zaxphi(self.handle, npoint, self._f(phi, "phi"), C.byref(d)), "chy_zaxphi")
        return d

    def phac(self):
        """PHAC(K), K=2..KMAX+1, of the last POT3/ZAXPHI on this grid."""
        out = np.zeros(self.grid.kmax)
        _check(lib().chy_phac(self.handle, out.ctypes.data_as(C.POINTER(C.c_double)), self.grid.kmax),
               "chy_phac")
        return out

    # -- host forms (numpy Fortran-ordered float64 arrays, updated in place) --
    def bdygen_host(self, rho, phi, maxtrm=10, isym=None, redge=0.0):
        _host_ok(self.grid.shape, rho=rho, phi=phi)
        d = Diag()
        dp = C.POINTER(C.c_double)
        _check(lib().chy_bdygen_host(self.handle, maxtrm, self.isym if isym is None else isym, redge,
                                     rho.ctypes.data_as(dp), phi.ctypes.data_as(dp), C.byref(d)),
               "chy_bdygen_host")
        return d

    def pot3_host(self, rho, phi, npoint=8, coef=None, coef_mode=COEF_NONE, klocat=0):
        _host_ok(self.grid.shape, rho=rho, phi=phi)
        if coef is not None:
            g = self.grid
            _host_ok((g.jmax + 2, g.kmax + 2, g.lmax // 2, 2), coef=coef)
        d = Diag(0.0, 0.0, klocat)
        dp = C.POINTER(C.c_double)
        cp = coef.ctypes.data_as(dp) if coef is not None else None
        _check(lib().chy_pot3_host(self.handle, npoint, rho.ctypes.data_as(dp), phi.ctypes.data_as(dp),
                                   cp, coef_mode, C.byref(d)), "chy_pot3_host")
        return d

    def close(self):
        if getattr(self, "handle", None):
            lib().chy_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _host_ok(shape, **arrs):
    """Host arrays must be Fortran-ordered float64 numpy arrays of `shape`
    (the COMMON block layout the library copies to and from)."""
    for name, a in arrs.items():
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
            raise ValueError(f"{name}: host arrays must be Fortran-ordered float64 numpy arrays")
        if a.shape != tuple(shape):
            raise ValueError(f"{name}: shape {a.shape}, expected {tuple(shape)}")


def to_device(a, device="cuda"):
    """Fortran-ordered (J+2,K+2,L) numpy array -> torch CUDA tensor (L,K+2,J+2)."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(device)


def to_host(t):
    """torch tensor (L,K+2,J+2) -> Fortran-ordered numpy array (J+2,K+2,L)."""
    return np.asfortranarray(t.detach().cpu().numpy().T)
