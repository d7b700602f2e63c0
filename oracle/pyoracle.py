"""ctypes binding of the CPU restatement in chyoracle.c (TEST INFRASTRUCTURE).

Only tests/, the smoke() entry and bench.py's cpu_baseline leg use this module,
and only as the checker.  The library is built by ``make -C oracle`` (also run
by ``__graft_entry__.build()``) into ``oracle/_build/libchyoracle.so``.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libchyoracle.so"
EXT_LIB_PATH = HERE / "_build" / "libchyoracle_ext.so"
_libs = {}
_extended = False


def use_extended(flag: bool):
    """Route the calls below to the 80-bit build (same algorithm, same FP64 inputs)."""
    global _extended
    _extended = bool(flag)


def lib():
    path = EXT_LIB_PATH if _extended else LIB_PATH
    if path not in _libs:
        if not path.exists():
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = C.CDLL(str(path))
        d = C.POINTER(C.c_double)
        i = C.POINTER(C.c_int)
        L.ora_rfft_fwd.argtypes = [C.c_int, d, d, d]
        L.ora_rfft_inv.argtypes = [C.c_int, d, d, d]
        L.ora_mode_solve.argtypes = [C.c_int, C.c_int, d, d, d, d, d, d, d]
        L.ora_cr_solve.argtypes = [C.c_int, C.c_int, d, d, d, d, d, d, d]
        L.ora_cr_variant.argtypes = [C.c_int, C.c_int]
        L.ora_mode_solver.argtypes = [C.c_int]
        grid = [C.c_int, C.c_int, C.c_int, d, d, d, d, C.c_double, C.c_double, C.c_double]
        L.ora_pot3.argtypes = grid + [C.c_int, d, d, d, d, i]
        L.ora_bdygen.argtypes = grid + [C.c_int, C.c_int, C.c_double, d, d, d]
        L.ora_residual.argtypes = [C.c_int, C.c_int, C.c_int, d, d, C.c_double, C.c_double,
                                   C.c_double, d, d, d, d]
        L.ora_terms.argtypes = [i, i, d, i, d]
        L.ora_source_grav.argtypes = [C.c_int] * 6 + [C.c_double] * 4 + [d] * 8
        L.ora_external_pot.argtypes = [C.c_int] * 5 + [C.c_double] * 5 + [d] * 4
        L.ora_find_force.argtypes = [C.c_int] * 4 + [C.c_double] * 4 + [d] * 3 + [C.c_double] * 2 + [d] * 2
        _libs[path] = L
    return _libs[path]


def terms():
    """The (l, m) multipole table: list of (l, m, dlm, poly[np]) with l+m even."""
    l = (C.c_int * 64)(); m = (C.c_int * 64)(); np_ = (C.c_int * 64)()
    dlm = np.zeros(64); poly = np.zeros(64 * 6)
    n = lib().ora_terms(l, m, _p(dlm), np_, _p(poly))
    return [(l[t], m[t], float(dlm[t]), poly[t * 6:t * 6 + np_[t]].copy()) for t in range(n)]


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def rfft_fwd(x):
    """Raw FFT(+1)+REALTR(+1) of a real lane: returns (A[0..N], B[0..N])."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    L = x.size
    A = np.zeros(L // 2 + 1)
    B = np.zeros(L // 2 + 1)
    rc = lib().ora_rfft_fwd(L, _p(x), _p(A), _p(B))
    if rc:
        raise ValueError(f"ora_rfft_fwd rc={rc}")
    return A, B


def rfft_inv(A, B):
    """Raw REALTR(-1)+FFT(-1) with B[0]=B[N]=0: returns the L interleaved samples."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    L = 2 * (A.size - 1)
    out = np.zeros(L)
    rc = lib().ora_rfft_inv(L, _p(A), _p(B), _p(out))
    if rc:
        raise ValueError(f"ora_rfft_inv rc={rc}")
    return out


def mode_solve(an, bn, cn, am, bm, cm, y):
    """Solve the BLKTRI system; y is Y(M,N) as a numpy (N,M) row-major array."""
    y = np.array(y, dtype=np.float64, order="C", copy=True)
    n, m = y.shape
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (an, bn, cn, am, bm, cm)]
    rc = lib().ora_mode_solve(m, n, *[_p(a) for a in args], _p(y))
    if rc:
        raise ValueError(f"ora_mode_solve rc={rc}")
    return y


def cr_solve(an, bn, cn, am, bm, cm, y, roots="reference", order="reference"):
    """The same system by BLKTRI's cyclic reduction (pot/blktri.f:60-260).
    roots: "reference" (TQLRAT's rule, the reference's values) or "exact"
    (80-bit eigenvalues rounded once); order: "reference" (r solves eliminate
    from the last row up) or "other" (from the first row down)."""
    y = np.array(y, dtype=np.float64, order="C", copy=True)
    n, m = y.shape
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (an, bn, cn, am, bm, cm)]
    L = lib()
    L.ora_cr_variant(0 if roots == "reference" else 1, 0 if order == "reference" else 1)
    try:
        rc = L.ora_cr_solve(m, n, *[_p(a) for a in args], _p(y))
    finally:
        L.ora_cr_variant(0, 0)
    if rc:
        raise ValueError(f"ora_cr_solve rc={rc}")
    return y


_modes = 0


def use_cyclic_reduction(flag: bool):
    """POT3's mode solves (solve(), pot3()) by BLKTRI's cyclic reduction, the
    reference's own arithmetic (True), or by the z eigenbasis (False, default)."""
    global _modes
    _modes = 1 if flag else 0


def _grid_args(g):
    r = np.ascontiguousarray(g.r())
    z = np.ascontiguousarray(g.z())
    rhf = np.ascontiguousarray(g.rhf())
    zhf = np.ascontiguousarray(g.zhf())
    keep = (r, z, rhf, zhf)
    return keep, [g.jmax, g.kmax, g.lmax, _p(r), _p(z), _p(rhf), _p(zhf), g.dtheta,
                  3.1415926535897932, 1.0]


def bdygen(g, rho, phi=None, maxtrm=10, isym=2, redge=0.0):
    """BDYGEN: returns (phi with boundary values, tmass)."""
    rho = np.asfortranarray(rho, dtype=np.float64)
    phi = np.zeros(g.shape, order="F") if phi is None else np.array(phi, order="F", copy=True)
    tm = np.zeros(1)
    keep, ga = _grid_args(g)
    rc = lib().ora_bdygen(*ga, maxtrm, isym, redge, _p(rho), _p(phi), _p(tm))
    if rc:
        raise ValueError(f"ora_bdygen rc={rc}")
    return phi, float(tm[0])


def pot3(g, rho, phi, npoint=8, coef=None, klocat=0):
    """POT3 on phi (holding boundary values): returns (phi, coef, phichk, klocat)."""
    rho = np.asfortranarray(rho, dtype=np.float64)
    phi = np.array(phi, order="F", copy=True, dtype=np.float64)
    if coef is None:
        coef = np.zeros((g.jmax + 2, g.kmax + 2, g.lmax // 2, 2), order="F")
    else:
        coef = np.array(coef, order="F", copy=True, dtype=np.float64)
    pc = np.zeros(1)
    kl = C.c_int(klocat)
    keep, ga = _grid_args(g)
    lib().ora_mode_solver(_modes)
    rc = lib().ora_pot3(*ga, npoint, _p(rho), _p(phi), _p(coef), _p(pc), C.byref(kl))
    if rc:
        raise ValueError(f"ora_pot3 rc={rc}")
    return phi, coef, float(pc[0]), int(kl.value)


def solve(g, rho, maxtrm=10, isym=2, redge=0.0, npoint=8):
    """SETBDY+BDYGEN+POT3 on a zero-initialised PHI, as the ITYPE=71 driver."""
    phi, tmass = bdygen(g, rho, None, maxtrm, isym, redge)
    phib = phi.copy()
    phi, coef, phichk, klocat = pot3(g, rho, phi, npoint)
    return dict(phi_b=phib, tmass=tmass, phi=phi, coef=coef, phichk=phichk, klocat=klocat)


def residual(g, phi, rho):
    """Residual A phi - 4 pi G rho of the discrete operator on J=2..JMAX, K=2..KMAX,
    and the per-cell sum of |terms| (the scale of its rounding floor)."""
    phi = np.asfortranarray(phi, dtype=np.float64)
    rho = np.asfortranarray(rho, dtype=np.float64)
    res = np.zeros(g.shape, order="F")
    scale = np.zeros(g.shape, order="F")
    rhf = np.ascontiguousarray(g.rhf())
    zhf = np.ascontiguousarray(g.zhf())
    lib().ora_residual(g.jmax, g.kmax, g.lmax, _p(rhf), _p(zhf), g.dtheta, 3.1415926535897932, 1.0,
                       _p(phi), _p(rho), _p(res), _p(scale))
    return res, scale


def source_grav(jmax, kmax, lmax, pjmax, pkmax, jstart, delt, dr, dz, dtheta, r, rho, omega, phi, p, s, t, a):
    """The first loop of SOURCE (radhydro/source.f:41-79); s, t, a (Fortran-
    ordered float64 arrays of shape (JMAX+2, KMAX+2, LMAX)) are updated in
    place; rho, phi have shape (PJMAX+2, PKMAX+2, LMAX), r length PJMAX+2."""
    hyd, pot = (jmax + 2, kmax + 2, lmax), (pjmax + 2, pkmax + 2, lmax)
    args = []
    for x, shape in ((r, None), (rho, pot), (omega, hyd), (phi, pot), (p, hyd), (s, hyd), (t, hyd), (a, hyd)):
        if shape is not None and (x.shape != shape or not x.flags.f_contiguous or x.dtype != np.float64):
            raise ValueError(f"expected a Fortran-ordered float64 array of shape {shape}")
        args.append(x)
    r = np.ascontiguousarray(r, dtype=np.float64)
    args[0] = r
    rc = lib().ora_source_grav(jmax, kmax, lmax, pjmax, pkmax, jstart, delt, dr, dz, dtheta, *[_p(x) for x in args])
    if rc:
        raise ValueError("ora_source_grav: bad arguments")


def external_pot(jmax, kmax, lmax, pjmax, pkmax, dtheta, mass_star, tmassacc, rpstar, phi_star, rhf, zhf, phi,
                 starphi):
    """ExternalPot (pot/ExternalPot.f:1-88, INDIRECT=0, BINARY=0): starphi (Fortran
    order, (JMAX+2, KMAX+2, LMAX)) is written and added to phi ((PJMAX+2,
    PKMAX+2, LMAX)) in place."""
    for x, shape in ((phi, (pjmax + 2, pkmax + 2, lmax)), (starphi, (jmax + 2, kmax + 2, lmax))):
        if x.shape != shape or not x.flags.f_contiguous or x.dtype != np.float64:
            raise ValueError(f"expected a Fortran-ordered float64 array of shape {shape}")
    rhf = np.ascontiguousarray(rhf, dtype=np.float64)
    zhf = np.ascontiguousarray(zhf, dtype=np.float64)
    rc = lib().ora_external_pot(jmax, kmax, lmax, pjmax, pkmax, dtheta, mass_star, tmassacc, rpstar, phi_star,
                                _p(rhf), _p(zhf), _p(phi), _p(starphi))
    if rc:
        raise ValueError("ora_external_pot: bad arguments")


def find_force(jmax, kmax, lmax, jmin, dr, dz, dphi, gsoft, rhf, zhf, rho, x, y):
    """find_force (pot/wiggle.F:99-135) in one thread's loop order (sums in
    80-bit arithmetic under use_extended(True)): returns
    ((fx, fy), (sum |x terms|, sum |y terms|)); rho Fortran-ordered (JMAX+2,
    KMAX+2, LMAX)."""
    if rho.shape != (jmax + 2, kmax + 2, lmax) or not rho.flags.f_contiguous or rho.dtype != np.float64:
        raise ValueError("rho: expected a Fortran-ordered float64 array of shape (JMAX+2, KMAX+2, LMAX)")
    rhf = np.ascontiguousarray(rhf, dtype=np.float64)
    zhf = np.ascontiguousarray(zhf, dtype=np.float64)
    f, a = np.zeros(2), np.zeros(2)
    if lib().ora_find_force(jmax, kmax, lmax, jmin, dr, dz, dphi, gsoft, _p(rhf), _p(zhf), _p(rho), x, y, _p(f),
                            _p(a)):
        raise ValueError("ora_find_force: bad arguments")
    return (float(f[0]), float(f[1])), (float(a[0]), float(a[1]))


def set_threads(n: int):
    os.environ["OMP_NUM_THREADS"] = str(n)
