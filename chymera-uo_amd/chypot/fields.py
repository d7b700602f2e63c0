"""Grids and seeded synthetic density fields for the potential solve.

Arrays follow the reference's Fortran layout ``PHI(JMAX+2, KMAX+2, LMAX)`` /
``RHO(...)`` (``headers/globals.h:55-58``): J (cylindrical radius) is the fastest
index, then K (height), then L (azimuth).  In numpy they are shaped
``(JMAX+2, KMAX+2, LMAX)`` with ``order='F'`` so that the raw buffer is exactly
the Fortran COMMON block byte layout.

The grid is the uniform one that the ITYPE=71 driver builds
(``radhydro/io.f:188-205``): ``R(J)=(J-2)*DELR``, ``RHF(J)=R(J)+DELR/2``,
``Z(K)=(K-2)*DELZ``, ``ZHF(K)=Z(K)+DELZ/2``; ``DTHETA=2*PI/LMAX``
(``io.f:140``).  Densities are placed on J=2..JMAX+1, K=2..KMAX+1 and floored at
``gridlim*max(rho)`` with ``gridlim=1e-10`` (``io.f:170-185``,
``headers/units.h`` ``gridlim``).  Cell-centre angles are ``psi=(L-1/2)*DTHETA``
(the convention of ``pot/boundary.f:32``).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

PI = 3.1415926535897932  # radhydro/io.f:137
GRIDLIM = 1.0e-10        # headers/units.h gridlim


@dataclasses.dataclass(frozen=True)
class Grid:
    """One potential-solve grid (sizes are the reference's JMAX, KMAX, LMAX)."""

    jmax: int
    kmax: int
    lmax: int
    delr: float
    delz: float

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.jmax + 2, self.kmax + 2, self.lmax)

    @property
    def dtheta(self) -> float:
        return 2.0 * PI / float(self.lmax)

    def r(self) -> np.ndarray:
        return (np.arange(1, self.jmax + 3) - 2) * self.delr

    def z(self) -> np.ndarray:
        return (np.arange(1, self.kmax + 3) - 2) * self.delz

    def rhf(self) -> np.ndarray:
        return self.r() + self.delr / 2.0

    def zhf(self) -> np.ndarray:
        return self.z() + self.delz / 2.0

    def psi(self) -> np.ndarray:
        return (np.arange(1, self.lmax + 1) - 0.5) * self.dtheta

    @property
    def name(self) -> str:
        return f"{self.jmax}x{self.kmax}x{self.lmax}"


def unit_grid(jmax: int, kmax: int, lmax: int) -> Grid:
    """Grid covering r in [0,1], z in [0,1/2]: DELR=1/JMAX, DELZ=0.5/KMAX."""
    return Grid(jmax, kmax, lmax, 1.0 / jmax, 0.5 / kmax)


def _centres(g: Grid):
    r = g.rhf()[:, None, None]
    z = g.zhf()[None, :, None]
    psi = g.psi()[None, None, :]
    return r, z, psi


def _place(g: Grid, field: np.ndarray) -> np.ndarray:
    """Copy field onto J=2..JMAX+1, K=2..KMAX+1, floor it, zero the rest."""
    out = np.zeros(g.shape, dtype=np.float64, order="F")
    inner = field[1:g.jmax + 1, 1:g.kmax + 1, :]
    floor = GRIDLIM * float(inner.max())
    out[1:g.jmax + 1, 1:g.kmax + 1, :] = np.maximum(inner, floor)
    return out


def torus(g: Grid, rc: float = 0.5, a: float = 0.3, n: float = 1.5) -> np.ndarray:
    """Polytropic torus profile max(0, 1-((r-rc)^2+z^2)/a^2)^n on cell centres."""
    r, z, _ = _centres(g)
    base = np.maximum(0.0, 1.0 - ((r - rc) ** 2 + z ** 2) / (a * a)) ** n
    return np.broadcast_to(base, g.shape).copy()


def density_small_torus(g: Grid, seed: int = 1) -> np.ndarray:
    """Torus (rc=0.5, a=0.3, n=1.5) times 1+1e-3*cos(2 psi) plus seeded uniform
    noise of amplitude 1e-4.  At 64x32x64 with DELR=DELZ=1/64 this is the
    CPU-runnable single-field workload."""
    _, _, psi = _centres(g)
    rng = np.random.default_rng(seed)
    noise = rng.uniform(-1.0, 1.0, size=g.shape)
    f = torus(g) * (1.0 + 1.0e-3 * np.cos(2.0 * psi)) + 1.0e-4 * noise
    return _place(g, f)


def density_spiral_torus(g: Grid, seed: int = 2) -> np.ndarray:
    """Torus (rc=0.5, a=0.3, n=1.5) times seeded log-normal noise exp(0.05 g),
    g~N(0,1), times 1 + 0.1*sum_{m=2,3,4} cos(m (psi - ln(r/0.5)/tan 15deg))
    (logarithmic spiral arms, pitch 15 degrees).  This is the 256x128x256
    plain-FP64 workload."""
    r, _, psi = _centres(g)
    rng = np.random.default_rng(seed)
    lognorm = np.exp(0.05 * rng.standard_normal(size=g.shape))
    with np.errstate(invalid="ignore", divide="ignore"):  # r<0 only in the J=1 ghost
        wind = np.log(r / 0.5) / math.tan(math.radians(15.0))
    arms = 1.0 + 0.1 * sum(np.cos(m * (psi - wind)) for m in (2, 3, 4))
    f = torus(g) * lognorm * arms
    return _place(g, f)


def random_field(seed: int, grids=((64, 128, 256), (32, 64, 128), (64, 128, 256))):
    """One field of the heterogeneous batch: JMAX, KMAX, LMAX drawn from the
    given choices, DELR=1/JMAX, DELZ=0.5/KMAX, torus centre rc~U(0.3,0.6),
    radius a~U(0.15,0.35), polytropic index n in {1,1.5,2.5}, log-normal noise
    with sigma 0.05.  Returns (grid, rho)."""
    rng = np.random.default_rng(seed)
    j = int(rng.choice(grids[0]))
    k = int(rng.choice(grids[1]))
    l = int(rng.choice(grids[2]))
    g = unit_grid(j, k, l)
    rc = float(rng.uniform(0.3, 0.6))
    a = float(rng.uniform(0.15, 0.35))
    n = float(rng.choice([1.0, 1.5, 2.5]))
    lognorm = np.exp(0.05 * rng.standard_normal(size=g.shape))
    f = torus(g, rc, a, n) * lognorm
    return g, _place(g, f)


def torus_device(g: Grid, rc: float, a: float, n: float, sigma: float, seed: int):
    """The torus field of random_field generated on the GPU as a torch tensor
    of shape (LMAX, KMAX+2, JMAX+2) (the Fortran layout): polytropic torus
    times exp(sigma * N(0,1)) noise from a seeded torch generator, floored at
    GRIDLIM times its maximum on the interior, zero outside it.  For large
    synthetic workloads (the noise differs from random_field's numpy stream)."""
    import torch
    dev = "cuda"
    r = torch.tensor(g.rhf(), dtype=torch.float64, device=dev)[None, :]
    z = torch.tensor(g.zhf(), dtype=torch.float64, device=dev)[:, None]
    base = torch.clamp(1.0 - ((r - rc) ** 2 + z ** 2) / (a * a), min=0.0) ** n
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    noise = torch.randn((g.lmax, g.kmax + 2, g.jmax + 2), generator=gen, dtype=torch.float64, device=dev)
    f = base[None] * torch.exp(sigma * noise)
    out = torch.zeros_like(f)
    inner = f[:, 1:g.kmax + 1, 1:g.jmax + 1]
    out[:, 1:g.kmax + 1, 1:g.jmax + 1] = torch.clamp(inner, min=GRIDLIM * float(inner.max()))
    return out


def random_field_device(seed: int, grids=((64, 128, 256), (32, 64, 128), (64, 128, 256))):
    """random_field's grid and torus parameters (same seed, same draws), the
    field generated on the GPU by torus_device.  Returns (grid, tensor)."""
    rng = np.random.default_rng(seed)
    j = int(rng.choice(grids[0]))
    k = int(rng.choice(grids[1]))
    l = int(rng.choice(grids[2]))
    g = unit_grid(j, k, l)
    rc = float(rng.uniform(0.3, 0.6))
    a = float(rng.uniform(0.15, 0.35))
    n = float(rng.choice([1.0, 1.5, 2.5]))
    return g, torus_device(g, rc, a, n, 0.05, seed)
