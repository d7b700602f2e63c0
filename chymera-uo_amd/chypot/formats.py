"""CHYMERA's on-disk formats around the potential solve, read and written
byte for byte as the reference's Fortran I/O does (unformatted sequential
records with 4-byte little-endian length markers before and after each record,
as gfortran and flang write them; list-directed text for fort.2).

  fort.2        the ITYPE=71 axisymmetric model (radhydro/io.f:157-167): one
                header line PINDEX, CON2, RRR2, OMCEN, DENCEN, TOVERW, ROF3N,
                ZOF3N, A1NEWZ, JREQ, KZPOL, then DENNY(JMAX+2, KMAX+2)
  fort.7 / saved.NNNNNN
                the restart model (written radhydro/io.f:1749-1758, read
                radhydro/io.f:464-475): records S, T, A, RHO, EPS, then
                ROF3N, ZOF3N, DELT, TIME, ELOST, DEN, SOUND, JREQ (int32), OMMAX,
                then (only when JMIN > 2) TMASSINI, TMASS, TMASSADD, TMASSOUT,
                TMASSACC, TOTCOOL, TOTDFLUX, TOTHEAT, TOTIRR, ETOTFL, EFLUFFTOT.
                S, T, A, EPS have the hydro grid's shape (JMAX+2, KMAX+2, LMAX),
                RHO the potential grid's (POT3JMAX+2, POT3KMAX+2, LMAX)
  phi3d.NNNNNN  the potential (radhydro/io.f:221-227 and the hydro loop's
                dumps): records PHI, TIME

Arrays are numpy float64 in Fortran order.
"""
from __future__ import annotations

import pathlib
import struct

import numpy as np

SCALARS = ("rof3n", "zof3n", "delt", "time", "elost", "den", "sound", "jreq", "ommax")
TAIL = ("tmassini", "tmass", "tmassadd", "tmassout", "tmassacc", "totcool", "totdflux", "totheat", "totirr",
        "etotfl", "eflufftot")
FORT2_HEADER = ("pindex", "con2", "rrr2", "omcen", "dencen", "toverw", "rof3n", "zof3n", "a1newz", "jreq",
                "kzpol")


# ---- unformatted sequential records ---------------------------------------

def write_records(path, records):
    """Write byte strings as unformatted sequential records."""
    with open(path, "wb") as f:
        for rec in records:
            n = struct.pack("<i", len(rec))
            f.write(n + rec + n)


def read_records(path):
    """All records of an unformatted sequential file as byte strings; raises
    ValueError when a record's leading and trailing markers disagree."""
    data = pathlib.Path(path).read_bytes()
    out, pos = [], 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {pos}")
        (n,) = struct.unpack_from("<i", data, pos)
        end = pos + 4 + n
        if n < 0 or end + 4 > len(data) or struct.unpack_from("<i", data, end)[0] != n:
            raise ValueError(f"{path}: bad record at byte {pos}")
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def _array(rec, shape, what):
    want = int(np.prod(shape)) * 8
    if len(rec) != want:
        raise ValueError(f"{what}: record of {len(rec)} bytes, expected {want} for shape {shape}")
    return np.frombuffer(rec, dtype="<f8").reshape(shape, order="F").copy(order="F")


def _bytes(a):
    return np.asarray(a, dtype="<f8").tobytes(order="F")


# ---- restart (fort.7 / saved.NNNNNN) ----------------------------------------

def write_restart(path, s, t, a, rho, eps, scalars: dict, tail: dict | None = None):
    """saved.NNNNNN as radhydro/io.f:1749-1758 writes it.  scalars: the nine
    values of SCALARS (JREQ an integer); tail: the eleven of TAIL, written
    only when given (the reference writes them when JMIN > 2)."""
    recs = [_bytes(x) for x in (s, t, a, rho, eps)]
    v = [float(scalars[k]) for k in SCALARS[:7]]
    recs.append(struct.pack("<7d", *v) + struct.pack("<i", int(scalars["jreq"])) +
                struct.pack("<d", float(scalars["ommax"])))
    if tail is not None:
        recs.append(struct.pack("<11d", *[float(tail[k]) for k in TAIL]))
    write_records(path, recs)


def read_restart(path, hydro_shape, pot_shape):
    """fort.7 as radhydro/io.f:464-475 reads it: dict with s, t, a, eps (hydro
    grid), rho (potential grid), the SCALARS and, when the file has the last
    record (JMIN > 2), the TAIL values."""
    recs = read_records(path)
    if len(recs) not in (6, 7):
        raise ValueError(f"{path}: {len(recs)} records, expected 6 or 7")
    out = {"s": _array(recs[0], hydro_shape, "S"), "t": _array(recs[1], hydro_shape, "T"),
           "a": _array(recs[2], hydro_shape, "A"), "rho": _array(recs[3], pot_shape, "RHO"),
           "eps": _array(recs[4], hydro_shape, "EPS")}
    if len(recs[5]) != 7 * 8 + 4 + 8:
        raise ValueError(f"{path}: scalar record of {len(recs[5])} bytes, expected 68")
    vals = struct.unpack("<7d", recs[5][:56]) + struct.unpack("<i", recs[5][56:60]) + \
        struct.unpack("<d", recs[5][60:68])
    out.update(dict(zip(SCALARS, vals)))
    if len(recs) == 7:
        if len(recs[6]) != 11 * 8:
            raise ValueError(f"{path}: last record of {len(recs[6])} bytes, expected 88")
        out.update(dict(zip(TAIL, struct.unpack("<11d", recs[6]))))
    return out


# ---- phi3d.NNNNNN -----------------------------------------------------------

def write_phi3d(path, phi, time: float):
    write_records(path, [_bytes(phi), struct.pack("<d", float(time))])


def read_phi3d(path, shape):
    """(PHI, TIME) of a phi3d file; shape (POT3JMAX+2, POT3KMAX+2, LMAX)."""
    recs = read_records(path)
    if len(recs) != 2 or len(recs[1]) != 8:
        raise ValueError(f"{path}: expected two records, PHI and TIME")
    return _array(recs[0], shape, "PHI"), struct.unpack("<d", recs[1])[0]


# ---- fort.2 -----------------------------------------------------------------

def write_fort2(path, header: dict, denny):
    """fort.2 for the ITYPE=71 read (radhydro/io.f:157-167): the header on one
    line, then DENNY(JMAX+2, KMAX+2) in Fortran order, list-directed."""
    h = [f"{float(header[k]):.17e}" for k in FORT2_HEADER[:9]] + [str(int(header[k])) for k in FORT2_HEADER[9:]]
    vals = np.asarray(denny, dtype=np.float64).ravel(order="F")
    lines = [" ".join(h)] + [" ".join(f"{v:.17e}" for v in vals[i:i + 4]) for i in range(0, vals.size, 4)]
    pathlib.Path(path).write_text("\n".join(lines) + "\n")


def read_fort2(path, shape):
    """(header dict, DENNY) of a fort.2 file; shape (JMAX+2, KMAX+2).  Read as
    the two list-directed READs of io.f:159-164 do: the header takes the first
    11 values and the rest of the line that completes it is skipped; DENNY
    then takes JMAX+2 * KMAX+2 values from the following lines.  Values are
    separated by blanks or commas, D is accepted as an exponent letter;
    repeat counts (r*c) and slashes are not supported."""
    lines = pathlib.Path(path).read_text().splitlines()
    def values(line):
        return line.replace(",", " ").replace("D", "E").replace("d", "e").split()
    head_tok, i = [], 0
    while len(head_tok) < 11:
        if i >= len(lines):
            raise ValueError(f"{path}: header has fewer than 11 values")
        head_tok += values(lines[i])
        i += 1
    head = {k: (int(float(v)) if k in ("jreq", "kzpol") else float(v))
            for k, v in zip(FORT2_HEADER, head_tok[:11])}
    need = shape[0] * shape[1]
    tok = []
    while len(tok) < need and i < len(lines):
        tok += values(lines[i])
        i += 1
    if len(tok) < need:
        raise ValueError(f"{path}: DENNY has {len(tok)} values, expected {need}")
    denny = np.array([float(v) for v in tok[:need]]).reshape(shape, order="F")
    return head, np.asfortranarray(denny)
