"""CHYMERA's files around the potential solve (chypot.formats): the restart
model fort.7 / saved.NNNNNN (radhydro/io.f:464-475 reads it, :1749-1758
writes it), phi3d.NNNNNN (io.f:221-227) and fort.2 (io.f:157-167).

CPU: round trips, and exchange with a Fortran compiler's own I/O -- a small
program compiled here with flang writes and reads the records in the order
io.f uses, so both the record markers and the record contents are checked
against a Fortran runtime, not against this module's own reading of them.
GPU: potdiag's restart input solves the same PHI as its raw-density input.
"""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from chypot import formats

FLANG = shutil.which("flang")
J2, K2, L, PJ2, PK2 = 6, 5, 4, 8, 7  # hydro (JMAX+2, KMAX+2, LMAX), potential grid (POT3JMAX+2, POT3KMAX+2)
SCAL = dict(rof3n=0.125, zof3n=0.0625, delt=1.5e-3, time=42.25, elost=-3.0e-7, den=2.5, sound=0.75, jreq=17,
            ommax=1.25)
TAIL = dict(zip(formats.TAIL, [0.5 + i for i in range(11)]))


def fields(seed=1):
    rng = np.random.default_rng(seed)
    h = [np.asfortranarray(rng.standard_normal((J2, K2, L))) for _ in range(4)]
    rho = np.asfortranarray(rng.standard_normal((PJ2, PK2, L)))
    return h[0], h[1], h[2], rho, h[3]


@pytest.mark.parametrize("tail", [None, TAIL])
def test_restart_round_trip(tmp_path, tail):
    s, t, a, rho, eps = fields()
    formats.write_restart(tmp_path / "saved.000001", s, t, a, rho, eps, SCAL, tail)
    back = formats.read_restart(tmp_path / "saved.000001", (J2, K2, L), (PJ2, PK2, L))
    for k, v in (("s", s), ("t", t), ("a", a), ("rho", rho), ("eps", eps)):
        assert np.array_equal(back[k], v)
    for k, v in SCAL.items():
        assert back[k] == v
    assert all(back.get(k) == v for k, v in (tail or {}).items())
    assert ("tmassini" in back) == (tail is not None)
    raw = (tmp_path / "saved.000001").read_bytes()
    assert int.from_bytes(raw[:4], "little") == J2 * K2 * L * 8  # the S record's marker


def test_restart_rejects_wrong_shapes(tmp_path):
    s, t, a, rho, eps = fields()
    formats.write_restart(tmp_path / "f7", s, t, a, rho, eps, SCAL)
    with pytest.raises(ValueError):
        formats.read_restart(tmp_path / "f7", (J2, K2, L), (J2, K2, L))  # RHO on the hydro grid: wrong size
    (tmp_path / "bad").write_bytes((tmp_path / "f7").read_bytes()[:-2])
    with pytest.raises(ValueError):
        formats.read_records(tmp_path / "bad")


def test_phi3d_and_fort2_round_trip(tmp_path):
    phi = fields(2)[3]
    formats.write_phi3d(tmp_path / "phi3d.000007", phi, 3.5)
    p, time = formats.read_phi3d(tmp_path / "phi3d.000007", phi.shape)
    assert np.array_equal(p, phi) and time == 3.5
    head = dict(zip(formats.FORT2_HEADER, [1.5, 0.1, 0.2, 0.3, 2.0, 0.05, 0.125, 0.0625, 0.9, 12, 3]))
    denny = np.asfortranarray(np.random.default_rng(3).uniform(0, 1, (J2, K2)))
    formats.write_fort2(tmp_path / "fort.2", head, denny)
    h, d = formats.read_fort2(tmp_path / "fort.2", (J2, K2))
    assert h == head and np.array_equal(d, denny)


def test_fort2_header_line_rest_is_skipped(tmp_path):
    """A list-directed READ skips the rest of the record that completes its
    list: extra values on the header line are not part of DENNY."""
    vals = " ".join(str(v) for v in range(1, J2 * K2 + 1))
    (tmp_path / "fort.2").write_text("1 2 3 4 5 6 7 8 9 10 11 99 98\n" + vals + "\n")
    h, d = formats.read_fort2(tmp_path / "fort.2", (J2, K2))
    assert h["kzpol"] == 11 and d[0, 0] == 1.0 and d[-1, -1] == J2 * K2


FPROG = """
      PROGRAM FMT
      IMPLICIT real*8 (a-h,o-z)
      INTEGER J2, K2, L, PJ2, PK2
      PARAMETER (J2={j2}, K2={k2}, L={l}, PJ2={pj2}, PK2={pk2})
      DIMENSION S(J2,K2,L), T(J2,K2,L), A(J2,K2,L), EPS(J2,K2,L)
      DIMENSION RHO(PJ2,PK2,L), PHI(PJ2,PK2,L), DENNY(J2,K2)
      INTEGER JREQ, KZPOL
      CHARACTER*8 MODE
      CALL GET_COMMAND_ARGUMENT(1, MODE)
      IF (MODE.EQ.'readrst') THEN
C       the read sequence of io.f:464-475
        OPEN(UNIT=7,FILE='fort.7',FORM='UNFORMATTED',STATUS='OLD')
        READ(7) S
        READ(7) T
        READ(7) A
        READ(7) RHO
        READ(7) EPS
        READ(7) ROF3N,ZOF3N,DELT,TIME,ELOST,DEN,SOUND,JREQ,OMMAX
        CLOSE(7)
        OPEN(UNIT=9,FILE='echo.bin',ACCESS='STREAM',FORM='UNFORMATTED')
        WRITE(9) S, T, A, RHO, EPS
        WRITE(9) ROF3N,ZOF3N,DELT,TIME,ELOST,DEN,SOUND,DBLE(JREQ),OMMAX
        CLOSE(9)
      ELSE IF (MODE.EQ.'writerst') THEN
C       the write sequence of io.f:1749-1758, JMIN > 2
        DO LL=1,L
          DO K=1,K2
            DO J=1,J2
              S(J,K,LL)=J+10*K+100*LL
              T(J,K,LL)=-S(J,K,LL)
              A(J,K,LL)=0.5D0*S(J,K,LL)
              EPS(J,K,LL)=0.25D0*S(J,K,LL)
            ENDDO
          ENDDO
          DO K=1,PK2
            DO J=1,PJ2
              RHO(J,K,LL)=J+10*K+100*LL+0.125D0
            ENDDO
          ENDDO
        ENDDO
        OPEN(UNIT=8,FILE='saved.000003',FORM='UNFORMATTED')
        WRITE(8) S
        WRITE(8) T
        WRITE(8) A
        WRITE(8) RHO
        WRITE(8) EPS
        WRITE(8) 0.125D0,0.0625D0,1.5D-3,42.25D0,-3.0D-7,2.5D0,0.75D0,
     &           17,1.25D0
        WRITE(8) (0.5D0+I, I=0,10)
        CLOSE(8)
        OPEN(UNIT=23,FILE='phi3d.000003',FORM='UNFORMATTED')
        WRITE(23) RHO
        WRITE(23) 42.25D0
        CLOSE(23)
      ELSE
C       the two list-directed reads of io.f:159-164
        OPEN(UNIT=2,FILE='fort.2',STATUS='OLD')
        READ(2,*) PINDEX,CON2,RRR2,OMCEN,DENCEN,TOVERW,ROF3N,ZOF3N,
     &            A1NEWZ,JREQ,KZPOL
        READ(2,*) DENNY
        CLOSE(2)
        OPEN(UNIT=9,FILE='echo.bin',ACCESS='STREAM',FORM='UNFORMATTED')
        WRITE(9) PINDEX,CON2,RRR2,OMCEN,DENCEN,TOVERW,ROF3N,ZOF3N,
     &           A1NEWZ,DBLE(JREQ),DBLE(KZPOL)
        WRITE(9) DENNY
        CLOSE(9)
      ENDIF
      END
"""


@pytest.fixture(scope="module")
def fprog(tmp_path_factory):
    if FLANG is None:
        pytest.skip("flang not installed")
    d = tmp_path_factory.mktemp("fmt")
    src = d / "fmt.f"
    src.write_text(FPROG.format(j2=J2, k2=K2, l=L, pj2=PJ2, pk2=PK2))
    subprocess.run([FLANG, "-ffixed-form", "-O1", "-o", str(d / "fmt"), str(src)], check=True,
                   capture_output=True, timeout=300)
    return d / "fmt"


def test_fortran_reads_our_restart(tmp_path, fprog):
    s, t, a, rho, eps = fields(4)
    formats.write_restart(tmp_path / "fort.7", s, t, a, rho, eps, SCAL)
    subprocess.run([str(fprog), "readrst"], cwd=tmp_path, check=True, timeout=60)
    raw = np.fromfile(tmp_path / "echo.bin", "<f8")
    n, p = J2 * K2 * L, PJ2 * PK2 * L
    got = np.split(raw, [n, 2 * n, 3 * n, 3 * n + p, 4 * n + p])
    for g, want in zip(got[:5], (s, t, a, rho, eps)):
        assert np.array_equal(g, want.ravel(order="F"))
    assert list(got[5]) == [SCAL[k] for k in formats.SCALARS]


def test_we_read_fortrans_restart_and_phi3d(tmp_path, fprog):
    subprocess.run([str(fprog), "writerst"], cwd=tmp_path, check=True, timeout=60)
    r = formats.read_restart(tmp_path / "saved.000003", (J2, K2, L), (PJ2, PK2, L))
    jj, kk, ll = np.meshgrid(np.arange(1, J2 + 1), np.arange(1, K2 + 1), np.arange(1, L + 1), indexing="ij")
    base = (jj + 10 * kk + 100 * ll).astype(float)
    assert np.array_equal(r["s"], base) and np.array_equal(r["t"], -base)
    assert np.array_equal(r["a"], 0.5 * base) and np.array_equal(r["eps"], 0.25 * base)
    pj, pk, pl = np.meshgrid(np.arange(1, PJ2 + 1), np.arange(1, PK2 + 1), np.arange(1, L + 1), indexing="ij")
    assert np.array_equal(r["rho"], pj + 10 * pk + 100 * pl + 0.125)
    assert {k: r[k] for k in formats.SCALARS} == SCAL
    assert [r[k] for k in formats.TAIL] == [0.5 + i for i in range(11)]
    phi, time = formats.read_phi3d(tmp_path / "phi3d.000003", (PJ2, PK2, L))
    assert np.array_equal(phi, r["rho"]) and time == 42.25


def test_fortran_reads_our_fort2(tmp_path, fprog):
    head = dict(zip(formats.FORT2_HEADER, [1.5, 0.1, 0.2, 0.3, 2.0, 0.05, 0.125, 0.0625, 0.9, 12, 3]))
    denny = np.asfortranarray(np.random.default_rng(5).uniform(0, 1, (J2, K2)))
    formats.write_fort2(tmp_path / "fort.2", head, denny)
    subprocess.run([str(fprog), "fort2"], cwd=tmp_path, check=True, timeout=60)
    raw = np.fromfile(tmp_path / "echo.bin", "<f8")
    assert list(raw[:11]) == [float(head[k]) for k in formats.FORT2_HEADER]
    assert np.array_equal(raw[11:], denny.ravel(order="F"))


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_potdiag_restart_input_equals_raw_density_input(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from chypot import fields as fl
    from test_fortran import nml, potdiag
    g = fl.unit_grid(32, 16, 32)
    exe = potdiag(g.name)
    rho = fl.density_spiral_torus(g, seed=6)
    zeros = np.zeros(g.shape, order="F")
    formats.write_restart(tmp_path / "fort.7", zeros, zeros, zeros, rho, zeros, dict(SCAL, rof3n=g.delr,
                                                                                      zof3n=g.delz))
    with open(tmp_path / "rho3d.bin", "wb") as f:
        f.write(np.array([g.delr, g.delz], "<f8").tobytes())
        f.write(np.asarray(rho, "<f8").tobytes(order="F"))
    for mode, fname, it in (("restart", "fort.7", 1), ("rho3d", "rho3d.bin", 2)):
        nml(tmp_path / "potdiag.nml", input=mode, rhofile=fname, iprint=0, itstop=it)
        print(f"[gpu] potdiag {mode} 32x16x32", flush=True)
        res = subprocess.run([str(exe), "potdiag.nml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr + res.stdout
    phi_r, time_r = formats.read_phi3d(tmp_path / "phi3d.000001", g.shape)
    phi_d, _ = formats.read_phi3d(tmp_path / "phi3d.000002", g.shape)
    assert np.array_equal(phi_r, phi_d)
    assert time_r == SCAL["time"]  # the restart's TIME
