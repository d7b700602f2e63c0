! Test harness around the reference potential solve (CHECKER ONLY).
!
! This program is compiled against the reference's own pot/pot3.f, pot/fft.f,
! pot/blktri.f and pot/boundary.f (see oracle/ref/Makefile) and is used only to
! write golden vectors for the tests and to time the reference on the CPU.  It is
! never linked into the product library.
!
! It reproduces what the ITYPE=71 potential-only driver does
! (radhydro/io.f:151-229): uniform grid R(J)=(J-2)*DELR, RHF=R+DELR/2 and the
! same for Z (io.f:188-205), DTHETA=2*PI/LMAX, GRAV=1, KWFW=log2(KMAX)-1
! (radhydro/3dhyd-main.f:175), SETBDY(0,ISYM) inside a parallel region
! (io.f:212-216), then BDYGEN(MAXTRM,ISYM,REDGE) and POT3(NPOINT,IPRINT)
! (io.f:217-218).  The density is read from a file instead of fort.2 so that
! 3-D fields can be used.
!
! Usage:  refpot solve  <in> <out>   -- one solve (+ optional timing loop)
!         refpot fft    <in> <out>   -- FFT/REALTR forward + inverse on vectors
!         refpot blktri <in> <out>   -- BLKTRI on explicit coefficient sets
!
! All files are raw little-endian stream files; layouts are documented at each
! read/write below and mirrored by tests/golden/make_golden.py.
program refpot
  use omp_lib
  implicit real*8 (a-h,o-z)
#include "hydroparam.h"
#include "globals.h"
  common /inside/ tmass,enew,elost,edif,phichk,klocat
  common /coefs/ coef(pot3jmax2,pot3kmax2,lmax2,2)

  character(len=256) :: mode, fin, fout
  integer(4) :: ihdr(8)
  real(8) :: rhdr(4)
  real(8), allocatable :: tim(:), phib(:,:,:)

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')

  if (trim(mode) == 'solve') then
     ! in : int32 [jmax,kmax,lmax,isym,maxtrm,npoint,nwarm,nrep]
     !      real8 [delr,delz,redge,unused]
     !      real8 rho(jmax+2,kmax+2,lmax)
     read(10) ihdr
     read(10) rhdr
     if (ihdr(1) /= pot3jmax .or. ihdr(2) /= pot3kmax .or. ihdr(3) /= lmax) then
        write(0,*) 'grid mismatch: binary built for', pot3jmax, pot3kmax, lmax
        stop 2
     end if
     isym = ihdr(4); maxtrm = ihdr(5); npoint = ihdr(6)
     nwarm = ihdr(7); nrep = ihdr(8)
     delr = rhdr(1); delz = rhdr(2); redge = rhdr(3)
     read(10) rho

     pi = 3.1415926535897932d0
     grav = 1.0d0
     dtheta = 2.0d0*pi/dble(lmax)
     kwfw = int(log10(dble(kmax))/log10(2.0d0)) - 1
     tmass = 0.0d0; enew = 0.0d0; elost = 0.0d0; edif = 0.0d0
     phichk = 0.0d0; klocat = 0
     do j = 1, pot3jmax2
        r(j) = (j-2)*delr
        rhf(j) = r(j) + delr/2.0d0
     end do
     do k = 1, pot3kmax2
        z(k) = (k-2)*delz
        zhf(k) = z(k) + delz/2.0d0
     end do
     phi = 0.0d0
     coef = 0.0d0

     ! The /BDY/ tables are threadprivate and BDYGEN copies the master's tables
     ! into every thread (COPYIN, boundary.f:314-315), so one call on the master
     ! thread is equivalent to io.f's call inside a parallel region.
     call setbdy(0, isym)
     call bdygen(maxtrm, isym, redge)
     allocate(phib(pot3jmax2,pot3kmax2,lmax))
     phib = phi
     tmassb = tmass
     call pot3(npoint, 0)

     ! out: real8 phi_after_bdygen(jmax+2,kmax+2,lmax), real8 tmass_after_bdygen,
     !      real8 phi(jmax+2,kmax+2,lmax), real8 coef(jmax+2,kmax+2,lmax/2,2),
     !      real8 [tmass, phichk, dble(klocat)]
     write(11) phib
     write(11) tmassb
     write(11) phi
     write(11) coef
     write(11) tmass, phichk, dble(klocat)

     ! optional timing loop: BDYGEN+POT3 per solve, wall clock per solve
     ! out: int32 nthreads, real8 seconds(nrep)
     if (nrep > 0) then
        allocate(tim(nrep))
        do it = 1, nwarm + nrep
           t0 = omp_get_wtime()
           call bdygen(maxtrm, isym, redge)
           call pot3(npoint, 0)
           t1 = omp_get_wtime()
           if (it > nwarm) tim(it-nwarm) = t1 - t0
        end do
        write(11) int(omp_get_max_threads(), 4)
        write(11) tim
     end if

  else if (trim(mode) == 'fft') then
     call run_fft()
  else if (trim(mode) == 'blktri') then
     call run_blktri()
  else
     write(0,*) 'unknown mode ', trim(mode)
     stop 2
  end if
  close(10)
  close(11)

contains

  subroutine run_fft()
    ! in : int32 [n, nvec], real8 x(2n, nvec)
    ! out: per vector real8 a(n+1), b(n+1) after FFT(+1),REALTR(+1)
    !      then real8 a(n), b(n) after REALTR(-1),FFT(-1) of those coefficients
    integer(4) :: hdr(2)
    real(8), allocatable :: x(:,:), a(:), b(:)
    read(10) hdr
    n = hdr(1); nvec = hdr(2)
    allocate(x(2*n, nvec), a(2*n+2), b(2*n+2))
    read(10) x
    do iv = 1, nvec
       a = 0.0d0; b = 0.0d0
       do l = 1, n
          a(l) = x(2*l-1, iv)
          b(l) = x(2*l, iv)
       end do
       call fft(a, b, n, n, n, 1)
       call realtr(a, b, n, 1)
       write(11) a(1:n+1)
       write(11) b(1:n+1)
       call realtr(a, b, n, -1)
       call fft(a, b, n, n, n, -1)
       write(11) a(1:n)
       write(11) b(1:n)
    end do
  end subroutine run_fft

  subroutine run_blktri()
    ! in : int32 [m, n, nsys]; per system real8 an(n),bn(n),cn(n),am(m),bm(m),cm(m),y(m,n)
    ! out: per system int32 ierror, real8 y(m,n)
    integer(4) :: hdr(3), ierr4
    real(8), allocatable :: an(:),bn(:),cn(:),am(:),bm(:),cm(:),y(:,:),w(:)
    read(10) hdr
    m = hdr(1); n = hdr(2); nsys = hdr(3)
    allocate(an(n),bn(n),cn(n),am(m),bm(m),cm(m),y(m,n),w(4*(n+2)*24+6*m+64))
    do is = 1, nsys
       read(10) an, bn, cn, am, bm, cm, y
       w = 0.0d0
       ierror = 0
       call blktri(0, 1, n, an, bn, cn, 1, m, am, bm, cm, m, y, ierror, w)
       if (ierror == 0) then
          call blktri(1, 1, n, an, bn, cn, 1, m, am, bm, cm, m, y, ierror, w)
       end if
       ierr4 = ierror
       write(11) ierr4
       write(11) y
    end do
  end subroutine run_blktri

end program refpot
