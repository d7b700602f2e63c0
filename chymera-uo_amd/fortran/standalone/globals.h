! The COMMON blocks the potential solve shares with its caller, with the
! member order and types of a CHYMERA build's headers/globals.h so that the
! byte layout is the same:
!   /blok6/  DTHETA, PI, GRAV, BGDEN, GSOFT, KWFW
!   /POIS/   INDIRECTX, INDIRECTY, PHI, RHO, RHOTOT (POT3JMAX2,POT3KMAX2,LMAX)
!   /GRID/   JREQ, KZPOL, R, Z, RHF, ZHF, ROF3N, ZOF3N
      integer kwfw
      real*8 dtheta, pi, grav, bgden, gsoft
      common /blok6/ dtheta, pi, grav, bgden, gsoft, kwfw

      real*8 phi, rho, rhotot, indirectx, indirecty
      common /pois/ indirectx, indirecty,
     &     phi(pot3jmax2,pot3kmax2,lmax),
     &     rho(pot3jmax2,pot3kmax2,lmax),
     &     rhotot(pot3jmax2,pot3kmax2,lmax)

      real*8 r, z, rhf, zhf, rof3n, zof3n
      integer jreq, kzpol
      common /grid/ jreq, kzpol,
     &     r(pot3jmax2), z(pot3kmax2),
     &     rhf(pot3jmax2), zhf(pot3kmax2),
     &     rof3n, zof3n
