! Grid PARAMETERs for building the drop-in potential solve outside a CHYMERA
! tree (tests and the potdiag driver).  The names and meanings are those the
! solver routines use from a CHYMERA build's headers/hydroparam.h; the sizes
! come from the preprocessor: -DCHY_JMAX=.. -DCHY_KMAX=.. -DCHY_LMAX=..
! (KMAX a power of two, LMAX even with LMAX/2 23-smooth).
      integer jmax, jmax1, jmax2
      integer kmax, kmax1, kmax2
      integer lmax, lmax2
      integer pot3jmax, pot3jmax1, pot3jmax2
      integer pot3kmax, pot3kmax1, pot3kmax2
      integer hj, hk, hj1, hk1, hj2, hk2
      parameter (jmax=CHY_JMAX, jmax1=jmax+1, jmax2=jmax+2)
      parameter (kmax=CHY_KMAX, kmax1=kmax+1, kmax2=kmax+2)
      parameter (lmax=CHY_LMAX, lmax2=lmax/2)
      parameter (pot3jmax=jmax, pot3jmax1=pot3jmax+1,
     &           pot3jmax2=pot3jmax+2)
      parameter (pot3kmax=kmax, pot3kmax1=pot3kmax+1,
     &           pot3kmax2=pot3kmax+2)
! size of the axisymmetric density table of the ITYPE=71 input (fort.2)
      parameter (hj=jmax, hk=kmax, hj1=hj+1, hk1=hk+1, hj2=hj+2,
     &           hk2=hk+2)
