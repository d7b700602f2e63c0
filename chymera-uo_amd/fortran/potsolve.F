C Drop-in replacements for the reference's potential-solve entry points
C
C     SUBROUTINE SETBDY(NCALL,ISYM)          pot/boundary.f:2
C     SUBROUTINE BDYGEN(MAXTRM,ISYM,REDGE)   pot/boundary.f:88
C     SUBROUTINE POT3(NPOINT,IPRINT)         pot/pot3.f:15
C     SUBROUTINE ZAXPHI(NPOINT,IPRINT)       pot/pot3.f:311
C
C Same names, arguments and COMMON data as the reference (PHI, RHO in /POIS/,
C R, Z, RHF, ZHF in /GRID/, DTHETA, PI, GRAV in /blok6/ from the host
C program's hydroparam.h and globals.h; COEF in /COEFS/; TMASS, PHICHK,
C KLOCAT in /INSIDE/), so a CHYMERA build links this file and libchypot
C in place of pot/pot3.f, pot/boundary.f, pot/fft.f and pot/blktri.f.  The
C solve itself runs on the GPU through the C-ABI (chypot_c.f90).
C
C Compile as fixed form with the C preprocessor, with the host program's
C headers on the include path (-I<tree>/headers), or with
C fortran/standalone/ for builds outside a CHYMERA tree.
C
C Threading: SETBDY may be called from every thread of a parallel region as
C radhydro/io.f:212-216 does (the calls are serialised and repeated calls
C with an unchanged grid do nothing); BDYGEN, POT3 and ZAXPHI are called from
C serial code, as in the reference.

      SUBROUTINE SETBDY(NCALL,ISYM)
      USE chypot_c, ONLY: chy_shim_setbdy
      IMPLICIT real*8 (a-h,o-z)
#include "hydroparam.h"
#include "globals.h"
      IF (JMAX.NE.POT3JMAX .OR. KMAX.NE.POT3KMAX) THEN
        WRITE(0,*) 'CHYPOT: JMAX/KMAX must equal POT3JMAX/POT3KMAX'
        STOP 1
      ENDIF
C$OMP CRITICAL (CHYPOT_SHIM)
      CALL chy_shim_setbdy(NCALL,ISYM,POT3JMAX,POT3KMAX,LMAX,
     &                     R,Z,RHF,ZHF,DTHETA,PI,GRAV)
C$OMP END CRITICAL (CHYPOT_SHIM)
      RETURN
      END


      SUBROUTINE BDYGEN(MAXTRM,ISYM,REDGE)
      USE chypot_c, ONLY: chy_shim_bdygen
      IMPLICIT real*8 (a-h,o-z)
#include "hydroparam.h"
#include "globals.h"
      COMMON /INSIDE/ TMASS,ENEW,ELOST,EDIF,PHICHK,KLOCAT
      CALL chy_shim_bdygen(MAXTRM,ISYM,REDGE,RHO,PHI,TMASS)
      RETURN
      END


      SUBROUTINE POT3(NPOINT,IPRINT)
      USE chypot_c, ONLY: chy_shim_pot3
      IMPLICIT real*8 (a-h,o-z)
#include "hydroparam.h"
#include "globals.h"
      COMMON /INSIDE/ TMASS,ENEW,ELOST,EDIF,PHICHK,KLOCAT
      COMMON /COEFS/ COEF(POT3JMAX2,POT3KMAX2,LMAX2,2)
      CALL CHYNPT(NPOINT)
      CALL chy_shim_pot3(NPOINT,RHO,PHI,COEF,PHICHK,KLOCAT)
      IF (IPRINT.EQ.1) CALL CHYREP
      RETURN
      END


      SUBROUTINE ZAXPHI(NPOINT,IPRINT)
      USE chypot_c, ONLY: chy_shim_zaxphi
      IMPLICIT real*8 (a-h,o-z)
#include "hydroparam.h"
#include "globals.h"
      COMMON /INSIDE/ TMASS,ENEW,ELOST,EDIF,PHICHK,KLOCAT
      CALL CHYNPT(NPOINT)
      CALL chy_shim_zaxphi(NPOINT,PHI,PHICHK,KLOCAT)
      IF (IPRINT.EQ.1) CALL CHYREP
      RETURN
      END


C NPOINT is limited to 10 (X and PHI2 hold 10 points); the reference reports
C it on unit 3 and lowers the argument (pot/pot3.f:323-328).
      SUBROUTINE CHYNPT(NPOINT)
      IF (NPOINT.GT.10) THEN
        WRITE(3,9000) NPOINT
        NPOINT=10
      ENDIF
 9000 FORMAT(///,' YOU ARE LIMITED TO A 10-POINT INTERPOLATION HERE IN',
     &       ' SUBROUTINE ZAXPHI.',/,' BUT YOU PUT NPOINT =',I4,///)
      RETURN
      END


C The IPRINT=1 report of ZAXPHI on unit 3 (pot/pot3.f:402-419): PHAC(K) for
C K=2..KMAX+1, then COEF(J,2,M,1:2) for M=1..8 and J=2..JMAX.
      SUBROUTINE CHYREP
      USE chypot_c, ONLY: chy_shim_phac
      IMPLICIT real*8 (a-h,o-z)
#include "hydroparam.h"
#include "globals.h"
      COMMON /COEFS/ COEF(POT3JMAX2,POT3KMAX2,LMAX2,2)
      DIMENSION PHAC(KMAX2)
      PHAC=0.0D0
      CALL chy_shim_phac(PHAC(2),KMAX)
      WRITE(3,9100) 'PHAC(1:KMAX1):'
      WRITE(3,9110) (PHAC(K),K=2,KMAX1)
      WRITE(3,9100) 'Fourier amplitudes and phases:'
      IF (LMAX.GT.8) THEN
        WRITE(3,9120) (J,((COEF(J,2,M,I),M=1,8),I=1,2),J=2,JMAX)
      ENDIF
 9100 FORMAT(A)
 9110 FORMAT(1P11E11.3)
 9120 FORMAT(20X,I5,' COEFS=',1P8E12.3,/,25X,' PHASE=',0P8F12.2)
      RETURN
      END
