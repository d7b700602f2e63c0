C Potential-only run: the ITYPE=71 path of the reference's model setup
C (radhydro/io.f:151-229) as a standalone program over the drop-in routines
C of potsolve.F.
C
C   potdiag [namelist-file]        (default potdiag.nml)
C
C   &POTDIAG
C     INPUT   = 'fort2'  axisymmetric model from fort.2 as io.f:157-181 reads
C                        it (header line, then DENNY(HJ2,HK2) list-directed;
C                        RHO(J,K,L)=DENNY(J,K) floored at GRIDLIM*DENCEN)
C             | 'rho3d'  stream file RHOFILE: real*8 DELR, DELZ, then
C                        RHO(JMAX2,KMAX2,LMAX) (3-D fields)
C             | 'restart' restart model RHOFILE (fort.7 / saved.NNNNNN) read
C                        as io.f:464-475 reads it: S, T, A, RHO, EPS, then
C                        ROF3N..OMMAX; RHO and the spacings are used and
C                        TIME replaces the namelist's
C     RHOFILE = 'rho3d.bin'
C     ISYM=2, MAXTRM=10, NPOINT=8, IPRINT=1, REDGE=0  (io.f:208-218 and
C                        radhydro/3dhyd-main.f:166-167)
C     ITSTOP=0, TIME=0   name and second record of the output file
C     OUTFILE = ''       optional stream file: TMASS, PHICHK, DBLE(KLOCAT),
C                        PHI, COEF (for tests)
C   /
C
C Writes phi3d.<ITSTOP as I6.6> with two unformatted sequential records,
C PHI then TIME, as io.f:221-227 does.  Grid: R(J)=(J-2)*DELR,
C RHF=R+DELR/2, the same for Z (io.f:188-205); PI, GRAV and DTHETA as
C io.f:136-140.
      PROGRAM POTDIAG
      IMPLICIT real*8 (a-h,o-z)
#include "hydroparam.h"
#include "globals.h"
      COMMON /INSIDE/ TMASS,ENEW,ELOST,EDIF,PHICHK,KLOCAT
      COMMON /COEFS/ COEF(POT3JMAX2,POT3KMAX2,LMAX2,2)
      PARAMETER (GRIDLIM=1.0D-10)
      CHARACTER*16 INPUT
      CHARACTER*256 RHOFILE, OUTFILE, NMLFILE, POTFILE
      CHARACTER*6 X1
      DIMENSION DENNY(HJ2,HK2)
      NAMELIST /POTDIAG/ INPUT, RHOFILE, ISYM, MAXTRM, NPOINT, IPRINT,
     &                   REDGE, ITSTOP, TIME, OUTFILE

      INPUT='fort2'
      RHOFILE='rho3d.bin'
      OUTFILE=' '
      ISYM=2
      MAXTRM=10
      NPOINT=8
      IPRINT=1
      REDGE=0.0D0
      ITSTOP=0
      TIME=0.0D0
      NMLFILE='potdiag.nml'
      IF (COMMAND_ARGUMENT_COUNT().GE.1) CALL GET_COMMAND_ARGUMENT(1,
     &    NMLFILE)
      OPEN(UNIT=7,FILE=TRIM(NMLFILE),STATUS='OLD')
      READ(7,NML=POTDIAG)
      CLOSE(7)

      PI=3.1415926535897932D0
      GRAV=1.0D0
      DTHETA=2.0D0*PI/DBLE(LMAX)
      KWFW=INT(LOG10(DBLE(KMAX))/LOG10(2.0D0))-1
      TMASS=0.0D0
      PHICHK=0.0D0
      KLOCAT=0
      PHI=0.0D0
      RHO=0.0D0
      COEF=0.0D0

      IF (TRIM(INPUT).EQ.'fort2') THEN
        OPEN(UNIT=2,FILE='fort.2',STATUS='OLD')
        READ(2,*) PINDEX,CON2,RRR2,OMCEN,DENCEN,TOVERW,ROF3N,ZOF3N,
     &            A1NEWZ,JREQ,KZPOL
        READ(2,*) DENNY
        CLOSE(2)
        DO L=1,LMAX
          DO K=2,KMAX1
            DO J=2,JMAX1
              RHO(J,K,L)=MAX(DENNY(J,K),GRIDLIM*DENCEN)
            ENDDO
          ENDDO
        ENDDO
        DELR=ROF3N
        DELZ=ZOF3N
      ELSE IF (TRIM(INPUT).EQ.'rho3d') THEN
        OPEN(UNIT=2,FILE=TRIM(RHOFILE),ACCESS='STREAM',
     &       FORM='UNFORMATTED',STATUS='OLD')
        READ(2) DELR,DELZ
        READ(2) RHO
        CLOSE(2)
        ROF3N=DELR
        ZOF3N=DELZ
      ELSE IF (TRIM(INPUT).EQ.'restart') THEN
        OPEN(UNIT=2,FILE=TRIM(RHOFILE),FORM='UNFORMATTED',STATUS='OLD')
        READ(2)
        READ(2)
        READ(2)
        READ(2) RHO
        READ(2)
        READ(2) ROF3N,ZOF3N,DELT,TIME,ELOST,DEN,SOUND,JREQ,OMMAX
        CLOSE(2)
        DELR=ROF3N
        DELZ=ZOF3N
      ELSE
        WRITE(0,*) 'potdiag: INPUT must be fort2, rho3d or restart'
        STOP 2
      ENDIF

      DO J=1,JMAX2
        R(J)=(J-2)*DELR
        RHF(J)=R(J)+DELR/2.0D0
      ENDDO
      DO K=1,KMAX2
        Z(K)=(K-2)*DELZ
        ZHF(K)=Z(K)+DELZ/2.0D0
      ENDDO

C$OMP PARALLEL DEFAULT(SHARED)
      CALL SETBDY(0,ISYM)
C$OMP END PARALLEL
      CALL BDYGEN(MAXTRM,ISYM,REDGE)
      CALL POT3(NPOINT,IPRINT)

      WRITE(X1,'(I6.6)') ITSTOP
      POTFILE='phi3d.'//TRIM(X1)
      OPEN(UNIT=23,FILE=TRIM(POTFILE),FORM='UNFORMATTED')
      WRITE(23) PHI
      WRITE(23) TIME
      CLOSE(23)

      IF (LEN_TRIM(OUTFILE).GT.0) THEN
        OPEN(UNIT=24,FILE=TRIM(OUTFILE),ACCESS='STREAM',
     &       FORM='UNFORMATTED',STATUS='REPLACE')
        WRITE(24) TMASS,PHICHK,DBLE(KLOCAT)
        WRITE(24) PHI
        WRITE(24) COEF
        CLOSE(24)
      ENDIF
      END
