C Test harness for the reference's own CIRSRADG_WAVE, built by
C oracle/Makefile into oracle/_ref/ (test infrastructure only).
C
C The reference routine takes most of its inputs from COMMON blocks that
C NEMESIS fills from run files.  These entry points fill the same blocks
C from arrays instead, so a caller can drive cirsradg_wave directly:
C
C   HRADINIT  switches off every optional physics flag
C             (/diagnostic/, /lbltable/, /nlte_flags/, /dust/, /solardat/,
C             /GASCONDEF/) and clears /ELECTRON/, as a fresh NEMESIS run
C             with zero flags would.
C   HSETGASCON sets the /GASCONDEF/ flags; HGETH2O copies out the CKD H2O
C             tables ngascon.f's BLOCK DATA holds.
C   HSETCIA   loads /CIATABLE/ as ciaread.f:63-103 would from a .tab file.
C   HREADK    registers .kta files through the reference's read_klist.
C   HREADLBL  registers .lta files through read_klbllist; HSETILBL sets
C             /LBLTABLE/ ILBL.
C   HCLOSEK   closes the units read_klist or read_klbllist opened.
C
C Compiled twice: as written (REAL = 4 bytes) and with REAL promoted to
C 8 bytes, which is the FP64 form of the reference.

      SUBROUTINE HRADINIT()
      IMPLICIT NONE
      INCLUDE 'arrdef.f'
      INCLUDE 'gascom.f'
      INCLUDE 'electron.f'
      INTEGER IDIAG, IQUIET, ILBL, INLTE, NSEC, NCONT
      INTEGER IREAD, IFORM, NSPT
      REAL NLTEK1, NLTEA, VSEC(MAXSEC), XSEC(2,MAXCON,MAXSEC)
      REAL SOLRAD, SWAVE(MAXBIN), SRAD(MAXBIN)
      COMMON /DIAGNOSTIC/ IDIAG, IQUIET
      COMMON /LBLTABLE/ ILBL
      COMMON /NLTE_FLAGS/ INLTE, NLTEK1, NLTEA
      COMMON /DUST/ VSEC, XSEC, NSEC, NCONT
      COMMON /SOLARDAT/ IREAD, IFORM, SOLRAD, SWAVE, SRAD, NSPT
      IDIAG = 0
      IQUIET = 1
      ILBL = 0
      INLTE = 0
      NLTEK1 = 0.0
      NLTEA = 0.0
      NSEC = 0
      NCONT = 0
      IREAD = 0
      IFORM = 0
      ICH4 = 0
      IH2O = 0
      IO3 = 0
      INH3 = 0
      IUVSCAT = 0
      PELEC = 0.0
      AMOUNTHMIN = 0.0
      PPRESSHMIN = 0.0
      EFLAG = 0
      RETURN
      END

C     /GASCONDEF/ (gascom.f): the gas-continuum flags ngascon.f reads
      SUBROUTINE HSETGASCON(JH2O, JCH4, JO3, JNH3, JUV)
      IMPLICIT NONE
      INCLUDE 'gascom.f'
      INTEGER JH2O, JCH4, JO3, JNH3, JUV
      IH2O = JH2O
      ICH4 = JCH4
      IO3 = JO3
      INH3 = JNH3
      IUVSCAT = JUV
      RETURN
      END

C     The CKD H2O continuum tables ngascon.f's BLOCK DATA fills
C     (/CH2OS0/ with its grid, /CH2OS1/, /CH2OF0/) and /COMCON/, copied out
C     as data: GRID = VB1, VT1, DV1; CON = TS, PS, C2, PI
      SUBROUTINE HGETH2O(GRID, NPT, S96, S60, F96, CON)
      IMPLICIT NONE
      REAL GRID(3), S96(2001), S60(2001), F96(2001), CON(4)
      INTEGER NPT, I
      REAL VB1, VT1, DV1, H2OS96, H2OS60, H2OF96, TS, PS, C2, PI
      INTEGER NPT1
      COMMON /CH2OF0/ H2OF96(2001)
      COMMON /CH2OS0/ VB1, VT1, DV1, NPT1, H2OS96(2001)
      COMMON /CH2OS1/ H2OS60(2001)
      COMMON /COMCON/ TS, PS, C2, PI
      GRID(1) = VB1
      GRID(2) = VT1
      GRID(3) = DV1
      NPT = NPT1
      DO I = 1, 2001
        S96(I) = H2OS96(I)
        S60(I) = H2OS60(I)
        F96(I) = H2OF96(I)
      ENDDO
      CON(1) = TS
      CON(2) = PS
      CON(3) = C2
      CON(4) = PI
      RETURN
      END

      SUBROUTINE HSETCIA(DNU, TEMPS, KC)
      IMPLICIT NONE
      INCLUDE 'ciacom.f'
      REAL DNU, TEMPS(NUMT), KC(NUMPAIRS,NUMT,NUMWN)
      INTEGER I, J, K
      IREAD1 = 0
      DNU1 = DNU
      DO I = 1, NUMT
        TEMPK1(I) = TEMPS(I)
      ENDDO
      DO I = 1, NUMWN
        VVK1(I) = (I-1)*DNU1
      ENDDO
      DO K = 1, NUMWN
        DO J = 1, NUMT
          DO I = 1, NUMPAIRS
            KCIA(I,J,K) = KC(I,J,K)
          ENDDO
        ENDDO
      ENDDO
      IT1C = 0
      IT2C = 0
      IVC = 0
      RETURN
      END

C     /CIATABLE/ to be read from the raddata table file NAME by ciaread.f
C     itself on first use (ciaread.f:63-103), as a run's .cia file sets it
      SUBROUTINE HSETCIAFILE(NAME, DNU)
      IMPLICIT NONE
      INCLUDE 'ciacom.f'
      CHARACTER*(*) NAME
      REAL DNU
      IREAD1 = 1
      ANAME1 = NAME
      DNU1 = DNU
      IT1C = 0
      IT2C = 0
      IVC = 0
      RETURN
      END

C     /DUST/ (cirsradg_wave.f:225,312): NSEC wavenumbers, NCONT particle
C     types, extinction cross-sections XEXT1(L,K); the second row of XSEC
C     (albedo times extinction) is not read by the emission branches.
      SUBROUTINE HSETDUST(NSEC1, NCONT1, VSEC1, XEXT1)
      IMPLICIT NONE
      INCLUDE 'arrdef.f'
      INTEGER NSEC1, NCONT1, NSEC, NCONT, K, L
      REAL VSEC1(NSEC1), XEXT1(NSEC1, NCONT1)
      REAL VSEC(MAXSEC), XSEC(2,MAXCON,MAXSEC)
      COMMON /DUST/ VSEC, XSEC, NSEC, NCONT
      NSEC = NSEC1
      NCONT = NCONT1
      DO L = 1, NSEC1
        VSEC(L) = VSEC1(L)
        DO K = 1, NCONT1
          XSEC(1,K,L) = XEXT1(L,K)
          XSEC(2,K,L) = 0.5*XEXT1(L,K)
        ENDDO
      ENDDO
      RETURN
      END

C     /NLTE_FLAGS/ (cirsradg_wave.f:328): flag, k1, A of calc_nlte_t.f
      SUBROUTINE HSETNLTE(IFLAG, XK1, XA)
      IMPLICIT NONE
      INTEGER IFLAG, INLTE
      REAL XK1, XA, NLTEK1, NLTEA
      COMMON /NLTE_FLAGS/ INLTE, NLTEK1, NLTEA
      INLTE = IFLAG
      NLTEK1 = XK1
      NLTEA = XA
      RETURN
      END

C     /SOLARDAT/ (cirsradg_wave.f:308-309): IFORM, with no solar file
      SUBROUTINE HSETIFORM(JFORM)
      IMPLICIT NONE
      INCLUDE 'arrdef.f'
      INTEGER JFORM, IREAD, IFORM, NSPT
      REAL SOLRAD, SWAVE(MAXBIN), SRAD(MAXBIN)
      COMMON /SOLARDAT/ IREAD, IFORM, SOLRAD, SWAVE, SRAD, NSPT
      IFORM = JFORM
      IREAD = 0
      RETURN
      END

C     /SOLARDAT/ with a stellar spectrum (iread = 999), as the caller's
C     solar-file read leaves it for get_solar_wave.f: N points W, R
      SUBROUTINE HSETSOLAR(N, W, R)
      IMPLICIT NONE
      INCLUDE 'arrdef.f'
      INTEGER N, I, IREAD, IFORM, NSPT
      DOUBLE PRECISION W(N), R(N)
      REAL SOLRAD, SWAVE(MAXBIN), SRAD(MAXBIN)
      COMMON /SOLARDAT/ IREAD, IFORM, SOLRAD, SWAVE, SRAD, NSPT
      IREAD = 999
      NSPT = N
      DO I = 1, N
        SWAVE(I) = W(I)
        SRAD(I) = R(I)
      ENDDO
      RETURN
      END

C     /FPTABLE/ (ciacom.f) in memory, so fpread_all.f reads no file:
C     NP para fractions FPV, KH2H2/KH2HE(NP,NUMT,NUMWN), KC5(5,NUMT,NUMWN)
      SUBROUTINE HSETPARA(DNU, TEMPS, NP, FPV, KHH, KHE, KC5)
      IMPLICIT NONE
      INCLUDE 'ciacom.f'
      INTEGER NP, I, J, K
      REAL DNU, TEMPS(NUMT), FPV(NP)
      REAL KHH(NP,NUMT,NUMWN), KHE(NP,NUMT,NUMWN), KC5(5,NUMT,NUMWN)
      IREAD2 = 0
      DNU2 = DNU
      IPARA2 = NP
      DO I = 1, NUMT
        TEMPK2(I) = TEMPS(I)
      ENDDO
      DO I = 1, NUMWN
        VVK2(I) = (I-1)*DNU2
      ENDDO
      DO I = 1, NP
        FP(I) = FPV(I)
      ENDDO
      DO K = 1, NUMWN
        DO J = 1, NUMT
          DO I = 1, NP
            KH2H2(I,J,K) = KHH(I,J,K)
            KH2HE(I,J,K) = KHE(I,J,K)
          ENDDO
          DO I = 1, 5
            KCIA5(I,J,K) = KC5(I,J,K)
          ENDDO
        ENDDO
      ENDDO
      IT1K = 0
      IT2K = 0
      IVK = 0
      RETURN
      END

C     /E3TABLE/ (e3interp.f:14-51) in memory, so no e3tab.dat is read:
C     the transmission grid, E3 and its gradient as e3interp.f forms it
      SUBROUTINE HSETE3(TT, TE)
      IMPLICIT NONE
      INTEGER MTAB, I
      PARAMETER (MTAB=1001)
      DOUBLE PRECISION TT(MTAB), TE(MTAB)
      DOUBLE PRECISION TABTRAN(MTAB), TABE3(MTAB), EGRAD(MTAB), DTRANS
      COMMON /E3TABLE/ TABTRAN, TABE3, EGRAD
      DTRANS = 1.0/FLOAT(MTAB-1)
      DO I = 1, MTAB
        TABTRAN(I) = TT(I)
        TABE3(I) = TE(I)
      ENDDO
      EGRAD(1) = (TABE3(2)-TABE3(1))/DTRANS
      EGRAD(MTAB) = (TABE3(MTAB)-TABE3(MTAB-1))/DTRANS
      DO I = 2, MTAB-1
        EGRAD(I) = 0.5*(TABE3(I+1)-TABE3(I-1))/DTRANS
      ENDDO
      RETURN
      END

      SUBROUTINE HREADK(KL, NGAS, IDGAS, ISOGAS, NWAVE, VWAVE, NKL)
      IMPLICIT NONE
      CHARACTER*(*) KL
      INTEGER NGAS, IDGAS(NGAS), ISOGAS(NGAS), NWAVE, NKL
      REAL VWAVE(NWAVE)
      CHARACTER*100 KLIST
      KLIST = KL
      CALL READ_KLIST(KLIST, NGAS, IDGAS, ISOGAS, NWAVE, VWAVE, NKL)
      RETURN
      END

      SUBROUTINE HCLOSEK(NKL)
      IMPLICIT NONE
      INTEGER NKL, I
      DO I = 1, NKL
        CLOSE(100 + I)
      ENDDO
      RETURN
      END

C     /LBLTABLE/ ILBL (cirsradg_wave.f:317): 2 runs get_klblg and
C     noverlapg1 on the tables read_klbllist registered
      SUBROUTINE HSETILBL(JLBL)
      IMPLICIT NONE
      INTEGER JLBL, ILBL
      COMMON /LBLTABLE/ ILBL
      ILBL = JLBL
      RETURN
      END

C     Registers .lta files through the reference's read_klbllist
C     (units 101.. as read_klist's; HCLOSEK closes them)
      SUBROUTINE HREADLBL(KL, NGAS, IDGAS, ISOGAS, NWAVE, VWAVE)
      IMPLICIT NONE
      CHARACTER*(*) KL
      INTEGER NGAS, IDGAS(NGAS), ISOGAS(NGAS), NWAVE
      REAL VWAVE(NWAVE)
      CHARACTER*100 KLIST
      KLIST = KL
      CALL READ_KLBLLIST(KLIST, NGAS, IDGAS, ISOGAS, NWAVE, VWAVE)
      RETURN
      END
