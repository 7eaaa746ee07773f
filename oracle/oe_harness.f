C Test harness for the reference's own optimal-estimation step, built by
C oracle/Makefile (target oeref) into oracle/_ref/liboeref64.so with REAL
C promoted to 8 bytes (test infrastructure only).
C
C HOESTEP runs, on compact arrays, the sequence coreret.f performs around
C one Levenberg-Marquardt proposal:
C   Sa^-1 by Cholesky inversion           dinvertm JMOD=2 (coreret.f:225-250)
C   Se^-1 diagonal, 1/se                  coreret.f:255-285
C   gain matrix and averaging kernels     calc_gain_matrix (coreret.f:554)
C   cost function at (xn, yn)             calc_phiret (coreret.f:557)
C   next state                            calcnextxn (coreret.f:609)
C
C Compact arguments (Fortran order): KKC(NY,NX), SAC(NX,NX), SEC(NY) (the
C diagonal measurement variances), XAC(NX), XNC(NX), YC(NY), YNC(NY).
C Outputs: XOUT(NX), DDC(NX,NY), AAC(NX,NX), SAIC(NX,NX), PHI, CHISQ,
C ICHK (dinvertm's check flag for Sa).

      SUBROUTINE HOESTEP(NX, NY, KKC, SAC, SEC, XAC, XNC, YC, YNC,
     1  XOUT, DDC, AAC, SAIC, PHI, CHISQ, ICHK)
      IMPLICIT NONE
      INCLUDE 'arrdef.f'
      INTEGER MX, MY
      PARAMETER (MY=IDIM, MX=MAXV)
      INTEGER NX, NY, ICHK, I, J, JMOD
      REAL KKC(NY,NX), SAC(NX,NX), SEC(NY), XAC(NX), XNC(NX)
      REAL YC(NY), YNC(NY), XOUT(NX), PHI, CHISQ
      DOUBLE PRECISION DDC(NX,NY), AAC(NX,NX), SAIC(NX,NX)
      REAL KK(MY,MX), SA(MX,MX), SE(MY,MY), XA(MX), XN(MX)
      REAL Y(MY), YN(MY), XO(MX), CALC_PHIRET
      DOUBLE PRECISION S1D(MX,MX), SAI(MX,MX), SEI(MY,MY)
      DOUBLE PRECISION DD(MX,MY), AA(MX,MX)
      INTEGER IDIAG, IQUIET
      COMMON /DIAGNOSTIC/ IDIAG, IQUIET
      SAVE KK, SA, SE, XA, XN, Y, YN, XO, S1D, SAI, SEI, DD, AA

      IDIAG = 0
      IQUIET = 1
      DO J = 1, MX
        DO I = 1, MX
          S1D(I,J) = 0.0
          SA(I,J) = 0.0
        ENDDO
      ENDDO
      DO J = 1, NX
        DO I = 1, NX
          SA(I,J) = SAC(I,J)
          S1D(I,J) = DBLE(SAC(I,J))
        ENDDO
        XA(J) = XAC(J)
        XN(J) = XNC(J)
        DO I = 1, NY
          KK(I,J) = KKC(I,J)
        ENDDO
      ENDDO
      JMOD = 2
      ICHK = 0
      CALL DINVERTM(JMOD, ICHK, S1D, NX, MX, SAI)
      DO J = 1, NY
        DO I = 1, NY
          SE(I,J) = 0.0
          SEI(I,J) = 0.0
        ENDDO
        SE(J,J) = SEC(J)
        SEI(J,J) = 1.0 / DBLE(SEC(J))
        Y(J) = YC(J)
        YN(J) = YNC(J)
      ENDDO
      CALL CALC_GAIN_MATRIX(NX, NY, KK, SA, SAI, SE, SEI, DD, AA)
      PHI = CALC_PHIRET(NY, Y, YN, SEI, NX, XN, XA, SAI, CHISQ)
      CALL CALCNEXTXN(NX, NY, XA, XN, Y, YN, DD, AA, XO)
      DO J = 1, NX
        XOUT(J) = XO(J)
        DO I = 1, NX
          AAC(J,I) = AA(J,I)
          SAIC(J,I) = SAI(J,I)
        ENDDO
        DO I = 1, NY
          DDC(J,I) = DD(J,I)
        ENDDO
      ENDDO
      RETURN
      END

C HMAP fills the COMMON blocks map2pro.f reads (/PATHVA/ npath, nlayin,
C imod, layinc, ngas, ncont; /MODCOM/ npro; /LGRAD/ DAM, DTE, DCO) and
C runs map2pro then map2xvec as cirsrtfg_wave.f:403-425 does for one
C wavenumber.  Compact arguments (Fortran order): NLAYC(NPTH),
C LAYC(NLYR,NPTH) (1-based layer numbers), DAMC(NLYR,NPR), DTEC(NLYR,NPR),
C DCOC(NLYR,NPR), DFPC(NLYR,NPR), XMAPC(NVC,NPAR,NPR), DQC(NPTH,NLYR,NPAR)
C with NPAR = NGS+1+NCT+NFP for NCT dust types and NFP (0 or 1) para-H2
C fractions; output DXC(NPTH,NVC).

      SUBROUTINE HMAP(NPTH, NLYR, NGS, NCT, NFP, NPR, NVC, NLAYC, LAYC,
     1  IMODC, DAMC, DTEC, DCOC, DFPC, XMAPC, DQC, DXC)
      IMPLICIT NONE
      INCLUDE 'arrdef.f'
      INCLUDE 'pathcom.f'
      INCLUDE 'laycom.f'
      INCLUDE 'laygrad.f'
      INTEGER NPTH, NLYR, NGS, NCT, NFP, NPR, NVC, I, J, K, NPAR
      INTEGER NLAYC(NPTH), LAYC(NLYR,NPTH), IMODC(NPTH)
      REAL DAMC(NLYR,NPR), DTEC(NLYR,NPR)
      REAL DCOC(NLYR,NPR), DFPC(NLYR,NPR)
      REAL XMAPC(NVC,NGS+1+NCT+NFP,NPR)
      REAL DQC(NPTH,NLYR,NGS+1+NCT+NFP), DXC(NPTH,NVC)
      REAL DQL(MAXPAT,MAXLAY,MAXGAS+2+MAXCON)
      REAL DPRF(MAXPAT,MAXGAS+2+MAXCON,MAXPRO)
      REAL XMV(MAXV,MAXGAS+2+MAXCON,MAXPRO), DXV(MAXPAT,MAXV)
      SAVE DQL, DPRF, XMV, DXV

      NPATH = NPTH
      NGAS = NGS
      NCONT = NCT
      NPRO = NPR
      NPAR = NGS + 1 + NCT + NFP
      DO J = 1, NPTH
        NLAYIN(J) = NLAYC(J)
        IMOD(J) = IMODC(J)
        DO I = 1, NLYR
          LAYINC(I,J) = LAYC(I,J)
        ENDDO
      ENDDO
      DO J = 1, NPR
        DO I = 1, NLYR
          DAM(I,J) = DAMC(I,J)
          DTE(I,J) = DTEC(I,J)
          DCO(I,J) = DCOC(I,J)
          DFP(I,J) = DFPC(I,J)
        ENDDO
      ENDDO
      DO K = 1, NPAR
        DO J = 1, NLYR
          DO I = 1, NPTH
            DQL(I,J,K) = DQC(I,J,K)
          ENDDO
        ENDDO
        DO J = 1, NPR
          DO I = 1, NVC
            XMV(I,K,J) = XMAPC(I,K,J)
          ENDDO
        ENDDO
      ENDDO
      CALL MAP2PRO(NPAR, DQL, DPRF)
      CALL MAP2XVEC(NPAR, NPR, NPTH, NVC, XMV, DPRF, DXV)
      DO J = 1, NVC
        DO I = 1, NPTH
          DXC(I,J) = DXV(I,J)
        ENDDO
      ENDDO
      RETURN
      END
