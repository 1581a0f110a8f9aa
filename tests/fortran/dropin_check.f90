!     dropin_check.f90 -- test driver for the Fortran drop-in entry points
!     (fem25_amd/fortran/f25_dropin.f90).
!
!         dropin_check <input file> <output file>
!
!     Reads a mesh, model, numbering, inversion structures and vectors (a
!     stream file written by tests/test_fortran_dropin.py), fills the
!     reference's MESH_INFO and INV_INFO the way its drivers do, and calls
!     ASMBLE_DRIVER (twice: the second call reuses the cached device mesh),
!     CFMAT_DIST, ADJFLAM_DIST and ASMB_PEFF, then ASMBLE_DRIVER once more
!     with a wrong NZLOC, which must fail with IERR /= 0, then MET225_MOD and
!     F25_CADJ_NU_ALPHA_RHS.  Writes the seven status words and the outputs
!     CDIST, FMAT_DIST, GPROD, PEFF, P1, P2, RHSBNU, RHSALP and P3.
      PROGRAM DROPIN_CHECK
      IMPLICIT NONE
      INCLUDE 'fwd_struc.h'
      INCLUDE 'mpif.h'
      TYPE (MESH_INFO) MSH
      TYPE (INV_INFO) INV
      CHARACTER(512) :: FIN, FOUT
      INTEGER*4 :: HDR(9), NEN, NDOF, NZ, NZF, NA35, NNPG, NELEM, I, MPIERR
      INTEGER*4 :: ISTAT(7)
      REAL*8 :: SC(12), VPVS, FREQ, PY
      CHARACTER(1) :: CH(4), CFTYPE
      COMPLEX*8, ALLOCATABLE :: CDIST(:), CDIST2(:), FMAT(:), WAVE(:), ADJLAM(:), UE(:), PEFF(:)
      COMPLEX*8, ALLOCATABLE :: NU(:), ALPHA(:), RHSBNU(:), RHSALP(:)
      REAL*4, ALLOCATABLE :: GPROD(:), V(:), P1(:), P2(:), P3(:)

      CALL MPI_INIT(MPIERR)
      CALL GET_COMMAND_ARGUMENT(1, FIN)
      CALL GET_COMMAND_ARGUMENT(2, FOUT)
      OPEN(10, FILE=TRIM(FIN), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
!     NORD, NELEM, NNPG, NDOF, NZ, NCON, NNPGL, NZ_FDIST, NA35
      READ(10) HDR
      READ(10) SC
      READ(10) VPVS
      READ(10) CH
      NELEM = HDR(2)
      NNPG = HDR(3)
      NDOF = HDR(4)
      NZ = HDR(5)
      NZF = HDR(8)
      NA35 = HDR(9)
      NEN = (HDR(1) + 1)**2
      msh%NORD = HDR(1)
      msh%NLXI = HDR(1) + 1
      msh%NLETA = HDR(1) + 1
      msh%NEN = NEN
      msh%NELEM = NELEM
      msh%NELEML = NELEM
      msh%NNPG = NNPG
      msh%NDOF = NDOF
      msh%NDOFL = NDOF
      msh%NZERO = NZ
      msh%NZLOC = NZ
      msh%IITYPE = 1
      msh%NCOEFF = 2
      msh%LISISO = .TRUE.
      msh%LDISP = .FALSE.
      msh%XWIDTH = SC(1)
      msh%ZWIDTH = SC(2)
      msh%RCOEFF_BDY = SC(3)
      msh%DMAX_BDY = SC(4)
      msh%KAPPA_BDY = SC(5)
      msh%FCENT_BDY = SC(6)
      msh%RCOEFF_SRF = SC(7)
      msh%DMAX_SRF = SC(8)
      msh%KAPPA_SRF = SC(9)
      msh%FCENT_SRF = SC(10)
      FREQ = SC(11)
      PY = SC(12)
      CFTYPE = CH(1)
      ALLOCATE(msh%IENG(4,NELEM), msh%XLOCS(NNPG), msh%ZLOCS(NNPG), msh%XD(NNPG), msh%ZD(NNPG))
      ALLOCATE(msh%DENS(NNPG), msh%ECOEFF(NNPG,2), msh%CDOMAIN(NELEM), msh%CNNPG(NNPG))
      ALLOCATE(msh%LM(3,NEN,NELEM), msh%IRPTR_LOC(NDOF+1), msh%JCPTR_LOC(NZ), msh%CNP(NDOF))
      ALLOCATE(msh%MYDOFS(NDOF), msh%MYELEM(NELEM))
      READ(10) msh%IENG
      READ(10) msh%XLOCS
      READ(10) msh%ZLOCS
      READ(10) msh%XD
      READ(10) msh%ZD
      READ(10) msh%DENS
      READ(10) msh%ECOEFF
      READ(10) msh%CDOMAIN
      READ(10) msh%CNNPG
      READ(10) msh%LM
      READ(10) msh%IRPTR_LOC
      READ(10) msh%JCPTR_LOC
      READ(10) msh%CNP
      DO I = 1, NDOF
         msh%MYDOFS(I) = I
      ENDDO
      DO I = 1, NELEM
         msh%MYELEM(I) = I
      ENDDO
      inv%NCON = HDR(6)
      inv%NNPGL = HDR(7)
      inv%NZ_FDIST = NZF
      inv%NA35 = NA35
      inv%VPVS = VPVS
      inv%CINVTYPE = CH(2)//CH(3)
      inv%NVINV = 1
      IF (inv%CINVTYPE == 'PS' .OR. inv%CINVTYPE == 'ps') inv%NVINV = 2
      ALLOCATE(inv%MCONN(NNPG,inv%NCON), inv%MASKG(NNPG), inv%MYGRAD(inv%NNPGL))
      ALLOCATE(inv%JCSC_FDIST(inv%NNPGL+1), inv%ICSC_FDIST(NZF), inv%ELEM_WTS(NELEM))
      READ(10) inv%MCONN
      READ(10) inv%MASKG
      READ(10) inv%MYGRAD
      READ(10) inv%JCSC_FDIST
      READ(10) inv%ICSC_FDIST
      READ(10) inv%ELEM_WTS
      ALLOCATE(WAVE(NDOF), ADJLAM(NDOF), UE(NDOF))
      READ(10) WAVE
      READ(10) ADJLAM
      READ(10) UE
      ALLOCATE(NU(NDOF), ALPHA(NDOF), V(NA35))
      READ(10) NU
      READ(10) ALPHA
      READ(10) V
      CLOSE(10)

      ALLOCATE(CDIST(NZ), CDIST2(NZ), FMAT(NZF), GPROD(NA35), PEFF(NDOF))
      ISTAT(:) = 0
      CALL ASMBLE_DRIVER(NZ, CFTYPE, FREQ, PY, MSH, CDIST, ISTAT(1))
      CALL ASMBLE_DRIVER(NZ, CFTYPE, FREQ, PY, MSH, CDIST2, ISTAT(2))
      ISTAT(5) = COUNT(CDIST /= CDIST2)
      CALL CFMAT_DIST(NZF, NDOF, FREQ, PY, WAVE, MSH, INV, FMAT, ISTAT(3))
      CALL ADJFLAM_DIST(0, MPI_COMM_WORLD, inv%NNPGL, NDOF, NZF, NA35, inv%MYGRAD,           &
                        inv%JCSC_FDIST, inv%ICSC_FDIST, FMAT, ADJLAM, GPROD)
      CALL ASMB_PEFF(0, MPI_COMM_WORLD, NDOF, NDOF, NZ, msh%CNP, msh%MYDOFS, msh%IRPTR_LOC,    &
                     msh%JCPTR_LOC, CDIST, UE, PEFF)
      CALL ASMBLE_DRIVER(NZ - 1, CFTYPE, FREQ, PY, MSH, CDIST2, ISTAT(4))
      ALLOCATE(P1(NA35), P2(NA35), P3(NA35), RHSBNU(NDOF), RHSALP(NDOF))
      CALL MET225_MOD(0, 0, MPI_COMM_WORLD, NDOF, FREQ, PY, ADJLAM, NU, ALPHA, WAVE, MSH, INV,   &
                      P1, P2, ISTAT(6))
      CALL F25_CADJ_NU_ALPHA_RHS(NDOF, NA35, FREQ, PY, V, ADJLAM, WAVE, MSH, INV, RHSBNU, RHSALP, &
                                 P3, ISTAT(7))

      OPEN(11, FILE=TRIM(FOUT), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(11) ISTAT
      WRITE(11) CDIST
      WRITE(11) FMAT
      WRITE(11) GPROD
      WRITE(11) PEFF
      WRITE(11) P1
      WRITE(11) P2
      WRITE(11) RHSBNU
      WRITE(11) RHSALP
      WRITE(11) P3
      CLOSE(11)
      CALL MPI_FINALIZE(MPIERR)
      END PROGRAM DROPIN_CHECK
