!     f25_cabi.f90 -- ISO_C_BINDING interface to libfem25hip.so (the C ABI
!     declared in include/fem25hip.h) and the state the Fortran drop-in
!     entry points (f25_dropin.f90) keep between calls: one mesh handle,
!     recreated when the caller's mesh arrays change.
!
!     Nothing here depends on the reference's derived types; the drop-in
!     routines unpack MESH_INFO / INV_INFO themselves and pass plain arrays.
      MODULE F25_CABI
      USE ISO_C_BINDING
      IMPLICIT NONE
      INTEGER(C_INT), PARAMETER :: F25_C64 = 0, F25_C128 = 1
      INTEGER(C_INT), PARAMETER :: F25_ERR_ARG = 1, F25_ERR_STATE = 8

!     f25_mesh_desc (include/fem25hip.h)
      TYPE, BIND(C) :: F25_MESH_DESC
         INTEGER(C_INT) :: NORD, NELEM, NNPG
         TYPE(C_PTR) :: IENG, XLOCS, ZLOCS, XD, ZD, CDOMAIN, CNNPG
         REAL(C_DOUBLE) :: XWIDTH, ZWIDTH
         REAL(C_DOUBLE) :: PML(4), PML_SRF(4)
      END TYPE F25_MESH_DESC

      INTERFACE
         INTEGER(C_INT) FUNCTION F25_MESH_CREATE_WITH_GRAPH(DESC, NDOF, LM, IRPTR, JCPTR, H)   &
                        BIND(C, NAME='f25_mesh_create_with_graph')
            IMPORT :: C_INT, C_PTR, F25_MESH_DESC
            TYPE(F25_MESH_DESC), INTENT(IN) :: DESC
            INTEGER(C_INT), VALUE :: NDOF
            TYPE(C_PTR), VALUE :: LM, IRPTR, JCPTR
            TYPE(C_PTR), INTENT(OUT) :: H
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_MESH_DESTROY(H) BIND(C, NAME='f25_mesh_destroy')
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: H
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_SET_MODEL(H, DENS, ECOEFF) BIND(C, NAME='f25_set_model')
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: H, DENS, ECOEFF
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_SET_ATTENUATION(H, QP, QS, FREQ0)                          &
                        BIND(C, NAME='f25_set_attenuation')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: H, QP, QS
            REAL(C_DOUBLE), VALUE :: FREQ0
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_ASMBLE(H, CFTYPE, FREQ, PY, PREC, OUT)                    &
                        BIND(C, NAME='f25_asmble')
            IMPORT :: C_INT, C_PTR, C_CHAR, C_DOUBLE
            TYPE(C_PTR), VALUE :: H
            CHARACTER(KIND=C_CHAR), VALUE :: CFTYPE
            REAL(C_DOUBLE), VALUE :: FREQ, PY
            INTEGER(C_INT), VALUE :: PREC
            TYPE(C_PTR), VALUE :: OUT
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_INV_ATTACH(H, CINVTYPE, VPVS, NVINV, NCON, MCONN, MASKG,  &
                        NNPGL, MYGRAD, JCSC, NZ_FDIST, ICSC, NA35, ELEM_WTS)                   &
                        BIND(C, NAME='f25_inv_attach')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: H, CINVTYPE
            REAL(C_DOUBLE), VALUE :: VPVS
            INTEGER(C_INT), VALUE :: NVINV, NCON, NNPGL, NZ_FDIST, NA35
            TYPE(C_PTR), VALUE :: MCONN, MASKG, MYGRAD, JCSC, ICSC, ELEM_WTS
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_CFMAT_DIST(H, FREQ, PY, WAVE, FMAT)                       &
                        BIND(C, NAME='f25_cfmat_dist')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: H
            REAL(C_DOUBLE), VALUE :: FREQ, PY
            TYPE(C_PTR), VALUE :: WAVE, FMAT
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_ASMB_PEFF_CSR(NDOF, NDOFL, NZLOC, CNP, MYDOFS, IRPTR_LOC,  &
                        JCPTR_LOC, CDIST, UE, BUFF) BIND(C, NAME='f25_asmb_peff_csr')
            IMPORT :: C_INT, C_PTR
            INTEGER(C_INT), VALUE :: NDOF, NDOFL, NZLOC
            TYPE(C_PTR), VALUE :: CNP, MYDOFS, IRPTR_LOC, JCPTR_LOC, CDIST, UE, BUFF
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_ADJFLAM_CSC(NNPGL, NDOF, NZ_FDIST, NA35, MYGRAD, JCSC,     &
                        ICSC, FMAT, ADJLAM, VWORK) BIND(C, NAME='f25_adjflam_csc')
            IMPORT :: C_INT, C_PTR
            INTEGER(C_INT), VALUE :: NNPGL, NDOF, NZ_FDIST, NA35
            TYPE(C_PTR), VALUE :: MYGRAD, JCSC, ICSC, FMAT, ADJLAM, VWORK
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_GN_JP(NNPGL, NA35, MYGRAD, NJAC, NOBS, JAC, P, JP)        &
                        BIND(C, NAME='f25_gn_jp')
            IMPORT :: C_INT, C_PTR
            INTEGER(C_INT), VALUE :: NNPGL, NA35, NJAC
            TYPE(C_PTR), VALUE :: MYGRAD, NOBS, JAC, P, JP
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_GN_JHZ(NNPGL, NA35, MYGRAD, NJAC, NOBS, ALPHA, JAC, JP, Z) &
                        BIND(C, NAME='f25_gn_jhz')
            IMPORT :: C_INT, C_PTR
            INTEGER(C_INT), VALUE :: NNPGL, NA35, NJAC
            TYPE(C_PTR), VALUE :: MYGRAD, NOBS, ALPHA, JAC, JP, Z
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_QP_NU_ALPHA(H, FREQ, PY, V, ADJLAM, WAVE, RHSBNU, RHSALP, P3) &
                        BIND(C, NAME='f25_qp_nu_alpha')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: H
            REAL(C_DOUBLE), VALUE :: FREQ, PY
            TYPE(C_PTR), VALUE :: V, ADJLAM, WAVE, RHSBNU, RHSALP, P3
         END FUNCTION
         INTEGER(C_INT) FUNCTION F25_MET225(H, FREQ, PY, ADJLAM, NU, ALPHA, WAVE, P1, P2)          &
                        BIND(C, NAME='f25_met225')
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: H
            REAL(C_DOUBLE), VALUE :: FREQ, PY
            TYPE(C_PTR), VALUE :: ADJLAM, NU, ALPHA, WAVE, P1, P2
         END FUNCTION
         TYPE(C_PTR) FUNCTION F25_LAST_ERROR() BIND(C, NAME='f25_last_error')
            IMPORT :: C_PTR
         END FUNCTION
         INTEGER(C_SIZE_T) FUNCTION C_STRLEN(S) BIND(C, NAME='strlen')
            IMPORT :: C_SIZE_T, C_PTR
            TYPE(C_PTR), VALUE :: S
         END FUNCTION
      END INTERFACE

!     cached mesh handle and what it was built from
      TYPE(C_PTR), SAVE :: F25_HMESH = C_NULL_PTR
      INTEGER(C_INTPTR_T), SAVE :: F25_KEY_LM = 0, F25_KEY_IENG = 0, F25_KEY_JC = 0
      INTEGER, SAVE :: F25_KEY_SIZES(5) = -1
!     cached inversion structures
      INTEGER(C_INTPTR_T), SAVE :: F25_KEY_JCSC = 0, F25_KEY_ICSC = 0, F25_KEY_MASKG = 0
      INTEGER, SAVE :: F25_KEY_INV(4) = -1
      CHARACTER(2), SAVE :: F25_KEY_CINV = '  '
      REAL(C_DOUBLE), SAVE :: F25_KEY_VPVS = -1.D0

      CONTAINS

!     Prints the library's last error message after a failed call.
      SUBROUTINE F25_REPORT(WHERE, IRC)
      CHARACTER(*), INTENT(IN) :: WHERE
      INTEGER(C_INT), INTENT(IN) :: IRC
      TYPE(C_PTR) :: P
      CHARACTER(KIND=C_CHAR), POINTER :: S(:)
      CHARACTER(256) :: MSG
      INTEGER :: N, I
      MSG = ' '
      P = F25_LAST_ERROR()
      IF (C_ASSOCIATED(P)) THEN
         N = INT(MIN(C_STRLEN(P), INT(256, C_SIZE_T)))
         CALL C_F_POINTER(P, S, [N])
         DO I = 1, N
            MSG(I:I) = S(I)
         ENDDO
      ENDIF
      WRITE(*,*) WHERE, ': fem25hip error ', IRC, ' ', TRIM(MSG)
      END SUBROUTINE F25_REPORT

      INTEGER(C_INTPTR_T) FUNCTION F25_ADDR(P)
      TYPE(C_PTR), INTENT(IN) :: P
      F25_ADDR = TRANSFER(P, F25_ADDR)
      END FUNCTION F25_ADDR

!     Returns the cached handle for this mesh, creating it when the LM, IENG
!     or column arrays or the sizes differ from the last call; then sets the
!     model (a no-op on the library side when it is unchanged).
      SUBROUTINE F25_MESH_FOR(DESC, NDOF, NZ, LM, IRPTR, JCPTR, DENS, ECOEFF, IERR)
      TYPE(F25_MESH_DESC), INTENT(IN) :: DESC
      INTEGER, INTENT(IN) :: NDOF, NZ
      TYPE(C_PTR), INTENT(IN) :: LM, IRPTR, JCPTR, DENS, ECOEFF
      INTEGER, INTENT(OUT) :: IERR
      INTEGER(C_INT) :: IRC
      INTEGER :: SIZES(5)
      SIZES = [DESC%NORD, DESC%NELEM, DESC%NNPG, NDOF, NZ]
      IERR = 0
      IF (.NOT. C_ASSOCIATED(F25_HMESH) .OR. F25_ADDR(LM) /= F25_KEY_LM .OR.                  &
          F25_ADDR(DESC%IENG) /= F25_KEY_IENG .OR. F25_ADDR(JCPTR) /= F25_KEY_JC .OR.          &
          ANY(SIZES /= F25_KEY_SIZES)) THEN
         IF (C_ASSOCIATED(F25_HMESH)) IRC = F25_MESH_DESTROY(F25_HMESH)
         F25_HMESH = C_NULL_PTR
         F25_KEY_INV = -1
         IRC = F25_MESH_CREATE_WITH_GRAPH(DESC, INT(NDOF, C_INT), LM, IRPTR, JCPTR, F25_HMESH)
         IF (IRC /= 0) THEN
            CALL F25_REPORT('f25_mesh_create_with_graph', IRC)
            F25_HMESH = C_NULL_PTR
            IERR = INT(IRC)
            RETURN
         ENDIF
         F25_KEY_LM = F25_ADDR(LM)
         F25_KEY_IENG = F25_ADDR(DESC%IENG)
         F25_KEY_JC = F25_ADDR(JCPTR)
         F25_KEY_SIZES = SIZES
      ENDIF
      IRC = F25_SET_MODEL(F25_HMESH, DENS, ECOEFF)
      IF (IRC /= 0) THEN
         CALL F25_REPORT('f25_set_model', IRC)
         IERR = INT(IRC)
      ENDIF
      END SUBROUTINE F25_MESH_FOR

!     Drops the cached handle (call after freeing or rebuilding a mesh whose
!     arrays might be reallocated at the same addresses).
      SUBROUTINE F25_DROPIN_RESET()
      INTEGER(C_INT) :: IRC
      IF (C_ASSOCIATED(F25_HMESH)) IRC = F25_MESH_DESTROY(F25_HMESH)
      F25_HMESH = C_NULL_PTR
      F25_KEY_SIZES = -1
      F25_KEY_INV = -1
      F25_KEY_CINV = '  '
      END SUBROUTINE F25_DROPIN_RESET

      END MODULE F25_CABI
