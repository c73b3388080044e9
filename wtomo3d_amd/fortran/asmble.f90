! Drop-in ASMBLE for the reference driver (elwave.f90:171): same name and
! argument list as asmble.f90:1-2 (interface.f90:4-12).  Instead of the
! Fortran stencil routines it passes the module state the reference reads
! (MODEL_MODULE, PML_MODULE, INIT_MODULE; modules.F90:1-76) to the GPU
! assembly in libwtomo3d_hip.so through the C ABI entry wt_asmble
! (include/wtomo3d.h).  Link this file and libwtomo3d_hip.so in place of
! asmble.f90, setcoefs3d.f, calcwts.f, fdfd3d.f, pfdfd3d.f, pml3d.f,
! fixed3d.f, tpfree.f and zeroas.f; compile it against the reference
! program's own modules.F90.
      SUBROUTINE ASMBLE(lunlog, nzero, nx, ny, nz, irptr, jcptr, a, ierr)
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, omega, dx, freesurf
      USE PML_MODULE, ONLY : pmld, pmlf, ipml
      USE INIT_MODULE, ONLY : usemin
      IMPLICIT NONE
      INTEGER, INTENT(IN) :: nzero, lunlog, nx, ny, nz
      INTEGER, INTENT(IN) :: irptr(3*nx*ny*nz+1), jcptr(nzero)
      COMPLEX, INTENT(OUT) :: a(nzero)
      INTEGER, INTENT(OUT) :: ierr
      INTERFACE
         SUBROUTINE wt_asmble(lunlog, nzero, nx, ny, nz, irptr, jcptr, a, ierr, &
                              da, mu, rho, om_re, om_im, dx, pmld, pmlf, ipml, &
                              fst, usemin) BIND(C, NAME='wt_asmble')
         USE ISO_C_BINDING
         INTEGER(C_INT), VALUE :: lunlog, nzero, nx, ny, nz, ipml, fst, usemin
         INTEGER(C_INT), INTENT(IN) :: irptr(*), jcptr(*)
         COMPLEX(C_FLOAT_COMPLEX), INTENT(OUT) :: a(*)
         INTEGER(C_INT), INTENT(OUT) :: ierr
         COMPLEX(C_FLOAT_COMPLEX), INTENT(IN) :: da(*), mu(*)
         REAL(C_FLOAT), INTENT(IN) :: rho(*)
         REAL(C_FLOAT), VALUE :: om_re, om_im, dx, pmld, pmlf
         END SUBROUTINE
      END INTERFACE
      INTEGER(C_INT) fst, um
      fst = 0
      IF (freesurf(1)) fst = 1
      um = 0
      IF (usemin) um = 1
      CALL wt_asmble(lunlog, nzero, nx, ny, nz, irptr, jcptr, a, ierr, &
                     da, mu, rho, REAL(omega), AIMAG(omega), dx, pmld, pmlf, &
                     ipml, fst, um)
      IF (ierr /= 0) THEN
         WRITE(*,*) 'ERROR calling fdfd3d from ASBMLE'
         WRITE(lunlog,*) 'ERROR calling fdfd3d from ASBMLE'
      ENDIF
      RETURN
      END SUBROUTINE ASMBLE
