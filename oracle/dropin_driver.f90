! Integration check of the drop-in boundary (test infrastructure only).
!
! A cut-down copy of the reference driver's call sequence (elwave.f90:
! 92-124, 138-143, 147-182) compiled against the reference's own
! modules.F90, interface.f90, srgraphs.f and sparse.f: SREL3D_GETSIZE /
! SREL3D / SPARSE_ADJ2CRS build the CSR, then the build's ASMBLE shim
! (wtomo3d_amd/fortran/asmble.f90) and STRUMPACK_SERIAL_FINTER jobs 1, 2, 3
! and -1 run from libwtomo3d_hip.so.  It solves A x = A 1 and prints the
! relative error of x against 1 and the job error flags.
      PROGRAM dropin_driver
      USE MODEL_MODULE
      USE PML_MODULE
      USE INIT_MODULE
      USE STRUMPACK_MODULE, ONLY : strumpack_serial_finter
      IMPLICIT NONE
      INTEGER, ALLOCATABLE :: xadj(:), adjncy(:), irptr(:), jcptr(:)
      COMPLEX, ALLOCATABLE :: a(:), rhs(:), sol(:)
      INTEGER numvert, nedge, nzero, ierr, i, j, ix, iy, iz, job
      REAL, PARAMETER :: pi = 3.141592653589793
      REAL err, vp, vs
      nx = 10
      ny = 9
      nz = 11
      dx = 250.0
      freesurf(:) = .FALSE.
      freesurf(1) = .TRUE.
      usemin = .TRUE.
      pmlr = 0.001
      pmld = 2.0*dx
      ipml = INT(pmld/dx) + 1
      pmlf = 3.0*LOG(1.0/pmlr)/(2.0*pmld*pmld*pmld)
      ALLOCATE(da(nz,nx,ny), mu(nz,nx,ny), rho(nz,nx,ny))
      DO iy = 1, ny
         DO ix = 1, nx
            DO iz = 1, nz
               vp = 5000.0 + 40.0*iz + 7.0*ix - 3.0*iy
               vs = vp/1.73
               rho(iz,ix,iy) = 310.0*vp**0.25
               mu(iz,ix,iy) = CMPLX(vs*vs*rho(iz,ix,iy), 0.0)
               da(iz,ix,iy) = CMPLX(vp*vp*rho(iz,ix,iy), 0.0) - 2.0*mu(iz,ix,iy)
            ENDDO
         ENDDO
      ENDDO
      CALL SREL3D_GETSIZE(nx, ny, nz, numvert, nedge)
      ALLOCATE(xadj(numvert+1), adjncy(2*nedge))
      CALL SREL3D(nx, ny, nz, numvert, nedge, xadj, adjncy)
      nzero = 2*nedge + numvert
      ALLOCATE(irptr(numvert+1), jcptr(nzero), a(nzero), rhs(numvert), sol(numvert))
      CALL SPARSE_ADJ2CRS(nzero, numvert, xadj, adjncy, irptr, jcptr)
      a(:) = CMPLX(0.0, 0.0)
      rhs(:) = CMPLX(0.0, 0.0)
      sol(:) = CMPLX(0.0, 0.0)
      job = 1
      CALL STRUMPACK_SERIAL_FINTER(job, numvert, nzero, irptr, jcptr, a, rhs, sol, ierr)
      WRITE(*,*) 'job 1 ierr', ierr
      omega = CMPLX(2.0*pi*0.3, 0.0)
      CALL ASMBLE(6, nzero, nx, ny, nz, irptr, jcptr, a, ierr)
      WRITE(*,*) 'asmble ierr', ierr
      job = 2
      CALL STRUMPACK_SERIAL_FINTER(job, numvert, nzero, irptr, jcptr, a, rhs, sol, ierr)
      WRITE(*,*) 'job 2 ierr', ierr
      DO i = 1, numvert
         DO j = irptr(i), irptr(i+1) - 1
            rhs(i) = rhs(i) + a(j)
         ENDDO
      ENDDO
      job = 3
      CALL STRUMPACK_SERIAL_FINTER(job, numvert, nzero, irptr, jcptr, a, rhs, sol, ierr)
      WRITE(*,*) 'job 3 ierr', ierr
      err = MAXVAL(ABS(sol - CMPLX(1.0, 0.0)))
      WRITE(*,*) 'max error vs ones', err
      job = -1
      CALL STRUMPACK_SERIAL_FINTER(job, numvert, nzero, irptr, jcptr, a, rhs, sol, ierr)
      WRITE(*,*) 'job -1 ierr', ierr
      IF (err < 1.0E-3) THEN
         WRITE(*,*) 'DROPIN OK'
      ELSE
         WRITE(*,*) 'DROPIN FAILED'
      ENDIF
      END PROGRAM
