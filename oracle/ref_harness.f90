! Test harness for the reference assembly path (test infrastructure only).
!
! Compiled together with the reference's own Fortran sources by
! oracle/Makefile into oracle/_ref/libwtref32.so (default REAL/COMPLEX) and
! oracle/_ref/libwtref64.so (-fdefault-real-8 promotion).  It sets the module
! globals the way the reference driver does and calls the reference routines
! unchanged:
!   readini.f90:81-84      ipml = INT(pmld/dx)+1, pmlf = 3 ln(1/pmlr)/(2 pmld^3)
!   modelio.f90:78-92      mu = vs^2 rho, lambda = vp^2 rho - 2 mu (complex)
!   elwave.f90:92-124      SREL3D_GETSIZE / SREL3D / SPARSE_ADJ2CRS
!   elwave.f90:151,164     omega = 2 pi f (+ i/tau)
!   elwave.f90:171         ASMBLE
! Only the oracle tests load the resulting library.
      SUBROUTINE wtref_graph_size(nx, ny, nz, numvert, nzero) &
                 BIND(C, NAME='wtref_graph_size')
      USE ISO_C_BINDING
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nx, ny, nz
      INTEGER(C_INT), INTENT(OUT) :: numvert, nzero
      INTEGER nedge
      CALL SREL3D_GETSIZE(nx, ny, nz, numvert, nedge)
      nzero = 2*nedge + numvert
      END SUBROUTINE

      SUBROUTINE wtref_graph(nx, ny, nz, irptr, jcptr) &
                 BIND(C, NAME='wtref_graph')
      USE ISO_C_BINDING
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nx, ny, nz
      INTEGER(C_INT), INTENT(OUT) :: irptr(*), jcptr(*)
      INTEGER, ALLOCATABLE :: xadj(:), adjncy(:)
      INTEGER numvert, nedge, nzero
      CALL SREL3D_GETSIZE(nx, ny, nz, numvert, nedge)
      ALLOCATE(xadj(numvert+1))
      ALLOCATE(adjncy(2*nedge))
      CALL SREL3D(nx, ny, nz, numvert, nedge, xadj, adjncy)
      nzero = 2*nedge + numvert
      CALL SPARSE_ADJ2CRS(nzero, numvert, xadj, adjncy, irptr, jcptr)
      DEALLOCATE(xadj)
      DEALLOCATE(adjncy)
      END SUBROUTINE

      SUBROUTINE wtref_assemble(nxi, nyi, nzi, dxi, freqi, taui, &
                                fst, usem, pmldi, pmlri, &
                                vp, vs, rh, nzero, irptr, jcptr, a, ierr) &
                 BIND(C, NAME='wtref_assemble')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, omega, dx, dy, dz, &
                               nx, ny, nz, freesurf
      USE PML_MODULE, ONLY : pmld, pmlr, pmlf, ipml
      USE INIT_MODULE, ONLY : usemin, tau
      USE INTERFACE_MODULE, ONLY : asmble
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nxi, nyi, nzi, fst, usem, nzero
      REAL, INTENT(IN) :: dxi, freqi, taui, pmldi, pmlri
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      INTEGER(C_INT), INTENT(IN) :: irptr(*), jcptr(*)
      COMPLEX, INTENT(OUT) :: a(*)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      COMPLEX vc
      REAL, PARAMETER :: pi = 3.141592653589793
      INTEGER ix, iy, iz
      nx = nxi
      ny = nyi
      nz = nzi
      dx = dxi
      dy = dxi
      dz = dxi
      freesurf(:) = .FALSE.
      freesurf(1) = (fst /= 0)
      usemin = (usem /= 0)
      tau = taui
      pmlr = pmlri
      pmld = pmldi
      ipml = INT(pmld/dx) + 1
      pmlf = 3.0*LOG(1.0/pmlr)/(2.0*pmld*pmld*pmld)
      IF (ALLOCATED(da)) DEALLOCATE(da, mu, rho, dap, mup, rhop)
      ALLOCATE(da(nz,nx,ny), mu(nz,nx,ny), rho(nz,nx,ny))
      ALLOCATE(dap(nz,nx,ny), mup(nz,nx,ny), rhop(nz,nx,ny))
      DO iy=1,ny
         DO ix=1,nx
            DO iz=1,nz
               rho(iz,ix,iy) = rh(iz,ix,iy)
               rhop(iz,ix,iy) = rh(iz,ix,iy)
               vc = CMPLX(vs(iz,ix,iy), 0.0)
               mu(iz,ix,iy) = vc*vc*rho(iz,ix,iy)
               vc = CMPLX(vp(iz,ix,iy), 0.0)
               da(iz,ix,iy) = vc*vc*rho(iz,ix,iy) - 2.0*mu(iz,ix,iy)
               mup(iz,ix,iy) = mu(iz,ix,iy)
               dap(iz,ix,iy) = da(iz,ix,iy)
            ENDDO
         ENDDO
      ENDDO
      omega = CMPLX(2.0*pi*freqi, 0.0)
      IF (tau /= 999.999) omega = omega + CMPLX(0.0, 1.0/tau)
      CALL ASMBLE(6, nzero, nx, ny, nz, irptr, jcptr, a, ierr)
      DEALLOCATE(da, mu, rho, dap, mup, rhop)
      END SUBROUTINE
