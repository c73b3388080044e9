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
!   haskgrn.f:31           HASKGRN (plane-wave depth profiles)
!   hetfldsrc.f90:17       HETFLDSRC (scattered-field source)
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

! Module globals of one model at one frequency, as readini / modelio /
! elwave set them; the background model (dap, mup, rhop) is a second copy
! of the given speeds.
      SUBROUTINE wtref_model(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, &
                             pmldi, pmlri, vp, vs, rh, vpb, vsb, rhb)
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, omega, dx, dy, dz, &
                               nx, ny, nz, freesurf
      USE PML_MODULE, ONLY : pmld, pmlr, pmlf, ipml
      USE INIT_MODULE, ONLY : usemin, tau
      IMPLICIT NONE
      INTEGER, INTENT(IN) :: nxi, nyi, nzi, fst, usem
      REAL, INTENT(IN) :: dxi, freqi, taui, pmldi, pmlri
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      REAL, INTENT(IN) :: vpb(nzi,nxi,nyi), vsb(nzi,nxi,nyi), rhb(nzi,nxi,nyi)
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
               vc = CMPLX(vs(iz,ix,iy), 0.0)
               mu(iz,ix,iy) = vc*vc*rho(iz,ix,iy)
               vc = CMPLX(vp(iz,ix,iy), 0.0)
               da(iz,ix,iy) = vc*vc*rho(iz,ix,iy) - 2.0*mu(iz,ix,iy)
               rhop(iz,ix,iy) = rhb(iz,ix,iy)
               vc = CMPLX(vsb(iz,ix,iy), 0.0)
               mup(iz,ix,iy) = vc*vc*rhop(iz,ix,iy)
               vc = CMPLX(vpb(iz,ix,iy), 0.0)
               dap(iz,ix,iy) = vc*vc*rhop(iz,ix,iy) - 2.0*mup(iz,ix,iy)
            ENDDO
         ENDDO
      ENDDO
      omega = CMPLX(2.0*pi*freqi, 0.0)
      IF (tau /= 999.999) omega = omega + CMPLX(0.0, 1.0/tau)
      END SUBROUTINE

      SUBROUTINE wtref_assemble(nxi, nyi, nzi, dxi, freqi, taui, &
                                fst, usem, pmldi, pmlri, &
                                vp, vs, rh, nzero, irptr, jcptr, a, ierr) &
                 BIND(C, NAME='wtref_assemble')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, nx, ny, nz
      USE INTERFACE_MODULE, ONLY : asmble
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nxi, nyi, nzi, fst, usem, nzero
      REAL, INTENT(IN) :: dxi, freqi, taui, pmldi, pmlri
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      INTEGER(C_INT), INTENT(IN) :: irptr(*), jcptr(*)
      COMPLEX, INTENT(OUT) :: a(*)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      CALL wtref_model(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, pmldi, pmlri, &
                       vp, vs, rh, vp, vs, rh)
      CALL ASMBLE(6, nzero, nx, ny, nz, irptr, jcptr, a, ierr)
      DEALLOCATE(da, mu, rho, dap, mup, rhop)
      END SUBROUTINE

! HETFLDSRC (hetfldsrc.f90:17) for one model against a background model and
! background fields ub, vb, wb given on the grid ((nz,nx,ny), z fastest).
! The output buffer is sized for the reference's own node numbering
! (hetfldsrc.f90:1020, (iy-1)*nx*ny + (ix-1)*nz + iz), which equals the
! unknown order when ny = nz; the first 3*nx*ny*nz entries are returned.
      SUBROUTINE wtref_hetfldsrc(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, &
                                 pmldi, pmlri, vp, vs, rh, vpb, vsb, rhb, &
                                 ub, vb, wb, srcr, srci, sfld, ierr) &
                 BIND(C, NAME='wtref_hetfldsrc')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, nx, ny, nz
      USE WAVEFIELD_MODULE, ONLY : utbkr, vtbkr, wtbkr
      USE INTERFACE_MODULE, ONLY : hetfldsrc
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nxi, nyi, nzi, fst, usem
      REAL, INTENT(IN) :: dxi, freqi, taui, pmldi, pmlri, srcr, srci
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      REAL, INTENT(IN) :: vpb(nzi,nxi,nyi), vsb(nzi,nxi,nyi), rhb(nzi,nxi,nyi)
      COMPLEX, INTENT(IN) :: ub(nzi,nxi,nyi), vb(nzi,nxi,nyi), wb(nzi,nxi,nyi)
      COMPLEX, INTENT(OUT) :: sfld(3*nzi*nxi*nyi)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      COMPLEX, ALLOCATABLE :: sf(:)
      INTEGER ierr1
      CALL wtref_model(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, pmldi, pmlri, &
                       vp, vs, rh, vpb, vsb, rhb)
      IF (ALLOCATED(utbkr)) DEALLOCATE(utbkr, vtbkr, wtbkr)
      ALLOCATE(utbkr(nz,nx,ny), vtbkr(nz,nx,ny), wtbkr(nz,nx,ny))
      utbkr = ub
      vtbkr = vb
      wtbkr = wb
      ALLOCATE(sf(3*((ny-1)*nx*ny + nx*nz)))
      CALL HETFLDSRC(CMPLX(srcr, srci), sf, 0, 0, ierr1)
      ierr = ierr1
      sfld = sf(1:3*nx*ny*nz)
      DEALLOCATE(sf, utbkr, vtbkr, wtbkr)
      DEALLOCATE(da, mu, rho, dap, mup, rhop)
      END SUBROUTINE

! HASKGRN (haskgrn.f:31) for source 1: the layer model goes into
! HASKELL_MODULE (top depths h1d, continuous and discrete speeds), the
! incidence angle and type into SRCPRM_MODULE ('pp' for P, 'ps' for S),
! the depth grid into MODEL_MODULE.
      SUBROUTINE wtref_haskgrn(nli, vp, vs, rh, top, vpd, vsd, ain1, iss, &
                               dfi, nzi, dzi, zorigi, ug, wg, ierr) &
                 BIND(C, NAME='wtref_haskgrn')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : dz, nz, zorig
      USE SRCPRM_MODULE, ONLY : ain, srctyp
      USE HASKELL_MODULE, ONLY : nl, ugrn1f, wgrn1f, vp1d, vs1d, rh1d, h1d, &
                                 vs1disc, vp1disc
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nli, iss, nzi
      REAL, INTENT(IN) :: vp(nli), vs(nli), rh(nli), top(nli), vpd(nli), vsd(nli)
      REAL, INTENT(IN) :: ain1, dfi, dzi, zorigi
      COMPLEX, INTENT(OUT) :: ug(nzi), wg(nzi)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      INTEGER ierr1
      nl = nli
      nz = nzi
      dz = dzi
      zorig = zorigi
      IF (ALLOCATED(vp1d)) DEALLOCATE(vp1d, vs1d, rh1d, h1d, vp1disc, vs1disc)
      ALLOCATE(vp1d(nl), vs1d(nl), rh1d(nl), h1d(nl), vp1disc(nl), vs1disc(nl))
      vp1d = vp
      vs1d = vs
      rh1d = rh
      h1d = top
      vp1disc = vpd
      vs1disc = vsd
      IF (ALLOCATED(ain)) DEALLOCATE(ain, srctyp)
      ALLOCATE(ain(1), srctyp(1))
      ain(1) = ain1
      srctyp(1) = 'pp'
      IF (iss /= 0) srctyp(1) = 'ps'
      IF (ALLOCATED(ugrn1f)) DEALLOCATE(ugrn1f, wgrn1f)
      ALLOCATE(ugrn1f(nz), wgrn1f(nz))
      ugrn1f = CMPLX(0.0, 0.0)
      wgrn1f = CMPLX(0.0, 0.0)
      CALL HASKGRN(dfi, 1, ierr1)
      ierr = ierr1
      ug = ugrn1f
      wg = wgrn1f
      DEALLOCATE(vp1d, vs1d, rh1d, h1d, vp1disc, vs1disc, ain, srctyp, ugrn1f, wgrn1f)
      END SUBROUTINE
